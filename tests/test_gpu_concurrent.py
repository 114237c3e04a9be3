"""Several contexts on one device at once, from several host threads: the
reference dispatches one test per worker thread, each with its own
FFTContext (P/dcore.cpp:125-192).  At most one context at a time runs the
one-launch outer pass (its workgroups wait on each other); the others take
the two-launch pass meanwhile.  Every result must equal GMP's."""
import json
import subprocess
import threading
from pathlib import Path

import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
SELFTEST = ROOT / "prothlogic_amd" / "prothhost_selftest"
CANDS = [(3, 20909), (3, 34350), (1023, 32768), (843, 120002)]


def run_threads(fn, items):
    out, errs = {}, []

    def body(i, item):
        try:
            out[i] = fn(item)
        except Exception as e:  # reported by the main thread
            errs.append((item, e))

    ts = [threading.Thread(target=body, args=(i, it)) for i, it in enumerate(items)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    assert not errs, errs
    return out


def test_four_proth_tests_from_four_threads(gpu):
    res = run_threads(lambda kn: gpu.proth_test(*kn), CANDS)
    for i, (k, n) in enumerate(CANDS):
        prime, r64, a = G.proth_verdict(k, n)
        assert res[i]["verdict"] == int(prime) and res[i]["res64"] == r64 and res[i]["base"] == a, (k, n)


def test_two_contexts_square_at_once(gpu):
    import random
    specs = [(433917, 1 << 21, 0), (11887, 300007, 0), (65535, 3000017, 0)]

    def work(spec):
        k, n, L = spec
        N = k * (1 << n) + 1
        x = random.Random(n).randrange(N)
        with gpu.Context(k, n, L) as c:
            c.set_residue(x)
            e = c.square(600)
            return c.get_residue() == G.powm(x, 1 << 600, N), e

    res = run_threads(work, specs)
    for i, spec in enumerate(specs):
        ok, e = res[i]
        assert ok, spec
        assert e < 0.3


def test_mirror_worker_threads(gpu, tmp_path):
    """The C++ mirror's execute_proth_test_fft on four std::threads, each with
    its own FFTContext, as the reference's dispatch runs its workers."""
    assert SELFTEST.exists(), "prothhost_selftest is not built"
    spec = ",".join(f"{k}:{n}" for k, n in CANDS)
    p = subprocess.run([str(SELFTEST), "threads", spec], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    lines = [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = {(d["k"], d["n"]): d for d in lines if d.get("case") == "thread"}
    assert sorted(got) == sorted(CANDS)
    for k, n in CANDS:
        prime, r64, _ = G.proth_verdict(k, n)
        assert got[(k, n)]["prime"] == int(prime) and int(got[(k, n)]["res64"], 16) == r64, (k, n)
