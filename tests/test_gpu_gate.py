"""The small-prime gate (P/proth_logic.cpp:27-50, used at :405-408) as the
GPU path applies it: pl_gate_rejects, pl_proth_test's `gated` result and the
pool's `gated` results must equal the closed form -- some odd prime
3 <= p <= 67 divides k*2^n+1 -- computed here with Python integers, on every
fixture candidate and on candidates chosen to hit and miss the gate."""
import random

import pytest

import golden_io
import gmp_ctypes as G

pytestmark = pytest.mark.gpu

PRIMES = (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67)


def closed_form(k, n):
    N = k * (1 << n) + 1
    return any(N % p == 0 for p in PRIMES)


def fixture_candidates():
    out = []
    for name in ("small_cases", "batch_cases", "c5_cases", "kats"):
        d = golden_io.load(name)
        if d:
            out += [(c["k"], c["n"]) for c in d["cases"]]
    return out


def mixed_candidates(count=48):
    rnd = random.Random(67)
    hit, miss = [], []
    while len(hit) < count // 2 or len(miss) < count // 2:
        k = 2 * rnd.randrange(1, 1 << 15) + 1
        n = rnd.randrange(9000, 40000)
        (hit if closed_form(k, n) else miss).append((k, n))
    return hit[: count // 2] + miss[: count // 2] + [(3, 34351)]


def test_gate_rejects_equals_closed_form(gpu):
    cands = fixture_candidates() + mixed_candidates() + [(5, 1947), (3, 2), (7, 1)]
    assert len(cands) > 100
    for k, n in cands:
        assert gpu.gate_rejects(k, n) == closed_form(k, n), (k, n)


def test_pool_gated_flags_equal_closed_form(gpu):
    cands = mixed_candidates()
    with gpu.Pool(k_max=1 << 16) as pool:
        for i, (k, n) in enumerate(cands):
            pool.submit(k, n, tag=i, stop_after=20)
        res = pool.drain(steps=64)
    assert sorted(r["tag"] for r in res) == list(range(len(cands)))
    for r in res:
        k, n = cands[r["tag"]]
        assert bool(r["gated"]) == closed_form(k, n), (k, n)
        assert r["error"] == 0
        if r["gated"]:
            assert r["verdict"] == 0 and r["squarings"] == 0
        else:
            a = G.reference_base(k, n)
            assert r["res64"] == G.residue_after(k, n, a, 20) & ((1 << 64) - 1), (k, n)


def test_proth_test_gated_result_equals_closed_form(gpu):
    for k, n in mixed_candidates(8):
        r = gpu.proth_test(k, n, stop_after=10)
        assert r["gated"] == closed_form(k, n), (k, n)
        if r["gated"]:
            assert r["verdict"] == 0 and r["iters_done"] == 0
        else:
            assert r["verdict"] == -1 and r["iters_done"] == 10
