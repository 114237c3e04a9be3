"""CPU checks of the measurement tooling around the full-length runs: the
config-5 segment summary (bench.py --c5-combine) and the GMP checker of
full-test results (tests/gmp_check_results.py)."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

import bench
import gmp_ctypes as G

ROOT = Path(__file__).resolve().parent.parent


def _segment(path, seg, rows, makespan):
    path.write_text(json.dumps(dict(segment=seg, segments=2, rate=0.5, makespan_s=makespan, tests=rows)))


def _row(tag, k, n, arrival, latency, verdict=0, res64="0123456789ABCDEF", error=0, max_err=0.01):
    return dict(tag=tag, k=k, n=n, arrival_s=arrival, latency_s=latency, verdict=verdict, error=error, res64=res64,
                max_err=max_err)


def test_c5_combine_replays_the_plain_loop_per_key(tmp_path):
    # two segments; tags index the key list as bench.py --plain-keys writes it
    rows0 = [_row(0, 3, 1000, 0.0, 5.0), _row(1, 5, 2000, 1.0, 7.0)]
    rows1 = [_row(2, 7, 1500, 0.0, 4.0, verdict=1, res64="0000000000000000"), _row(3, 9, 3000, 10.0, 9.0)]
    _segment(tmp_path / "s0.json", 0, rows0, 20.0)
    _segment(tmp_path / "s1.json", 1, rows1, 30.0)
    keys = ["4096/8", "4096/16", "4096/8", "8192/8"]
    plain = {"4096/8": dict(k=3, n=1000, s=1e-3, plan="x"), "4096/16": dict(k=5, n=2000, s=2e-3, plan="x"),
             "8192/8": dict(k=9, n=3000, s=3e-3, plan="x")}
    (tmp_path / "pk.json").write_text(json.dumps(dict(keys=keys, keys_c4=[], plain=plain)))
    txt = bench.c5_combine(str(tmp_path / "pk.json"), "-", [str(tmp_path / "s1.json"), str(tmp_path / "s0.json")])

    # the plain replay by hand: FIFO per segment, service = s(key) * (n - 1 + bits(k) - 1)
    def svc(k, n, s):
        return s * (n - 1 + k.bit_length() - 1)
    s0a, s0b = svc(3, 1000, 1e-3), svc(5, 2000, 2e-3)
    s1a, s1b = svc(7, 1500, 1e-3), svc(9, 3000, 3e-3)
    lat = [s0a, max(1.0, s0a) + s0b - 1.0, s1a, max(10.0, s1a) + s1b - 10.0]
    busy = s0a + s0b + s1a + s1b
    assert "tests                      4\n" in txt
    assert "'prime': 1, 'composite': 3, 'void': 0" in txt
    assert f"plain loop busy time       {busy:.1f} s" in txt
    assert f"plain p50 / p99 / max      {np.percentile(lat, 50):.1f} / {np.percentile(lat, 99):.1f} / {max(lat):.1f} s" in txt
    assert f"pool p50 / p99 / max       {np.percentile([5, 7, 4, 9], 50):.2f}" in txt
    tph, ptph = 3600 * 4 / 50.0, 3600 * 4 / busy
    assert f"pool / plain throughput    {tph / ptph:.2f}x" in txt
    assert "segment 0: 2 tests, makespan 20.0 s" in txt and "segment 1: 2 tests, makespan 30.0 s" in txt
    assert txt.index("segment 0:") < txt.index("segment 1:")


def test_gmp_checker_accepts_the_right_result_and_flags_a_wrong_one(tmp_path):
    # 3*2^20909+1 is prime (Res64 0); a composite's Res64 from GMP directly, then one bit flipped
    k, n = 45, 4001
    prime, r64, _ = G.proth_verdict(k, n)
    good = format(r64, "016X")
    bad = format(r64 ^ 1, "016X")
    rows = [dict(k=3, n=20909, verdict=1, error=0, res64="0000000000000000"),
            dict(k=k, n=n, verdict=1 if prime else 0, error=0, res64=good)]
    (tmp_path / "ok.json").write_text(json.dumps(rows))
    (tmp_path / "bad.json").write_text(json.dumps(dict(tests=[dict(k=k, n=n, verdict=1 if prime else 0, error=0,
                                                                   res64=bad)])))
    tool = ROOT / "tests" / "gmp_check_results.py"
    r = subprocess.run([sys.executable, str(tool), "--procs", "2", str(tmp_path / "ok.json")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = [json.loads(line) for line in r.stdout.splitlines()]
    assert sorted((o["k"], o["n"], o["match"]) for o in out) == [(3, 20909, True), (k, n, True)]
    assert next(o for o in out if o["k"] == 3)["gmp_prime"] is True
    r = subprocess.run([sys.executable, str(tool), str(tmp_path / "bad.json")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 1
    assert json.loads(r.stdout)["match"] is False
