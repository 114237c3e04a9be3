"""Benchmark: microseconds per exact modular squaring of a Proth residue on
one MI355X, with the residue resident in HBM.

Workload (default): the n = 2^21 single large test -- k drawn odd in
[3, 2^20) from PCG64(seed=2) (gate survivors only), base a by the reference's
rule, seed a^k mod N built on the device.  A "step" is one squaring
x <- x^2 mod N.  W untimed warm-up squarings run on a scratch context; then
the residue is reset to the seed and exactly K squarings are timed with the
device synchronised on both sides.  Res64 after K squarings is checked
against the GMP fixture when K matches one.

The same line reports the n = 2^24 test (1000 squarings), the per-kernel
roofline (HIP events on the engine's stream), and the reference CPU engine
timed on this host for a bounded sample.

  python bench.py [--gpus 1] [--steps 10000] [--warmup 200] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "us/squaring at n=2^21 and 2^24; batched tests/hour vs plain loop, Res64 exact"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def golden_res64(k, n, i):
    p = ROOT / "tests" / "golden" / "large_cases.json"
    if not p.exists():
        return None
    for c in json.loads(p.read_text())["cases"]:
        if c["k"] == k and c["n"] == n and str(i) in c["res64_after"]:
            return int(c["res64_after"][str(i)], 16)
    return None


def seed_context(core, k, n):
    """Context holding a^k mod N, powered on the device (left-to-right)."""
    a = core.choose_base(k, n)
    c = core.Context(k, n)
    c.set_residue(a)
    for bit in bin(k)[3:]:
        c.square(1)
        if bit == "1":
            c.mul_small(a)
    return c, a


def time_squarings(core, k, n, steps, warmup):
    c, a = seed_context(core, k, n)
    seed = c.get_residue()
    if warmup:
        c.square(warmup)
    c.set_residue(seed)
    c.square(0)
    t0 = time.perf_counter()
    emax = c.square(steps)
    t1 = time.perf_counter()
    r64 = c.res64()
    kt, _ = c.kernel_times(reps=20)
    info = dict(fft_len=c.fft_len, rows=c.info.rows, cols=c.info.cols, bits_per_digit=round(c.info.bits_per_digit, 4),
                kernels_per_squaring=c.info.passes)
    c.close()
    nt = 0 if info["kernels_per_squaring"] == 5 else (8 if kt.get("inv_carry", 0) else 0)
    return dict(seconds=t1 - t0, max_err=emax, res64=r64, kernel_ms=kt, a=a, nt=nt, **info)


def kernel_bytes(L: int, nt: int) -> dict:
    """Algorithmic HBM bytes one launch of each kernel moves for a length-L
    residue: complex FP64 points are 16 B (read and/or written once); the
    split outer pass hands 4 B int32 digits and 24 B of run carries per run
    of nt digits from its first to its second kernel."""
    run = 24.0 / nt if nt else 0.0
    return {"fwd_col": 32 * L, "row_square": 32 * L, "inv_col": 32 * L, "carry": 32 * L,
            "inv_carry": int((16 + 4 + run) * L), "adjust_fwd": int((4 + run + 16) * L)}


STEADY = ("inv_carry", "adjust_fwd", "row_square")
UNFUSED = ("inv_col", "carry", "fwd_col", "row_square")


def roofline(kt: dict, L: int, nt: int):
    """Dominant steady-state kernel's achieved algorithmic GB/s vs the HBM peak."""
    names = STEADY if nt else UNFUSED
    kb = kernel_bytes(L, nt)
    dom = max(names, key=lambda k: kt[k])
    achieved = kb[dom] / (kt[dom] * 1e-3) / 1e9
    per_sq = sum(kb[k] for k in names)
    return dom, dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None), kb[dom], per_sq


def measured_traffic(kernel: str, workload: str):
    p = ROOT / "profiles" / f"traffic_{kernel}_{workload}.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def cpu_baseline(k, n, budget_s=12.0):
    """The reference CPU engine (oracle/_ref, built from the reference's own
    sources) timed on this host; the C restatement when that is unavailable."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    from oracle import Oracle, RefLib, int_to_limbs
    import random
    o = Oracle()
    L = o.auto_tune_fft_len.__self__ and None
    L = 16
    while not o.fits(L, n, k) or n / L > 18.0:
        L *= 2
    N = k * (1 << n) + 1
    x = random.Random(1).randrange(N)
    v = int_to_limbs(x, L // 4)
    kind = "reference"
    try:
        ref = RefLib()
        h = ref.ctx(L)
        step = lambda w: ref.square_limbs(h, w, L // 4, n, k)
    except Exception as e:  # not built here, or no AVX-512 on this host
        log(f"[bench] reference build unavailable ({e}); timing the C restatement")
        kind = "port"
        oc = o.ctx(L)
        step = lambda w: oc.square_limbs(w, L // 4, n, k)
    cnt = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        v = step(v)
        cnt += 1
    dt = time.perf_counter() - t0
    return dict(value=round(dt / cnt * 1e6, 1), unit="us/squaring", cores=1, kind=kind,
                sample=f"{cnt} squarings of a random residue mod {k}*2^{n}+1 with fft_square_karatsuba "
                       f"at L={L} (the shortest length whose limb layout holds the residue), single thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the n=2^24 line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every configuration fits one MI355X and candidates are "
              "independent (they shard by candidate with no communication); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from prothlogic_amd import core, workloads
    if core.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)

    (k2, n2), = workloads.c2().candidates
    log(f"[bench] c2: k={k2} n={n2} steps={args.steps} warmup={args.warmup}")
    r2 = time_squarings(core, k2, n2, args.steps, args.warmup)
    us2 = r2["seconds"] / args.steps * 1e6
    g2 = golden_res64(k2, n2, args.steps)
    dom, roof, dom_bytes, sq_bytes = roofline(r2["kernel_ms"], r2["fft_len"], r2["nt"])
    roof["traffic"] = measured_traffic(dom, "c2_n2p21")
    log(f"[bench] c2: {us2:.2f} us/squaring, max_err={r2['max_err']:.4g}, kernels(ms)={r2['kernel_ms']}")

    extra = {}
    if not args.no_c3:
        (k3, n3), = workloads.c3().candidates
        s3 = 1000
        log(f"[bench] c3: k={k3} n={n3} steps={s3}")
        r3 = time_squarings(core, k3, n3, s3, 20)
        us3 = r3["seconds"] / s3 * 1e6
        g3 = golden_res64(k3, n3, s3)
        dom3, roof3, dom3_bytes, sq3_bytes = roofline(r3["kernel_ms"], r3["fft_len"], r3["nt"])
        roof3["traffic"] = measured_traffic(dom3, "c3_n2p24")
        roofline_us = sq3_bytes / (HBM_PEAK_GBS * 1e9) * 1e6
        extra["n2p24"] = dict(k=k3, n=n3, fft_len=r3["fft_len"], us_per_squaring=round(us3, 3), squarings=s3,
                              max_err=r3["max_err"], res64=format(r3["res64"], "016X"),
                              res64_matches_gmp=(None if g3 is None else r3["res64"] == g3),
                              kernel_ms=r3["kernel_ms"], dominant_kernel=dom3, roofline=roof3,
                              kernels_per_squaring=r3["kernels_per_squaring"], bytes_per_squaring=sq3_bytes,
                              hbm_roofline_us=round(roofline_us, 3), x_of_roofline=round(us3 / roofline_us, 2))
        log(f"[bench] c3: {us3:.2f} us/squaring, max_err={r3['max_err']:.4g}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(k2, n2)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:
            log(f"[bench] cpu baseline failed: {e}")

    line = {
        "metric": METRIC,
        "value": round(us2, 3),
        "unit": "us/squaring",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(us2 / 1000.0, 6),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Proth candidate (PCG64 seed 2), residue a^k mod N built on the device",
        "config": {"workload": "c2_n2p21 single test: n=2^21, k odd in [3,2^20), 1 squaring per step",
                   "k": k2, "n": n2, "base": r2["a"], "fft_len": r2["fft_len"], "rows": r2["rows"],
                   "cols": r2["cols"], "bits_per_digit": r2["bits_per_digit"],
                   "kernels_per_squaring": r2["kernels_per_squaring"]},
        "res64": format(r2["res64"], "016X"),
        "res64_matches_gmp": None if g2 is None else r2["res64"] == g2,
        "max_err": r2["max_err"],
        "kernel_ms": r2["kernel_ms"],
        "dominant_kernel": dom,
        "dominant_kernel_bytes": dom_bytes,
        "bytes_per_squaring": sq_bytes,
        "hbm_roofline_us": round(sq_bytes / (HBM_PEAK_GBS * 1e9) * 1e6, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
