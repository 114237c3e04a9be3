"""Short, fixed program for hardware-counter passes over the squaring kernels.

    python -m prothlogic_amd.profile_run --workload c3 --iters 256 [--copy-mib 256]

Runs `iters` plain squarings (one graph-captured chunk per 256) of the named
single-test workload (c2: n=2^21, c3: n=2^24) from a random residue, or, for
c4, the 256 batched candidates through the pool, each stopped after `iters`
squarings past its seed.  Then one plain device-to-device copy of `copy-mib`
MiB (reads and writes exactly that many bytes, pl_copy_probe).  The copy calibrates the per-dispatch
FETCH_SIZE / WRITE_SIZE counters: HBM bytes of a squaring kernel are its
counter reading scaled by (copy bytes / the copy's counter reading).

--tune runs the launch-plan autotuner first (the pool's group plans for c4,
the single context's plan for c2/c3), as bench.py does before it times, so
the profiled kernels are the ones the bench runs; the tuning trials' own
launches then appear in the trace too.  --plans FILE loads the table a bench
run saved (bench.py --save-plans FILE) instead, with no trials.

Prints one line per step before it runs, so a counter pass names what it ran.
"""
from __future__ import annotations

import argparse
import random
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c3")
    ap.add_argument("--iters", type=int, default=256)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--tune", action="store_true", help="run the launch-plan autotuner first")
    ap.add_argument("--plans", metavar="FILE", default=None,
                    help="load a tuned-plan table (bench.py --save-plans) instead of timing alternatives")
    args = ap.parse_args()

    from prothlogic_amd import core, workloads
    if args.plans:
        core.tune_load(args.plans)
        print(f"[profile_run] plans loaded from {args.plans}", flush=True)
    if args.workload == "c4":
        cands = workloads.c4().candidates
        print(f"[profile_run] c4: {len(cands)} candidates through the pool, {args.iters} squarings each "
              "past the seed", flush=True)
        with core.Pool(k_max=1 << 16) as pool:
            if args.tune:
                print("[profile_run] c4: tuning the pool's group plans", flush=True)
                for k, n in cands:
                    r = pool.autotune(k, n)
                    if r["trials"]:
                        b = r["best"]
                        print(f"[profile_run]   L={core.plan_fft_len(n, k)}: {b['rows']}x{b['cols']} "
                              f"row radix {b['row_radix']} column radix {b['col_radix']} run {b['run']}", flush=True)
            for i, (k, n) in enumerate(cands):
                pool.submit(k, n, tag=i, stop_after=args.iters)
            res = pool.drain(steps=64)
        print(f"[profile_run] c4: {len(res)} results, {sum(r['squarings'] for r in res)} squarings, "
              f"max_err={max(r['max_err'] for r in res):.4g}", flush=True)
    else:
        (k, n), = (workloads.c2() if args.workload == "c2" else workloads.c3()).candidates
        N = k * (1 << n) + 1
        if args.tune:
            b = core.autotune_plan(k, n)["best"]
            print(f"[profile_run] {args.workload}: tuned plan {b['rows']}x{b['cols']} row radix {b['row_radix']} "
                  f"column radix {b['col_radix']} run {b['run']} one-launch {b['one_launch']}", flush=True)
        c = core.Context(k, n)
        c.set_residue(random.Random(7).randrange(N))
        print(f"[profile_run] {args.workload}: k={k} n={n} L={c.fft_len} squarings={args.iters}", flush=True)
        e = c.square(args.iters)
        print(f"[profile_run] max_err={e:.4g} res64={c.res64():016X}", flush=True)
        c.close()
    if args.copy_mib:
        nbytes = args.copy_mib << 20
        print(f"[profile_run] calibration copy: {args.copy_mib} MiB device-to-device, 1 timed launch", flush=True)
        ms = core.copy_probe(nbytes, reps=1)
        print(f"[profile_run] copy: {ms:.4f} ms, {2 * nbytes / ms / 1e6:.1f} GB/s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
