"""Short, fixed program for hardware-counter passes over the squaring kernels.

    python -m prothlogic_amd.profile_run --workload c3 --iters 256 [--copy-mib 256]

Runs `iters` plain squarings (one graph-captured chunk per 256) of the named
single-test workload (c2: n=2^21, c3: n=2^24) from a random residue, or, for
c4, the 256 batched candidates through the pool, each stopped after `iters`
squarings past its seed.  Then one plain device-to-device copy of `copy-mib`
MiB (reads and writes exactly that many bytes, pl_copy_probe).  The copy calibrates the per-dispatch
FETCH_SIZE / WRITE_SIZE counters: HBM bytes of a squaring kernel are its
counter reading scaled by (copy bytes / the copy's counter reading).

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
    args = ap.parse_args()

    from prothlogic_amd import core, workloads
    if args.workload == "c4":
        cands = workloads.c4().candidates
        print(f"[profile_run] c4: {len(cands)} candidates through the pool, {args.iters} squarings each "
              "past the seed", flush=True)
        with core.Pool(k_max=1 << 16) as pool:
            for i, (k, n) in enumerate(cands):
                pool.submit(k, n, tag=i, stop_after=args.iters)
            res = pool.drain(steps=64)
        print(f"[profile_run] c4: {len(res)} results, {sum(r['squarings'] for r in res)} squarings, "
              f"max_err={max(r['max_err'] for r in res):.4g}", flush=True)
    else:
        (k, n), = (workloads.c2() if args.workload == "c2" else workloads.c3()).candidates
        N = k * (1 << n) + 1
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
