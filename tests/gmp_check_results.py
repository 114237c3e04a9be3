"""Checks full-test results (verdict and Res64) against GMP, one line of JSON
per test.  Test infrastructure: GMP's mpz_powm is the exactness arbiter.

  python3 tests/gmp_check_results.py [--nmax N] [--every M [--offset O]] [--procs P] [--out OUT.jsonl] RESULTS.json ...

RESULTS.json is a file that bench.py's full-length modes write
(`--c4-full OUT.json`: a list of per-test rows; `--c5-full ... OUT.json`:
an object whose "tests" hold the rows).  Every row with n < N (default
120000: GMP takes about a minute per test there on one core) is checked;
with --every M, instead every M-th row in order of n from the O-th on
(O defaults to M).  Rows already present in OUT are skipped, so an
interrupted check resumes.
Each output line: k, n, the device's Res64 and verdict, GMP's, and match.
The verdict rule is the reference's (P/proth_logic.cpp:336-355: N is prime
when a^((N-1)/2) = -1 mod N, with a the base of P/proth_logic.cpp:423-435);
Res64 is the low 64 bits of that power (0 for a prime with n >= 64, whose
power is N-1 = k*2^n).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent))
import gmp_ctypes as G  # noqa: E402


def rows_of(path: str) -> list:
    d = json.loads(Path(path).read_text())
    return d["tests"] if isinstance(d, dict) else d


def check(row: dict) -> dict:
    t0 = time.time()
    prime, r64, a = G.proth_verdict(row["k"], row["n"])
    gmp = format(r64, "016X")
    return dict(k=row["k"], n=row["n"], base=a, device=row["res64"], gmp=gmp, device_verdict=row["verdict"],
                gmp_prime=bool(prime), seconds=round(time.time() - t0, 1),
                match=row["res64"] == gmp and (row["verdict"] == 1) == bool(prime))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("results", nargs="+")
    ap.add_argument("--nmax", type=int, default=120000)
    ap.add_argument("--every", type=int, default=0, help="every M-th test in order of n instead of n < nmax")
    ap.add_argument("--offset", type=int, default=None)
    ap.add_argument("--procs", type=int, default=4)
    ap.add_argument("--out", default=None, help="append the lines here (and skip tests already in it)")
    args = ap.parse_args()
    done = set()
    if args.out and Path(args.out).exists():
        done = {(o["k"], o["n"]) for o in map(json.loads, Path(args.out).read_text().splitlines())}
    rows, seen = [], set()
    for p in args.results:
        for r in rows_of(p):
            kn = (r["k"], r["n"])
            if r.get("error", 0) == 0 and kn not in seen:
                seen.add(kn)
                rows.append(r)
    rows.sort(key=lambda r: r["n"])
    if args.every:
        rows = rows[(args.every if args.offset is None else args.offset)::args.every]
    else:
        rows = [r for r in rows if r["n"] < args.nmax]
    todo = [r for r in rows if (r["k"], r["n"]) not in done]
    print(f"{len(todo)} tests to check", file=sys.stderr, flush=True)
    out = open(args.out, "a") if args.out else sys.stdout
    bad = 0
    try:
        with ProcessPoolExecutor(args.procs) as ex:
            for o in ex.map(check, todo):
                bad += not o["match"]
                out.write(json.dumps(o) + "\n")
                out.flush()
    finally:
        if out is not sys.stdout:
            out.close()
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
