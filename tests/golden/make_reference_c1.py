"""Writes tests/golden/c1_reference.json: the reference engine's own full
Proth test of the n = 32768 workload's candidate (k = 1023, PCG64 seed 1),
run through oracle/_ref (the reference's sources compiled in this container
by oracle/Makefile) at the FFT length its autotuner picks and at the next
longer one, beside GMP's verdict and Res64.

  make -C oracle ref && python tests/golden/make_reference_c1.py

The reference driver writes its checkpoint and primes.txt into the working
directory, so it runs in a temporary one.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
sys.path.insert(0, str(HERE.parent.parent))

import gmp_ctypes as G  # noqa: E402
from oracle import RefLib  # noqa: E402
from prothlogic_amd import workloads as W  # noqa: E402


def main():
    (k, n), = W.c1().candidates
    ref = RefLib()
    prime, r64, a = G.proth_verdict(k, n)
    out = {"generator": "tests/golden/make_reference_c1.py", "k": k, "n": n, "a": a,
           "gmp": {"prime": bool(prime), "res64": format(r64, "016X")},
           "reference_autotune_fft_len": int(ref.lib.ref_auto_tune_fft_len(n, k)), "reference": {}}
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            for L in (out["reference_autotune_fft_len"], 2 * out["reference_autotune_fft_len"]):
                buf = (ctypes.c_uint64 * L)()
                v = ref.lib.ref_execute_proth_test(k, n, L, buf, L)
                out["reference"][str(L)] = {"prime": v == 1, "res64": format(buf[0], "016X")}
                print(L, out["reference"][str(L)], flush=True)
        finally:
            os.chdir(cwd)
    p = HERE / "c1_reference.json"
    p.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print("wrote", p)


if __name__ == "__main__":
    main()
