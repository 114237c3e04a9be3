"""Builds the native libraries in-tree with hipcc (gfx950) and g++.

  python -m prothlogic_amd.build          # or __graft_entry__.build()

Outputs (git-ignored, shipped with the tree):
  prothlogic_amd/libprothcore.so   HIP kernels + C ABI (include/prothcore.h)
  prothlogic_amd/libprothhost.so   C++ host mirror of the reference API
                                   (fft_engine / proth_logic / worker /
                                   checkpoint_manager) over the C ABI
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PL_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wno-unused-result", "-Wno-unused-command-line-argument"]

CORE_SRCS = [CSRC / "engine.hip", CSRC / "capi.cpp", CSRC / "bignum.cpp", CSRC / "prewrap.cpp", CSRC / "pool.cpp",
             CSRC / "tune.cpp", CSRC / "sieve.hip", CSRC / "noalloc.hip"]
HOST_SRCS = [HOST / "fft_engine.cpp", HOST / "proth_logic.cpp", HOST / "checkpoint_manager.cpp",
             HOST / "worker.cpp", HOST / "autotune.cpp", HOST / "presieve_core.cpp"]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_core(defines: tuple = (), out: Path | None = None) -> Path:
    """hipcc -> libprothcore.so.  `defines` (e.g. ("PL_MAX_RADIX=8",)) and
    `out` build a variant library for A/B measurements."""
    out = Path(out) if out else PKG / "libprothcore.so"
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-shared", f"-I{INCLUDE}", f"-I{CSRC}", "-o", tmp,
           *CORE_SRCS]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_host() -> Path | None:
    srcs = [s for s in HOST_SRCS if s.exists()]
    if not srcs:
        return None
    out = PKG / "libprothhost.so"
    tmp = out.with_suffix(".so.tmp")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", f"-I{HOST}", "-o", tmp, *srcs,
           f"-L{PKG}", "-lprothcore", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(tmp, out)
    # the mirror's self-test program (tests/test_host_mirror.py runs it)
    exe = PKG / "prothhost_selftest"
    tmp = PKG / "prothhost_selftest.tmp"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-pthread", f"-I{INCLUDE}", f"-I{HOST}", "-o", tmp,
          HOST / "selftest.cpp", f"-L{PKG}", "-lprothhost", "-lprothcore", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, exe)
    # the command-line front end
    exe = PKG / "prothlogic"
    tmp = PKG / "prothlogic.tmp"
    _run([cxx, "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{HOST}", "-o", tmp, HOST / "cli.cpp",
          f"-L{PKG}", "-lprothhost", "-lprothcore", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, exe)
    return out


def build_all() -> None:
    build_core()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
