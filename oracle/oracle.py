"""ctypes bindings of the test-only CPU checkers.

TEST INFRASTRUCTURE ONLY: importable from tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never imports this module.

  Oracle   -- oracle/liboracle.so, the C restatement of the reference path
  RefLib   -- oracle/_ref/libprothref.so, the reference's own sources built
              here by `make -C oracle ref` (absent where the reference tree
              or AVX-512 is absent)
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def int_to_limbs(x: int, nl: int) -> np.ndarray:
    assert x >= 0
    b = x.to_bytes(nl * 8, "little")
    return np.frombuffer(b, dtype=np.uint64).copy()


def limbs_to_int(a: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint64).tobytes(), "little")


class Oracle:
    def __init__(self, path: str | os.PathLike | None = None):
        path = Path(path) if path else HERE / "liboracle.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(path))
        vp, sz, u64, u32, dbl, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint, ctypes.c_double, ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_fft_init.argtypes = [sz]; L.or_fft_init.restype = vp
        L.or_fft_destroy.argtypes = [vp]
        L.or_fft_len.argtypes = [vp]; L.or_fft_len.restype = sz
        L.or_last_max_diff.argtypes = [vp]; L.or_last_max_diff.restype = dbl
        L.or_fft_inplace.argtypes = [vp, dp, dp, i32]
        L.or_pack.argtypes = [vp, _u64p, sz]
        L.or_unpack.argtypes = [vp, _u64p, sz, u32, u64]
        L.or_square.argtypes = [vp, _u64p, sz, u32, u64]
        L.or_gate_rejects.argtypes = [u64, u32]; L.or_gate_rejects.restype = i32
        L.or_choose_base.argtypes = [u64, u32]; L.or_choose_base.restype = u64
        L.or_jacobi_small.argtypes = [u64, u64, u32]; L.or_jacobi_small.restype = i32
        L.or_auto_tune_fft_len.argtypes = [u32, u64, i32, dp]; L.or_auto_tune_fft_len.restype = sz
        L.or_fits.argtypes = [sz, u32, u64]; L.or_fits.restype = i32
        L.or_seed.argtypes = [vp, _u64p, u64, u64, u32]
        L.or_is_nm1.argtypes = [_u64p, sz, u32, u64]; L.or_is_nm1.restype = i32
        L.or_save_checkpoint.argtypes = [u64, sz, _u64p, ctypes.c_char_p]; L.or_save_checkpoint.restype = i32
        L.or_load_checkpoint.argtypes = [_u64p, sz, _u64p, ctypes.c_char_p]; L.or_load_checkpoint.restype = i32
        L.or_execute_proth_test.argtypes = [u64, u32, _u64p, vp, ctypes.c_char_p, u64, _u64p, dp, _u64p]
        L.or_execute_proth_test.restype = i32
        L.or_mod_N.argtypes = [_u64p, sz, u32, u64]
        self.lib = L

    # -- engine -----------------------------------------------------------
    def ctx(self, fft_len: int) -> "OracleCtx":
        return OracleCtx(self, fft_len)

    def gate_rejects(self, k: int, n: int) -> bool:
        return bool(self.lib.or_gate_rejects(k, n))

    def choose_base(self, k: int, n: int) -> int:
        return int(self.lib.or_choose_base(k, n))

    def jacobi_small(self, b: int, k: int, n: int) -> int:
        return int(self.lib.or_jacobi_small(b, k, n))

    def auto_tune_fft_len(self, n: int, k: int, fit_fix: bool = True):
        e = ctypes.c_double(0.0)
        L = self.lib.or_auto_tune_fft_len(n, k, 1 if fit_fix else 0, ctypes.byref(e))
        return int(L), e.value

    def fits(self, L: int, n: int, k: int) -> bool:
        return bool(self.lib.or_fits(L, n, k))

    def mod_N(self, x: int, nl: int, n: int, k: int) -> int:
        a = int_to_limbs(x, nl)
        self.lib.or_mod_N(_ptr(a), nl, n, k)
        return limbs_to_int(a)

    def is_nm1(self, limbs: np.ndarray, n: int, k: int) -> bool:
        return bool(self.lib.or_is_nm1(_ptr(limbs), limbs.size, n, k))

    def save_checkpoint(self, i: int, limbs: np.ndarray, path: str) -> bool:
        return bool(self.lib.or_save_checkpoint(i, limbs.size, _ptr(limbs), str(path).encode()))

    def load_checkpoint(self, n_limbs: int, path: str):
        out = np.zeros(n_limbs, dtype=np.uint64)
        i = ctypes.c_uint64(0)
        ok = self.lib.or_load_checkpoint(ctypes.byref(i), n_limbs, _ptr(out), str(path).encode())
        return (int(i.value), out) if ok else (None, None)

    def proth_test(self, k: int, n: int, fft_len: int | None = None, checkpoint_dir: str | None = None,
                   stop_after: int = 0):
        """P/proth_logic.cpp:384-580 restated.  Returns dict(verdict, a, max_err, iters, limbs)."""
        if fft_len is None:
            fft_len, _ = self.auto_tune_fft_len(n, k, True)
        c = self.ctx(fft_len)
        cap = max(fft_len, 2 * n // 64 + 64) + 8
        limbs = np.zeros(cap, dtype=np.uint64)
        a = ctypes.c_uint64(0)
        e = ctypes.c_double(0.0)
        it = ctypes.c_uint64(0)
        v = self.lib.or_execute_proth_test(k, n, _ptr(limbs), c.h,
                                           checkpoint_dir.encode() if checkpoint_dir else None,
                                           stop_after, ctypes.byref(a), ctypes.byref(e), ctypes.byref(it))
        return dict(verdict=int(v), a=int(a.value), max_err=e.value, iters=int(it.value),
                    limbs=limbs[: fft_len // 4].copy(), fft_len=fft_len)


class OracleCtx:
    def __init__(self, o: Oracle, L: int):
        self.o = o
        self.L = L
        self.h = o.lib.or_fft_init(L)
        if not self.h:
            raise MemoryError("or_fft_init failed")

    def __del__(self):
        try:
            self.o.lib.or_fft_destroy(self.h)
        except Exception:
            pass

    @property
    def last_max_diff(self) -> float:
        return float(self.o.lib.or_last_max_diff(self.h))

    def fft(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        self.o.lib.or_fft_inplace(self.h, _dptr(re), _dptr(im), 1 if inverse else 0)
        return re + 1j * im

    def square_limbs(self, v: np.ndarray, n_limbs: int, n: int, k: int) -> np.ndarray:
        """In-place P/fft_engine.cpp:986 on a copy; v must hold >= round_up8 chunks."""
        w = np.zeros(max(v.size, n_limbs + 4), dtype=np.uint64)
        w[: v.size] = v
        self.o.lib.or_square(self.h, _ptr(w), n_limbs, n, k)
        return w[:n_limbs].copy()

    def square(self, x: int, n: int, k: int) -> int:
        nl = self.L // 4
        return limbs_to_int(self.square_limbs(int_to_limbs(x, nl), nl, n, k))

    def seed(self, a: int, k: int, n: int) -> int:
        nl = max(self.L // 4, (n + 128) // 64 + 2) + 8
        w = np.zeros(nl, dtype=np.uint64)
        self.o.lib.or_seed(self.h, _ptr(w), a, k, n)
        return limbs_to_int(w)


def avx512_available() -> bool:
    try:
        txt = Path("/proc/cpuinfo").read_text()
    except OSError:
        return False
    flags = set()
    for line in txt.splitlines():
        if line.startswith("flags"):
            flags = set(line.split(":", 1)[1].split())
            break
    return {"avx512f", "avx512dq", "avx512bw", "avx512vl", "bmi2", "adx", "fma"} <= flags


class RefLib:
    """The reference engine (oracle/_ref/libprothref.so).  Raises when it is
    not built or the CPU lacks the AVX-512 subsets it was compiled for."""

    def __init__(self, path: str | os.PathLike | None = None):
        path = Path(path) if path else HERE / "_ref" / "libprothref.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference tree is present")
        if not avx512_available():
            raise RuntimeError("reference build needs AVX-512F/DQ/BW/VL")
        L = ctypes.CDLL(str(path))
        vp, sz, u64, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint
        L.ref_fft_init.argtypes = [sz]; L.ref_fft_init.restype = vp
        L.ref_fft_destroy.argtypes = [vp]
        L.ref_fft_square.argtypes = [vp, _u64p, sz, u32, u64]
        L.ref_last_max_diff.argtypes = [vp]; L.ref_last_max_diff.restype = ctypes.c_double
        L.ref_auto_tune_fft_len.argtypes = [u32, u64]; L.ref_auto_tune_fft_len.restype = sz
        L.ref_execute_proth_test.argtypes = [u64, u32, sz, _u64p, sz]; L.ref_execute_proth_test.restype = ctypes.c_int
        self.lib = L

    def ctx(self, L: int):
        h = self.lib.ref_fft_init(L)
        if not h:
            raise MemoryError("ref fft_init failed")
        return h

    def destroy(self, h):
        self.lib.ref_fft_destroy(h)

    def square_limbs(self, h, v: np.ndarray, n_limbs: int, n: int, k: int) -> np.ndarray:
        w = np.zeros(max(v.size, n_limbs + 4), dtype=np.uint64)
        w[: v.size] = v
        self.lib.ref_fft_square(h, _ptr(w), n_limbs, n, k)
        return w[:n_limbs].copy()

    def last_max_diff(self, h) -> float:
        return float(self.lib.ref_last_max_diff(h))
