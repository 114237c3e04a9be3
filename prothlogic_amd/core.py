"""ctypes binding of libprothcore.so (include/prothcore.h).

Every call goes to the HIP engine; there is no CPU fallback.  Loading fails
with an ImportError-like RuntimeError when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PL_CORE_LIB", PKG / "libprothcore.so"))

PL_OK = 0
PL_ERR_ARG = -1
PL_ERR_HIP = -2
PL_ERR_ROUNDOFF = -3
PL_ERR_STATE = -4
PL_ERR_IO = -5

_u64p = ctypes.POINTER(ctypes.c_uint64)
_dp = ctypes.POINTER(ctypes.c_double)


class PlError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"prothcore error {code}: {msg}")
        self.code = code


class PlInfo(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint64), ("n", ctypes.c_uint32), ("fft_len", ctypes.c_uint32),
                ("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32), ("bits_per_digit", ctypes.c_double),
                ("device_bytes", ctypes.c_uint64), ("passes", ctypes.c_uint32),
                ("handoff_fallbacks", ctypes.c_uint32), ("row_radix", ctypes.c_uint32),
                ("col_radix", ctypes.c_uint32), ("run", ctypes.c_uint32)]


class TuneTrial(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32), ("row_radix", ctypes.c_uint32),
                ("col_radix", ctypes.c_uint32), ("run", ctypes.c_uint32), ("one_launch", ctypes.c_uint32),
                ("us_per_squaring", ctypes.c_double), ("max_err", ctypes.c_double), ("exact", ctypes.c_int32),
                ("chosen", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double)


class PlTestArgs(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint64), ("n", ctypes.c_uint32), ("fft_len", ctypes.c_uint32),
                ("device", ctypes.c_int), ("checkpoint_path", ctypes.c_char_p),
                ("checkpoint_every", ctypes.c_uint64), ("stop_after", ctypes.c_uint64),
                ("err_guard", ctypes.c_double), ("report_step", ctypes.c_int), ("progress", PROGRESS_FN),
                ("progress_user", ctypes.c_void_p), ("final_limbs", ctypes.POINTER(ctypes.c_uint64)),
                ("final_n_limbs", ctypes.c_size_t), ("gerbicz_block", ctypes.c_uint32),
                ("gerbicz_blocks", ctypes.c_uint32), ("inject_error_at", ctypes.c_uint64)]


class PlTestResult(ctypes.Structure):
    _fields_ = [("verdict", ctypes.c_int), ("gated", ctypes.c_int), ("base", ctypes.c_uint64),
                ("res64", ctypes.c_uint64), ("iters_done", ctypes.c_uint64), ("resumed_at", ctypes.c_uint64),
                ("max_err", ctypes.c_double), ("fft_len", ctypes.c_uint32), ("gerbicz_checks", ctypes.c_uint32),
                ("gerbicz_rollbacks", ctypes.c_uint32)]


class PoolResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("k", ctypes.c_uint64), ("n", ctypes.c_uint32), ("fft_len", ctypes.c_uint32),
                ("verdict", ctypes.c_int), ("gated", ctypes.c_int), ("base", ctypes.c_uint64),
                ("res64", ctypes.c_uint64), ("squarings", ctypes.c_uint64), ("max_err", ctypes.c_double),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("error", ctypes.c_int)]


_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m prothlogic_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, sz, u64, u32, i32, dbl = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
    sig = {
        "pl_last_error": ([], ctypes.c_char_p),
        "pl_abi_version": ([], i32),
        "pl_device_count": ([ctypes.POINTER(i32)], i32),
        "pl_plan_fft_len": ([u32, u64, ctypes.POINTER(u32)], i32),
        "pl_ctx_create": ([u64, u32, u32, i32, ctypes.POINTER(vp)], i32),
        "pl_ctx_destroy": ([vp], None),
        "pl_ctx_info": ([vp, ctypes.POINTER(PlInfo)], i32),
        "pl_set_residue": ([vp, _u64p, sz], i32),
        "pl_get_residue": ([vp, _u64p, sz], i32),
        "pl_square": ([vp, u64, dbl, _dp, _dp], i32),
        "pl_mul_small": ([vp, u64], i32),
        "pl_res64": ([vp, _u64p], i32),
        "pl_is_n_minus_1": ([vp, ctypes.POINTER(i32)], i32),
        "pl_fft_square_karatsuba": ([vp, _u64p, sz, ctypes.c_uint, u64, _dp], i32),
        "pl_fft_square_karatsuba_noalloc": ([i32, _u64p, sz, ctypes.c_uint, u64, _dp], i32),
        "pl_kernel_times": ([vp, i32, _dp, _u64p], i32),
        "pl_copy_probe": ([i32, u64, i32, _dp], i32),
        "pl_loop_time": ([vp, i32, _dp, _u64p, _u64p], i32),
        "pl_autotune_fft_len": ([u32, u64, i32, dbl, ctypes.POINTER(u32), _dp, _dp], i32),
        "pl_autotune_plan": ([i32, u64, u32, ctypes.POINTER(TuneTrial), u32, ctypes.POINTER(u32),
                              ctypes.POINTER(TuneTrial)], i32),
        "pl_tune_pin": ([i32, u64, u32, ctypes.POINTER(TuneTrial)], i32),
        "pl_tune_clear": ([], i32),
        "pl_tune_save": ([ctypes.c_char_p], i32),
        "pl_tune_load": ([ctypes.c_char_p], i32),
        "pl_pool_autotune": ([vp, u64, u32, ctypes.POINTER(TuneTrial), u32, ctypes.POINTER(u32),
                              ctypes.POINTER(TuneTrial)], i32),
        "pl_sieve_range": ([u64, u64, u32, u64, i32, _u64p, sz, ctypes.POINTER(sz)], i32),
        "pl_sieve_list": ([_u64p, sz, u32, u64, u64, i32, ctypes.POINTER(ctypes.c_uint8)], i32),
        "pl_sieve_last_error": ([], ctypes.c_char_p),
        "pl_gate_rejects": ([u64, u32], i32),
        "pl_choose_base": ([u64, u32], u64),
        "pl_proth_test": ([ctypes.POINTER(PlTestArgs), ctypes.POINTER(PlTestResult)], i32),
        "pl_save_checkpoint": ([ctypes.c_char_p, u64, u32, u64, u64, _u64p, sz], i32),
        "pl_load_checkpoint": ([ctypes.c_char_p, u64, u32, _u64p, _u64p, _u64p, sz], i32),
        "pl_pool_create": ([i32, u64, u32, u64, ctypes.POINTER(vp)], i32),
        "pl_pool_destroy": ([vp], None),
        "pl_pool_submit": ([vp, u64, u32, u64, u64], i32),
        "pl_pool_run": ([vp, u32, ctypes.POINTER(PoolResult), sz, ctypes.POINTER(sz)], i32),
        "pl_pool_status": ([vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz), _u64p], i32),
        "pl_pool_last_error": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def _check(rc: int):
    if rc != PL_OK:
        raise PlError(rc, lib().pl_last_error().decode(errors="replace"))


def _limbs(x: int, nl: int):
    """Low nl 64-bit limbs of x >= 0, least significant first.  One
    to_bytes pass (linear in the size of x); a shift per limb would be
    quadratic, minutes at tens of millions of bits."""
    if x < 0:
        raise ValueError("residue must be non-negative")
    x &= (1 << (64 * nl)) - 1
    return (ctypes.c_uint64 * nl).from_buffer_copy(x.to_bytes(8 * nl, "little"))


def _int(arr) -> int:
    return int.from_bytes(bytes(arr), "little")


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(lib().pl_device_count(ctypes.byref(c)))
    return c.value


def plan_fft_len(n: int, k: int) -> int:
    L = ctypes.c_uint32(0)
    _check(lib().pl_plan_fft_len(n, k, ctypes.byref(L)))
    return L.value


def gate_rejects(k: int, n: int) -> bool:
    return bool(lib().pl_gate_rejects(k, n))


def choose_base(k: int, n: int) -> int:
    return int(lib().pl_choose_base(k, n))


def copy_probe(nbytes: int, reps: int = 10, device: int = 0) -> float:
    """ms per plain device-to-device copy of nbytes (reads and writes nbytes)."""
    ms = ctypes.c_double(0.0)
    _check(lib().pl_copy_probe(device, nbytes, reps, ctypes.byref(ms)))
    return ms.value


class Context:
    """One candidate N = k*2^n+1 with its residue resident on the device."""

    def __init__(self, k: int, n: int, fft_len: int = 0, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().pl_ctx_create(k, n, fft_len, device, ctypes.byref(h)))
        self.h = h
        self.k, self.n = k, n
        self.N = k * (1 << n) + 1
        self.nl = (n + 128) // 64 + 2
        info = PlInfo()
        _check(lib().pl_ctx_info(self.h, ctypes.byref(info)))
        self.info = info

    def close(self):
        if getattr(self, "h", None):
            lib().pl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def fft_len(self) -> int:
        return self.info.fft_len

    def refresh_info(self) -> PlInfo:
        info = PlInfo()
        _check(lib().pl_ctx_info(self.h, ctypes.byref(info)))
        self.info = info
        return info

    def set_residue(self, x: int):
        nl = max(1, (x.bit_length() + 63) // 64)
        _check(lib().pl_set_residue(self.h, _limbs(x, nl), nl))

    def get_residue(self) -> int:
        arr = (ctypes.c_uint64 * self.nl)()
        _check(lib().pl_get_residue(self.h, arr, self.nl))
        return _int(arr)

    def square(self, iters: int, err_guard: float = 0.0, trace: bool = False):
        e = ctypes.c_double(0.0)
        tr = (ctypes.c_double * iters)() if trace and iters else None
        _check(lib().pl_square(self.h, iters, err_guard, ctypes.byref(e), tr))
        return (e.value, list(tr)) if trace else e.value

    def mul_small(self, a: int):
        _check(lib().pl_mul_small(self.h, a))

    def res64(self) -> int:
        r = ctypes.c_uint64(0)
        _check(lib().pl_res64(self.h, ctypes.byref(r)))
        return r.value

    def is_n_minus_1(self) -> bool:
        r = ctypes.c_int(0)
        _check(lib().pl_is_n_minus_1(self.h, ctypes.byref(r)))
        return bool(r.value)

    KERNELS = ("fwd_col", "row_square", "inv_col", "carry", "fixup", "inv_carry", "adjust_fwd", "carry_turn")

    def kernel_times(self, reps: int = 20):
        """{kernel: ms per launch} timed with HIP events on the context's stream,
        plus the algorithmic bytes of one full-array pass."""
        ms = (ctypes.c_double * len(self.KERNELS))()
        tb = ctypes.c_uint64(0)
        _check(lib().pl_kernel_times(self.h, reps, ms, ctypes.byref(tb)))
        return dict(zip(self.KERNELS, list(ms))), tb.value

    def loop_time(self, reset: bool = False) -> dict:
        """Device time of the resident launches the context's squarings ran
        (HIP events around each launch), with launch and squaring counts."""
        ms = ctypes.c_double(0.0)
        nl, ns = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(lib().pl_loop_time(self.h, 1 if reset else 0, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(ns)))
        return dict(ms=ms.value, launches=nl.value, squarings=ns.value)

    def fft_square_karatsuba(self, limbs, n_limbs: int):
        arr = (ctypes.c_uint64 * n_limbs)(*limbs)
        e = ctypes.c_double(0.0)
        _check(lib().pl_fft_square_karatsuba(self.h, arr, n_limbs, self.n, self.k, ctypes.byref(e)))
        return list(arr), e.value


def fft_square_karatsuba_noalloc(limbs, n_limbs: int, n: int, k: int, device: int = 0):
    """v <- v^2 mod k*2^n+1 with no device allocation (the static scratch
    array; pl_fft_square_karatsuba_noalloc).  Returns (limbs, round-off)."""
    arr = (ctypes.c_uint64 * n_limbs)(*limbs)
    e = ctypes.c_double(0.0)
    _check(lib().pl_fft_square_karatsuba_noalloc(device, arr, n_limbs, n, k, ctypes.byref(e)))
    return list(arr), e.value


def proth_test(k: int, n: int, fft_len: int = 0, device: int = 0, checkpoint_path: str | None = None,
               checkpoint_every: int = 0, stop_after: int = 0, err_guard: float = 0.0, report_step: int = 0,
               progress=None, final_n_limbs: int = 0, gerbicz_block: int = 0, gerbicz_blocks: int = 0,
               inject_error_at: int = 0) -> dict:
    """Full Proth test on the device (pl_proth_test).  progress(done, total,
    max_err) is called at every report_step percent; final_n_limbs > 0
    returns the final residue as `residue`."""
    cb = PROGRESS_FN(lambda _u, done, total, err: progress(done, total, err)) if progress else PROGRESS_FN()
    fl = (ctypes.c_uint64 * final_n_limbs)() if final_n_limbs else None
    a = PlTestArgs(k, n, fft_len, device, checkpoint_path.encode() if checkpoint_path else None,
                   checkpoint_every, stop_after, err_guard, report_step, cb, None,
                   ctypes.cast(fl, ctypes.POINTER(ctypes.c_uint64)) if fl is not None else None, final_n_limbs,
                   gerbicz_block, gerbicz_blocks, inject_error_at)
    r = PlTestResult()
    _check(lib().pl_proth_test(ctypes.byref(a), ctypes.byref(r)))
    out = dict(verdict=r.verdict, gated=bool(r.gated), base=r.base, res64=r.res64, iters_done=r.iters_done,
               resumed_at=r.resumed_at, max_err=r.max_err, fft_len=r.fft_len, gerbicz_checks=r.gerbicz_checks,
               gerbicz_rollbacks=r.gerbicz_rollbacks)
    if fl is not None:
        out["residue"] = _int(fl)
    return out


def _scheck(rc: int):
    if rc != PL_OK:
        raise PlError(rc, lib().pl_sieve_last_error().decode(errors="replace"))


def sieve_range(k_start: int, k_end: int, n: int, pmax: int, device: int = 0) -> list:
    """Odd k in [k_start, k_end] with no prime factor p <= pmax of k*2^n+1."""
    cap = max(1, (k_end - k_start) // 2 + 1) if k_end >= k_start else 1  # odd k in the range
    out = (ctypes.c_uint64 * cap)()
    cnt = ctypes.c_size_t(0)
    _scheck(lib().pl_sieve_range(k_start, k_end, n, pmax, device, out, cap, ctypes.byref(cnt)))
    return list(out[:cnt.value])


def sieve_list(ks, n: int, pstart: int, pmax: int, device: int = 0) -> list:
    """keep flags of ks against the primes in [pstart, pmax]."""
    arr = (ctypes.c_uint64 * max(1, len(ks)))(*ks)
    keep = (ctypes.c_uint8 * max(1, len(ks)))()
    _scheck(lib().pl_sieve_list(arr, len(ks), n, pstart, pmax, device, keep))
    return [bool(v) for v in keep[:len(ks)]]


def autotune_fft_len(n: int, k: int, device: int = 0, err_guard: float = 0.0) -> dict:
    """Measured FFT length (pl_autotune_fft_len)."""
    L = ctypes.c_uint32(0)
    e = ctypes.c_double(0.0)
    us = ctypes.c_double(0.0)
    _check(lib().pl_autotune_fft_len(n, k, device, err_guard, ctypes.byref(L), ctypes.byref(e), ctypes.byref(us)))
    return dict(fft_len=L.value, max_err=e.value, us_per_squaring=us.value)


def _tune_out(call):
    trials = (TuneTrial * 64)()
    nt = ctypes.c_uint32(0)
    best = TuneTrial()
    call(trials, 64, ctypes.byref(nt), ctypes.byref(best))
    return dict(best=best.as_dict(), trials=[trials[i].as_dict() for i in range(min(nt.value, 64))])


def autotune_plan(k: int, n: int, device: int = 0) -> dict:
    """Times the launch plans of a single context for (k, n) and records the
    fastest exact one (pl_autotune_plan)."""
    return _tune_out(lambda t, m, c, b: _check(lib().pl_autotune_plan(device, k, n, t, m, c, b)))


def tune_pin(k: int, n: int, rows: int, row_radix: int = 8, col_radix: int = 8, run: int = 8,
             one_launch: bool = True, device: int = 0):
    """Records a plan choice for single contexts of (k, n) (pl_tune_pin)."""
    c = TuneTrial(rows=rows, cols=0, row_radix=row_radix, col_radix=col_radix, run=run,
                  one_launch=1 if one_launch else 0)
    _check(lib().pl_tune_pin(device, k, n, ctypes.byref(c)))


def tune_clear():
    _check(lib().pl_tune_clear())


def tune_save(path: str):
    _check(lib().pl_tune_save(str(path).encode()))


def tune_load(path: str):
    _check(lib().pl_tune_load(str(path).encode()))


def save_checkpoint(path: str, k: int, n: int, base: int, next_iter: int, x: int, nl: int):
    _check(lib().pl_save_checkpoint(path.encode(), k, n, base, next_iter, _limbs(x, nl), nl))


def load_checkpoint(path: str, k: int, n: int, nl: int):
    b = ctypes.c_uint64(0)
    it = ctypes.c_uint64(0)
    arr = (ctypes.c_uint64 * nl)()
    _check(lib().pl_load_checkpoint(path.encode(), k, n, ctypes.byref(b), ctypes.byref(it), arr, nl))
    return b.value, it.value, _int(arr)


def _pcheck(rc: int):
    if rc != PL_OK:
        raise PlError(rc, lib().pl_pool_last_error().decode(errors="replace"))


class Pool:
    """Continuous batching of Proth tests on one GPU (pl_pool_*)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_slots_per_length: int = 0, k_max: int = 0):
        h = ctypes.c_void_p()
        _pcheck(lib().pl_pool_create(device, hbm_budget, max_slots_per_length, k_max, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pl_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, k: int, n: int, tag: int = 0, stop_after: int = 0):
        _pcheck(lib().pl_pool_submit(self.h, k, n, tag, stop_after))

    def autotune(self, k: int, n: int) -> dict:
        """Tunes the plan of the group (k, n) would join (pl_pool_autotune);
        call before that group is created."""
        return _tune_out(lambda t, m, c, b: _pcheck(lib().pl_pool_autotune(self.h, k, n, t, m, c, b)))

    def run(self, steps: int = 64, cap: int = 4096):
        buf = (PoolResult * cap)()
        got = ctypes.c_size_t(0)
        _pcheck(lib().pl_pool_run(self.h, steps, buf, cap, ctypes.byref(got)))
        return [{f: getattr(buf[i], f) for f, _ in PoolResult._fields_} for i in range(got.value)]

    def status(self):
        q, a, f = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        b = ctypes.c_uint64()
        _pcheck(lib().pl_pool_status(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(f), ctypes.byref(b)))
        return dict(queued=q.value, active=a.value, finished_unread=f.value, hbm_bytes=b.value)

    def drain(self, steps: int = 64, on_result=None):
        """Run until every submitted test has finished; returns all results."""
        out = []
        while True:
            got = self.run(steps)
            for r in got:
                if on_result:
                    on_result(r)
            out.extend(got)
            st = self.status()
            if st["queued"] == 0 and st["active"] == 0 and st["finished_unread"] == 0:
                return out
