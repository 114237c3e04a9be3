"""GMP through its public C ABI (libgmp.so.10) via ctypes -- the exactness
arbiter of the test-suite.

GMP's mpz_powm / mpz_jacobi are mathematically defined, so any GMP version
(or Python's own int) is a valid oracle.  Python ints are used when libgmp
cannot be loaded; they are exact, only slower.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os


class _mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def _load():
    cands = ["libgmp.so.10", "/usr/lib/x86_64-linux-gnu/libgmp.so.10", "/opt/conda/lib/libgmp.so.10"]
    found = ctypes.util.find_library("gmp")
    if found:
        cands.insert(0, found)
    for c in cands:
        try:
            return ctypes.CDLL(c)
        except OSError:
            continue
    return None


_lib = _load()
if _lib is not None:
    _P = ctypes.POINTER(_mpz)

    def _fn(name, args, res=None):
        f = getattr(_lib, name)
        f.argtypes = args
        f.restype = res
        return f

    _z_init = _fn("__gmpz_init", [_P])
    _z_clear = _fn("__gmpz_clear", [_P])
    _z_set_str = _fn("__gmpz_set_str", [_P, ctypes.c_char_p, ctypes.c_int], ctypes.c_int)
    _z_get_str = _fn("__gmpz_get_str", [ctypes.c_char_p, ctypes.c_int, _P], ctypes.c_void_p)
    _z_powm = _fn("__gmpz_powm", [_P, _P, _P, _P])
    _z_jacobi = _fn("__gmpz_jacobi", [_P, _P], ctypes.c_int)
    _libc = ctypes.CDLL(None)
    _libc.free.argtypes = [ctypes.c_void_p]

HAVE_GMP = _lib is not None


class _Z:
    def __init__(self, v: int = 0):
        self.z = _mpz()
        _z_init(ctypes.byref(self.z))
        if v:
            self.set(v)

    def set(self, v: int):
        assert v >= 0
        s = format(v, "x").encode()
        if _z_set_str(ctypes.byref(self.z), s, 16) != 0:
            raise ValueError("mpz_set_str failed")

    def get(self) -> int:
        p = _z_get_str(None, 16, ctypes.byref(self.z))
        s = ctypes.string_at(p).decode()
        _libc.free(p)
        return int(s, 16)

    def __del__(self):
        try:
            _z_clear(ctypes.byref(self.z))
        except Exception:
            pass


def powm(base: int, exp: int, mod: int) -> int:
    """base^exp mod mod (GMP mpz_powm when available)."""
    if not HAVE_GMP:
        return pow(base, exp, mod)
    r, b, e, m = _Z(), _Z(base), _Z(exp), _Z(mod)
    _z_powm(ctypes.byref(r.z), ctypes.byref(b.z), ctypes.byref(e.z), ctypes.byref(m.z))
    return r.get()


def jacobi(a: int, n: int) -> int:
    if not HAVE_GMP:
        return _jacobi_py(a, n)
    x, y = _Z(a), _Z(n)
    return _z_jacobi(ctypes.byref(x.z), ctypes.byref(y.z))


def _jacobi_py(a: int, n: int) -> int:
    assert n > 0 and n % 2 == 1
    a %= n
    t = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                t = -t
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            t = -t
        a %= n
    return t if n == 1 else 0


def proth_N(k: int, n: int) -> int:
    return k * (1 << n) + 1


def reference_base(k: int, n: int) -> int:
    """The reference's base rule (P/proth_logic.cpp:423-435) with GMP's jacobi."""
    N = proth_N(k, n)
    for b in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if jacobi(b, N) == -1:
            return b
    for b in range(2, 100):
        if jacobi(b, N) == -1:
            return b
    return 0


def residue_after(k: int, n: int, a: int, iters: int) -> int:
    """a^(k*2^iters) mod N: the residue after `iters` squarings of the seed a^k."""
    N = proth_N(k, n)
    return powm(a, k << iters, N)


def proth_verdict(k: int, n: int, a: int | None = None):
    """(is_prime, res64, a) for the full Proth test a^((N-1)/2) mod N."""
    if a is None:
        a = reference_base(k, n)
    N = proth_N(k, n)
    r = powm(a, (N - 1) // 2, N)
    return r == N - 1, r & 0xFFFFFFFFFFFFFFFF, a
