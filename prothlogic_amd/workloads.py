"""Seeded synthetic Proth candidates for the five benchmark configurations.

Each workload draws odd k = 2*rng.integers(1, 2**(b-1)) + 1 in [3, 2**b) from
numpy's PCG64 with a fixed seed, and redraws any k whose N = k*2^n+1 has a
prime factor p <= 67 (the reference's silent gate, P/proth_logic.cpp:405-408)
so that every drawn candidate reaches the squaring loop.  The number of
redraws is reported with the workload.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GATE_PRIMES = (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67)


def gate_rejects(k: int, n: int) -> bool:
    """True when some odd prime p <= 67 divides k*2^n+1 (P/proth_logic.cpp:41-50)."""
    return any((k % p) * pow(2, n, p) % p == p - 1 for p in GATE_PRIMES)


@dataclass
class Workload:
    name: str
    seed: int
    candidates: list = field(default_factory=list)   # list of (k, n)
    redraws: int = 0


def _draw_k(rng, b: int) -> int:
    return 2 * int(rng.integers(1, 2 ** (b - 1))) + 1


def _survivor(rng, b: int, n_of):
    redraws = 0
    while True:
        n = n_of()
        k = _draw_k(rng, b)
        if not gate_rejects(k, n):
            return k, n, redraws
        redraws += 1


def single(name: str, seed: int, b: int, n: int) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    k, n, r = _survivor(rng, b, lambda: n)
    return Workload(name, seed, [(k, n)], r)


def c1() -> Workload:
    """n=32768, k in [3, 2^12): the reference's CPU-runnable full test."""
    return single("c1_n32768", 1, 12, 32768)


def c2() -> Workload:
    """n=2^21, k in [3, 2^20): single large test, timed over 10^4 squarings."""
    return single("c2_n2p21", 2, 20, 1 << 21)


def c3() -> Workload:
    """n=2^24, k in [3, 2^16): single very large test."""
    return single("c3_n2p24", 3, 16, 1 << 24)


def c4(count: int = 256) -> Workload:
    """count candidates, k in [3, 2^16), n uniform in [2^18, 2^20]."""
    rng = np.random.Generator(np.random.PCG64(4))
    w = Workload("c4_batch", 4)
    for _ in range(count):
        k, n, r = _survivor(rng, 16, lambda: int(rng.integers(1 << 18, (1 << 20) + 1)))
        w.candidates.append((k, n))
        w.redraws += r
    return w


def c5(count: int = 2000) -> Workload:
    """count candidates, k in [3, 2^20), n log-uniform in [5e4, 4e6]."""
    rng = np.random.Generator(np.random.PCG64(5))
    lo, hi = math.log(5e4), math.log(4e6)
    w = Workload("c5_serving", 5)
    for _ in range(count):
        k, n, r = _survivor(rng, 20, lambda: int(math.floor(math.exp(rng.uniform(lo, hi)))))
        w.candidates.append((k, n))
        w.redraws += r
    return w


# Known Proth primes usable as injected "prime" labels (GMP-verified).
KNOWN_PRIMES = ((3, 20909), (3, 34350), (843, 120002), (3, 303093))
