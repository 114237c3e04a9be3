"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the reference's candidate
sieve, used by tests/ as the checker of the device sieve (pl_sieve_range,
pl_sieve_list).  The product never imports it.

Reference behaviour restated (paths relative to the reference tree,
P/ = prothlogic4.2/):
  * execute_full_streaming_sieve, P/presieve_core.cpp:15-88: every prime
    3 <= p <= pmax (windows over [0, pmax], P/presieve_core.cpp:48-66) strikes
    k == -2^(-n) (mod p) in [k_start, k_end] (P/gpu_engine.cu:290-318:
    inv2 = (p+1)/2, target = p - inv2^n mod p); the result keeps odd k only
    (P/presieve_core.cpp:78-83).  The reference's window sieve marks the
    multiples of only the first primes (one prime per block,
    P/gpu_engine.cu:89 with the grid of :144), so its "prime" list also
    holds composites; a
    composite m striking k means m | k*2^n+1, so one of m's prime factors
    (also in the list) strikes the same k: the full result is the one
    restated here.
  * execute_sparse_streaming_sieve, P/presieve_core.cpp:93-167: the same
    test for a list of k and primes from pstart.  The reference compares k
    itself, not k mod p, with the target (P/gpu_engine.cu:397), so it strikes
    only k < p; restated here is the intended test (k mod p == target), and
    that divergence is deliberate.

Pinned by: brute force (k*2^n+1) mod p with Python integers on small
inputs, and the closed form p | t_p*2^n + 1 for the targets
(tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def primes_upto(m: int) -> np.ndarray:
    """Primes 2 <= p <= m (plain sieve of Eratosthenes)."""
    if m < 2:
        return np.zeros(0, dtype=np.int64)
    s = np.ones(m + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(m ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0].astype(np.int64)


def target(p: int, n: int) -> int:
    """-2^(-n) mod p: the k residue p strikes (P/gpu_engine.cu:304-308)."""
    return (p - pow((p + 1) // 2, n, p)) % p


def sieve_range(k_start: int, k_end: int, n: int, pmax: int) -> list:
    """Odd survivors of [k_start, k_end] (P/presieve_core.cpp:15-88)."""
    if k_end < k_start:
        return []
    size = k_end - k_start + 1
    keep = np.ones(size, dtype=bool)
    for p in primes_upto(pmax):
        p = int(p)
        if p < 3:
            continue
        off = (target(p, n) - k_start) % p
        keep[off::p] = False
    idx = np.nonzero(keep)[0]
    ks = [k_start + int(i) for i in idx]
    return [k for k in ks if k & 1]


def sieve_list(ks, n: int, pstart: int, pmax: int) -> list:
    """keep flags of ks against the primes in [max(pstart, 3), pmax]
    (the intended test of P/presieve_core.cpp:93-167)."""
    keep = [True] * len(ks)
    for p in primes_upto(pmax):
        p = int(p)
        if p < 3 or p < pstart:
            continue
        t = target(p, n)
        for i, k in enumerate(ks):
            if keep[i] and k % p == t:
                keep[i] = False
    return keep
