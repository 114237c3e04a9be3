"""Device candidate sieve (pl_sieve_range / pl_sieve_list) against the
oracle's restatement of the reference sieve (oracle/sieve_oracle.py, itself
pinned by brute force in tests/test_oracle.py).  Integer work: bit-exact."""
import random

import pytest

import sieve_oracle as SO


@pytest.mark.gpu
@pytest.mark.parametrize("n,k0,k1,pmax", [
    (1000, 1, 100_000, 10_000),              # one tile, small primes only
    (1 << 21, 3, 2_000_001, 1_000_000),      # 8 tiles; primes above 2^18 take the global path
    (120002, 10**9, 10**9 + 700_000, 3_000_000),
    (33, 1, 5000, 100_000),                  # primes far above the range (<= 1 strike each)
    (303093, 1, 1, 1_000_000),               # a single k
    (97, 10, 9, 1000),                       # empty range
    (500, 1, 999, 2),                        # no odd prime up to pmax: every odd k survives
    (500, 1, 999, 3),
])
def test_sieve_range_matches_oracle(gpu, n, k0, k1, pmax):
    print(f"[sieve] n={n} k=[{k0},{k1}] pmax={pmax}", flush=True)
    assert gpu.sieve_range(k0, k1, n, pmax) == SO.sieve_range(k0, k1, n, pmax)


@pytest.mark.gpu
def test_sieve_range_many_windows(gpu, monkeypatch):
    """Prime windows of 2^16 odd numbers: pmax = 10^6 spans 8 of them."""
    monkeypatch.setenv("PL_SIEVE_WINDOW_ODD", str(1 << 16))
    n, k0, k1, pmax = 4096, 1, 300_000, 1_000_000
    print(f"[sieve] windows of 2^16: n={n} pmax={pmax}", flush=True)
    assert gpu.sieve_range(k0, k1, n, pmax) == SO.sieve_range(k0, k1, n, pmax)


@pytest.mark.gpu
def test_sieve_list_matches_oracle(gpu):
    rnd = random.Random(11)
    ks = [rnd.randrange(3, 1 << 48) | 1 for _ in range(5000)]  # two list chunks of 4096
    n, pstart, pmax = 120002, 1000, 200_000
    print(f"[sieve] list of {len(ks)} n={n} p in [{pstart},{pmax}]", flush=True)
    got = gpu.sieve_list(ks, n, pstart, pmax)
    assert got == SO.sieve_list(ks, n, pstart, pmax)
    assert 0 < sum(got) < len(ks)


@pytest.mark.gpu
def test_sieve_keeps_the_known_primes(gpu):
    """The known primes of the workloads survive any sieve depth below them."""
    for k, n in [(3, 20909), (3, 34350), (843, 120002), (3, 303093)]:
        assert k in gpu.sieve_range(k, k + 2, n, 10_000_000)
