"""Parity of the HIP engine (through the C ABI) with exact arithmetic, the
GMP-generated golden fixtures, and the oracle (CPU restatement of the
reference engine).  Integer results must be bit-exact."""
import random

import numpy as np
import pytest

import golden_io
import gmp_ctypes as G

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def mont_inv(n, N):
    return pow(pow(2, 2 * n, N), -1, N)


# (k, n, forced L or 0): small lengths through 2^20, bits/digit 8..19, k up to 2^32
SQUARE_CASES = [
    (3, 9000, 1024), (843, 18000, 1024), (1023, 32768, 2048), (1048573, 40000, 4096),
    (433917, 70000, 4096), (65535, 150001, 8192), (11887, 300007, 16384), (12345, 262144, 16384),
    (268435399, 500000, 32768), (4294967291, 500000, 0), (3, 1000003, 65536), (999999, 2097152, 131072),
    (24675, 484807, 32768), (7, 4000037, 262144),
]


@pytest.mark.parametrize("k,n,L", SQUARE_CASES)
def test_square_exact_vs_python_int(gpu, k, n, L):
    rnd = random.Random(k ^ n)
    N = k * (1 << n) + 1
    with gpu.Context(k, n, L) as c:
        if L == 0:
            # k near 2^32 bounds the digit size (the top scalar's share of the
            # round-off): 15.3 bits at L = 2^15 is too many, 7.6 at 2^16 fits
            assert k < 1 << 26 or c.fft_len == 65536
        for _ in range(2):
            x = rnd.randrange(N)
            c.set_residue(x)
            assert c.get_residue() == x
            e = c.square(1)
            assert c.get_residue() == x * x % N, (k, n, L)
            assert 0 <= e < 0.25


# Steady-state modes: "resident" (default: the whole steady state of a chunk
# as one resident launch where the length has such a shape, else as
# "turn"), "turn" (PL_RESIDENT=0: one launch for the outer pass with the
# carry hand-off between column blocks, 2 kernels per squaring), "split"
# (PL_CARRY_TURN=0: two launches; 3 kernels), "unfused" (PL_FUSED=0: 5).
@pytest.mark.parametrize("mode", ["resident", "turn", "split", "unfused"])
@pytest.mark.parametrize("k,n,L", [(3, 20000, 2048), (1048573, 60000, 4096), (11887, 300007, 16384),
                                   (433917, 2097152, 131072), (65535, 3000017, 262144), (3, 4000037, 262144)])
def test_fused_and_unfused_paths(gpu, monkeypatch, mode, k, n, L):
    monkeypatch.setenv("PL_FUSED", "0" if mode == "unfused" else "1")
    monkeypatch.setenv("PL_CARRY_TURN", "0" if mode == "split" else "1")
    monkeypatch.setenv("PL_RESIDENT", "1" if mode == "resident" else "0")
    rnd = random.Random(n)
    N = k * (1 << n) + 1
    x = rnd.randrange(N)
    # shorter lengths: one 256-squaring graph chunk plus a partial one; the
    # long lengths' full chunks are covered by test_large_golden_c2_c3
    iters = 300 if n < 10 ** 6 else 70
    print(f"[parity] mode={mode} k={k} n={n} L={L} iters={iters}", flush=True)
    with gpu.Context(k, n, L) as c:
        if mode == "unfused":
            assert c.info.passes == 5
        elif n / L >= 12:
            # runs of 16 digits (large k for the digit size) keep two launches
            want = {"resident": (0, 2, 3), "turn": (2, 3), "split": (3,)}[mode]
            assert c.info.passes in want, "split outer pass expected for this length"
            if mode == "turn" and n >= 10 ** 6:
                assert c.info.passes == 2
            if mode == "resident" and L in (131072, 262144):
                assert c.info.passes == 0
        c.set_residue(x)
        e, tr = c.square(iters, trace=True)
        assert 0 < e < 0.3 and len(tr) == iters
        assert c.get_residue() == G.powm(x, 1 << iters, N)


@pytest.mark.parametrize("k,n,L", [(3, 20000, 2048), (1048573, 60000, 4096), (11887, 300007, 16384)])
def test_square_chain_vs_powm(gpu, k, n, L):
    rnd = random.Random(7)
    N = k * (1 << n) + 1
    x = rnd.randrange(N)
    with gpu.Context(k, n, L) as c:
        c.set_residue(x)
        e, tr = c.square(300, trace=True)
        assert len(tr) == 300 and max(tr) == e and min(tr) > 0
        assert c.get_residue() == G.powm(x, 1 << 300, N)


def test_edge_residues(gpu):
    k, n = 1023, 32768
    N = k * (1 << n) + 1
    with gpu.Context(k, n) as c:
        for x in (0, 1, 2, N - 1, N - 2, (1 << n) - 1, 1 << n, k << (n - 1), N // 2):
            c.set_residue(x)
            c.square(1)
            assert c.get_residue() == x * x % N, x
        # all-ones pattern of the reference autotune (P/dcore.cpp:52-56), reduced mod N
        x = ((1 << (n - 1)) - 1) | 0
        c.set_residue(x)
        c.square(5)
        assert c.get_residue() == G.powm(x, 32, N)


def test_mul_small_and_res64(gpu):
    k, n = 843, 20000
    N = k * (1 << n) + 1
    with gpu.Context(k, n) as c:
        c.set_residue(12345678901234567890)
        c.mul_small(97)
        assert c.get_residue() == 12345678901234567890 * 97 % N
        assert c.res64() == (12345678901234567890 * 97 % N) & M64
        c.set_residue(N - 1)
        assert c.is_n_minus_1()
        c.set_residue(N - 2)
        assert not c.is_n_minus_1()


def test_small_golden_res64_and_verdicts(gpu):
    data = golden_io.load("small_cases")
    assert data is not None
    for case in data["cases"]:
        k, n, a = case["k"], case["n"], case["a"]
        if n < 8192 or case["gated"]:
            continue
        N = k * (1 << n) + 1
        with gpu.Context(k, n) as c:
            c.set_residue(pow(a, k, N))
            done = 0
            for i_s, r64 in sorted(case["res64_after"].items(), key=lambda t: int(t[0])):
                i = int(i_s)
                c.square(i - done)
                done = i
                assert c.res64() == int(r64, 16), (k, n, i)
        if "prime" in case and n <= 40000:
            r = gpu.proth_test(k, n)
            assert r["base"] == a
            assert r["verdict"] == (1 if case["prime"] else 0), (k, n)
            assert r["res64"] == int(case["res64_final"], 16), (k, n)
            assert r["iters_done"] == n - 1


def test_known_primes_full_test(gpu):
    for k, n in [(3, 20909), (3, 34350)]:
        r = gpu.proth_test(k, n)
        assert r["verdict"] == 1 and r["res64"] == 0 and r["base"] == 5
    # 5 divides 3*2^34351+1: the small-prime gate rejects it before any squaring
    r = gpu.proth_test(3, 34351)
    assert r["gated"] and r["verdict"] == 0 and r["iters_done"] == 0
    # composites that pass the gate reach the verdict through the squarings
    for k, n in [(1023, 32768), (3, 34353)]:
        assert not gpu.gate_rejects(k, n)
        prime, r64, a = G.proth_verdict(k, n)
        r = gpu.proth_test(k, n)
        assert not prime and r["verdict"] == 0 and not r["gated"]
        assert r["res64"] == r64 and r["base"] == a and r["iters_done"] == n - 1


def test_fft_square_karatsuba_shim_matches_oracle(gpu):
    from oracle import Oracle, int_to_limbs, limbs_to_int
    o = Oracle()
    k, n = 1023, 32768
    L = 4096  # reference layout that fits (16*L >= n + 2*bitlen(k))
    N = k * (1 << n) + 1
    oc = o.ctx(L)
    rnd = random.Random(3)
    with gpu.Context(k, n) as c:
        for _ in range(3):
            x = rnd.randrange(N)
            ref = oc.square_limbs(int_to_limbs(x, L // 4), L // 4, n, k)
            got, e = c.fft_square_karatsuba([int(v) for v in int_to_limbs(x, L // 4)], L // 4)
            assert np.array_equal(np.array(got, dtype=np.uint64), ref)
            assert limbs_to_int(ref) == x * x % N
            assert 0 < e < 0.25


def test_large_golden_c2_c3(gpu):
    data = golden_io.load("large_cases")
    if data is None:
        pytest.skip("large golden fixtures not generated")
    for case in data["cases"]:
        k, n, a = case["k"], case["n"], case["a"]
        N = k * (1 << n) + 1
        seed = G.powm(a, k, N)
        with gpu.Context(k, n) as c:
            c.set_residue(seed)
            done = 0
            for i_s, r64 in sorted(case["res64_after"].items(), key=lambda t: int(t[0])):
                i = int(i_s)
                e = c.square(i - done)
                done = i
                assert c.res64() == int(r64, 16), (k, n, i)
                assert e < 0.35


# Lengths above 2^18: 2^19 and 2^21 on the split pass, 2^22 on the unfused
# pass (column DFTs of 1024 points, two butterflies per thread).  At 2^19
# 66 squarings cross one graph-captured chunk; the longer lengths run a few
# (GMP, the arbiter, needs about a second per squaring at 60M bits).
@pytest.mark.parametrize("k,n,L,iters", [(3, 9000001, 1 << 19, 66), (5, 33000001, 1 << 21, 8),
                                         (3, 60000001, 1 << 22, 5)])
def test_square_long_lengths_vs_powm(gpu, k, n, L, iters):
    rnd = random.Random(L)
    N = k * (1 << n) + 1
    x = rnd.getrandbits(n) % N
    print(f"[parity] k={k} n={n} L={L} iters={iters}", flush=True)
    with gpu.Context(k, n, L) as c:
        assert c.fft_len == L
        c.set_residue(x)
        e = c.square(iters)
        print(f"[parity] squared, max round-off {e:.4f}; GMP check", flush=True)
        assert 0 < e < 0.3
        assert c.res64() == G.powm(x, 1 << iters, N) & M64


@pytest.mark.parametrize("k,n", [(433917, 2097152), (11887, 560001)])
def test_resident_chunks_and_their_event_times(gpu, monkeypatch, k, n):
    """Whole and partial chunks of the resident loop in one call (the chunks
    are pipelined: each one is enqueued before the one ahead of it is read
    back), against GMP; pl_loop_time counts one resident launch per chunk
    of at least two squarings, carrying all but each chunk's first and last
    squaring, and its event time is positive."""
    monkeypatch.setenv("PL_RESIDENT", "1")
    N = k * (1 << n) + 1
    x = random.Random(n).randrange(N)
    iters = 256 * 2 + 77
    print(f"[parity] resident chunks k={k} n={n} iters={iters}", flush=True)
    with gpu.Context(k, n) as c:
        assert c.info.passes == 0
        c.set_residue(x)
        c.loop_time(reset=True)
        e, tr = c.square(iters, trace=True)
        lt = c.loop_time(reset=True)
        assert c.get_residue() == G.powm(x, 1 << iters, N)
        assert 0 < e < 0.3 and len(tr) == iters and all(0 <= t < 0.3 for t in tr)
        assert lt["launches"] == 3 and lt["squarings"] == 255 * 2 + 76 and lt["ms"] > 0
        assert c.loop_time()["launches"] == 0
