"""Continuous batching (pl_pool_*): every result of a test run in a batch of
uneven-length candidates must equal GMP's, and equal the same candidate run
alone; requests arriving while others run must join the batch."""
import random

import pytest

import golden_io
import gmp_ctypes as G

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def gmp_res64(k, n, a, s):
    return G.residue_after(k, n, a, s) & M64


def test_pool_full_tests_mixed_lengths(gpu):
    cands = [(3, 20909), (3, 34350), (3, 34351), (1023, 32768), (13, 12000), (4095, 9001), (65535, 60001),
             (1048573, 70001), (12345, 17000)]
    with gpu.Pool(k_max=1 << 21) as pool:
        for i, (k, n) in enumerate(cands):
            pool.submit(k, n, tag=i)
        res = pool.drain(steps=256)
    assert len(res) == len(cands)
    for r in res:
        k, n = cands[r["tag"]]
        if r["gated"]:
            assert gpu.gate_rejects(k, n)
            continue
        prime, r64, a = G.proth_verdict(k, n)
        assert r["base"] == a
        assert r["verdict"] == (1 if prime else 0), (k, n)
        assert r["res64"] == r64, (k, n)
        assert r["max_err"] < 0.3


def test_pool_stop_after_matches_gmp_and_c4_fixture(gpu):
    data = golden_io.load("batch_cases")
    if data is None:
        pytest.skip("batch fixtures not generated")
    cases = [c for c in data["cases"] if not c["gated"]]
    with gpu.Pool(k_max=1 << 16) as pool:
        for i, c in enumerate(cases):
            pool.submit(c["k"], c["n"], tag=i, stop_after=200)
        res = pool.drain(steps=64)
    assert len(res) == len(cases)
    for r in res:
        c = cases[r["tag"]]
        assert r["verdict"] == -1
        assert r["base"] == c["a"]
        assert r["res64"] == int(c["res64_after"]["200"], 16), (c["k"], c["n"])


def test_pool_late_arrivals_and_solo_equivalence(gpu):
    rnd = random.Random(9)
    cands = []
    while len(cands) < 12:
        k = 2 * rnd.randrange(1, 1 << 15) + 1
        n = rnd.choice([9000, 15000, 33000, 70000, 140000, 270000])
        if not gpu.gate_rejects(k, n):
            cands.append((k, n))
    S = 300
    with gpu.Pool(k_max=1 << 16, max_slots_per_length=4) as pool:
        for i, (k, n) in enumerate(cands[:6]):
            pool.submit(k, n, tag=i, stop_after=S)
        res = pool.run(steps=50)                      # first batch is running
        for i, (k, n) in enumerate(cands[6:], start=6):
            pool.submit(k, n, tag=i, stop_after=S)   # arrives mid-run, queues for slots
        res += pool.drain(steps=50)
    assert sorted(r["tag"] for r in res) == list(range(len(cands)))
    mixed = {r["tag"]: r["res64"] for r in res}
    for i, (k, n) in enumerate(cands):
        a = G.reference_base(k, n)
        assert mixed[i] == gmp_res64(k, n, a, S), (k, n)
    # the same candidate alone gives the same residue
    k, n = cands[3]
    with gpu.Pool(k_max=1 << 16, max_slots_per_length=1) as pool:
        pool.submit(k, n, tag=0, stop_after=S)
        alone = pool.drain(steps=64)
    assert alone[0]["res64"] == mixed[3]


def test_proth_test_seed_on_device(gpu):
    # seeds with many set bits in k exercise the multiplier schedule
    for k, n in [(0xFFFF, 40000), (1048575 ^ 2, 30001), (999999, 25000), (3, 20909)]:
        if gpu.gate_rejects(k, n):
            continue
        r = gpu.proth_test(k, n, stop_after=150)
        a = G.reference_base(k, n)
        assert r["res64"] == gmp_res64(k, n, a, 150), (k, n)


@pytest.mark.parametrize("k_max", [1 << 11, 1 << 16])
def test_pool_launch_span_follows_occupancy(gpu, k_max):
    # A step chunk launches only the slots up to the highest occupied one,
    # rounded up to a power of two (at least 8).  Results must not depend on
    # the span a step ran with while it grows past 8 and 16 and holes open
    # below it.  n = 200000 gives L = 2^14; k_max = 2^11 keeps the group's
    # carry runs at 8 digits, 2^16 makes them 16 (both on the split outer
    # pass, which the pool uses).
    rnd = random.Random(11)
    n = 200000

    def cand():
        while True:
            k = 2 * rnd.randrange(1, k_max // 2) + 1
            if not gpu.gate_rejects(k, n):
                return k

    first = [(cand(), 60) for _ in range(5)]      # retire early, leaving holes at the bottom
    second = [(cand(), 400) for _ in range(20)]   # grows the span past 16
    third = [(cand(), 200) for _ in range(6)]     # refills the holes
    allc = first + second + third
    print(f"[pool span] k_max={k_max} n={n} tests={len(allc)}", flush=True)
    with gpu.Pool(k_max=k_max) as pool:
        for i, (k, s) in enumerate(first):
            pool.submit(k, n, tag=i, stop_after=s)
        res = pool.run(steps=20)
        for i, (k, s) in enumerate(second, start=len(first)):
            pool.submit(k, n, tag=i, stop_after=s)
        res += pool.run(steps=64)
        res += pool.run(steps=64)
        for i, (k, s) in enumerate(third, start=len(first) + len(second)):
            pool.submit(k, n, tag=i, stop_after=s)
        res += pool.drain(steps=64)
    assert sorted(r["tag"] for r in res) == list(range(len(allc)))
    for r in res:
        k, s = allc[r["tag"]]
        assert r["fft_len"] == 1 << 14
        a = G.reference_base(k, n)
        assert r["res64"] == gmp_res64(k, n, a, s), (k, n, s)
        assert r["max_err"] < 0.3
