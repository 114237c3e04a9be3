"""The first squarings of a test, while its exact value is small, run in
normal form (the slot holds x, not x*2^(2n)): for k whose powers of two
repeat soon modulo k^2 the Montgomery form of a small x has periodic,
coherent digits, and squaring those at the top of a length's digit size
gave round-off up to 0.44 (3*2^303093+1 at L = 2^14).  With the normal form
the round-off of these tests stays at the random-digit level, and every
Res64 equals GMP's -- through pl_proth_test, the pool, and contexts whose
residue is set small by the caller."""
import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
# n at the top of each length's digit size (18.5 bits at 2^14, 18 at 2^16)
SMALL_PERIOD = [(3, 303093), (5, 303091), (15, 303095), (17, 1179641), (63, 1179647), (255, 303089),
                (1023, 1179643), (65535, 303097), (65537, 1179631)]


@pytest.fixture(scope="module")
def survivors():
    from prothlogic_amd import workloads
    out = []
    for k, n in SMALL_PERIOD:
        while workloads.gate_rejects(k, n):
            n -= 2
        out.append((k, n))
    return out


def test_proth_test_small_period_k(gpu, survivors):
    for k, n in survivors:
        a = G.reference_base(k, n)
        r = gpu.proth_test(k, n, stop_after=400)
        print(f"[prewrap] k={k} n={n} L={r['fft_len']} max_err={r['max_err']:.4f}", flush=True)
        assert r["res64"] == G.residue_after(k, n, a, 400) & M64, (k, n)
        assert r["max_err"] < 0.12, (k, n, r["max_err"])


def test_pool_small_period_k(gpu, survivors):
    with gpu.Pool(k_max=1 << 17) as pool:
        for i, (k, n) in enumerate(survivors):
            pool.submit(k, n, tag=i, stop_after=400)
        res = pool.drain(steps=64)
    assert sorted(r["tag"] for r in res) == list(range(len(survivors)))
    for r in res:
        k, n = survivors[r["tag"]]
        a = G.reference_base(k, n)
        assert r["error"] == 0
        assert r["res64"] == G.residue_after(k, n, a, 400) & M64, (k, n)
        assert r["max_err"] < 0.12, (k, n, r["max_err"])


def test_pool_test_shorter_than_its_prewrap(gpu):
    # 3 squarings past the seed: the value never wraps before the test ends
    with gpu.Pool(k_max=1 << 16) as pool:
        pool.submit(3, 303093, tag=0, stop_after=3)
        pool.submit(65535, 300007, tag=1, stop_after=1)
        res = pool.drain(steps=64)
    for r in res:
        k, n, s = [(3, 303093, 3), (65535, 300007, 1)][r["tag"]]
        a = G.reference_base(k, n)
        assert r["error"] == 0 and r["res64"] == G.residue_after(k, n, a, s) & M64


def test_context_small_residues_square_exactly(gpu):
    k, n = 3, 303093
    N = G.proth_N(k, n)
    with gpu.Context(k, n) as c:
        assert c.fft_len == 1 << 14
        # 2^E stays structured after it wraps (2^n == -1/k), so x = 2 keeps
        # a higher round-off; no Proth base is a power of two
        for x, it, emax in ((125, 40, 0.12), (2, 30, 0.3), (1, 100, 0.12), (N - 1, 5, 0.12), (N - 7, 30, 0.12),
                            (0, 3, 0.12)):
            c.set_residue(x)
            e = c.square(it)
            assert c.get_residue() == G.powm(x, 1 << it, N), (x, it)
            assert e < emax, (x, e)
        # the seed's multiply in normal form
        c.set_residue(5)
        c.mul_small(5)
        c.square(20)
        assert c.get_residue() == G.powm(25, 1 << 20, N)
