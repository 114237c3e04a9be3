"""pl_fft_square_karatsuba_noalloc: the per-call squaring with every device
buffer in the library's static scratch array.  Its output must equal that
of pl_fft_square_karatsuba (the context-based shim) and exact x^2 mod N, on
calls in a row whose second input is smaller (a shorter transform laid out
over what the first call left in the array), and it must refuse lengths
above the array's limit."""
import random

import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu


def limbs_of(x, nl):
    return [(x >> (64 * i)) & ((1 << 64) - 1) for i in range(nl)]


def int_of(limbs):
    return sum(int(v) << (64 * i) for i, v in enumerate(limbs))


def both(gpu, k, n, x):
    nl = (n + 63) // 64 + 2
    got, e = gpu.fft_square_karatsuba_noalloc(limbs_of(x, nl), nl, n, k)
    with gpu.Context(k, n) as c:
        want, e_ctx = c.fft_square_karatsuba(limbs_of(x, nl), nl)
    return got, want, e, e_ctx


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_matches_the_context_shim_on_successive_calls(gpu, order):
    rnd = random.Random(77 + order)
    cases = [(433917, 1 << 21), (1023, 32768), (1023, 32768), (3, 303093)]
    if order:
        cases = cases[1:] + cases[:1]     # a short layout first, then the longest
    for k, n in cases:
        N = G.proth_N(k, n)
        x = rnd.randrange(N)
        print(f"[noalloc] k={k} n={n}", flush=True)
        got, want, e, e_ctx = both(gpu, k, n, x)
        assert got == want, (k, n)
        assert int_of(got) == x * x % N, (k, n)
        assert 0 <= e < 0.3 and 0 <= e_ctx < 0.3
    # small and edge values
    k, n = 3, 303093
    N = G.proth_N(k, n)
    for x in (0, 1, 125, N - 1, N - 2):
        got, want, _, _ = both(gpu, k, n, x)
        assert got == want and int_of(got) == x * x % N, x


def test_noalloc_refuses_lengths_beyond_its_scratch(gpu):
    n = 1 << 23        # needs L = 2^19
    nl = n // 64 + 2
    with pytest.raises(gpu.PlError) as ei:
        gpu.fft_square_karatsuba_noalloc([5] + [0] * (nl - 1), nl, n, 3)
    assert ei.value.code == gpu.PL_ERR_ARG and "2^18" in str(ei.value)
    # and the next call in range still works on the same array
    k, n = 65535, 300007
    N = G.proth_N(k, n)
    x = random.Random(5).randrange(N)
    nl = n // 64 + 2
    got, _ = gpu.fft_square_karatsuba_noalloc(limbs_of(x, nl), nl, n, k)
    assert int_of(got) == x * x % N
