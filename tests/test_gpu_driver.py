"""pl_proth_test's call-site behaviour (progress at report_step percent, the
final residue handed back as limbs) and the measured autotuner
(pl_autotune_fft_len), through the C ABI."""
import pytest

import gmp_ctypes as G


@pytest.mark.gpu
def test_progress_reports_follow_the_percentage_rule(gpu):
    k, n = 3, 20909
    calls = []
    r = gpu.proth_test(k, n, report_step=25, progress=lambda d, t, e: calls.append((d, t, e)))
    total = n - 1
    # P/proth_logic.cpp:541-548: report when floor((i+1)*100/(n-1)) reaches a
    # new multiple of report_step; the first such squaring count for p% is
    # ceil(p*(n-1)/100)
    assert [c[0] for c in calls] == [-(-p * total // 100) for p in (25, 50, 75, 100)]
    assert all(c[1] == total for c in calls)
    assert all(0 <= c[2] < 0.4 for c in calls)
    assert r["verdict"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("k,n", [(3, 20909), (1023, 32768), (843, 120002)])
def test_final_residue_limbs(gpu, k, n):
    r = gpu.proth_test(k, n, final_n_limbs=n // 64 + 4)
    N = G.proth_N(k, n)
    a = G.reference_base(k, n)
    want = G.powm(a, (N - 1) // 2, N)
    assert r["residue"] == want
    assert r["res64"] == want & 0xFFFFFFFFFFFFFFFF
    assert r["verdict"] == int(want == N - 1)


@pytest.mark.gpu
def test_final_residue_too_small_buffer_is_an_error(gpu):
    with pytest.raises(gpu.PlError):
        gpu.proth_test(1023, 32768, final_n_limbs=4)


@pytest.mark.gpu
@pytest.mark.parametrize("k,n", [(843, 120002), (433917, 1 << 21)])
def test_autotune_picks_a_passing_length(gpu, k, n):
    t = gpu.autotune_fft_len(n, k)
    L0 = gpu.plan_fft_len(n, k)
    assert t["fft_len"] in (L0 // 2, L0, 2 * L0)
    assert 0 <= t["max_err"] <= 0.30
    assert t["us_per_squaring"] > 0
    # the chosen length computes exact squarings
    c = gpu.Context(k, n, fft_len=t["fft_len"])
    N = G.proth_N(k, n)
    x = pow(5, 12345, N)
    c.set_residue(x)
    c.square(3)
    assert c.get_residue() == pow(x, 8, N)
    c.close()
