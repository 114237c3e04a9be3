"""The launch-plan autotuner on the device: every plan the table can name
(shape R x C, row and column radix 8 or 16, carry runs of 8 or 16 digits,
one-launch or split outer pass) squares exactly -- each against the rule
plan's result from the same start, and the rule plan against GMP -- and a
tuning run records an exact choice that later contexts and pool groups
take."""
import random

import pytest

import golden_io
import gmp_ctypes as G

pytestmark = pytest.mark.gpu

SQUARINGS = 260   # one whole graph-captured chunk and a partial one


def variants(logL):
    shapes = [lr for lr in range(6, 10) if 5 <= logL - lr <= 12]
    for lr in shapes:
        for rr in (8, 16):
            for cr in (8, 16):
                for run, one in ((8, True), (8, False), (16, False)):
                    yield 1 << lr, rr, cr, run, one


@pytest.fixture(autouse=True)
def empty_table(gpu):
    gpu.tune_clear()
    yield
    gpu.tune_clear()


@pytest.mark.parametrize("k,n", [(65533, 1 << 19), (433917, 1 << 21), (11887, 1 << 24)])
def test_every_pinned_plan_squares_exactly(gpu, k, n):
    N = k * (1 << n) + 1
    x = random.Random(n ^ k).randrange(2, N - 1)
    with gpu.Context(k, n) as c:
        L = c.fft_len
        rule = (c.info.rows, c.info.row_radix, c.info.col_radix, c.info.run, c.info.passes)
        c.set_residue(x)
        c.square(SQUARINGS)
        want = c.get_residue()
    assert rule[1:4] == (8, 8, 8) and rule[4] in (0, 2)
    if n <= 1 << 21:
        assert want == G.powm(x, 1 << SQUARINGS, N)
    logL = L.bit_length() - 1
    seen = 0
    for rows, rr, cr, run, one in variants(logL):
        gpu.tune_pin(k, n, rows, rr, cr, run=run, one_launch=one)
        with gpu.Context(k, n) as c:
            i = c.info
            case = (k, n, rows, rr, cr, run, one)
            assert (i.rows, i.cols, i.row_radix, i.col_radix, i.run) == (rows, L // rows, rr, cr, run), case
            assert i.passes in (((0, 2) if rr == cr == 8 else (2,)) if one else (3,)), case
            c.set_residue(x)
            e = c.square(SQUARINGS)
            assert c.get_residue() == want, case
            assert e < 0.3, case
        seen += 1
    assert seen >= 24


def test_autotune_plan_records_an_exact_fastest_choice(gpu):
    k, n = 433917, 1 << 21
    r = gpu.autotune_plan(k, n)
    trials, best = r["trials"], r["best"]
    assert len(trials) >= 6
    assert all(t["exact"] for t in trials), trials
    assert sum(t["chosen"] for t in trials) == 1
    assert best["exact"] and best["max_err"] <= 0.3
    assert best["us_per_squaring"] <= trials[0]["us_per_squaring"]
    assert best["us_per_squaring"] == min(t["us_per_squaring"] for t in trials if t["max_err"] <= 0.3)
    with gpu.Context(k, n) as c:
        i = c.info
        assert (i.rows, i.row_radix, i.col_radix, i.run) == (best["rows"], best["row_radix"], best["col_radix"],
                                                             best["run"])
        assert i.passes in ((0, 2) if best["one_launch"] else (3,))
        N = k * (1 << n) + 1
        x = random.Random(3).randrange(2, N - 1)
        c.set_residue(x)
        c.square(SQUARINGS)
        assert c.get_residue() == G.powm(x, 1 << SQUARINGS, N)


def test_autotuned_fft_len_also_tunes_the_plan(gpu, tmp_path):
    k, n = 1023, 1 << 17
    r = gpu.autotune_fft_len(n, k)
    assert r["fft_len"] in (4096, 8192, 16384)
    p = tmp_path / "plans.txt"
    gpu.tune_save(p)
    lines = p.read_text().strip().splitlines()
    assert len(lines) == 2 and lines[1].split()[1] == str(r["fft_len"].bit_length() - 1)
    # the saved table reloads into an empty one and a context takes it
    gpu.tune_clear()
    gpu.tune_load(p)
    with gpu.Context(k, n, fft_len=r["fft_len"]) as c:
        logR = int(lines[1].split(":")[1].split()[0])
        assert c.info.rows == 1 << logR


def test_pool_autotune_then_batch_matches_fixture(gpu):
    data = golden_io.load("batch_cases")
    if data is None:
        pytest.skip("batch fixtures not generated")
    cases = [c for c in data["cases"] if not c["gated"]][:24]
    with gpu.Pool(k_max=1 << 16, max_slots_per_length=32) as pool:
        tuned = set()
        for c in cases:
            L = gpu.plan_fft_len(c["n"], c["k"])
            if L in tuned:
                continue
            r = pool.autotune(c["k"], c["n"])
            assert all(t["exact"] for t in r["trials"]), r["trials"]
            assert all(t["one_launch"] == 0 for t in r["trials"])
            tuned.add(L)
        for i, c in enumerate(cases):
            pool.submit(c["k"], c["n"], tag=i, stop_after=200)
        res = pool.drain(steps=64)
    assert len(res) == len(cases)
    for r in res:
        c = cases[r["tag"]]
        assert r["res64"] == int(c["res64_after"]["200"], 16), (c["k"], c["n"])
