"""CPU-side checks of the launch-plan table: pl_tune_pin validates every
field against the plans the kernels have, and the table survives a save /
clear / load round trip; a malformed or foreign file is refused."""
import pytest

from prothlogic_amd import core

C2 = (433917, 1 << 21)     # L = 2^17, runs of 8 digits
C4 = (65533, 1 << 19)      # L = 2^15


@pytest.fixture(autouse=True)
def empty_table():
    core.tune_clear()
    yield
    core.tune_clear()


def test_pin_accepts_every_kernel_plan():
    k, n = C2
    for rows in (64, 128, 256, 512):
        for rr in (8, 16):
            for cr in (8, 16):
                core.tune_pin(k, n, rows, rr, cr, run=8, one_launch=True)
                core.tune_pin(k, n, rows, rr, cr, run=8, one_launch=False)
                core.tune_pin(k, n, rows, rr, cr, run=16, one_launch=False)


@pytest.mark.parametrize("kw,msg", [
    (dict(rows=96), "power of two"),
    (dict(rows=1 << 17), "power of two"),
    (dict(rows=32), "no kernels"),          # R = 32 has no split outer pass
    (dict(rows=1024), "no kernels"),        # nor R = 1024
    (dict(rows=256, row_radix=4), "radices"),
    (dict(rows=256, col_radix=32), "radices"),
    (dict(rows=256, run=4), "run"),
    (dict(rows=256, run=16, one_launch=True), "one-launch"),
])
def test_pin_rejects_plans_the_kernels_lack(kw, msg):
    k, n = C2
    args = dict(rows=256, row_radix=8, col_radix=8, run=8, one_launch=True)
    args.update(kw)
    with pytest.raises(core.PlError) as e:
        core.tune_pin(k, n, **args)
    assert e.value.code == core.PL_ERR_ARG and msg in str(e.value)


def test_pin_rejects_runs_the_carries_need_more_than():
    # k near 2^32 at this length needs runs of 16 digits: 8 is refused
    k, n = 4294967291, 500000
    with pytest.raises(core.PlError):
        core.tune_pin(k, n, 256, run=8)
    core.tune_pin(k, n, 256, run=16, one_launch=False)


def test_save_clear_load_round_trip(tmp_path):
    k, n = C2
    core.tune_pin(k, n, 128, 16, 8, run=16, one_launch=False)
    core.tune_pin(C4[0], C4[1], 64, 8, 16, run=8, one_launch=True)
    a = tmp_path / "plans.txt"
    core.tune_save(a)
    text = a.read_text()
    assert text.startswith("# prothlogic tuned plans v1\n")
    assert len(text.strip().splitlines()) == 3
    core.tune_clear()
    b = tmp_path / "empty.txt"
    core.tune_save(b)
    assert len(b.read_text().strip().splitlines()) == 1
    core.tune_load(a)
    c = tmp_path / "again.txt"
    core.tune_save(c)
    assert c.read_text() == text


def test_load_refuses_bad_files(tmp_path):
    missing = tmp_path / "none.txt"
    with pytest.raises(core.PlError) as e:
        core.tune_load(missing)
    assert e.value.code == core.PL_ERR_IO
    foreign = tmp_path / "foreign.txt"
    foreign.write_text("hello\n0 17 8 0 0 : 8 8 8 8 1\n")
    with pytest.raises(core.PlError):
        core.tune_load(foreign)
    bad = tmp_path / "bad.txt"
    for line in ("0 17 8 0 0 : 8 8", "0 17 8 0 0 : 3 8 8 8 1", "0 17 8 0 0 : 8 4 8 8 1",
                 "0 17 8 0 0 : 8 8 8 4 1", "0 30 8 0 0 : 8 8 8 8 1"):
        bad.write_text("# prothlogic tuned plans v1\n" + line + "\n")
        with pytest.raises(core.PlError):
            core.tune_load(bad)
    # nothing from a refused file reaches the table
    out = tmp_path / "out.txt"
    core.tune_save(out)
    assert len(out.read_text().strip().splitlines()) == 1
