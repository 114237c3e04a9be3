"""Gerbicz check of pl_proth_test (gerbicz_block > 0): epochs verified on the
device through the running product of block-end residues.  A clean run must
verify every epoch (which also pins the device product d <- d*u: a wrong
product fails the check); an error planted in the residue from the host
must be caught and recomputed; verdicts and Res64 match GMP throughout."""
import pytest

import gmp_ctypes as G


@pytest.mark.gpu
@pytest.mark.parametrize("k,n,B,M", [(3, 20909, 500, 4), (843, 120002, 1000, 16), (1023, 32768, 256, 64)])
def test_clean_run_verifies_every_epoch(gpu, k, n, B, M):
    r = gpu.proth_test(k, n, gerbicz_block=B, gerbicz_blocks=M)
    prime, r64, a = G.proth_verdict(k, n)
    assert r["verdict"] == int(prime) and r["res64"] == r64
    assert r["iters_done"] == n - 1
    assert r["gerbicz_rollbacks"] == 0
    assert r["gerbicz_checks"] >= (n - 1) // (B * M)


@pytest.mark.gpu
def test_planted_error_is_caught_and_recomputed(gpu):
    k, n = 3, 20909
    r = gpu.proth_test(k, n, gerbicz_block=500, gerbicz_blocks=4, inject_error_at=5000)
    assert r["gerbicz_rollbacks"] == 1
    assert r["verdict"] == 1 and r["res64"] == 0
    assert r["iters_done"] == n - 1


@pytest.mark.gpu
def test_planted_error_without_the_check_goes_unnoticed(gpu):
    """The same planted error with the check off: nothing catches it, and the
    prime is reported composite -- what the check protects against."""
    r = gpu.proth_test(3, 20909, gerbicz_block=0, inject_error_at=5000)
    assert r["verdict"] == 0


@pytest.mark.gpu
def test_stop_and_resume_at_a_verified_epoch(gpu, tmp_path):
    k, n = 843, 120002
    ck = str(tmp_path / "g.ckpt")
    r1 = gpu.proth_test(k, n, checkpoint_path=ck, stop_after=7000, gerbicz_block=500, gerbicz_blocks=4)
    a = G.reference_base(k, n)
    assert r1["verdict"] == -1 and r1["iters_done"] == 7000
    assert r1["res64"] == G.residue_after(k, n, a, 7000) & 0xFFFFFFFFFFFFFFFF
    r2 = gpu.proth_test(k, n, checkpoint_path=ck, gerbicz_block=500, gerbicz_blocks=4)
    prime, r64, _ = G.proth_verdict(k, n)
    assert r2["resumed_at"] == 7000 and r2["iters_done"] == n - 1 - 7000
    assert r2["verdict"] == int(prime) and r2["res64"] == r64
