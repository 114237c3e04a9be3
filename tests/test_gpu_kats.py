"""The reference's own known answers, run to their verdicts on the GPU path.

* 843, 853, 895 * 2^120002 + 1 (the reference transcript, Run test:41, :81,
  :88: PRIME, COMPOSITE, COMPOSITE), with the Res64 that GMP gives in
  tests/golden/kats.json;
* 3 * 2^303093 + 1, a known prime in the batched workload's range (GMP
  verified, Res64 0);
* 199971 * 2^5435752 + 1 (Run test:183: COMPOSITE), 5.4 million squarings at
  L = 2^19 with the Gerbicz check on.  GMP would need hours for it, so its
  Res64 is not pinned; the transcript pins the verdict, and the check pins
  the chain of squarings to itself."""
import pytest

import golden_io

pytestmark = pytest.mark.gpu


def test_transcript_kats_n120002(gpu):
    data = golden_io.load("kats")
    assert data is not None
    want_prime = {843: True, 853: False, 895: False}
    for c in data["cases"]:
        k, n = c["k"], c["n"]
        print(f"[kats] k={k} n={n}", flush=True)
        r = gpu.proth_test(k, n)
        assert c["prime"] == want_prime[k]
        assert r["base"] == c["a"]
        assert r["verdict"] == int(c["prime"]), (k, n)
        assert r["res64"] == int(c["res64_final"], 16), (k, n)
        assert r["iters_done"] == n - 1 and r["max_err"] < 0.3


def test_known_prime_3_2_303093(gpu):
    r = gpu.proth_test(3, 303093)
    assert r["verdict"] == 1 and r["res64"] == 0 and r["base"] == 5
    assert r["iters_done"] == 303092


@pytest.mark.timeout(900)
def test_transcript_composite_n5435752_with_gerbicz(gpu, capsys):
    k, n = 199971, 5435752

    def progress(done, total, err):
        with capsys.disabled():
            print(f"[kats] k={k} n={n}: {done}/{total} squarings, max err {err:.4f}", flush=True)

    r = gpu.proth_test(k, n, gerbicz_block=1000, gerbicz_blocks=64, report_step=10, progress=progress)
    assert r["fft_len"] == 1 << 19
    assert r["verdict"] == 0 and not r["gated"]
    assert r["iters_done"] == n - 1
    assert r["gerbicz_rollbacks"] == 0 and r["gerbicz_checks"] >= (n - 1) // 64000
    assert r["max_err"] < 0.35
    with capsys.disabled():
        print(f"[kats] {k}*2^{n}+1 COMPOSITE, Res64 {r['res64']:016X}, max err {r['max_err']:.4f}", flush=True)
