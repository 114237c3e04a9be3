"""The serving workload (config 5: k odd in [3, 2^20), n log-uniform in
[5e4, 4e6], PCG64 seed 5) through the pool with continuous batching: its
first 32 candidates arrive in three staggered waves while earlier ones are
on the device, beside the known prime 843*2^120002+1 run to its verdict in
the same mix.  Every Res64 after 200 squarings must equal GMP's
(tests/golden/c5_cases.json), the prime must come out PRIME with Res64 0,
and a candidate run alone must give the Res64 it gave in the mix."""
import pytest

import golden_io

pytestmark = pytest.mark.gpu

S = 200


def test_c5_waves_match_gmp_and_prime_in_the_mix(gpu):
    data = golden_io.load("c5_cases")
    assert data is not None, "tests/golden/c5_cases.json missing (make_golden.py c5)"
    cases = data["cases"]
    assert len(cases) == 32 and not any(c["gated"] for c in cases)
    waves = [cases[:12], cases[12:22], cases[22:]]
    prime_tag = 1000
    got = {}
    with gpu.Pool(k_max=1 << 20) as pool:
        tag = 0
        for w, wave in enumerate(waves):
            for c in wave:
                pool.submit(c["k"], c["n"], tag=tag, stop_after=S)
                tag += 1
            if w == 1:
                pool.submit(843, 120002, tag=prime_tag)      # full test, in the mix
            print(f"[c5] wave {w}: {len(wave)} arrivals, status {pool.status()}", flush=True)
            for r in pool.run(steps=64):
                got[r["tag"]] = r
        for r in pool.drain(steps=64):
            got[r["tag"]] = r
    assert sorted(got) == list(range(32)) + [prime_tag]
    lengths = set()
    for i, c in enumerate(cases):
        r = got[i]
        assert r["error"] == 0 and r["verdict"] == -1, (c["k"], c["n"])
        assert r["base"] == c["a"]
        assert r["res64"] == int(c["res64_after"][str(S)], 16), (c["k"], c["n"])
        assert r["max_err"] < 0.35
        lengths.add(r["fft_len"])
    assert len(lengths) >= 5           # uneven n: several transform lengths at once
    p = got[prime_tag]
    assert p["verdict"] == 1 and p["res64"] == 0 and p["base"] == 5 and p["error"] == 0
    # the largest candidate alone gives the Res64 it gave in the mix
    big = max(range(32), key=lambda i: cases[i]["n"])
    with gpu.Pool(k_max=1 << 20, max_slots_per_length=1) as pool:
        pool.submit(cases[big]["k"], cases[big]["n"], tag=0, stop_after=S)
        alone = pool.drain(steps=64)
    assert alone[0]["res64"] == got[big]["res64"]
