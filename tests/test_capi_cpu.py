"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and its host-only functions (gate, base rule, FFT-length
rule, checkpoints) agree with closed forms, GMP and the oracle."""
import re
from pathlib import Path

import pytest

import gmp_ctypes as G
from prothlogic_amd import core, workloads

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    txt = (ROOT / "include" / "prothcore.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = core.lib()
    names = header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    declared = int(re.search(r"#define PL_ABI_VERSION (\d+)", (ROOT / "include" / "prothcore.h").read_text()).group(1))
    assert lib.pl_abi_version() == declared


def test_gate_matches_closed_form():
    for k in range(3, 4001, 2):
        for n in (1000, 32768, 120002, 2097152):
            want = any((k * pow(2, n, p) + 1) % p == 0 for p in workloads.GATE_PRIMES)
            assert core.gate_rejects(k, n) == want, (k, n)
            assert workloads.gate_rejects(k, n) == want


def test_base_rule_matches_gmp_jacobi():
    import random
    rnd = random.Random(11)
    for _ in range(200):
        k = 2 * rnd.randrange(1, 1 << 19) + 1
        n = rnd.randrange(100, 5000)
        assert core.choose_base(k, n) == G.reference_base(k, n), (k, n)
    for k, n, a in [(843, 120002, 5), (853, 120002, 3), (895, 120002, 3), (3, 34351, 11)]:
        assert core.choose_base(k, n) == a


def test_plan_fft_len_fits_and_limits():
    for n in (9000, 32768, 120002, 1 << 18, 1 << 20, 1 << 21, 5435752, 1 << 24):
        L = core.plan_fft_len(n, 65535)
        assert L & (L - 1) == 0 and 1024 <= L <= 1 << 22
        assert 8 <= n / L <= 18.5
    assert core.plan_fft_len(1 << 21, 433917) == 131072
    assert core.plan_fft_len(1 << 24, 11887) == 1 << 20
    # k near 2^32: the top scalar's share of the round-off caps the digit size
    assert core.plan_fft_len(500000, 4294967291) == 65536
    assert core.plan_fft_len(500000, 268435399) == 32768
    with pytest.raises(core.PlError):
        core.plan_fft_len(4000, 3)       # below 7 bits per digit at the shortest length
    with pytest.raises(core.PlError):
        core.plan_fft_len(10000, 4)      # even k


def test_checkpoint_v2_roundtrip_and_corruption(tmp_path):
    p = str(tmp_path / "ck.bin")
    k, n = 843, 120002
    x = (1 << 100000) + 12345
    core.save_checkpoint(p, k, n, 5, 20001, x, 1900)
    b, it, y = core.load_checkpoint(p, k, n, 1900)
    assert (b, it, y) == (5, 20001, x)
    with pytest.raises(core.PlError):
        core.load_checkpoint(p, k + 2, n, 1900)          # foreign candidate
    raw = bytearray(Path(p).read_bytes())
    raw[100] ^= 1
    Path(p).write_bytes(bytes(raw))
    with pytest.raises(core.PlError):
        core.load_checkpoint(p, k, n, 1900)              # checksum
    Path(p).write_bytes(bytes(raw[:-8]))
    with pytest.raises(core.PlError):
        core.load_checkpoint(p, k, n, 1900)              # truncated


def test_checkpoint_header_is_covered_by_the_checksum(tmp_path):
    """A flipped next_iter, base or n_limbs field in the header must be caught
    like a flipped limb (the checksum covers header and limbs)."""
    import struct
    p = str(tmp_path / "ck.bin")
    k, n = 843, 120002
    x = (1 << 90000) + 999
    core.save_checkpoint(p, k, n, 5, 30001, x, 1900)
    good = Path(p).read_bytes()
    assert good[:8] == b"PLCKPT02" and len(good) == 64 + 8 * 1900
    # header fields: magic[8] k@8 n@16 base@20 next_iter@24 n_limbs@32 checksum@40
    for off, fmt in ((24, "<Q"), (20, "<I")):
        raw = bytearray(good)
        (v,) = struct.unpack_from(fmt, raw, off)
        struct.pack_into(fmt, raw, off, v ^ 1)
        Path(p).write_bytes(bytes(raw))
        with pytest.raises(core.PlError) as ei:
            core.load_checkpoint(p, k, n, 1900)
        assert ei.value.code == core.PL_ERR_IO
    Path(p).write_bytes(good)
    assert core.load_checkpoint(p, k, n, 1900) == (5, 30001, x)
    assert not Path(p + ".tmp").exists()


def test_checkpoint_write_failure_is_reported(tmp_path):
    with pytest.raises(core.PlError) as ei:
        core.save_checkpoint(str(tmp_path / "missing_dir" / "ck.bin"), 843, 120002, 5, 1, 7, 4)
    assert ei.value.code == core.PL_ERR_IO


def test_reads_reference_checkpoint_layout(tmp_path):
    from oracle import Oracle, int_to_limbs
    o = Oracle()
    p = str(tmp_path / "check_k843_n120002.bin")
    x = (1 << 64000) + 77
    o.save_checkpoint(10000, int_to_limbs(x, 2048), p)   # reference: index of the last applied squaring
    b, it, y = core.load_checkpoint(p, 843, 120002, 2048)
    assert it == 10001 and y == x


def test_limb_conversion_is_exact_and_linear():
    """Residues cross the ABI as little-endian 64-bit limbs; the conversion
    must truncate to nl limbs and stay linear in the size (60M-bit residues
    are converted at the longest transform lengths)."""
    import random
    import time
    assert list(core._limbs(5, 2)) == [5, 0]
    assert list(core._limbs((1 << 70) + 7, 1)) == [7]
    assert list(core._limbs((1 << 70) + 7, 2)) == [7, 64]
    with pytest.raises(ValueError):
        core._limbs(-1, 1)
    x = random.Random(3).getrandbits(60_000_001)
    t = time.perf_counter()
    a = core._limbs(x, 60_000_001 // 64 + 2)
    assert time.perf_counter() - t < 2.0
    assert core._int(a) == x
