"""The C++ host mirror (prothlogic_amd/host/, libprothhost.so): the
reference's own entry points -- fft_init / fft_destroy / fft_square_karatsuba
(P/fft_engine.h:10-22), execute_proth_test_fft (P/proth_logic.h:12),
save_checkpoint / load_checkpoint (P/checkpoint_manager.h:15-16),
auto_tune_fft_len (P/dcore.cpp:20), run_task_with_ctx (P/worker.h) -- with
their C++ signatures, over the C ABI.

CPU: the library exports those symbols with the reference's exact mangled
names (checked against the reference build when it is present).
GPU: the self-test program drives the mirror as the reference's callers do;
every verdict, Res64 and residue is checked against GMP.
"""
import json
import os
import re
import subprocess
from pathlib import Path

import pytest

import gmp_ctypes as G

ROOT = Path(__file__).resolve().parent.parent
HOST_LIB = ROOT / "prothlogic_amd" / "libprothhost.so"
SELFTEST = ROOT / "prothlogic_amd" / "prothhost_selftest"
REF_LIB = ROOT / "oracle" / "_ref" / "libprothref.so"

# demangled signatures of the reference's declarations
SIGNATURES = [
    "fft_init(FFTContext&, FFTBackend, unsigned long)",
    "fft_destroy(FFTContext&)",
    "fft_square_karatsuba(FFTContext&, unsigned long*, unsigned long, unsigned int, unsigned long)",
    "execute_proth_test_fft(unsigned long, unsigned int, unsigned long*, FFTContext&, int)",
    "save_checkpoint(unsigned long, unsigned long, unsigned long*, std::__cxx11::basic_string<char, "
    "std::char_traits<char>, std::allocator<char> > const&)",
    "load_checkpoint(unsigned long&, unsigned long, unsigned long*, std::__cxx11::basic_string<char, "
    "std::char_traits<char>, std::allocator<char> > const&)",
    "auto_tune_fft_len(unsigned int, unsigned long)",
    "run_task_with_ctx[abi:cxx11](Task const&, FFTContext&, unsigned long*)",
    "execute_full_streaming_sieve(unsigned long, unsigned long, unsigned int, unsigned long)",
    "execute_sparse_streaming_sieve(std::vector<unsigned long, std::allocator<unsigned long> > const&, "
    "unsigned int, unsigned long, unsigned long)",
]
REF_TREE = Path("/root/reference/prothlogic4.2")


def _defined(lib: Path, demangle: bool):
    out = subprocess.run(["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [str(lib)],
                         capture_output=True, text=True, check=True).stdout
    return {line.split(" ", 2)[2] for line in out.splitlines() if line.count(" ") >= 2}


def test_mirror_exports_reference_signatures():
    assert HOST_LIB.exists(), "libprothhost.so is not built"
    names = _defined(HOST_LIB, demangle=True)
    for sig in SIGNATURES:
        assert sig in names, sig


@pytest.mark.skipif(not REF_LIB.exists(), reason="reference build (oracle/_ref) not present on this host")
def test_mirror_mangled_names_equal_reference_build():
    """Every mirror entry point is the very symbol the reference's own
    objects define (same name, same parameter types), so the reference's
    callers link against the mirror unchanged."""
    ours = {m for m in _defined(HOST_LIB, demangle=False) if m.startswith("_Z")}
    ref = {m for m in _defined(REF_LIB, demangle=False) if m.startswith("_Z")}
    for stem in ("8fft_init", "11fft_destroy", "20fft_square_karatsuba", "22execute_proth_test_fft",
                 "15save_checkpoint", "15load_checkpoint", "17auto_tune_fft_len", "17run_task_with_ctx"):
        o = {m for m in ours if m.startswith("_Z" + stem)}
        r = {m for m in ref if m.startswith("_Z" + stem)}
        assert o and o == r, (stem, o, r)


@pytest.mark.skipif(not (REF_TREE / "presieve_core.h").exists(), reason="reference headers not present on this host")
def test_sieve_mirror_names_equal_reference_declarations(tmp_path):
    """The sieve entry points, declared by the reference's own presieve_core.h
    (compiled here, only to read the symbols it references), are the symbols
    the mirror defines."""
    src = tmp_path / "use.cpp"
    src.write_text('#include "presieve_core.h"\n'
                   'void* used[] = {(void*)&execute_full_streaming_sieve, (void*)&execute_sparse_streaming_sieve};\n')
    obj = tmp_path / "use.o"
    subprocess.run(["g++", "-std=c++17", "-c", f"-I{REF_TREE}", str(src), "-o", str(obj)], check=True)
    und = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout.split()
    wanted = {m for m in und if "streaming_sieve" in m}
    assert len(wanted) == 2
    assert wanted <= _defined(HOST_LIB, demangle=False)


@pytest.mark.gpu
def test_selftest_program(gpu, tmp_path):
    assert SELFTEST.exists(), "prothhost_selftest is not built"
    print(f"[host mirror] running {SELFTEST} in {tmp_path}", flush=True)
    p = subprocess.run([str(SELFTEST)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    lines = [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    cases = {}
    for c in lines:
        cases.setdefault(c["case"], []).append(c)
    assert "done" in cases

    for c in cases["full"]:
        prime, r64, a = G.proth_verdict(c["k"], c["n"])
        assert c["prime"] == int(prime), c
        assert int(c["res64"], 16) == r64, c
        assert c["fft_len"] >= 1024 and c["max_err"] < 0.3

    for c in cases["square"]:
        N = G.proth_N(c["k"], c["n"])
        x = int(c["x0"], 16) % N
        for _ in range(c["steps"]):
            x = x * x % N
        assert int(c["x"], 16) == x, c

    (c,) = cases["resume"]
    prime, r64, _ = G.proth_verdict(c["k"], c["n"])
    assert c["roundtrip"] == 1 and c["checkpoint_removed"] == 1
    assert c["prime"] == int(prime) and int(c["res64"], 16) == r64
    # the saved index i is the last squaring applied: the run continues at i+1
    assert c["squarings_run"] == (c["n"] - 1) - (c["saved_i"] + 1)

    (w,) = cases["worker"]
    n = w["n"]
    assert w["single"] == ("prime" if G.proth_verdict(w["single_k"], n)[0] else "PRP")
    assert w["batch"] == "BATCH" and w["other"] == "fail"
    found = set()
    for line in (tmp_path / "primes.txt").read_text().splitlines():
        m = re.match(r"(\d+) \* 2\^(\d+) \+ 1 \(a=(\d+)\)", line)
        assert m, line
        k, nn, a = int(m.group(1)), int(m.group(2)), int(m.group(3))
        assert a == G.reference_base(k, nn)
        found.add((k, nn))
    expect = {(c["k"], c["n"]) for c in cases["full"] if c["prime"]}
    expect |= {(int(k), n) for k in w["k_list"].split(",") if G.proth_verdict(int(k), n)[0]}
    if w["single"] == "prime":
        expect.add((w["single_k"], n))
    assert found == expect
    assert not [f for f in os.listdir(tmp_path) if f.startswith("check_")]
