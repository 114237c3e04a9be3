"""Command-line front end (prothlogic_amd/prothlogic, flags of P/cli.cpp:39-68):
sieve -> survivor file -> device pool -> primes file, checked against the
sieve oracle and GMP."""
import re
import subprocess
from pathlib import Path

import pytest

import gmp_ctypes as G
import sieve_oracle as SO

EXE = Path(__file__).resolve().parent.parent / "prothlogic_amd" / "prothlogic"


def run(args, cwd, timeout=240):
    print(f"[cli] prothlogic {' '.join(args)}", flush=True)
    return subprocess.run([str(EXE), *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_usage_and_missing_parameters(tmp_path):
    assert EXE.exists(), "prothlogic front end is not built"
    p = run([], tmp_path)
    assert p.returncode == 0 and "Usage" in p.stdout
    p = run(["--k-start", "5"], tmp_path)
    assert p.returncode == 1 and "Missing required parameters" in p.stderr
    p = run(["--n", "20909", "--bogus"], tmp_path)
    assert p.returncode == 1


@pytest.mark.gpu
def test_range_sieve_then_tests(gpu, tmp_path):
    n, k0, k1, pmax = 20909, 1, 400, 100_000
    p = run(["--n", str(n), "--k-start", str(k0), "--k-end", str(k1), "--pmax", str(pmax),
             "--out-primes", "found.txt", "--out-sieve", "sieve.txt"], tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    survivors = [int(x) for x in (tmp_path / "sieve.txt").read_text().split()]
    assert survivors == SO.sieve_range(k0, k1, n, pmax)
    primes = set()
    if (tmp_path / "found.txt").exists():
        for line in (tmp_path / "found.txt").read_text().splitlines():
            m = re.match(r"(\d+) \* 2\^(\d+) \+ 1 \(a=(\d+)\)", line)
            assert m and int(m.group(2)) == n
            primes.add(int(m.group(1)))
    assert primes == {k for k in survivors if G.proth_verdict(k, n)[0]}
    assert 3 in primes  # 3*2^20909+1


@pytest.mark.gpu
def test_sieve_only_and_list_sieve(gpu, tmp_path):
    n = 120002
    p = run(["--n", str(n), "--k-start", "1", "--k-end", "20000", "--pmax", "50000", "--sieve-only",
             "--out-sieve", "s.txt"], tmp_path)
    assert p.returncode == 0 and "--sieve-only" in p.stdout
    s = [int(x) for x in (tmp_path / "s.txt").read_text().split()]
    assert s == SO.sieve_range(1, 20000, n, 50000)
    p = run(["--n", str(n), "--in-file", "s.txt", "--pstart", "50000", "--pmax", "400000", "--sieve-only",
             "--out-sieve", "s2.txt"], tmp_path)
    assert p.returncode == 0
    s2 = [int(x) for x in (tmp_path / "s2.txt").read_text().split()]
    keep = SO.sieve_list(s, n, 50000, 400000)
    assert s2 == [k for k, kp in zip(s, keep) if kp]
    assert not (tmp_path / "primes.txt").exists()


@pytest.mark.gpu
def test_single_candidate(gpu, tmp_path):
    p = run(["--n", "20909", "--k-start", "3", "--k-end", "3", "--report", "50"], tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "FOUND PRIME: 3 * 2^20909 + 1" in p.stdout
    assert "Res64: 0000000000000000" in p.stdout
    assert "[STEP] k=3 -> 50%" in p.stdout and "-> 100%" in p.stdout
