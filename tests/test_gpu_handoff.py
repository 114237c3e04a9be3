"""Carry hand-off between column blocks of the one-launch outer pass.

A wait that never sees its predecessor's carry words gives up after a bounded
number of polls: the launch still ends, the call fails with PL_ERR_STATE and
says why, and a new context on the same device squares exactly.  The
condition is forced from the host (PL_HANDOFF_GIVE_UP=1 makes every wait
expect a tag no run writes, with a short poll limit), so every kernel exits
normally."""
import random

import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,n,L", [(433917, 2097152, 131072), (11887, 300007, 16384)])
def test_wait_gives_up_and_reports(gpu, monkeypatch, k, n, L):
    N = k * (1 << n) + 1
    monkeypatch.setenv("PL_HANDOFF_GIVE_UP", "1")
    print(f"[handoff] give-up path k={k} n={n} L={L}", flush=True)
    with gpu.Context(k, n, L) as c:
        assert c.info.passes == 2
        c.set_residue(random.Random(1).randrange(N))
        with pytest.raises(gpu.PlError) as ei:
            c.square(3)
        assert ei.value.code == gpu.PL_ERR_STATE
        assert "hand-off" in str(ei.value)
    monkeypatch.delenv("PL_HANDOFF_GIVE_UP")
    print(f"[handoff] normal path after it k={k} n={n} L={L}", flush=True)
    x = random.Random(2).randrange(N)
    with gpu.Context(k, n, L) as c:
        c.set_residue(x)
        e = c.square(70)
        assert 0 < e < 0.3
        assert c.get_residue() == G.powm(x, 1 << 70, N)
