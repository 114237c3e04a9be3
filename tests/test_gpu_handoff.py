"""Carry hand-off between column blocks of the one-launch outer pass.

A wait that never sees its predecessor's carry words gives up after a bounded
number of polls: the launch still ends, and the chunk it belonged to is
recomputed from its saved start on the two-launch pass, so the call still
returns the exact result and the context counts the recomputation.  The
condition is forced from the host (PL_HANDOFF_GIVE_UP=1 makes every wait
expect a tag no run writes, with a short poll limit), so every kernel exits
normally."""
import random

import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,n,L", [(433917, 2097152, 131072), (11887, 300007, 16384)])
def test_wait_gives_up_and_the_chunk_is_recomputed(gpu, monkeypatch, k, n, L):
    N = k * (1 << n) + 1
    monkeypatch.setenv("PL_HANDOFF_GIVE_UP", "1")
    print(f"[handoff] give-up path k={k} n={n} L={L}", flush=True)
    x = random.Random(1).randrange(N)
    with gpu.Context(k, n, L) as c:
        assert c.info.passes == 2
        c.set_residue(x)
        e = c.square(300)          # one whole chunk and a partial one
        assert 0 < e < 0.3
        assert c.get_residue() == G.powm(x, 1 << 300, N)
        assert c.refresh_info().handoff_fallbacks == 2
    monkeypatch.delenv("PL_HANDOFF_GIVE_UP")
    print(f"[handoff] normal path after it k={k} n={n} L={L}", flush=True)
    x = random.Random(2).randrange(N)
    with gpu.Context(k, n, L) as c:
        c.set_residue(x)
        e = c.square(70)
        assert 0 < e < 0.3
        assert c.get_residue() == G.powm(x, 1 << 70, N)
        assert c.refresh_info().handoff_fallbacks == 0
