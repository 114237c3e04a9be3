"""Carry hand-off between column blocks of the one-launch outer pass, in the
two-kernel steady state (k_carry_turn + k_row) and in the resident loop
(k_loop), whose workgroups also wait on each other's points.

A wait that never sees its predecessor's carry words gives up after a bounded
number of polls: the launch still ends (in the resident loop every other
workgroup sees the give-up and stops), and the chunk it belonged to is
recomputed from its saved start on the two-launch pass, so the call still
returns the exact result and the context counts the recomputation.  The
condition is forced from the host (PL_HANDOFF_GIVE_UP=1 makes every carry
wait expect a tag no run writes, with a short poll limit), so every kernel
exits normally."""
import random

import pytest

import gmp_ctypes as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("resident", ["1", "0"])
@pytest.mark.parametrize("k,n,L", [(433917, 2097152, 131072), (11887, 300007, 16384)])
def test_wait_gives_up_and_the_chunk_is_recomputed(gpu, monkeypatch, k, n, L, resident):
    N = k * (1 << n) + 1
    monkeypatch.setenv("PL_RESIDENT", resident)
    monkeypatch.setenv("PL_HANDOFF_GIVE_UP", "1")
    print(f"[handoff] give-up path k={k} n={n} L={L} resident={resident}", flush=True)
    x = random.Random(1).randrange(N)
    with gpu.Context(k, n, L) as c:
        # 0: resident loop (where the length has one), 2: two-kernel steady state
        assert c.info.passes == (0 if resident == "1" and L == 131072 else 2)
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
