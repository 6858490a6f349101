"""hipGraph replays: the captured loop is replayed several times with eager
work and host synchronisation between replays, and every replay must equal
the eager execution of the same loop (EXACT + CSR: bit-identical; the
persistent kernel's self-resetting queue counters must also survive replays)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1], ids=["atomic-fused", "csr"])
@pytest.mark.parametrize("persistent", [0, 1])
def test_graph_replays_match_eager(C, mode, persistent):
    params = C.GenParams(4, 512, 2048, 80, 77, C.PROFILE_STABLE)
    graph_states, eager_states = [], []
    with C.Context(0) as g, C.Context(0) as e:
        for ctx in (g, e):
            ctx.generate(params)
            ctx.set_option(C.OPT_PERSISTENT, persistent)
            ctx.set_option(C.OPT_CNC_BLOCK, 128)
        for _ in range(4):
            g.run_loops(1, policy=C.POLICY_EXACT, mode=mode, use_graph=True)
            graph_states.append(g.store())          # eager D2H + sync between replays
            g.sanity_check_wires()                   # an eager launch between replays
            e.run_loops(1, policy=C.POLICY_EXACT, mode=mode, use_graph=False)
            eager_states.append(e.store())
    for (gn, gw), (en, ew) in zip(graph_states, eager_states):
        if mode == 1:
            assert gn.tobytes() == en.tobytes()
            assert gw.tobytes() == ew.tobytes()
        else:
            # float atomics: same values up to the order of the additions
            import numpy as np
            assert np.allclose(gn["voltage"], en["voltage"], rtol=1e-6, atol=0)
            assert gw["current"].tobytes() == ew["current"].tobytes() or \
                np.allclose(gw["current"], ew["current"], rtol=1e-5, atol=1e-12)
