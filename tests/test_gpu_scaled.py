"""The SCALED policy's per-wire choice between its scaled-deviation step and
FAST's tridiagonal step, and its handling of step counts that are not a
multiple of its 4-step block.

Wires are given parameters from four regimes (through the accessor-form
entry point, which takes wire records as they are):
  * contracting (the stable profile): the scaled step, within 1e-5 of the
    oracle, and not bit-equal to FAST (so the scaled path really ran);
  * marginal (L/(dt R) = 1, an undamped mode): FAST's step, bit-equal to FAST;
  * tiny coupling (wire capacitance 1e4, d < 2^-25): FAST's step, bit-equal;
  * diverging (the reference inductance): all-NaN, as in the oracle.
"""
import numpy as np
import pytest

from helpers import TOL, assert_same_class, max_rel_err, max_wire_rel_err

pytestmark = pytest.mark.gpu


def _run_accessor(C, circ, flag):
    dw = C.DeviceArray(circ.wires)
    pvt, owned, ghost = (C.DeviceArray(circ.nodes) for _ in range(3))
    for p in circ.pieces:
        C.calc_new_currents_gpu(p, dw, pvt, owned, ghost, flag=flag)
    return dw.download()


def test_scaled_guard_paths(C, O):
    c = O.generate(4, 256, 1024, 95, 31, O.PROFILE_STABLE)
    w = c.wires
    k = np.arange(len(w)) % 4
    dt = np.float32(1e-6)
    w["inductance"][k == 1] = w["resistance"][k == 1] * dt          # L/(dt R) = 1
    w["capacitance"][k == 2] = np.float32(1e4)                       # d ~ 1e-10
    rng = np.random.default_rng(3)
    w["inductance"][k == 3] = (0.01 * rng.random((k == 3).sum()) + 0.1).astype(np.float32)
    ref = c.copy()
    for p in c.pieces:
        O.calc_new_currents(ref, p)
    scaled = _run_accessor(C, c, C.POLICY_SCALED)
    fast = _run_accessor(C, c, C.POLICY_FAST)
    for f in ("current", "voltage"):
        # contracting: within tolerance of the oracle, computed by the scaled step
        g0 = k == 0
        assert np.isfinite(scaled[f][g0]).all()
        assert max_wire_rel_err(scaled[f][g0], ref.wires[f][g0]) <= TOL
        # marginal and tiny coupling: FAST's step, bit for bit
        for grp in (1, 2):
            sel = k == grp
            assert scaled[f][sel].tobytes() == fast[f][sel].tobytes(), (f, grp)
        # diverging: all NaN on both sides
        g3 = k == 3
        assert np.isnan(ref.wires[f][g3]).all()
        assert_same_class(scaled[f][g3], ref.wires[f][g3], f"diverging {f}")
    assert scaled["current"][k == 0].tobytes() != fast["current"][k == 0].tobytes()


@pytest.mark.parametrize("steps", [2, 3, 5, 6, 9, 13, 17, 67, 200])
def test_scaled_short_step_counts(C, steps):
    # Remainder steps (steps - 1 mod 4) run first, then whole blocks; every
    # count agrees with EXACT + CSR (bit-identical to the oracle) to 1e-5.
    gp = C.GenParams(4, 1024, 4096, 95, 12345, C.PROFILE_STABLE)
    out = {}
    for name, policy, mode in (("exact", C.POLICY_EXACT, C.REDUCE_CSR),
                               ("fast", C.POLICY_FAST, C.REDUCE_CSR),
                               ("scaled", C.POLICY_SCALED, C.REDUCE_CSR)):
        with C.Context(0) as ctx:
            ctx.generate(gp)
            ctx.set_option(C.OPT_STEPS, steps)
            ctx.run_loops(1, policy=policy, mode=mode)
            out[name] = ctx.store()
    for name in ("fast", "scaled"):
        nodes, wires = out[name]
        en, ew = out["exact"]
        assert np.isfinite(wires["current"]).all()
        assert max_wire_rel_err(wires["current"], ew["current"]) <= TOL, name
        assert max_wire_rel_err(wires["voltage"], ew["voltage"]) <= TOL, name
        assert max_rel_err(nodes["voltage"], en["voltage"]) <= TOL, name
