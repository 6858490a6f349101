"""The SCALED policy's per-wire choice between its scaled-deviation step and
FAST's tridiagonal step, and its handling of step counts that are not a
multiple of its 4-step block.

Wires are given parameters from four regimes (through the accessor-form
entry point, which takes wire records as they are):
  * contracting (the stable profile): the scaled step, within 1e-5 of the
    oracle, and not bit-equal to FAST (so the scaled path really ran);
  * marginal (L/(dt R) = 1, an undamped mode): the whole loop in the
    reference's order, as FAST runs it, bit-equal to FAST;
  * tiny coupling (wire capacitance 1e4, d < 2^-25): FAST's step, bit-equal;
  * diverging (the reference inductance): all-NaN, as in the oracle.
"""
import numpy as np
import pytest

from helpers import TOL, assert_bit_identical, assert_same_class, max_rel_err, max_wire_rel_err

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


def test_scaled_matches_fast_on_random_parameters(C, O):
    # Wire parameters and states spread over many decades, so that
    # g = L/(dt R) runs over 1e-8..1e8 and d = dt/(R C) over 1e-11..1e13:
    # contracting, diverging, marginal and tiny-coupling wires side by side,
    # mixed within waves.  SCALED must give FAST's non-finite class on every
    # element and FAST's finite values to 1e-5 (per wire, relative to the
    # wire's largest segment value), bit for bit where it falls back; both,
    # and RECIP, the oracle's classes on every element and its finite values
    # to 1e-5.
    c = O.generate(4, 256, 1024, 95, 41, O.PROFILE_STABLE)
    w, n = c.wires, c.nodes
    rng = np.random.default_rng(11)
    W = len(w)
    w["resistance"] = (10.0 ** rng.uniform(-2, 2, W)).astype(np.float32)
    w["inductance"] = (10.0 ** rng.uniform(-12, 0, W)).astype(np.float32)
    w["capacitance"] = (10.0 ** rng.uniform(-9, 3, W)).astype(np.float32)
    w["current"] = (rng.uniform(-1, 1, (W, 10)) * 10.0 ** rng.uniform(-3, 3, (W, 1))).astype(np.float32)
    w["voltage"] = (rng.uniform(-1, 1, (W, 9)) * 10.0 ** rng.uniform(-2, 2, (W, 1))).astype(np.float32)
    n["voltage"] = (rng.uniform(-1, 1, len(n)) * 10.0 ** rng.uniform(-2, 2, len(n))).astype(np.float32)
    scaled = _run_accessor(C, c, C.POLICY_SCALED)
    fast = _run_accessor(C, c, C.POLICY_FAST)
    recip = _run_accessor(C, c, C.POLICY_RECIP)
    ref = c.copy()
    for p in c.pieces:
        O.calc_new_currents(ref, p)
    for f in ("current", "voltage"):
        assert_same_class(scaled[f], fast[f], f)
        fin = np.isfinite(fast[f]).all(axis=1)
        assert max_wire_rel_err(scaled[f][fin], fast[f][fin]) <= TOL, f
        fin = np.isfinite(ref.wires[f]).all(axis=1)
        for name, got in (("scaled", scaled), ("fast", fast), ("recip", recip)):
            assert_same_class(got[f], ref.wires[f], f"{name} {f} vs oracle")
            assert max_wire_rel_err(got[f][fin], ref.wires[f][fin]) <= TOL, (name, f)
    # the regimes are all present
    dt = 1e-6
    R, L, Cw = (w[k].astype(np.float64) for k in ("resistance", "inductance", "capacitance"))
    g, d = L / dt / R, dt / (R * Cw)
    assert (g - 4 * d >= 2).sum() > 100 and (np.maximum(np.abs(2 * d - g) + 2 * d, 0) <= 0.5).sum() > 100
    assert (d < 2.0 ** -25).sum() > 100
    nonfin = ~np.isfinite(fast["current"]).all(axis=1)
    assert 0 < nonfin.sum() < W


def test_degenerate_wire_parameters(C, O):
    # Zero wire capacitance (the coupling d = dt/(R C) is infinite), zero
    # inductance (no inductive term, g = 0), a huge resistance (d and g near
    # zero) and ordinary wires, mixed in every wave.  Every policy gives the
    # oracle's non-finite class on every element and its finite values to
    # 1e-5 per wire; EXACT is bit-identical.
    c = O.generate(2, 256, 1024, 95, 23, O.PROFILE_STABLE)
    w = c.wires
    k = np.arange(len(w)) % 4
    w["capacitance"][k == 0] = np.float32(0)
    w["inductance"][k == 1] = np.float32(0)
    w["resistance"][k == 2] = np.float32(1e30)
    ref = c.copy()
    for p in c.pieces:
        O.calc_new_currents(ref, p)
    assert np.isnan(ref.wires["current"][k == 0]).all()
    assert np.isfinite(ref.wires["current"][k != 0]).all()
    for policy in (C.POLICY_EXACT, C.POLICY_RECIP, C.POLICY_FAST, C.POLICY_SCALED):
        got = _run_accessor(C, c, policy)
        for f in ("current", "voltage"):
            if policy == C.POLICY_EXACT:
                assert_bit_identical(got[f], ref.wires[f], f"EXACT {f}")
                continue
            assert_same_class(got[f], ref.wires[f], f"policy {policy} {f}")
            assert max_wire_rel_err(got[f], ref.wires[f]) <= TOL, (policy, f)
