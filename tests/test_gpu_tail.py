"""Values that leave the finite range near the end of a loop.

Where the reference recurrence (circuit_gpu.cu:114-129) overflows, its NaN
spreads one segment per step, so a wire that overflows in the last ~10 steps
of a loop ends with a mix of finite, +-inf and NaN values, fixed by the exact
step at which each value overflowed.  Every policy runs the last 16 steps in
the reference's order, and wires whose modes grow or decay slowly (gain
magnitude in [0.95, 1000]) or that start from values below 2^-96 run the whole
loop in it.  Each case below is compared element by element with the oracle
(circuit.cc:352-413 with the GPU seeding): EXACT bit-identical, RECIP, FAST
and SCALED with the same non-finite class on every current and internal
voltage and within 1e-5 (per wire) where finite.
"""
import sys

import numpy as np
import pytest

from helpers import TOL, assert_bit_identical, assert_same_class, classes, max_wire_rel_err

pytestmark = pytest.mark.gpu

POLICIES = [(0, "exact"), (1, "recip"), (2, "fast"), (3, "scaled")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _relax_gpu(C, circ, policy, steps):
    """One calc_new_currents launch over every piece with `steps` inner steps."""
    with C.Context(0) as ctx:
        ctx.load(C.HostCircuit(circ.nodes, circ.wires, circ.locator, circ.pieces))
        ctx.set_option(C.OPT_STEPS, steps)
        ctx.calc_new_currents(ctx.pieces(), policy)
        return ctx.store()[1]


def _relax_oracle(O, circ, steps):
    ref = circ.copy()
    for p in circ.pieces:
        O.calc_new_currents(ref, p, steps=steps)
    return ref.wires


def _mixed_wires(w):
    c = classes(w["current"])
    return int(((c == 0).any(1) & (c != 0).any(1) | (c == 3).any(1) & (c != 3).any(1)).sum())


def _check_all_policies(C, circ, want, steps, label):
    for policy, name in POLICIES:
        got = _relax_gpu(C, circ, policy, steps)
        for f in ("current", "voltage"):
            if policy == 0:
                assert_bit_identical(got[f], want[f], f"{label} {name} {f}")
                continue
            assert_same_class(got[f], want[f], f"{label} {name} {f}")
            fin = np.isfinite(want[f]).all(axis=1)
            err = max_wire_rel_err(got[f][fin], want[f][fin])
            assert err <= TOL, (label, name, f, err)


@pytest.mark.parametrize("steps", [9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 24, 32])
def test_reference_profile_overflow_near_the_end(C, O, steps):
    # The 4-piece circuit in the reference profile: its wires overflow after
    # 8-10 steps, so counts 9-16 end with mixed classes on most wires.
    circ = O.generate(4, 1024, 4096, 95, 12345, O.PROFILE_REFERENCE)
    want = _relax_oracle(O, circ, steps)
    say(f"[tail] reference profile, {steps} steps: {_mixed_wires(want)} wires of "
        f"{len(want)} end with mixed classes")
    _check_all_policies(C, circ, want, steps, f"{steps} steps")


def _growing_circuit(O, gain, steps, seed, log2_start=None):
    """4096 single-piece wires whose slowest mode grows by `gain` per step
    (R = 1, L = gain * dt, so g = L / (dt R) = gain; C = 1, so d = 1e-6 and
    every mode's gain is ~ -g), all nodes at 0 V.  The wires start from random
    currents of magnitude e0: by default chosen so that FLT_MAX / gain^n = e0
    for n uniform in [steps - 25, steps + 5] (overflow in the last 25 steps
    for most); with log2_start = (lo, hi), e0 = 2^U(lo, hi)."""
    c = O.generate(1, 256, 4096, 100, 3, O.PROFILE_STABLE)
    w, n = c.wires, c.nodes
    W = len(w)
    rng = np.random.default_rng(seed)
    n["voltage"] = 0
    w["resistance"] = 1.0
    w["inductance"] = np.float32(gain * 1e-6)
    w["capacitance"] = 1.0
    if log2_start is None:
        n_over = rng.uniform(steps - 25, steps + 5, W)
        e0 = np.exp(np.log(3.4e38) - n_over * np.log(gain))
    else:
        e0 = 2.0 ** rng.uniform(log2_start[0], log2_start[1], W)
    sign = rng.choice([-1.0, 1.0], (W, 10))
    w["current"] = (sign * rng.uniform(0.5, 1.0, (W, 10)) * e0[:, None]).astype(np.float32)
    w["voltage"] = 0
    return c


@pytest.mark.parametrize("gain,steps,start", [
    (1.015, 10000, None),        # slow growth, overflow in the last 25 of 10000 steps
    (1.3, 300, None),
    (2.2, 120, None),
    (10.0, 60, None),
    (1100.0, 20, (-90, 0)),       # fast growth from normal-range starts
    (5000.0, 18, (-90, 0)),
    (1100.0, 30, (-150, -96)),    # starts below 2^-96 (partly subnormal)
], ids=["g1.015", "g1.3", "g2.2", "g10", "g1100", "g5000", "g1100-tiny"])
def test_growth_regimes_overflowing_near_the_end(C, O, gain, steps, start):
    circ = _growing_circuit(O, gain, steps, seed=int(gain * 10) + steps, log2_start=start)
    want = _relax_oracle(O, circ, steps)
    nonfinite = int((~np.isfinite(want["current"])).any(1).sum())
    say(f"[tail] gain {gain}, {steps} steps: {nonfinite} of {len(want)} wires non-finite, "
        f"{_mixed_wires(want)} mixed")
    assert 0 < nonfinite < len(want) or start == (-150, -96)
    _check_all_policies(C, circ, want, steps, f"gain {gain}")


@pytest.mark.parametrize("steps", [17, 40, 10000])
def test_wires_that_skip_the_tail(C, O, steps):
    # Contracting wires with parameters and starting values at the edges of
    # the region where the policies skip the reference-order tail (R at
    # 2^-20..2^20, values up to 2^40, step norms up to 1/2), mixed with
    # wires just outside it (values of 2^41, R of 2^21), which take the tail:
    # every policy against the oracle, as above.
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    from test_tail_guard import draw
    c = O.generate(1, 256, 4096, 100, 3, O.PROFILE_STABLE)
    w, n = c.wires, c.nodes
    W = len(w)
    rng = np.random.default_rng(steps)
    I, V, vin, vout, L, R, Cap = draw(rng, W)
    w["current"], w["voltage"] = I, V
    w["inductance"], w["resistance"], w["capacitance"] = L, R, Cap
    n["voltage"] = (rng.choice([-1, 1], len(n)) * 2.0 ** rng.uniform(0, 40, len(n))).astype(np.float32)
    out = rng.random(W) < 0.2
    w["current"][out, 0] = np.float32(2.0**41)
    w["resistance"][rng.random(W) < 0.1] = np.float32(2.0**21)
    want = _relax_oracle(O, c, steps)
    nonfinite = int((~np.isfinite(want["current"])).any(1).sum())
    say(f"[tail] guard edges, {steps} steps: {nonfinite} of {W} wires non-finite")
    _check_all_policies(C, c, want, steps, f"guard edges {steps}")
