"""Circuit shapes drawn at random (fixed seeds) and three edge cases: 1-7
pieces, node and wire counts that are not multiples of the workgroup, wave or
LDS block sizes, in-piece percentages from 0 to 100, both generator profiles.  Each runs two
loops (circuit.cc:232-307) through the resident context and is compared with
the oracle's two loops of the reference task bodies: EXACT + CSR bit for bit
on every node and wire; RECIP, FAST and SCALED with the fused ATOMIC
reduction by non-finite class on every value and within 1e-5 where finite
(node voltages elementwise, wire values per wire); SCALED with the separate
LDS-privatised ATOMIC reduction likewise."""
import numpy as np
import pytest

from helpers import TOL, assert_bit_identical, assert_same_class, max_rel_err, max_wire_rel_err

pytestmark = pytest.mark.gpu


def _shape(k):
    rng = np.random.default_rng(500 + k)
    P = int(rng.integers(1, 8))
    npp = int(rng.integers(1, 700))
    wpp = int(rng.integers(1, 3000))
    pct = int(rng.choice([0, 17, 50, 95, 100]))
    return P, npp, wpp, pct, 700 + k, int(k % 2)


SHAPES = [_shape(k) for k in range(12)] + [
    (1, 1, 1, 50, 720, 0),     # one node, one wire, one piece (the wrap rule)
    (1, 5, 70, 0, 721, 1),     # every wire "crosses" to the only piece
    (3, 1, 64, 95, 722, 1),    # one node per piece
]


def _gpu(C, shape, policy, mode, **opts):
    P, npp, wpp, pct, seed, profile = shape
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, pct, seed, profile))
        for k, v in opts.items():
            ctx.set_option(getattr(C, k), v)
        ctx.run_loops(2, policy=policy, mode=mode)
        return ctx.store()


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}x{s[2]}-pct{s[3]}-prof{s[5]}"
                                               for s in SHAPES])
def test_random_shape_against_oracle(C, O, shape):
    P, npp, wpp, pct, seed, profile = shape
    ref = O.generate(P, npp, wpp, pct, seed, profile)
    O.run(ref, 2)
    nodes, wires = _gpu(C, shape, C.POLICY_EXACT, C.REDUCE_CSR)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "EXACT node voltage")
    assert_bit_identical(nodes["charge"], ref.nodes["charge"], "EXACT node charge")
    assert_bit_identical(wires["current"], ref.wires["current"], "EXACT wire current")
    assert_bit_identical(wires["voltage"], ref.wires["voltage"], "EXACT wire voltage")
    runs = [(C.POLICY_RECIP, {}), (C.POLICY_FAST, {}), (C.POLICY_SCALED, {}),
            (C.POLICY_SCALED, {"OPT_FUSE_DSC": 0, "OPT_DSC_ATOMIC_IMPL": 2})]
    for policy, opts in runs:
        if opts.get("OPT_DSC_ATOMIC_IMPL") == 2 and npp * 4 > 160 * 1024:
            continue
        n, w = _gpu(C, shape, policy, C.REDUCE_ATOMIC, **opts)
        what = f"policy {policy} {opts}"
        assert_same_class(n["voltage"], ref.nodes["voltage"], what + " node voltage")
        assert_same_class(w["current"], ref.wires["current"], what + " wire current")
        assert_same_class(w["voltage"], ref.wires["voltage"], what + " wire voltage")
        assert max_rel_err(n["voltage"], ref.nodes["voltage"]) <= TOL, what
        fin = np.isfinite(ref.wires["current"]).all(axis=1) & np.isfinite(ref.wires["voltage"]).all(axis=1)
        if fin.any():
            assert max_wire_rel_err(w["current"][fin], ref.wires["current"][fin]) <= TOL, what
            assert max_wire_rel_err(w["voltage"][fin], ref.wires["voltage"][fin]) <= TOL, what
