"""GPU parity: the HIP kernels, called through the C ABI, against the oracle.

Sizes: the 4-piece circuit (4 pieces x 1024 nodes x 4096 wires, 95% in-piece,
seed 12345, 5 loops of 10000 inner steps) and its pct-50 variant, in the reference profile
(values diverge to inf/NaN, parity by class and bit pattern) and the stable
profile (finite values).  EXACT + CSR must be bit-identical to the oracle;
EXACT + KAHAN64 bit-identical to the oracle's Kahan restatement; the other
policy/mode combinations within 1e-5 relative error on node voltages.
"""
import sys

import numpy as np
import pytest

from helpers import (TOL, assert_bit_identical, assert_same_class, max_rel_err,
                     max_wire_rel_err)

pytestmark = pytest.mark.gpu

SMALL = dict(P=4, npp=1024, wpp=4096, pct=95, seed=12345)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_gpu(C, P, npp, wpp, pct, seed, profile, loops, policy, mode, **opts):
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, pct, seed, profile))
        for k, v in opts.items():
            ctx.set_option(getattr(C, k), v)
        ctx.run_loops(loops, policy=policy, mode=mode)
        return ctx.store()


_oracle_cache = {}


def run_oracle(O, P, npp, wpp, pct, seed, profile, loops, reduce=0):
    key = (P, npp, wpp, pct, seed, profile, loops, reduce)
    if key not in _oracle_cache:
        c = O.generate(P, npp, wpp, pct, seed, profile)
        O.run(c, loops, reduce=reduce)
        _oracle_cache[key] = c
    return _oracle_cache[key]


@pytest.mark.parametrize("profile", [0, 1], ids=["reference", "stable"])
def test_small_exact_csr_bit_identical(C, O, profile):
    say(f"[4-piece exact csr] profile={profile}")
    nodes, wires = run_gpu(C, **SMALL, profile=profile, loops=5, policy=C.POLICY_EXACT,
                           mode=C.REDUCE_CSR)
    ref = run_oracle(O, **SMALL, profile=profile, loops=5)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(nodes["charge"], ref.nodes["charge"], "node charge")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")
    assert_bit_identical(wires["voltage"], ref.wires["voltage"], "wire voltage")
    say(f"  NaN fraction of node voltages: {np.isnan(nodes['voltage']).mean():.4f}")


def test_small_pct50_exact_csr_bit_identical(C, O):
    cfg = dict(SMALL, pct=50)
    nodes, wires = run_gpu(C, **cfg, profile=1, loops=5, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
    ref = run_oracle(O, **cfg, profile=1, loops=5)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")


def test_small_kahan64_bit_identical(C, O):
    nodes, wires = run_gpu(C, **SMALL, profile=1, loops=5, policy=C.POLICY_EXACT,
                           mode=C.REDUCE_CSR_KAHAN64)
    ref = run_oracle(O, **SMALL, profile=1, loops=5, reduce=1)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")


@pytest.mark.parametrize("policy", [0, 1, 2, 3], ids=["exact", "recip", "fast", "scaled"])
@pytest.mark.parametrize("mode", [0, 1], ids=["atomic", "csr"])
def test_small_stable_policies_within_tolerance(C, O, policy, mode):
    nodes, wires = run_gpu(C, **SMALL, profile=1, loops=5, policy=policy, mode=mode)
    ref = run_oracle(O, **SMALL, profile=1, loops=5)
    e_v = max_rel_err(nodes["voltage"], ref.nodes["voltage"])
    e_i = max_wire_rel_err(wires["current"], ref.wires["current"])
    e_w = max_wire_rel_err(wires["voltage"], ref.wires["voltage"])
    say(f"[4-piece stable policy={policy} mode={mode}] node V {e_v:.2e} wire I {e_i:.2e} wire V {e_w:.2e}")
    assert np.isfinite(nodes["voltage"]).all()
    assert e_v <= TOL
    assert e_i <= TOL
    assert e_w <= TOL


@pytest.mark.parametrize("loops", [1, 5])
@pytest.mark.parametrize("policy", [1, 2, 3], ids=["recip", "fast", "scaled"])
def test_small_reference_profile_fast_policies_match_nan_classes(C, O, policy, loops):
    # After one loop the currents carry the divergence themselves (the node
    # voltages are still finite going in); after five, NaN node voltages
    # feed every wire.
    nodes, wires = run_gpu(C, **SMALL, profile=0, loops=loops, policy=policy, mode=C.REDUCE_ATOMIC)
    ref = run_oracle(O, **SMALL, profile=0, loops=loops)
    assert_same_class(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_same_class(wires["current"], ref.wires["current"], "wire current")
    assert_same_class(wires["voltage"], ref.wires["voltage"], "wire voltage")
    fin = np.isfinite(ref.nodes["voltage"])
    assert max_rel_err(nodes["voltage"][fin], ref.nodes["voltage"][fin]) <= TOL


def test_run_loop_modes_agree(C):
    results = []
    for sync, graph, per_piece in [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0), (0, 0, 1)]:
        with C.Context(0) as ctx:
            ctx.generate(C.GenParams(4, 256, 1024, 90, 3, C.PROFILE_STABLE))
            rep = ctx.run_loops(3, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR, sync=sync,
                                use_graph=bool(graph), per_piece=bool(per_piece))
            assert rep["elapsed_s"] > 0 and rep["gflops"] > 0
            results.append(ctx.store())
    for nodes, wires in results[1:]:
        assert nodes.tobytes() == results[0][0].tobytes()
        assert wires.tobytes() == results[0][1].tobytes()


def test_task_variant_table_and_nops(C, O):
    params = C.GenParams(4, 128, 512, 80, 21, C.PROFILE_STABLE)
    with C.Context(0) as ctx:
        ctx.generate(params)
        pieces = ctx.pieces()
        for task in (C.CALC_NEW_CURRENTS_NOP, C.CALC_NEW_CURRENTS, C.DISTRIBUTE_CHARGE_NOP,
                     C.DISTRIBUTE_CHARGE, C.UPDATE_VOLTAGES_NOP, C.UPDATE_VOLTAGES,
                     C.SANITY_CHECK_WIRES):
            ctx.execute_index_space(task, pieces, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
        nodes, wires = ctx.store()
        with pytest.raises(C.CircuitError):
            ctx.execute_index_space(C.REGION_MAIN, pieces)
        with pytest.raises(C.CircuitError):
            ctx.execute_index_space(99, pieces)
    ref = O.generate(4, 128, 512, 80, 21, O.PROFILE_STABLE)
    O.run(ref, 1)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")


def test_per_piece_calls_and_errors(C, O):
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(4, 128, 512, 80, 22, C.PROFILE_STABLE))
        pieces = ctx.pieces()
        # calc_new_currents one piece at a time, in reverse order
        for p in pieces[::-1]:
            ctx.calc_new_currents(p[None], C.POLICY_EXACT)
        ctx.distribute_charge(pieces, C.REDUCE_CSR)
        ctx.update_voltages(pieces)
        nodes, wires = ctx.store()
        # CSR needs the whole index space in one launch
        with pytest.raises(C.CircuitError) as e:
            ctx.distribute_charge(pieces[:2], C.REDUCE_CSR)
        assert e.value.status == 8
        bogus = pieces[:1].copy()
        bogus["first_wire"] += 1
        with pytest.raises(C.CircuitError) as e:
            ctx.calc_new_currents(bogus)
        assert e.value.status == 2
        with pytest.raises(C.CircuitError):
            ctx.calc_new_currents(pieces, policy=7)
        # a piece named twice in one call would be processed twice at once
        twice = np.concatenate([pieces, pieces[1:2]])
        for call in (lambda: ctx.calc_new_currents(twice), lambda: ctx.update_voltages(twice),
                     lambda: ctx.distribute_charge(twice, C.REDUCE_ATOMIC)):
            with pytest.raises(C.CircuitError) as e:
                call()
            assert e.value.status == 1
        assert ctx.store()[0].tobytes() == nodes.tobytes()
    ref = O.generate(4, 128, 512, 80, 22, O.PROFILE_STABLE)
    O.run(ref, 1)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")


def test_one_inner_step_is_ohms_law_on_gpu(C):
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(2, 64, 256, 80, 5, C.PROFILE_STABLE))
        ctx.set_option(C.OPT_STEPS, 1)
        n0, w0 = ctx.store()
        for policy in (C.POLICY_EXACT, C.POLICY_RECIP, C.POLICY_FAST, C.POLICY_SCALED):
            ctx.generate(C.GenParams(2, 64, 256, 80, 5, C.PROFILE_STABLE))
            ctx.calc_new_currents(None, policy)
            _, w = ctx.store()
            vin = n0["voltage"][w0["in_ptr"]]
            vout = n0["voltage"][w0["out_ptr"]]
            want0 = -vin / w0["resistance"]
            want9 = vout / w0["resistance"]
            if policy == C.POLICY_EXACT:
                assert np.array_equal(w["current"][:, 0], want0)
                assert np.array_equal(w["current"][:, 9], want9)
            else:
                assert np.allclose(w["current"][:, 0], want0, rtol=1e-6)
                assert np.allclose(w["current"][:, 9], want9, rtol=1e-6)
            assert not w["current"][:, 1:9].any()


def test_exact_subnormal_quotients_bit_identical(C, O):
    # Currents and voltages around 1e-38..1e-44 keep the relaxation in the
    # subnormal range, so EXACT's divisions return subnormal and zero
    # quotients (from nonzero and zero dividends): the correctly rounded
    # division's rare path, bit for bit against the oracle's IEEE division.
    c = O.generate(2, 256, 1024, 95, 17, O.PROFILE_STABLE)
    rng = np.random.default_rng(5)
    W, N = len(c.wires), len(c.nodes)
    c.wires["current"] = (rng.uniform(-1, 1, (W, 10)) * 1e-38).astype(np.float32)
    c.wires["voltage"] = (rng.uniform(-1, 1, (W, 9)) * 1e-38).astype(np.float32)
    c.nodes["voltage"] = (rng.uniform(-1, 1, N) * 1e-38).astype(np.float32)
    ref = c.copy()
    dw = C.DeviceArray(c.wires)
    pvt, owned, ghost = (C.DeviceArray(c.nodes) for _ in range(3))
    for p in c.pieces:
        C.calc_new_currents_gpu(p, dw, pvt, owned, ghost, flag=C.POLICY_EXACT)
        O.calc_new_currents(ref, p)
    got = dw.download()
    assert_bit_identical(got["current"], ref.wires["current"], "wire current")
    assert_bit_identical(got["voltage"], ref.wires["voltage"], "wire voltage")
    cur = ref.wires["current"]
    sub = (cur != 0) & (np.abs(cur) < np.float32(2.0 ** -126))
    assert sub.mean() > 0.5


def test_accessor_form_dropins(C, O):
    """The _gpu entry points on reference-layout instances (separate pvt,
    owned and ghost instances, fold buffers, int locator)."""
    c = O.generate(4, 256, 1024, 70, 31, O.PROFILE_STABLE)
    N = len(c.nodes)
    dw = C.DeviceArray(c.wires)
    pvt, owned, ghost = (C.DeviceArray(c.nodes) for _ in range(3))
    ref = c.copy()
    for p in c.pieces:
        C.calc_new_currents_gpu(p, dw, pvt, owned, ghost, flag=C.POLICY_EXACT)
        O.calc_new_currents(ref, p)
    assert_bit_identical(dw.download()["current"], ref.wires["current"], "wire current")
    assert_bit_identical(dw.download()["voltage"], ref.wires["voltage"], "wire voltage")
    # distribute_charge: private into pvt, shared/ghost into fold buffers
    zeros = np.zeros(N, np.float32)
    ofold, gfold = C.DeviceArray(zeros), C.DeviceArray(zeros)
    C.fold_init_gpu(ofold, 0, N)
    C.fold_init_gpu(gfold, 0, N)
    want_pvt = ref.nodes.copy()
    want_o, want_g = zeros.copy(), zeros.copy()
    for p in c.pieces:
        C.distribute_charge_gpu(p, dw, pvt, ofold, gfold)
        O.distribute_charge_fold(ref.wires, want_pvt, want_o, want_g, int(p["first_wire"]),
                                 int(p["num_wires"]))
    assert max_rel_err(pvt.download()["charge"], want_pvt["charge"], floor=1e-12) <= TOL
    assert max_rel_err(ofold.download(), want_o, floor=1e-12) <= TOL
    assert max_rel_err(gfold.download(), want_g, floor=1e-12) <= TOL
    # apply the fold buffers to the owner (shared) instance
    C.fold_apply_gpu(owned, ofold, 0, N)
    C.fold_apply_gpu(owned, gfold, 0, N)
    # update_voltages from an exactly known pre-state
    pre = ref.nodes.copy()
    pre["charge"] = np.random.default_rng(0).uniform(-1e-3, 1e-3, N).astype(np.float32)
    pvt.upload(pre)
    owned.upload(pre)
    locator = C.DeviceArray(c.locator)
    want_p, want_s = pre.copy(), pre.copy()
    for p in c.pieces:
        C.update_voltages_gpu(p, pvt, owned, locator)
        O.update_voltages(ref, p, pvt=want_p, shr=want_s)
    assert pvt.download().tobytes() == want_p.tobytes()
    assert owned.download().tobytes() == want_s.tobytes()
    # sanity check: clean wires pass; an in-node moved out of its piece fails
    assert C.sanity_check_wires_gpu(c.pieces[0], dw, N) == 0
    bad = c.wires.copy()
    bad["in_ptr"][5] = int(c.pieces[1]["first_node"])
    bad["out_loc"][7] = 3
    dw.upload(bad)
    assert C.sanity_check_wires_gpu(c.pieces[0], dw, N) == 2
    # an invalid PointerLocation is reported, not followed
    with pytest.raises(C.CircuitError):
        C.calc_new_currents_gpu(c.pieces[0], dw, pvt, owned, ghost)


def _uneven_pieces(O, W, N, cuts_w, cuts_n):
    """A piece table over all W wires and N nodes in uneven ranges, with an
    empty piece among them."""
    bw = np.unique(np.clip(np.concatenate([[0], cuts_w, [W]]), 0, W))
    bn = np.linspace(0, N, len(bw)).astype(np.int64)
    for i, c in enumerate(cuts_n[:len(bn) - 2]):
        bn[i + 1] = min(max(c, bn[i]), N)
    pieces = np.zeros(len(bw) + 1, O.PIECE)
    pieces[1:]["first_wire"], pieces[1:]["num_wires"] = bw, np.diff(np.append(bw, W))
    pieces[1:]["first_node"], pieces[1:]["num_nodes"] = bn, np.diff(np.append(bn, N))
    pieces[0] = (0, 0, 0, 0)
    return pieces


def test_accessor_index_form(C, O):
    """The index-space accessor entry points over uneven pieces (sizes 0, 1,
    255, 256, 257, ... elements, an empty piece first) give the per-piece
    results: the relaxation and update bit-identical to the oracle, the
    charge reduction within its float-atomic reordering."""
    c = O.generate(4, 256, 1024, 70, 33, O.PROFILE_STABLE)
    N, W = len(c.nodes), len(c.wires)
    pieces = _uneven_pieces(O, W, N, [1, 256, 512, 513, 1300, 4000], [1, 2, 258, 600, 700])
    assert pieces["num_wires"].sum() == W and pieces["num_nodes"].sum() == N
    assert pieces["num_wires"][0] == 0 and 1 in pieces["num_wires"] and 256 in pieces["num_wires"]
    dw = C.DeviceArray(c.wires)
    pvt, owned, ghost = (C.DeviceArray(c.nodes) for _ in range(3))
    ref = c.copy()
    C.calc_new_currents_index_gpu(pieces, dw, pvt, owned, ghost, flag=C.POLICY_EXACT)
    O.calc_new_currents(ref)
    got = dw.download()
    assert_bit_identical(got["current"], ref.wires["current"], "wire current")
    assert_bit_identical(got["voltage"], ref.wires["voltage"], "wire voltage")
    zeros = np.zeros(N, np.float32)
    ofold, gfold = C.DeviceArray(zeros), C.DeviceArray(zeros)
    C.distribute_charge_index_gpu(pieces, dw, pvt, ofold, gfold)
    want_pvt, want_o, want_g = ref.nodes.copy(), zeros.copy(), zeros.copy()
    O.distribute_charge_fold(ref.wires, want_pvt, want_o, want_g, 0, W)
    assert max_rel_err(pvt.download()["charge"], want_pvt["charge"], floor=1e-12) <= TOL
    assert max_rel_err(ofold.download(), want_o, floor=1e-12) <= TOL
    assert max_rel_err(gfold.download(), want_g, floor=1e-12) <= TOL
    pre = ref.nodes.copy()
    pre["charge"] = np.random.default_rng(1).uniform(-1e-3, 1e-3, N).astype(np.float32)
    pvt.upload(pre)
    owned.upload(pre)
    locator = C.DeviceArray(c.locator)
    want_p, want_s = pre.copy(), pre.copy()
    C.update_voltages_index_gpu(pieces, pvt, owned, locator)
    O.update_voltages(ref, None, pvt=want_p, shr=want_s)
    assert pvt.download().tobytes() == want_p.tobytes()
    assert owned.download().tobytes() == want_s.tobytes()
    # an empty table is a no-op; a bad location in any piece is reported
    C.calc_new_currents_index_gpu(pieces[:0], dw, pvt, owned, ghost)
    bad = got.copy()
    bad["in_loc"][W - 1] = 7
    dw.upload(bad)
    with pytest.raises(C.CircuitError):
        C.calc_new_currents_index_gpu(pieces, dw, pvt, owned, ghost)


@pytest.mark.parametrize("flag", [1, 2, 3], ids=["recip", "fast", "scaled"])
def test_accessor_form_fast_policies(C, O, flag):
    c = O.generate(4, 256, 1024, 70, 32, O.PROFILE_STABLE)
    dw = C.DeviceArray(c.wires)
    pvt, owned, ghost = (C.DeviceArray(c.nodes) for _ in range(3))
    ref = c.copy()
    for p in c.pieces:
        C.calc_new_currents_gpu(p, dw, pvt, owned, ghost, flag=flag)
        O.calc_new_currents(ref, p)
    got = dw.download()
    assert max_wire_rel_err(got["current"], ref.wires["current"]) <= TOL
    assert max_wire_rel_err(got["voltage"], ref.wires["voltage"]) <= TOL


def test_sanity_check_on_context(C):
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(**{"num_pieces": 8, "nodes_per_piece": 512,
                                    "wires_per_piece": 2048, "pct_wire_in_piece": 60}))
        assert ctx.sanity_check_wires() == 0


def test_load_validation(C):
    host = C.generate(C.GenParams(2, 32, 64, 50, 1))
    with C.Context(0) as ctx:
        bad = host.copy()
        bad.wires["out_ptr"][0] = 10**6
        with pytest.raises(C.CircuitError) as e:
            ctx.load(bad)
        assert e.value.status == 2
        bad = host.copy()
        bad.wires["in_loc"][3] = 5
        with pytest.raises(C.CircuitError):
            ctx.load(bad)
        bad = host.copy()
        # a private node of piece 1 reached from piece 0
        priv1 = np.flatnonzero(bad.locator[32:] == 1)[0] + 32
        bad.wires["out_ptr"][0] = priv1
        with pytest.raises(C.CircuitError):
            ctx.load(bad)
        ctx.load(host)
        assert ctx.sizes() == (64, 128, 2)


def test_context_calls_from_several_threads(C, O):
    # Calls on one context from several host threads run one at a time: four
    # threads each run 2 loops of EXACT + CSR (one call each, holding the
    # context for both loops) while a fifth keeps reading the piece table and
    # sizes; the circuit ends after 8 loops, bit-identical to the oracle's.
    import threading
    s = SMALL
    errors = []
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(s["P"], s["npp"], s["wpp"], s["pct"], s["seed"], C.PROFILE_STABLE))
        done = threading.Event()

        def loops():
            try:
                ctx.run_loops(2, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
            except Exception as e:  # noqa: BLE001 -- reported below
                errors.append(e)

        def reader():
            try:
                while not done.is_set():
                    assert len(ctx.pieces()) == s["P"]
            except Exception as e:  # noqa: BLE001
                errors.append(e)

        r = threading.Thread(target=reader)
        r.start()
        ts = [threading.Thread(target=loops) for _ in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        done.set()
        r.join()
        assert not errors, errors
        nodes, wires = ctx.store()
    ref = run_oracle(O, s["P"], s["npp"], s["wpp"], s["pct"], s["seed"], O.PROFILE_STABLE, 8)
    assert nodes.tobytes() == ref.nodes.tobytes()
    assert wires.tobytes() == ref.wires.tobytes()
