"""Full-size (64 pieces x 16384 nodes x 65536 wires, ~4.2M wires) checks
through size-independent properties.

  * the generator is bit-identical to the oracle at this size;
  * one EXACT loop: a random sample of wires matches the oracle's relaxation
    bit for bit (the relaxation is per wire, so a sample is a full check of
    those wires);
  * node voltages equal the float64 scatter-add of the GPU's own currents
    followed by the update (checks distribute_charge + update_voltages at
    full size, any reduction order);
  * CSR and ATOMIC reductions agree to 1e-5, and the reference profile gives
    the oracle's NaN pattern on the sample;
  * the benchmark's configuration (FAST, fused ATOMIC) within 1e-5 of the
    oracle on a sample of wires and on every node voltage, and with the
    oracle's non-finite classes under the reference parameters;
  * all 10 loops of the workload, loop by loop, for every policy in both
    profiles: each GPU loop against the oracle's task bodies run from the
    GPU's state before it, on a sample of wires and on every node.
"""
import numpy as np
import pytest

from helpers import (TOL, assert_bit_identical, assert_same_class, max_rel_err,
                     max_wire_rel_err)

pytestmark = pytest.mark.gpu

FULL = (64, 16384, 65536, 95, 12345)  # 4,194,304 wires, 1,048,576 nodes


@pytest.mark.parametrize("profile", [1, 0], ids=["stable", "reference"])
def test_full_one_loop_properties(C, O, profile):
    P, npp, wpp, pct, seed = FULL
    host = C.generate(C.GenParams(P, npp, wpp, pct, seed, profile))
    ref = O.generate(P, npp, wpp, pct, seed, profile)
    assert host.wires.tobytes() == ref.wires.tobytes()
    assert host.nodes.tobytes() == ref.nodes.tobytes()
    del host
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, pct, seed, profile))
        ctx.run_loops(1, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
        nodes, wires = ctx.store()
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(len(ref.wires), 1500, replace=False))
    for k in sample:
        O.lib().orc_calc_new_currents_f32(0, O.STEPS, O._p(ref.wires), O._p(ref.nodes),
                                          O._p(ref.nodes), O._p(ref.nodes), int(k), 1)
    assert_bit_identical(wires["current"][sample], ref.wires["current"][sample], "sampled currents")
    assert_bit_identical(wires["voltage"][sample], ref.wires["voltage"][sample], "sampled voltages")
    # distribute_charge + update_voltages from the GPU's currents, in float64
    dt = np.float32(1e-6)
    q = np.zeros(len(nodes))
    np.add.at(q, wires["in_ptr"], (-dt * wires["current"][:, 0]).astype(np.float64))
    np.add.at(q, wires["out_ptr"], (dt * wires["current"][:, 9]).astype(np.float64))
    v0 = ref.nodes["voltage"].astype(np.float64)
    cap = ref.nodes["capacitance"].astype(np.float64)
    leak = ref.nodes["leakage"].astype(np.float64)
    want = (v0 + q / cap) * (1.0 - leak)
    if profile == 1:
        assert np.isfinite(nodes["voltage"]).all()
        assert max_rel_err(nodes["voltage"], want) <= TOL
        assert not nodes["charge"].any()
    else:
        assert_same_class(nodes["voltage"], want.astype(np.float32), "node voltage class")


def test_full_atomic_and_csr_agree(C):
    P, npp, wpp, pct, seed = FULL
    out = {}
    for mode in (C.REDUCE_ATOMIC, C.REDUCE_CSR):
        with C.Context(0) as ctx:
            ctx.generate(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE))
            ctx.set_option(C.OPT_STEPS, 200)
            ctx.run_loops(2, policy=C.POLICY_FAST, mode=mode)
            out[mode] = ctx.store()[0]
    assert max_rel_err(out[C.REDUCE_ATOMIC]["voltage"], out[C.REDUCE_CSR]["voltage"]) <= TOL


@pytest.mark.parametrize("policy", [1, 2, 3], ids=["recip", "fast", "scaled"])
def test_full_fast_sampled_wires_within_tolerance(C, O, policy):
    # The benchmark's configuration (FAST, fused ATOMIC reduction) at full
    # size, stable profile: one loop, a random sample of wires against the
    # oracle's reference-order relaxation (per wire, relative to the wire's
    # largest segment value), and every node voltage against the float64
    # scatter-add of the GPU's own currents.
    P, npp, wpp, pct, seed = FULL
    ref = O.generate(P, npp, wpp, pct, seed, C.PROFILE_STABLE)
    with C.Context(0) as ctx:
        ctx.generate_device(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE))
        ctx.run_loops(1, policy=policy, mode=C.REDUCE_ATOMIC)
        nodes, wires = ctx.store()
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(len(ref.wires), 1500, replace=False))
    for k in sample:
        O.lib().orc_calc_new_currents_f32(0, O.STEPS, O._p(ref.wires), O._p(ref.nodes),
                                          O._p(ref.nodes), O._p(ref.nodes), int(k), 1)
    assert np.isfinite(wires["current"][sample]).all()
    assert max_wire_rel_err(wires["current"][sample], ref.wires["current"][sample]) <= TOL
    assert max_wire_rel_err(wires["voltage"][sample], ref.wires["voltage"][sample]) <= TOL
    dt = np.float32(1e-6)
    q = np.zeros(len(nodes))
    np.add.at(q, wires["in_ptr"], (-dt * wires["current"][:, 0]).astype(np.float64))
    np.add.at(q, wires["out_ptr"], (dt * wires["current"][:, 9]).astype(np.float64))
    want = (ref.nodes["voltage"].astype(np.float64) + q / ref.nodes["capacitance"]) * \
        (1.0 - ref.nodes["leakage"].astype(np.float64))
    assert np.isfinite(nodes["voltage"]).all()
    assert max_rel_err(nodes["voltage"], want) <= TOL


@pytest.mark.parametrize("policy", [1, 2, 3], ids=["recip", "fast", "scaled"])
def test_full_fast_reference_profile_classes_after_one_loop(C, O, policy):
    # The benchmark's configuration and profile (reference parameters, FAST,
    # fused ATOMIC), one loop: every sampled wire's currents and internal
    # voltages have the oracle's non-finite classes, and every node voltage
    # has the class of EXACT + CSR's (bit-identical to the oracle at this size).
    P, npp, wpp, pct, seed = FULL
    gp = C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_REFERENCE)
    got = {}
    for name, pol, mode in (("exact+csr", C.POLICY_EXACT, C.REDUCE_CSR),
                            ("fast+atomic", policy, C.REDUCE_ATOMIC)):
        with C.Context(0) as ctx:
            ctx.generate_device(gp)
            ctx.run_loops(1, policy=pol, mode=mode)
            got[name] = ctx.store()
    ref = O.generate(P, npp, wpp, pct, seed, O.PROFILE_REFERENCE)
    rng = np.random.default_rng(2)
    sample = np.sort(rng.choice(len(ref.wires), 1500, replace=False))
    for k in sample:
        O.lib().orc_calc_new_currents_f32(0, O.STEPS, O._p(ref.wires), O._p(ref.nodes),
                                          O._p(ref.nodes), O._p(ref.nodes), int(k), 1)
    wires = got["fast+atomic"][1]
    assert_same_class(wires["current"][sample], ref.wires["current"][sample], "sampled currents")
    assert_same_class(wires["voltage"][sample], ref.wires["voltage"][sample], "sampled voltages")
    assert_same_class(got["fast+atomic"][0]["voltage"], got["exact+csr"][0]["voltage"],
                      "node voltage")


def test_full_ten_loops_against_exact_and_fp64(C, O):
    # The full-size circuit for 10 loops, stable profile.  FAST and SCALED
    # with the fused ATOMIC reduction (the benchmark's configurations) against
    # EXACT + CSR -- bit-identical to the CPU restatement on the 4-piece
    # circuits, on the sampled wires above and over this whole circuit for 10
    # loops (tests/full_size_exact_vs_oracle.py, ~1.7 core-hours of CPU), so it
    # stands in for the CPU run -- on every node voltage; all against the FP64
    # oracle (the relaxation's FP64 limit), printed with -rP.
    P, npp, wpp, pct, seed = FULL
    gp = C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE)
    out = {}
    for name, policy, mode in (("exact+csr", C.POLICY_EXACT, C.REDUCE_CSR),
                               ("fast+atomic", C.POLICY_FAST, C.REDUCE_ATOMIC),
                               ("scaled+atomic", C.POLICY_SCALED, C.REDUCE_ATOMIC)):
        with C.Context(0) as ctx:
            ctx.generate_device(gp)
            ctx.run_loops(10, policy=policy, mode=mode, use_graph=True)
            out[name] = ctx.store()[0]["voltage"]
        assert np.isfinite(out[name]).all()
    truth, _ = O.run_f64_limit(O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE), 10)
    tv = truth["voltage"]
    errs = {k: max_rel_err(v, tv) for k, v in out.items()}
    diff = {k: max_rel_err(out[k], out["exact+csr"]) for k in ("fast+atomic", "scaled+atomic")}
    print(f"\nC2, 10 loops, stable profile: max relative node-voltage error vs FP64 "
          f"exact+csr {errs['exact+csr']:.3e}, fast+atomic {errs['fast+atomic']:.3e}, "
          f"scaled+atomic {errs['scaled+atomic']:.3e}; vs exact+csr: fast+atomic "
          f"{diff['fast+atomic']:.3e}, scaled+atomic {diff['scaled+atomic']:.3e}")
    assert max(diff.values()) <= TOL
    assert max(errs.values()) <= 1e-3


@pytest.mark.parametrize("profile", [1, 0], ids=["stable", "reference"])
def test_full_scaled_matches_exact_other_seed_pct50(C, profile):
    # Another seed and half the wires crossing pieces (86% shared nodes):
    # SCALED + fused ATOMIC against EXACT + CSR (bit-identical to the oracle)
    # after 3 loops, every node and wire, by non-finite class and to 1e-5.
    gp = C.GenParams(64, 16384, 65536, 50, 7, profile)
    out = {}
    for name, policy, mode in (("exact", C.POLICY_EXACT, C.REDUCE_CSR),
                               ("scaled", C.POLICY_SCALED, C.REDUCE_ATOMIC)):
        with C.Context(0) as ctx:
            ctx.generate_device(gp)
            ctx.run_loops(3, policy=policy, mode=mode)
            out[name] = ctx.store()
    (en, ew), (sn, sw) = out["exact"], out["scaled"]
    assert_same_class(sn["voltage"], en["voltage"], "node voltage")
    assert_same_class(sw["current"], ew["current"], "wire current")
    assert_same_class(sw["voltage"], ew["voltage"], "wire voltage")
    assert max_rel_err(sn["voltage"], en["voltage"]) <= TOL
    assert max_wire_rel_err(sw["current"], ew["current"]) <= TOL
    assert max_wire_rel_err(sw["voltage"], ew["voltage"]) <= TOL


def _node_update_f64(before_nodes, wires):
    """distribute_charge + update_voltages (circuit.cc:415-460) in float64
    from the wires' currents onto the nodes' state before the loop."""
    dt = np.float32(1e-6)
    N = len(before_nodes)
    q = np.bincount(wires["in_ptr"], (-dt * wires["current"][:, 0]).astype(np.float64), N)
    q += np.bincount(wires["out_ptr"], (dt * wires["current"][:, 9]).astype(np.float64), N)
    q += before_nodes["charge"].astype(np.float64)
    v0 = before_nodes["voltage"].astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        return ((v0 + q / before_nodes["capacitance"]) *
                (1.0 - before_nodes["leakage"].astype(np.float64)))


@pytest.mark.parametrize("policy", [0, 1, 2, 3], ids=["exact+csr", "recip", "fast", "scaled"])
@pytest.mark.parametrize("profile", [1, 0], ids=["stable", "reference"])
def test_full_ten_loops_loop_by_loop_against_oracle(C, O, profile, policy):
    # The full-size workload for its 10 loops.  Before each GPU loop the
    # state is read back; the oracle's task bodies then run that loop from
    # it: calc_new_currents (circuit.cc:352-413) on 1500 sampled wires, and
    # distribute_charge + update_voltages (circuit.cc:415-460) over every
    # node from the GPU's new currents.  EXACT with the CSR reduction is
    # compared bit for bit (wires and every node); RECIP, FAST and SCALED with
    # the fused ATOMIC reduction by non-finite class on every compared value
    # and within 1e-5 where finite (the node update against a float64 sum of
    # the contributions).
    P, npp, wpp, pct, seed = FULL
    gp = C.GenParams(P, npp, wpp, pct, seed, profile)
    mode = C.REDUCE_CSR if policy == C.POLICY_EXACT else C.REDUCE_ATOMIC
    rng = np.random.default_rng(10 + policy)
    threads = 16
    worst = {"wire": 0.0, "node": 0.0}
    with C.Context(0) as ctx:
        ctx.generate_device(gp)
        before_n, before_w = ctx.store()
        sample = np.sort(rng.choice(len(before_w), 1500, replace=False)).astype(np.uint32)
        pieces = np.zeros(P, O.PIECE)
        pieces["first_wire"] = np.arange(P) * wpp
        pieces["num_wires"] = wpp
        pieces["first_node"] = np.arange(P) * npp
        pieces["num_nodes"] = npp
        private = np.ones(P * npp, np.int32)  # one node array serves every instance
        for loop in range(10):
            ctx.run_loops(1, policy=policy, mode=mode)
            after_n, after_w = ctx.store()
            # calc_new_currents on the sample, from the state before the loop
            w = before_w.copy()
            O.lib().orc_cnc_sample_f32(0, O.STEPS, threads, O._p(w), O._p(before_n),
                                       O._p(sample), len(sample))
            for f in ("current", "voltage"):
                if policy == C.POLICY_EXACT:
                    assert_bit_identical(after_w[f][sample], w[f][sample], f"loop {loop} {f}")
                    continue
                assert_same_class(after_w[f][sample], w[f][sample], f"loop {loop} {f}")
                fin = np.isfinite(w[f][sample]).all(axis=1)
                e = max_wire_rel_err(after_w[f][sample][fin], w[f][sample][fin])
                worst["wire"] = max(worst["wire"], e)
                assert e <= TOL, (loop, f, e)
            # distribute_charge + update_voltages over every node
            if policy == C.POLICY_EXACT:
                ref = O.Circuit(before_n.copy(), after_w, private, pieces)
                for p in pieces:
                    O.distribute_charge(ref, p)
                for p in pieces:
                    O.update_voltages(ref, p)
                assert_bit_identical(after_n["voltage"], ref.nodes["voltage"], f"loop {loop} nodes")
            else:
                want = _node_update_f64(before_n, after_w)
                assert_same_class(after_n["voltage"], want.astype(np.float32),
                                  f"loop {loop} node voltage class")
                e = max_rel_err(after_n["voltage"], want)
                worst["node"] = max(worst["node"], e)
                assert e <= TOL, (loop, e)
            assert not after_n["charge"].any()
            before_n, before_w = after_n, after_w
    print(f"\n10 loops, profile {profile}, policy {policy}: worst sampled-wire error "
          f"{worst['wire']:.2e}, worst node-voltage error {worst['node']:.2e}; "
          f"NaN node voltages at the end: {np.isnan(before_n['voltage']).mean():.4f}")
