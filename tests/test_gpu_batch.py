"""Batching on the GPU: circuits of uneven size advanced by shared launches
give the same state as each circuit run alone (EXACT or SCALED with CSR:
bit-identical; ATOMIC: within 1e-5) and as the oracle's loop, up to the full
batched workload (512 circuits of 2^10..2^18 wires)."""
import numpy as np
import pytest

from helpers import (TOL, assert_bit_identical, assert_same_class, max_rel_err,
                     max_wire_rel_err)

pytestmark = pytest.mark.gpu

SHAPES = [  # (pieces, nodes/piece, wires/piece, pct, seed, loops)
    (1, 256, 1024, 95, 101, 3),
    (3, 1024, 4096, 95, 102, 1),
    (2, 700, 2900, 80, 103, 2),
    (5, 333, 1300, 95, 104, 3),
    (1, 64, 250, 50, 105, 2),
    (4, 1024, 4096, 50, 106, 1),
]


def alone(C, shape, policy, mode):
    P, npp, wpp, pct, seed, loops = shape
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE))
        ctx.run_loops(loops, policy=policy, mode=mode)
        return ctx.store()


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("policy", [0, 3], ids=["exact", "scaled"])
def test_batch_equals_each_circuit_alone(C, use_graph, persistent, policy):
    # With the deterministic CSR reduction a batched circuit is bit-identical
    # to the same circuit run alone under the same policy.
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), loops)
           for (P, npp, wpp, pct, seed, loops) in SHAPES]
    cfg = C.Batch.config(policy=policy, mode=C.REDUCE_CSR, use_graph=use_graph,
                         block_threads=128 if persistent else 256, persistent=persistent,
                         waves_per_simd=4)
    rep = b.run(cfg)
    assert rep["loops_run"] == 3
    assert rep["wire_updates"] == sum(P * wpp * loops for (P, _, wpp, _, _, loops) in SHAPES)
    for cid, shape in zip(ids, SHAPES):
        P, npp, wpp, *_ = shape
        assert b.loops_left(cid) == 0
        nodes, wires = b.store(cid, P * npp, P * wpp)
        n1, w1 = alone(C, shape, policy, C.REDUCE_CSR)
        assert nodes.tobytes() == n1.tobytes()
        assert wires.tobytes() == w1.tobytes()
    b.close()


@pytest.mark.parametrize("policy", [2, 3], ids=["fast", "scaled"])
def test_batch_fast_atomic_within_tolerance(C, policy):
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), loops)
           for (P, npp, wpp, pct, seed, loops) in SHAPES]
    b.run(C.Batch.config(policy=policy, mode=C.REDUCE_ATOMIC))
    for cid, shape in zip(ids, SHAPES):
        P, npp, wpp, *_ = shape
        nodes, _ = b.store(cid, P * npp, P * wpp)
        n1, _ = alone(C, shape, C.POLICY_EXACT, C.REDUCE_CSR)
        assert max_rel_err(nodes["voltage"], n1["voltage"]) <= TOL
    b.close()


def test_batch_step_remove_and_budget(C):
    b = C.Batch(0, budget_bytes=C.footprint_bytes(4096, 16384, 4) * 2 + 1)
    a = b.add_generated(C.GenParams(4, 1024, 4096, 95, 1, C.PROFILE_STABLE), 2)
    bb = b.add_generated(C.GenParams(4, 1024, 4096, 95, 2, C.PROFILE_STABLE), 1)
    with pytest.raises(C.CircuitError) as e:
        b.add_generated(C.GenParams(4, 1024, 4096, 95, 3, C.PROFILE_STABLE), 1)
    assert e.value.status == 4  # over budget
    cfg = C.Batch.config(policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
    assert b.step(cfg) == 1
    b.remove(bb)
    assert b.count()[0] == 1
    c = b.add_generated(C.GenParams(4, 1024, 4096, 95, 3, C.PROFILE_STABLE), 1)
    assert b.step(cfg) == 0
    na, _ = b.store(a, 4096, 16384)
    nc, _ = b.store(c, 4096, 16384)
    ra, _ = alone(C, (4, 1024, 4096, 95, 1, 2), C.POLICY_EXACT, C.REDUCE_CSR)
    rc, _ = alone(C, (4, 1024, 4096, 95, 3, 1), C.POLICY_EXACT, C.REDUCE_CSR)
    assert na.tobytes() == ra.tobytes() and nc.tobytes() == rc.tobytes()
    b.close()


def test_autotune_restores_state(C):
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), loops)
           for (P, npp, wpp, pct, seed, loops) in SHAPES]
    best, ms = b.autotune(C.POLICY_EXACT, probe_steps=50)
    assert best.block_threads in (64, 128, 256, 512) and ms > 0
    assert best.wires_per_lane in (1, 2, 4) and best.tile_span in (0, 1)
    assert best.tile_order in (0, 1, 2)
    assert b.probe(best, probe_steps=50) > 0
    best.policy, best.reduce_mode = C.POLICY_EXACT, C.REDUCE_CSR
    b.run(best)
    for cid, shape in zip(ids, SHAPES):
        P, npp, wpp, *_ = shape
        nodes, wires = b.store(cid, P * npp, P * wpp)
        n1, w1 = alone(C, shape, C.POLICY_EXACT, C.REDUCE_CSR)
        assert_bit_identical(nodes["voltage"], n1["voltage"], "node voltage")
        assert_bit_identical(wires["current"], w1["current"], "wire current")
    b.close()


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("sched", [(2, 0, 0), (4, 1, 1), (1, 1, 2), (2, 1, 2)],
                         ids=["2-per-lane", "4-per-lane-span-longest", "span-interleaved",
                              "2-per-lane-span-interleaved"])
def test_tile_schedules_bit_identical(C, O, sched, persistent):
    # calc_new_currents' tile schedule (wires per lane, tiles spanning
    # adjacent pieces, tile order) moves work between workgroups only: with
    # EXACT + CSR every circuit stays bit-identical to the oracle's loop.
    lane, span, order = sched
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), loops)
           for (P, npp, wpp, pct, seed, loops) in SHAPES]
    cfg = C.Batch.config(policy=C.POLICY_EXACT, mode=C.REDUCE_CSR, block_threads=128,
                         persistent=persistent, waves_per_simd=4, wires_per_lane=lane,
                         tile_span=span, tile_order=order)
    b.run(cfg)
    for cid, (P, npp, wpp, pct, seed, loops) in zip(ids, SHAPES):
        nodes, wires = b.store(cid, P * npp, P * wpp)
        ref = O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE)
        O.run(ref, loops)
        assert nodes.tobytes() == ref.nodes.tobytes()
        assert wires.tobytes() == ref.wires.tobytes()
    b.close()


@pytest.mark.parametrize("shape", [(8, 2048, 8192), (3, 5000, 20000), (600, 64, 256)],
                         ids=["8-pieces", "3-pieces-ragged-subranges", "600-pieces"])
def test_atomic_implementations_agree(C, shape):
    # Global atomics, the LDS-privatised reduction (its pieces split into
    # 4 and 9 wire subranges with partial sums in the first two shapes, one
    # workgroup per piece in the third) and the atomics fused into
    # calc_new_currents, against the deterministic CSR reduction.
    P, npp, wpp = shape
    outs = []
    for fuse, impl in ((0, 1), (0, 2), (1, 0)):
        with C.Context(0) as ctx:
            ctx.generate(C.GenParams(P, npp, wpp, 80, 9, C.PROFILE_STABLE))
            ctx.set_option(C.OPT_FUSE_DSC, fuse)
            ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, impl)
            ctx.run_loops(2, policy=C.POLICY_FAST, mode=C.REDUCE_ATOMIC)
            outs.append(ctx.store()[0])
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, 80, 9, C.PROFILE_STABLE))
        ctx.run_loops(2, policy=C.POLICY_FAST, mode=C.REDUCE_CSR)
        csr = ctx.store()[0]
    for o in outs:
        assert max_rel_err(o["voltage"], csr["voltage"]) <= TOL
        assert not o["charge"].any()


def _c4_shapes(C, profile):
    # The batched workload: 512 circuits, wires log-uniform in [2^10, 2^18],
    # nodes = wires/4, pieces = ceil(nodes/4096), 95% in-piece, seeds 1000+i.
    rng = np.random.default_rng(4)
    return [C.shape_for_wires(int(np.exp(rng.uniform(np.log(2**10), np.log(2**18)))), 95,
                              1000 + i, profile) for i in range(512)]


@pytest.mark.parametrize("profile", [0, 1], ids=["reference", "stable"])
def test_batched_workload_512_circuits(C, O, profile):
    # The full batched workload in one batch, 5 loops, with the benchmark's
    # schedule (autotuned) and configuration (SCALED, fused ATOMIC), and in
    # the stable profile also EXACT + CSR.  Every circuit runs its 5 loops;
    # the ten smallest are checked against the oracle's 5-loop run: EXACT +
    # CSR bit for bit, SCALED + ATOMIC by non-finite class on every node
    # voltage, wire current and internal voltage and within 1e-5 where finite.
    shapes = _c4_shapes(C, profile)
    sizes = [(s.num_pieces * s.nodes_per_piece, s.num_pieces * s.wires_per_piece) for s in shapes]
    smallest = sorted(range(len(shapes)), key=lambda i: sizes[i][1])[:10]
    refs = {}
    for i in smallest:
        g = shapes[i]
        ref = O.generate(g.num_pieces, g.nodes_per_piece, g.wires_per_piece,
                         g.pct_wire_in_piece, g.random_seed, g.profile)
        O.run(ref, 5)
        refs[i] = ref
    runs = [(C.POLICY_SCALED, C.REDUCE_ATOMIC)]
    if profile == 1:
        runs.append((C.POLICY_EXACT, C.REDUCE_CSR))
    for policy, mode in runs:
        b = C.Batch(0)
        ids = b.add_generated_many(shapes, [5] * len(shapes), threads=8)
        best, _ = b.autotune(policy, probe_steps=200)
        best.policy, best.reduce_mode, best.use_graph = policy, mode, 1
        rep = b.run(best)
        assert rep["loops_run"] == 5
        assert rep["wire_updates"] == 5 * sum(w for _, w in sizes)
        assert all(b.loops_left(cid) == 0 for cid in ids)
        for i in smallest:
            N, W = sizes[i]
            nodes, wires = b.store(ids[i], N, W)
            ref = refs[i]
            if policy == C.POLICY_EXACT:
                assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], f"{i} node voltage")
                assert_bit_identical(wires["current"], ref.wires["current"], f"{i} wire current")
                continue
            assert_same_class(nodes["voltage"], ref.nodes["voltage"], f"{i} node voltage")
            assert_same_class(wires["current"], ref.wires["current"], f"{i} wire current")
            assert_same_class(wires["voltage"], ref.wires["voltage"], f"{i} wire voltage")
            assert max_rel_err(nodes["voltage"], ref.nodes["voltage"]) <= TOL
            fin = np.isfinite(ref.wires["current"]).all(axis=1)
            assert max_wire_rel_err(wires["current"][fin], ref.wires["current"][fin]) <= TOL
        b.close()


def test_batch_small_circuits_equal_oracle(C, O):
    # EXACT + CSR in a batch against the oracle's loop on the same circuits
    # (one-piece circuits with the wrap rule among them).
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), loops)
           for (P, npp, wpp, pct, seed, loops) in SHAPES]
    b.run(C.Batch.config(policy=C.POLICY_EXACT, mode=C.REDUCE_CSR))
    for cid, (P, npp, wpp, pct, seed, loops) in zip(ids, SHAPES):
        nodes, wires = b.store(cid, P * npp, P * wpp)
        ref = O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE)
        O.run(ref, loops)
        assert nodes.tobytes() == ref.nodes.tobytes()
        assert wires.tobytes() == ref.wires.tobytes()
    b.close()


def test_batch_steps_from_several_threads(C, O):
    # Two host threads each take two batch steps at once; the batch's mutex
    # runs the four steps one at a time, so every circuit runs 4 loops,
    # bit-identical (EXACT + CSR) to the oracle's 4-loop run.
    import threading
    shapes = [(P, npp, wpp, pct, seed) for (P, npp, wpp, pct, seed, _) in SHAPES[:4]]
    b = C.Batch(0)
    ids = [b.add_generated(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE), 4)
           for (P, npp, wpp, pct, seed) in shapes]
    cfg = C.Batch.config(policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
    errors = []

    def steps():
        try:
            for _ in range(2):
                b.step(cfg)
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)

    ts = [threading.Thread(target=steps) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for cid, (P, npp, wpp, pct, seed) in zip(ids, shapes):
        assert b.loops_left(cid) == 0
        nodes, wires = b.store(cid, P * npp, P * wpp)
        ref = O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE)
        O.run(ref, 4)
        assert nodes.tobytes() == ref.nodes.tobytes()
        assert wires.tobytes() == ref.wires.tobytes()
    b.close()
