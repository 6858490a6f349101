"""Batching on the GPU: circuits of uneven size advanced by shared launches
give the same state as each circuit run alone (EXACT or SCALED with CSR:
bit-identical; ATOMIC: within 1e-5)."""
import pytest

from helpers import TOL, assert_bit_identical, max_rel_err

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
    best.policy, best.reduce_mode = C.POLICY_EXACT, C.REDUCE_CSR
    b.run(best)
    for cid, shape in zip(ids, SHAPES):
        P, npp, wpp, *_ = shape
        nodes, wires = b.store(cid, P * npp, P * wpp)
        n1, w1 = alone(C, shape, C.POLICY_EXACT, C.REDUCE_CSR)
        assert_bit_identical(nodes["voltage"], n1["voltage"], "node voltage")
        assert_bit_identical(wires["current"], w1["current"], "wire current")
    b.close()


def test_atomic_implementations_agree(C):
    outs = []
    for fuse, impl in ((0, 1), (0, 2), (1, 0)):
        with C.Context(0) as ctx:
            ctx.generate(C.GenParams(8, 2048, 8192, 80, 9, C.PROFILE_STABLE))
            ctx.set_option(C.OPT_FUSE_DSC, fuse)
            ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, impl)
            ctx.run_loops(2, policy=C.POLICY_FAST, mode=C.REDUCE_ATOMIC)
            outs.append(ctx.store()[0])
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(8, 2048, 8192, 80, 9, C.PROFILE_STABLE))
        ctx.run_loops(2, policy=C.POLICY_FAST, mode=C.REDUCE_CSR)
        csr = ctx.store()[0]
    for o in outs:
        assert max_rel_err(o["voltage"], csr["voltage"]) <= TOL
