"""Serving on the GPU: requests arriving while others run are admitted as
budget and slots free up, share iterations, and each result equals the same
circuit run alone (CSR: bit-identical under the same policy; ATOMIC: within
1e-5 of EXACT + CSR)."""
import numpy as np
import pytest

from helpers import TOL, max_rel_err

pytestmark = pytest.mark.gpu


def alone(C, gen, loops, policy, mode):
    with C.Context(0) as ctx:
        ctx.generate(gen)
        ctx.run_loops(loops, policy=policy, mode=mode)
        return ctx.store()


def trace(C):
    rng = np.random.default_rng(7)
    reqs, t = [], 0.0
    for i in range(10):
        t += rng.exponential(0.01)
        W = int(np.exp(rng.uniform(np.log(2**10), np.log(2**15))))
        reqs.append(C.Request(C.shape_for_wires(W, 95, 500 + i, C.PROFILE_STABLE),
                              int(rng.integers(1, 5)), t))
    return reqs


@pytest.mark.parametrize("policy,mode", [(0, 1), (2, 0), (3, 1), (3, 0)],
                         ids=["exact-csr", "fast-atomic", "scaled-csr", "scaled-atomic"])
def test_served_results_equal_alone(C, policy, mode):
    reqs = trace(C)
    # a small budget forces queueing behind resident circuits
    budget = 3 * max(C.footprint_bytes(r.gen.num_pieces * r.gen.nodes_per_piece,
                                       r.gen.num_pieces * r.gen.wires_per_piece,
                                       r.gen.num_pieces) for r in reqs)
    s = C.Server(0, budget_bytes=budget, max_active=4, policy=policy, mode=mode, gen_threads=3)
    s.set_keep_state(True)
    results, rep = s.run_trace(reqs)
    assert rep["completed"] == len(reqs) and rep["rejected"] == 0
    assert rep["max_resident"] <= 4
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0
    for i, (r, res) in enumerate(zip(reqs, results)):
        assert res["status"] == 0
        assert res["arrival_s"] <= res["admit_s"] <= res["finish_s"]
        g = r.gen
        N, W = g.num_pieces * g.nodes_per_piece, g.num_pieces * g.wires_per_piece
        nodes, wires = s.fetch(i, N, W)
        if mode == 1:
            # CSR: bit-identical to the same circuit alone under the same policy
            n1, w1 = alone(C, g, r.num_loops, policy, mode)
            assert nodes.tobytes() == n1.tobytes()
            assert wires.tobytes() == w1.tobytes()
        else:
            n1, w1 = alone(C, g, r.num_loops, 0, 1)
            assert max_rel_err(nodes["voltage"], n1["voltage"]) <= TOL
        assert np.isclose(res["checksum"], nodes["voltage"].astype(np.float64).sum())
    s.close()


def test_served_checksums_with_zero_loop_requests(C):
    # Without kept state, the path the benchmark serves: finishers are summed
    # from the pinned buffer of their last iteration while the next one runs;
    # zero-loop requests finish without running, on their generated state.
    rng = np.random.default_rng(11)
    reqs, t = [], 0.0
    for i in range(12):
        t += rng.exponential(0.005)
        W = int(np.exp(rng.uniform(np.log(2**10), np.log(2**14))))
        loops = 0 if i % 4 == 1 else int(rng.integers(1, 6))
        reqs.append(C.Request(C.shape_for_wires(W, 95, 700 + i, C.PROFILE_STABLE), loops, t))
    s = C.Server(0, max_active=5, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR, gen_threads=3)
    results, rep = s.run_trace(reqs)
    assert rep["completed"] == len(reqs) and rep["rejected"] == 0 and rep["iterations"] > 0
    for r, res in zip(reqs, results):
        assert res["status"] == 0
        n1, _ = alone(C, r.gen, r.num_loops, C.POLICY_EXACT, C.REDUCE_CSR)
        v = n1["voltage"].astype(np.float64)
        assert abs(res["checksum"] - v.sum()) <= 1e-12 * np.abs(v).sum(), (r.num_loops, res)
    s.close()


def test_server_sheds_requests_past_their_admission_deadline(C):
    # One slot: the first request occupies it for 20 loops; the others, all
    # arriving at t=0 with a 1 ms admission deadline, time out in the queue.
    s = C.Server(0, max_active=1, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
    reqs = [C.Request(C.GenParams(4, 1024, 4096, 95, 10 + i, C.PROFILE_STABLE), 20, 0.0,
                      0.001 if i else 0.0) for i in range(4)]
    results, rep = s.run_trace(reqs)
    assert results[0]["status"] == 0
    assert [r["status"] for r in results[1:]] == [11, 11, 11]  # CKT_ERR_TIMEOUT
    assert rep["timed_out"] == 3 and rep["completed"] == 1
    s.close()


def test_server_rejects_oversized_and_invalid(C):
    s = C.Server(0, budget_bytes=C.footprint_bytes(4096, 16384, 4), policy=C.POLICY_FAST)
    reqs = [C.Request(C.GenParams(4, 1024, 4096, 95, 1, C.PROFILE_STABLE), 1, 0.0),
            C.Request(C.GenParams(16, 1024, 4096, 95, 2, C.PROFILE_STABLE), 1, 0.0),
            C.Request(C.GenParams(0, 1024, 4096, 95, 3, C.PROFILE_STABLE), 1, 0.0)]
    results, rep = s.run_trace(reqs)
    assert results[0]["status"] == 0
    assert results[1]["status"] == 4  # over budget
    assert results[2]["status"] == 1  # invalid generator arguments
    assert rep["completed"] == 1 and rep["rejected"] == 2
    s.close()
