"""Serving on the GPU: requests arriving while others run are admitted as
budget and slots free up, share iterations, and each result equals the same
circuit run alone (CSR: bit-identical under the same policy; ATOMIC: within
1e-5 of EXACT + CSR) and the oracle's loop.  Requests come from replayed
traces (the served workload's size mix among them) and from the online API
(submit, poll, wait, cancel)."""
import sys
import time

import numpy as np
import pytest

from helpers import TOL, assert_same_class, max_rel_err, max_wire_rel_err

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
def test_served_results_equal_alone(C, O, policy, mode):
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
            if policy == 0 and i < 2:
                # EXACT + CSR: bit-identical to the oracle's loop
                ref = O.generate(g.num_pieces, g.nodes_per_piece, g.wires_per_piece,
                                 g.pct_wire_in_piece, g.random_seed, g.profile)
                O.run(ref, r.num_loops)
                assert nodes.tobytes() == ref.nodes.tobytes()
                assert wires.tobytes() == ref.wires.tobytes()
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CANCELLED, TIMEOUT, STATE, BUDGET = 10, 11, 7, 4


def test_online_submit_wait_and_cancel(C, O):
    # One slot.  Request a (200 loops) holds it; b and c queue behind it.
    # b is cancelled while queued, a while running; c then runs and matches
    # the same circuit alone and the oracle; d shows wait's time-out.
    big = C.GenParams(4, 1024, 4096, 95, 900, C.PROFILE_STABLE)
    small = [C.GenParams(1, 128, 512, 95, 901 + i, C.PROFILE_STABLE) for i in range(3)]
    s = C.Server(0, max_active=1, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR, gen_threads=2)
    s.set_keep_state(True)
    s.start()
    a = s.submit(big, 200)
    b = s.submit(small[0], 2)
    c = s.submit(small[1], 3)
    assert (a, b, c) == (0, 1, 2)
    t_end = time.time() + 30
    while s.poll(a)[1]["admit_s"] <= 0 and time.time() < t_end:
        time.sleep(0.001)
    assert s.poll(a)[1]["admit_s"] > 0 and not s.poll(a)[0], "request a never started"
    assert not s.poll(b)[0]
    s.cancel(b)
    assert s.wait(b, 30)["status"] == CANCELLED
    s.cancel(a)
    ra = s.wait(a, 30)
    assert ra["status"] == CANCELLED and ra["finish_s"] >= ra["admit_s"] > 0
    rc = s.wait(c, 60)
    assert rc["status"] == 0 and rc["admit_s"] >= ra["finish_s"]
    nodes, wires = s.fetch(c, 128, 512)
    n1, w1 = alone(C, small[1], 3, C.POLICY_EXACT, C.REDUCE_CSR)
    assert nodes.tobytes() == n1.tobytes() and wires.tobytes() == w1.tobytes()
    ref = O.generate(1, 128, 512, 95, small[1].random_seed, O.PROFILE_STABLE)
    O.run(ref, 3)
    assert nodes.tobytes() == ref.nodes.tobytes()
    # wait with a limit on a request that outlasts it, then cancel it
    d = s.submit(big, 500)
    with pytest.raises(C.CircuitError) as e:
        s.wait(d, 0.05)
    assert e.value.status == TIMEOUT
    s.cancel(d)
    assert s.wait(d, 30)["status"] == CANCELLED
    # errors: cancelling a finished request, an unknown id, an oversized request
    with pytest.raises(C.CircuitError) as e:
        s.cancel(c)
    assert e.value.status == STATE
    with pytest.raises(C.CircuitError):
        s.poll(999)
    with pytest.raises(C.CircuitError) as e:
        s.submit(C.GenParams(0, 1, 1, 95, 1), 1)
    s.stop(drain=True)
    s.close()


def test_online_stop_without_drain_and_concurrent_submitters(C):
    # Requests from four host threads at once; stop(drain=False) then ends
    # the run: every request has ended, each either finished (checksum of the
    # circuit alone) or cancelled.
    import threading
    s = C.Server(0, max_active=3, policy=C.POLICY_SCALED, mode=C.REDUCE_CSR, gen_threads=3)
    s.start()
    ids, gens, lock = [], {}, threading.Lock()

    def client(k):
        for j in range(4):
            g = C.GenParams(2, 256, 1024, 95, 1000 + 10 * k + j, C.PROFILE_STABLE)
            rid = s.submit(g, 5 + 40 * (j % 2))
            with lock:
                ids.append(rid)
                gens[rid] = (g, 5 + 40 * (j % 2))

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert sorted(ids) == list(range(16))
    s.wait(ids[0], 60)
    s.stop(drain=False)
    statuses = {}
    for rid in ids:
        done, res = s.poll(rid)
        assert done
        statuses[rid] = res["status"]
        assert res["status"] in (0, CANCELLED)
        if res["status"] == 0:
            g, loops = gens[rid]
            n1, _ = alone(C, g, loops, C.POLICY_SCALED, C.REDUCE_CSR)
            v = n1["voltage"].astype(np.float64)
            assert abs(res["checksum"] - v.sum()) <= 1e-12 * np.abs(v).sum()
    assert list(statuses.values()).count(0) >= 1
    with pytest.raises(C.CircuitError):
        s.submit(C.GenParams(2, 256, 1024, 95, 5, C.PROFILE_STABLE), 1)
    s.close()


def test_served_poisson_workload_mix(C, O):
    # The served workload's size mix (wires log-uniform in [2^12, 2^20],
    # loops uniform in [1, 20], 95% in-piece) at 200 requests/s for 2.5 s,
    # stable profile, SCALED + CSR (deterministic, so a request's result in
    # the mix equals the same circuit alone bit for bit).  Checked: every
    # request completes; the kept requests (up to 8192 wires) and every 10th
    # other request against the same circuit alone; the five smallest against
    # the oracle's loop (non-finite class and 1e-5 on every node voltage and
    # wire current).
    trace = C.poisson_trace(200.0, 2.5, seed=8, profile=C.PROFILE_STABLE)
    s = C.Server(0, budget_bytes=200 * 10**9, max_active_wires=8 << 20,
                 policy=C.POLICY_SCALED, mode=C.REDUCE_CSR, gen_threads=12)
    s.set_keep_state(True, max_wires=8192)
    results, rep = s.run_trace(trace)
    say(f"[served mix] {len(trace)} requests: {rep['wire_updates_per_s']:.3g} wire-updates/s, "
        f"p50 {rep['latency_p50_s']:.3f} s, p99 {rep['latency_p99_s']:.3f} s, "
        f"max resident {rep['max_resident']}")
    assert rep["completed"] == len(trace) and rep["rejected"] == 0
    assert rep["max_resident"] > 1
    sizes = [r.gen.num_pieces * r.gen.wires_per_piece for r in trace]
    kept = [i for i, w in enumerate(sizes) if w <= 8192]
    assert len(kept) >= 5
    for i in sorted(set(kept) | set(range(0, len(trace), 10))):
        r, res = trace[i], results[i]
        assert res["status"] == 0
        n1, w1 = alone(C, r.gen, r.num_loops, C.POLICY_SCALED, C.REDUCE_CSR)
        v = n1["voltage"].astype(np.float64)
        assert abs(res["checksum"] - v.sum()) <= 1e-12 * np.abs(v).sum(), (i, res)
        if i in kept:
            g = r.gen
            nodes, wires = s.fetch(i, g.num_pieces * g.nodes_per_piece,
                                   g.num_pieces * g.wires_per_piece)
            assert nodes.tobytes() == n1.tobytes() and wires.tobytes() == w1.tobytes(), i
    for i in sorted(kept, key=lambda k: sizes[k])[:5]:
        g, loops = trace[i].gen, trace[i].num_loops
        ref = O.generate(g.num_pieces, g.nodes_per_piece, g.wires_per_piece,
                         g.pct_wire_in_piece, g.random_seed, g.profile)
        O.run(ref, loops)
        nodes, wires = s.fetch(i, len(ref.nodes), len(ref.wires))
        assert_same_class(nodes["voltage"], ref.nodes["voltage"], f"request {i} node voltage")
        assert max_rel_err(nodes["voltage"], ref.nodes["voltage"]) <= TOL, i
        assert max_wire_rel_err(wires["current"], ref.wires["current"]) <= TOL, i
    s.close()
