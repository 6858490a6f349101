"""Device-side generator (load_circuit, circuit.cc:1071-1288, built in HBM)
against the oracle's generator and the library's host generator.

ckt_ctx_generate_device must load the same circuit bit for bit as
ckt_ctx_generate: node and wire fields, locations, the locator and the
node -> (wire, end) index.  Cases: the 4-piece circuit (4 x 1024 nodes x
4096 wires) in both profiles, pct 0 / 50 / 100, one piece (the wrap rule), no
wires, and a circuit of 12.6 M wires whose wire start states cross several
host -> device chunks (2^22 wires each).  A device-generated 4-piece circuit
then runs 5 loops of EXACT + CSR bit-identical to the oracle.
"""
import sys
import time

import numpy as np
import pytest

from helpers import assert_bit_identical

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected_index(wires, num_nodes):
    """Node -> entry index in (wire, in-before-out) order (the definition of
    the CSR reduction order): entry 2w+e, keyed by end e's node."""
    keys = np.stack([wires["in_ptr"], wires["out_ptr"]], axis=1).reshape(-1).astype(np.int64)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    row = np.zeros(num_nodes + 1, np.uint32)
    row[1:] = np.cumsum(np.bincount(keys, minlength=num_nodes)).astype(np.uint32)
    return row, order


def load_both(C, gp):
    with C.Context(0) as a, C.Context(0) as b:
        t0 = time.perf_counter()
        a.generate(gp)
        t1 = time.perf_counter()
        b.generate_device(gp)
        t2 = time.perf_counter()
        say(f"  host generate + upload {t1 - t0:.3f} s, device generate {t2 - t1:.3f} s")
        return (a.store(), a.store_topology(), a.pieces()), (b.store(), b.store_topology(),
                                                             b.pieces())


def same_bytes(a, b, chunk=1 << 26):
    x, y = a.reshape(-1).view(np.uint8), b.reshape(-1).view(np.uint8)
    if x.size != y.size:
        return False
    return all(np.array_equal(x[i:i + chunk], y[i:i + chunk]) for i in range(0, x.size, chunk))


def assert_same_load(host, dev):
    (hn, hw), (hl, hr, he), hp = host
    (dn, dw), (dl, dr, de), dp = dev
    assert same_bytes(hn, dn), "nodes differ"
    assert same_bytes(hw, dw), "wires differ"
    assert np.array_equal(hl, dl), "locator differs"
    assert np.array_equal(hr, dr), "index rows differ"
    assert np.array_equal(he, de), "index entries differ"
    assert hp.tobytes() == dp.tobytes(), "pieces differ"


CASES = [
    dict(P=4, npp=1024, wpp=4096, pct=95, seed=12345, profile=0),
    dict(P=4, npp=1024, wpp=4096, pct=95, seed=12345, profile=1),
    dict(P=4, npp=1024, wpp=4096, pct=50, seed=12345, profile=1),
    dict(P=5, npp=300, wpp=1000, pct=0, seed=7, profile=0),
    dict(P=3, npp=100, wpp=500, pct=100, seed=99, profile=1),
    dict(P=1, npp=64, wpp=256, pct=50, seed=3, profile=0),
    dict(P=2, npp=4, wpp=8, pct=95, seed=12345, profile=0),
    dict(P=3, npp=700, wpp=0, pct=95, seed=1, profile=0),
    dict(P=7, npp=333, wpp=1001, pct=95, seed=-5, profile=1),
]


@pytest.mark.parametrize("case", CASES, ids=[
    f"P{c['P']}_npp{c['npp']}_wpp{c['wpp']}_pct{c['pct']}_prof{c['profile']}" for c in CASES])
def test_device_generator_matches_host_and_oracle(C, O, case):
    say(f"[devgen] {case}")
    gp = C.GenParams(case["P"], case["npp"], case["wpp"], case["pct"], case["seed"],
                     case["profile"])
    host, dev = load_both(C, gp)
    assert_same_load(host, dev)
    ref = O.generate(case["P"], case["npp"], case["wpp"], case["pct"], case["seed"],
                     case["profile"])
    (dn, dw), (dl, dr, de), dp = dev
    assert same_bytes(dn, ref.nodes), "nodes differ from the oracle"
    assert same_bytes(dw, ref.wires), "wires differ from the oracle"
    assert np.array_equal(dl, ref.locator), "locator differs from the oracle"
    assert dp.tobytes() == ref.pieces.tobytes(), "pieces differ from the oracle"
    row, ent = expected_index(ref.wires, len(ref.nodes))
    assert np.array_equal(dr, row) and np.array_equal(de, ent), "index differs"


def test_device_generator_several_chunks(C):
    # 3 x 4 194 305 wires: four start-state chunks, so both buffers are reused.
    gp = C.GenParams(3, 1 << 20, (1 << 22) + 1, 95, 12345, 1)
    say("[devgen] P=3 npp=2^20 wpp=2^22+1")
    host, dev = load_both(C, gp)
    assert_same_load(host, dev)


def test_device_generated_c1_runs_bit_identical(C, O):
    gp = C.GenParams(4, 1024, 4096, 95, 12345, 1)
    say("[devgen] 4-piece circuit, stable, 5 loops EXACT + CSR")
    with C.Context(0) as ctx:
        ctx.generate_device(gp)
        ctx.run_loops(5, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
        nodes, wires = ctx.store()
    ref = O.generate(4, 1024, 4096, 95, 12345, 1)
    O.run(ref, 5, reduce=0)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")
    assert_bit_identical(wires["voltage"], ref.wires["voltage"], "wire voltage")


def test_batch_device_generated_matches_host(C):
    say("[devgen] batch")
    params = [C.GenParams(2 + i, 256 + 64 * i, 1024 + 512 * i, 95, 100 + i, 1) for i in range(4)]
    b = C.Batch(0)
    try:
        ids_h = [b.add_generated(p, 1) for p in params]
        ids_d = [b.add_generated_device(p, 1) for p in params]

        def check():
            for p, h, d in zip(params, ids_h, ids_d):
                n, w = p.num_pieces * p.nodes_per_piece, p.num_pieces * p.wires_per_piece
                hn, hw = b.store(h, n, w)
                dn, dw = b.store(d, n, w)
                assert same_bytes(hn, dn) and same_bytes(hw, dw)

        check()
        b.run(C.Batch.config(policy=C.POLICY_EXACT, mode=C.REDUCE_CSR))
        check()
    finally:
        b.close()


def test_device_generator_rejects_bad_params(C):
    with C.Context(0) as ctx:
        with pytest.raises(C.CircuitError):
            ctx.generate_device(C.GenParams(0, 16, 16, 95, 1, 0))
        with pytest.raises(C.CircuitError):
            ctx.generate_device(C.GenParams(1 << 16, 1 << 15, 16, 95, 1, 0))
