"""update_voltages (circuit_gpu.cu:306-350, circuit.cc:1299-1302) on node
arrays larger than the last-level cache, where the kernel switches to
nontemporal 128-bit accesses, and on a small circuit, where it keeps ordinary
ones.  Both must equal the per-node closed form v' = (v + q/C) * (1 - leak),
q' = 0, bit for bit (FP32, IEEE division, no contraction), computed with
numpy.  Pieces of 4 194 305 nodes make every tile of pieces 1 and 2 start off
a 16-byte boundary, so the one-node-at-a-time path is covered too.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def closed_form(nodes):
    v, q = nodes["voltage"], nodes["charge"]
    C, lk = nodes["capacitance"], nodes["leakage"]
    return (v + q / C) * (np.float32(1) - lk)


def make(C, P, npp, seed):
    rng = np.random.default_rng(seed)
    N = P * npp
    nodes = np.zeros(N, C.NODE)
    nodes["voltage"] = rng.uniform(-1, 1, N).astype(np.float32)
    nodes["charge"] = rng.uniform(-1e-3, 1e-3, N).astype(np.float32)
    nodes["capacitance"] = rng.uniform(1, 2, N).astype(np.float32)
    nodes["leakage"] = rng.uniform(0, 0.1, N).astype(np.float32)
    pieces = np.zeros(P, C.PIECE)
    pieces["first_node"] = np.arange(P, dtype=np.uint32) * npp
    pieces["num_nodes"] = npp
    return C.HostCircuit(nodes, np.zeros(0, C.WIRE), np.ones(N, np.int32), pieces)


@pytest.mark.parametrize("P,npp", [(3, (1 << 22) + 1), (4, 1021)],
                         ids=["12.6M_nodes_nontemporal", "4084_nodes_cached"])
def test_update_voltages_closed_form(C, P, npp):
    say(f"[upv] P={P} npp={npp} ({P * npp * 24 / 2**20:.0f} MiB of node arrays)")
    host = make(C, P, npp, seed=P)
    want = closed_form(host.nodes)
    with C.Context(0) as ctx:
        ctx.load(host)
        ctx.update_voltages()
        nodes, _ = ctx.store()
    assert np.array_equal(nodes["voltage"].view(np.uint32), want.view(np.uint32))
    assert not nodes["charge"].any()
    assert np.array_equal(nodes["capacitance"], host.nodes["capacitance"])
    assert np.array_equal(nodes["leakage"], host.nodes["leakage"])
