"""The C ABI library on the CPU: exports, layouts, host-side logic.

No kernel is launched here: the generator is host code, and the device
entry points are only exercised for their error behaviour without a GPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "circuit_abi.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void\*|const char\*)\s+(ckt_\w+)\(", text, re.M)))


def test_every_declared_symbol_is_exported(C):
    lib = C.lib()
    names = declared_functions()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n
    # and the binding declares exactly the header's functions
    assert sorted(C.lib().__dict__.keys() & set(names)) == names


def test_abi_version_and_status_strings(C):
    L = C.lib()
    assert L.ckt_abi_version() == 1
    assert L.ckt_status_string(0) == b"ok"
    assert L.ckt_status_string(4) == b"over HBM budget"


def test_record_layouts_match_reference(C, O):
    # circuit.h:35-55: 16-byte nodes, 128-byte wires
    assert C.NODE.itemsize == O.NODE.itemsize == O.lib().orc_sizeof_node() == 16
    assert C.WIRE.itemsize == O.WIRE.itemsize == O.lib().orc_sizeof_wire() == 128
    for name in C.WIRE.names:
        assert C.WIRE.fields[name][1] == O.WIRE.fields[name][1], name
    assert C.WIRE.fields["current"][1] == 24 and C.WIRE.fields["voltage"][1] == 68


@pytest.mark.parametrize("P,npp,wpp,pct,seed,profile", [
    (4, 1024, 4096, 95, 12345, 0),  # 4 pieces, pct 95
    (4, 1024, 4096, 50, 12345, 0),  # 4 pieces, pct 50
    (4, 1024, 4096, 95, 12345, 1),  # stable profile
    (1, 300, 1200, 80, 7, 0),       # one piece, cross branch wraps
    (7, 33, 129, 0, -5, 1),         # every wire crosses pieces
    (3, 5, 0, 95, 1, 0),            # no wires
])
def test_product_generator_is_bit_identical_to_oracle(C, O, P, npp, wpp, pct, seed, profile):
    host = C.generate(C.GenParams(P, npp, wpp, pct, seed, profile))
    ref = O.generate(P, npp, wpp, pct, seed, profile)
    assert host.nodes.tobytes() == ref.nodes.tobytes()
    assert host.wires.tobytes() == ref.wires.tobytes()
    assert np.array_equal(host.locator, ref.locator)
    assert host.pieces.tobytes() == ref.pieces.tobytes()


def test_generator_rejects_bad_sizes(C):
    for bad in (C.GenParams(0, 4, 4), C.GenParams(2, 0, 4), C.GenParams(2, 4, -1),
                C.GenParams(1 << 16, 1 << 15, 1), C.GenParams(2, 4, 4, profile=9)):
        with pytest.raises(C.CircuitError) as e:
            C.generate(bad)
        assert e.value.status in (1, 2)


def test_footprint(C):
    small = C.footprint_bytes(4096, 16384, 4)
    big = C.footprint_bytes(1 << 20, 1 << 22, 64)
    # ~110 B per wire and ~21 B per node of SoA state, CSR index and tables
    assert 100 * 16384 < small < 140 * 16384 + 64 * 4096
    assert 100 * (1 << 22) < big < 120 * (1 << 22) + 32 * (1 << 20)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is visible")
def test_no_device_is_an_error_not_an_abort(C):
    L = C.lib()
    h = ctypes.c_void_p()
    st = L.ckt_ctx_create(0, ctypes.byref(h))
    assert st == 6  # CKT_ERR_NO_DEVICE
    assert not h.value
    st = L.ckt_batch_create(0, 0, ctypes.byref(h))
    assert st == 6
    # null arguments are rejected before any device call
    assert L.ckt_calc_new_currents_gpu(None, None, None, None, None, 0, None) == 1
    assert L.ckt_ctx_synchronize(None) == 1
    assert L.ckt_run_loops(None, 1, None, None) == 1


def test_index_form_rejects_overlapping_pieces(C):
    # The pieces of one index-space call must not share wires (calc_new_currents,
    # distribute_charge) or nodes (update_voltages): the check runs on the host
    # before any device call, so it needs no GPU (the instance pointers are
    # never dereferenced).
    L = C.lib()
    dummy = ctypes.c_void_p(256)
    pieces = np.zeros(3, C.PIECE)
    pieces["first_wire"] = [0, 4096, 4000]       # third overlaps the first
    pieces["num_wires"] = [4096, 4096, 200]
    pieces["first_node"] = [0, 1024, 2048]
    pieces["num_nodes"] = [1024, 1024, 1024]
    p = pieces.ctypes.data_as(ctypes.c_void_p)
    assert L.ckt_calc_new_currents_index_gpu(p, 3, dummy, dummy, dummy, dummy, 0, None) == 1
    assert "share wires" in L.ckt_last_error().decode()
    assert L.ckt_calc_new_currents_index_gpu_noalloc(p, 3, dummy, dummy, dummy, dummy, 0, None) == 1
    assert L.ckt_distribute_charge_index_gpu(p, 3, dummy, dummy, dummy, dummy, 0, None) == 1
    nodes = pieces.copy()
    nodes["first_wire"] = [0, 4096, 8192]
    nodes["first_node"] = [0, 1024, 1500]        # third overlaps the second's nodes
    q = nodes.ctypes.data_as(ctypes.c_void_p)
    assert L.ckt_update_voltages_index_gpu(q, 3, dummy, dummy, dummy, 0, None) == 1
    assert "share nodes" in L.ckt_last_error().decode()
