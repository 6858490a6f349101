"""The allocation-free entry point ckt_calc_new_currents_index_gpu_noalloc
against the allocating entry point it varies, ckt_calc_new_currents_index_gpu,
both called through the C ABI on the same reference-layout instances.

Both launch the same kernel, so their results must be bit-identical; EXACT is
also checked against the oracle.  Two calls follow each other on the shared
static scratch, the second over fewer pieces and wires, and a call that finds
an invalid PointerLocation is followed by a clean one (the scratch's counter
must start at zero on every call).
"""
import ctypes

import numpy as np
import pytest

from helpers import assert_bit_identical

pytestmark = pytest.mark.gpu

ALLOC = "ckt_calc_new_currents_index_gpu"
NOALLOC = "ckt_calc_new_currents_index_gpu_noalloc"


def _call(C, fn, pieces, wires, nodes, flag):
    pa = np.ascontiguousarray(pieces)
    return getattr(C.lib(), fn)(pa.ctypes.data_as(ctypes.c_void_p), len(pa), wires.ptr,
                                nodes.ptr, nodes.ptr, nodes.ptr, flag, None)


def _relax(C, fn, circ, flag, pieces=None):
    wires, nodes = C.DeviceArray(circ.wires), C.DeviceArray(circ.nodes)
    st = _call(C, fn, circ.pieces if pieces is None else pieces, wires, nodes, flag)
    assert st == 0, (fn, st, C.lib().ckt_last_error())
    return wires.download()


@pytest.mark.parametrize("flag", [0, 3], ids=["exact", "scaled"])
def test_noalloc_matches_allocating_entry_point(C, O, flag):
    # 16 pieces x 1024 wires, then 3 pieces x 300 wires: the second call
    # reuses the scratch the first one filled with a longer table.
    for P, npp, wpp, seed in ((16, 256, 1024, 3), (3, 128, 300, 4)):
        circ = O.generate(P, npp, wpp, 95, seed, O.PROFILE_STABLE)
        want = _relax(C, ALLOC, circ, flag)
        got = _relax(C, NOALLOC, circ, flag)
        for f in ("current", "voltage"):
            assert got[f].tobytes() == want[f].tobytes(), (P, f)
        if flag == 0:
            ref = circ.copy()
            O.calc_new_currents(ref)
            assert_bit_identical(got["current"], ref.wires["current"], "current vs oracle")
            assert_bit_identical(got["voltage"], ref.wires["voltage"], "voltage vs oracle")


def test_noalloc_counter_and_size_limit(C, O):
    circ = O.generate(4, 64, 256, 95, 5, O.PROFILE_STABLE)
    wires, nodes = C.DeviceArray(circ.wires), C.DeviceArray(circ.nodes)
    bad = circ.wires.copy()
    bad["out_loc"][17] = 9
    wires.upload(bad)
    assert _call(C, NOALLOC, circ.pieces, wires, nodes, 0) == 1  # CKT_ERR_INVALID_ARG
    assert b"invalid PointerLocation" in C.lib().ckt_last_error()
    # the next call starts from a cleared counter
    wires.upload(circ.wires)
    assert _call(C, NOALLOC, circ.pieces, wires, nodes, 0) == 0
    # more pieces than the scratch holds: rejected before any device work
    wires.upload(circ.wires)
    many = np.zeros(4097, circ.pieces.dtype)
    many["first_wire"] = 0
    many["num_wires"] = 1
    assert _call(C, NOALLOC, many, wires, nodes, 0) == 2  # CKT_ERR_OUT_OF_RANGE
    assert b"CKT_NOALLOC_MAX_PIECES" in C.lib().ckt_last_error()
    assert wires.download().tobytes() == circ.wires.tobytes()
    # exactly the maximum is accepted: 4096 one-wire pieces over 4096 wires,
    # with the per-piece results
    circ = O.generate(4, 64, 1024, 95, 6, O.PROFILE_STABLE)
    many = many[:4096]
    many["first_wire"] = np.arange(4096)
    got = _relax(C, NOALLOC, circ, 0, pieces=many)
    want = _relax(C, ALLOC, circ, 0)
    assert got.tobytes() == want.tobytes()
