"""ctypes binding of the CPU oracle (``libcircuit_oracle.so``).

TEST INFRASTRUCTURE ONLY: the oracle that GPU results are checked against.  The
record layouts below are written out independently of the product package
(``apps/circuit/circuit.h`` lines 35-55 give them) so that a layout mistake in
one cannot hide in the other.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIBRARY selects another build (libcircuit_oracle_san.so: under
# AddressSanitizer and UndefinedBehaviorSanitizer, `make -C oracle san`).
LIB_PATH = os.environ.get("ORACLE_LIBRARY") or os.path.join(HERE, "libcircuit_oracle.so")

SEG = 10
STEPS = 10000
SEED_GPU, SEED_CPU_LITERAL = 0, 1
PROFILE_REFERENCE, PROFILE_STABLE = 0, 1

NODE = np.dtype([("charge", "<f4"), ("voltage", "<f4"), ("capacitance", "<f4"),
                 ("leakage", "<f4")])
WIRE = np.dtype([("in_ptr", "<u4"), ("out_ptr", "<u4"), ("in_loc", "<i4"), ("out_loc", "<i4"),
                 ("inductance", "<f4"), ("resistance", "<f4"), ("current", "<f4", (SEG,)),
                 ("capacitance", "<f4"), ("voltage", "<f4", (SEG - 1,)), ("pad", "<f4", (6,))])
NODE64 = np.dtype([("charge", "<f8"), ("voltage", "<f8"), ("capacitance", "<f8"),
                   ("leakage", "<f8")])
WIRE64 = np.dtype([("in_ptr", "<u4"), ("out_ptr", "<u4"), ("in_loc", "<i4"), ("out_loc", "<i4"),
                   ("inductance", "<f8"), ("resistance", "<f8"), ("current", "<f8", (SEG,)),
                   ("capacitance", "<f8"), ("voltage", "<f8", (SEG - 1,))])
PIECE = np.dtype([("first_wire", "<u4"), ("num_wires", "<u4"), ("first_node", "<u4"),
                  ("num_nodes", "<u4")])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all", "san"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _declare(L):
    vp, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64
    L.orc_generate.argtypes = [i32, i32, i32, i32, i32, i32, vp, vp, vp, vp,
                               ctypes.POINTER(u64)]
    L.orc_generate.restype = i32
    L.orc_drand48_stream.argtypes = [i32, vp, i32]
    L.orc_calc_new_currents_f32.argtypes = [i32, i32, vp, vp, vp, vp, u32, u32]
    L.orc_calc_new_currents_f64.argtypes = [i32, i32, vp, vp, u32, u32]
    L.orc_distribute_charge_f32.argtypes = [vp, vp, vp, vp, u32, u32]
    L.orc_distribute_charge_fold_f32.argtypes = [vp, vp, vp, vp, u32, u32]
    L.orc_fold_apply_f32.argtypes = [vp, vp, u32, u32]
    L.orc_distribute_charge_f64.argtypes = [vp, vp, u32, u32]
    L.orc_distribute_charge_kahan64.argtypes = [vp, u32, vp, vp, u32]
    L.orc_update_voltages_f32.argtypes = [vp, vp, vp, u32, u32]
    L.orc_update_voltages_f64.argtypes = [vp, u32, u32]
    L.orc_run_f32.argtypes = [i32, i32, i32, i32, i32, i32, vp, u32, vp, u32, vp, vp, u32]
    L.orc_run_f32.restype = i32
    L.orc_run_f64.argtypes = [i32, i32, i32, i32, vp, u32, vp, u32, vp, u32]
    L.orc_calc_new_currents_f64_limit.argtypes = [vp, vp, u32, u32]
    L.orc_run_f64_limit.argtypes = [i32, i32, vp, u32, vp, u32, vp, u32]
    L.orc_cnc_sample_f32.argtypes = [i32, i32, i32, vp, vp, vp, u32]
    L.orc_cnc_sample_f32.restype = ctypes.c_double
    L.orc_promote.argtypes = [vp, u32, vp, u32, vp, vp]


class Circuit:
    """A circuit in the reference layouts (host numpy arrays)."""

    def __init__(self, nodes, wires, locator, pieces, draws=0):
        self.nodes, self.wires, self.locator, self.pieces = nodes, wires, locator, pieces
        self.draws = draws

    def copy(self):
        return Circuit(self.nodes.copy(), self.wires.copy(), self.locator.copy(),
                       self.pieces.copy(), self.draws)

    @property
    def num_nodes(self):
        return len(self.nodes)

    @property
    def num_wires(self):
        return len(self.wires)


def generate(num_pieces, nodes_per_piece, wires_per_piece, pct=95, seed=12345,
             profile=PROFILE_REFERENCE) -> Circuit:
    N, W = num_pieces * nodes_per_piece, num_pieces * wires_per_piece
    nodes = np.zeros(N, NODE)
    wires = np.zeros(W, WIRE)
    loc = np.zeros(N, np.int32)
    pieces = np.zeros(num_pieces, PIECE)
    draws = ctypes.c_uint64(0)
    rc = lib().orc_generate(num_pieces, nodes_per_piece, wires_per_piece, pct, seed, profile,
                            _p(nodes), _p(wires), _p(loc), _p(pieces), ctypes.byref(draws))
    if rc != 0:
        raise ValueError(f"orc_generate failed rc={rc}")
    return Circuit(nodes, wires, loc, pieces, draws.value)


def drand48_stream(seed, count):
    out = np.zeros(count, np.float64)
    lib().orc_drand48_stream(seed, _p(out), count)
    return out


def calc_new_currents(c: Circuit, piece=None, steps=STEPS, seeding=SEED_GPU, pvt=None, shr=None,
                      ghost=None):
    fw, nw = (0, c.num_wires) if piece is None else (int(piece["first_wire"]),
                                                     int(piece["num_wires"]))
    n = c.nodes
    return lib().orc_calc_new_currents_f32(seeding, steps, _p(c.wires), _p(pvt if pvt is not None else n),
                                           _p(shr if shr is not None else n),
                                           _p(ghost if ghost is not None else n), fw, nw)


def distribute_charge(c: Circuit, piece=None):
    fw, nw = (0, c.num_wires) if piece is None else (int(piece["first_wire"]),
                                                     int(piece["num_wires"]))
    return lib().orc_distribute_charge_f32(_p(c.wires), _p(c.nodes), _p(c.nodes), _p(c.nodes), fw, nw)


def distribute_charge_fold(wires, pvt_nodes, shr_fold, ghost_fold, first_wire, num_wires):
    return lib().orc_distribute_charge_fold_f32(_p(wires), _p(pvt_nodes), _p(shr_fold),
                                                _p(ghost_fold), first_wire, num_wires)


def fold_apply(owner, fold, first, count):
    lib().orc_fold_apply_f32(_p(owner), _p(fold), first, count)


def distribute_charge_kahan64(c: Circuit):
    lib().orc_distribute_charge_kahan64(_p(c.wires), c.num_wires, _p(c.nodes), _p(c.locator),
                                        c.num_nodes)


def update_voltages(c: Circuit, piece=None, pvt=None, shr=None):
    fn, nn = (0, c.num_nodes) if piece is None else (int(piece["first_node"]),
                                                     int(piece["num_nodes"]))
    lib().orc_update_voltages_f32(_p(pvt if pvt is not None else c.nodes),
                                  _p(shr if shr is not None else c.nodes), _p(c.locator), fn, nn)


def run(c: Circuit, loops, steps=STEPS, seeding=SEED_GPU, reduce=0, nthreads=None, grain=64):
    """region_main loop in FP32 on the circuit (in place)."""
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    return lib().orc_run_f32(seeding, steps, loops, reduce, nthreads, grain, _p(c.nodes),
                             c.num_nodes, _p(c.wires), c.num_wires, _p(c.locator), _p(c.pieces),
                             len(c.pieces))


def promote(c: Circuit):
    n64 = np.zeros(c.num_nodes, NODE64)
    w64 = np.zeros(c.num_wires, WIRE64)
    lib().orc_promote(_p(c.nodes), c.num_nodes, _p(c.wires), c.num_wires, _p(n64), _p(w64))
    return n64, w64


def run_f64(c: Circuit, loops, steps=STEPS, seeding=SEED_GPU, nthreads=None):
    """Full-FP64 twin from the same FP32 inputs; returns (nodes64, wires64)."""
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    n64, w64 = promote(c)
    lib().orc_run_f64(seeding, steps, loops, nthreads, _p(n64), c.num_nodes, _p(w64), c.num_wires,
                      _p(c.pieces), len(c.pieces))
    return n64, w64


def run_f64_limit(c: Circuit, loops, nthreads=None):
    """FP64 loops with each relaxation replaced by its limit (the backward-
    Euler solve); valid where the relaxation converges (stable profile)."""
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    n64, w64 = promote(c)
    lib().orc_run_f64_limit(loops, nthreads, _p(n64), c.num_nodes, _p(w64), c.num_wires,
                            _p(c.pieces), len(c.pieces))
    return n64, w64


def calc_new_currents_f64(n64, w64, first_wire, num_wires, steps=STEPS, seeding=SEED_GPU):
    lib().orc_calc_new_currents_f64(seeding, steps, _p(w64), _p(n64), first_wire, num_wires)


def cnc_sample(c: Circuit, wire_idx, steps=STEPS, nthreads=1, seeding=SEED_GPU):
    """calc_new_currents over the listed wires; returns wall seconds."""
    idx = np.ascontiguousarray(wire_idx, dtype=np.uint32)
    return lib().orc_cnc_sample_f32(seeding, steps, nthreads, _p(c.wires), _p(c.nodes), _p(idx),
                                    len(idx))
