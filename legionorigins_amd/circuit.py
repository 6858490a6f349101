"""Host-side binding of the gfx950 circuit library (``libckt.so``).

Mirrors the circuit application's task interface (``apps/circuit/circuit.h``
task IDs and records, ``circuit.cc:352-460`` task bodies,
``circuit_gpu.cu:186-368`` GPU entry points) over the C ABI declared in
``include/circuit_abi.h``.  All compute runs in the HIP library; this module
only marshals arguments.  A missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CKT_LIBRARY selects another build of the library: libckt_checked.so (kernel
# bounds checks) or libckt_san.so (bounds checks, host code under
# AddressSanitizer and UndefinedBehaviorSanitizer; load it with the sanitizer
# runtime preloaded).
LIB_PATH = os.environ.get("CKT_LIBRARY") or os.path.join(HERE, "libckt.so")

# apps/circuit/circuit.h:42-44
WIRE_SEGMENTS = 10
STEPS = 10000
DELTAT = 1e-6

# circuit.h:13-17
PRIVATE_PTR, SHARED_PTR, GHOST_PTR = 0, 1, 2
# circuit.h:19-32
(REGION_MAIN, CALC_NEW_CURRENTS, DISTRIBUTE_CHARGE, UPDATE_VOLTAGES, CALC_NEW_CURRENTS_NOP,
 DISTRIBUTE_CHARGE_NOP, UPDATE_VOLTAGES_NOP, SANITY_CHECK_WIRES) = range(8)
REDUCE_ID = 1

POLICY_EXACT, POLICY_RECIP, POLICY_FAST, POLICY_SCALED = 0, 1, 2, 3
REDUCE_ATOMIC, REDUCE_CSR, REDUCE_CSR_KAHAN64 = 0, 1, 2
PROFILE_REFERENCE, PROFILE_STABLE = 0, 1
(OPT_STEPS, OPT_CNC_BLOCK, OPT_PERSISTENT, OPT_WAVES_PER_SIMD, OPT_DSC_ATOMIC_IMPL,
 OPT_FUSE_DSC, OPT_WIRES_PER_LANE, OPT_TILE_SPAN, OPT_TILE_ORDER) = range(9)

# CircuitNode / CircuitWire / CircuitPiece records (circuit.h:35-70).
NODE = np.dtype([("charge", "<f4"), ("voltage", "<f4"), ("capacitance", "<f4"),
                 ("leakage", "<f4")])
WIRE = np.dtype([("in_ptr", "<u4"), ("out_ptr", "<u4"), ("in_loc", "<i4"), ("out_loc", "<i4"),
                 ("inductance", "<f4"), ("resistance", "<f4"),
                 ("current", "<f4", (WIRE_SEGMENTS,)), ("capacitance", "<f4"),
                 ("voltage", "<f4", (WIRE_SEGMENTS - 1,)), ("pad", "<f4", (6,))])
PIECE = np.dtype([("first_wire", "<u4"), ("num_wires", "<u4"), ("first_node", "<u4"),
                  ("num_nodes", "<u4")])
assert NODE.itemsize == 16 and WIRE.itemsize == 128 and PIECE.itemsize == 16


class CircuitError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


class _GenParams(ctypes.Structure):
    _fields_ = [("num_pieces", ctypes.c_int32), ("nodes_per_piece", ctypes.c_int32),
                ("wires_per_piece", ctypes.c_int32), ("pct_wire_in_piece", ctypes.c_int32),
                ("random_seed", ctypes.c_int32), ("profile", ctypes.c_int32)]


class _TaskConfig(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_int32), ("reduce_mode", ctypes.c_int32)]


class _LoopConfig(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_int32), ("reduce_mode", ctypes.c_int32),
                ("sync", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("per_piece", ctypes.c_int32)]


class _LoopReport(ctypes.Structure):
    _fields_ = [("elapsed_s", ctypes.c_double), ("gflops", ctypes.c_double),
                ("wire_updates_per_s", ctypes.c_double), ("kernel_ms", ctypes.c_double * 3)]


class _BatchConfig(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_int32), ("reduce_mode", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("persistent", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
                ("wires_per_lane", ctypes.c_int32), ("tile_span", ctypes.c_int32),
                ("tile_order", ctypes.c_int32)]

    def describe(self):
        return {k: getattr(self, k) for k in ("block_threads", "persistent", "waves_per_simd",
                                              "wires_per_lane", "tile_span", "tile_order")}


class _BatchReport(ctypes.Structure):
    _fields_ = [("elapsed_s", ctypes.c_double), ("wire_updates", ctypes.c_double),
                ("wire_updates_per_s", ctypes.c_double), ("cnc_ms", ctypes.c_double),
                ("loops_run", ctypes.c_int32)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_active_wires", ctypes.c_uint64),
                ("max_active", ctypes.c_int32), ("policy", ctypes.c_int32),
                ("reduce_mode", ctypes.c_int32), ("gen_threads", ctypes.c_int32)]


class _Request(ctypes.Structure):
    _fields_ = [("gen", _GenParams), ("num_loops", ctypes.c_int32), ("arrival_s", ctypes.c_double),
                ("timeout_s", ctypes.c_double)]


class _RequestResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("num_wires", ctypes.c_uint32),
                ("num_loops", ctypes.c_int32), ("arrival_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("final_voltage_checksum", ctypes.c_double)]


class _ServerReport(ctypes.Structure):
    _fields_ = [("elapsed_s", ctypes.c_double), ("wire_updates", ctypes.c_double),
                ("wire_updates_per_s", ctypes.c_double),
                ("offered_wire_updates_per_s", ctypes.c_double),
                ("latency_p50_s", ctypes.c_double), ("latency_p99_s", ctypes.c_double),
                ("latency_mean_s", ctypes.c_double), ("latency_max_s", ctypes.c_double),
                ("completed", ctypes.c_int32), ("rejected", ctypes.c_int32),
                ("iterations", ctypes.c_int32), ("max_resident", ctypes.c_int32),
                ("timed_out", ctypes.c_int32)]


_lib = None


def lib():
    """The loaded C library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i32, u32, u64, sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                             ctypes.c_size_t)
    P = ctypes.POINTER
    sig = {
        "ckt_abi_version": ([], i32),
        "ckt_status_string": ([i32], ctypes.c_char_p),
        "ckt_last_error": ([], ctypes.c_char_p),
        "ckt_device_info": ([i32, P(i32), P(i32), P(sz), P(sz)], i32),
        "ckt_device_malloc": ([i32, P(vp), sz], i32),
        "ckt_device_free": ([vp], i32),
        "ckt_memcpy_htod": ([vp, vp, sz], i32),
        "ckt_memcpy_dtoh": ([vp, vp, sz], i32),
        "ckt_device_synchronize": ([i32], i32),
        "ckt_generate_sizes": ([P(_GenParams), P(u32), P(u32), P(u32)], i32),
        "ckt_generate": ([P(_GenParams), vp, vp, vp, vp], i32),
        "ckt_calc_new_currents_gpu": ([vp, vp, vp, vp, vp, i32, vp], i32),
        "ckt_calc_new_currents_index_gpu": ([vp, u32, vp, vp, vp, vp, i32, vp], i32),
        "ckt_calc_new_currents_index_gpu_noalloc": ([vp, u32, vp, vp, vp, vp, i32, vp], i32),
        "ckt_distribute_charge_index_gpu": ([vp, u32, vp, vp, vp, vp, i32, vp], i32),
        "ckt_update_voltages_index_gpu": ([vp, u32, vp, vp, vp, i32, vp], i32),
        "ckt_distribute_charge_gpu": ([vp, vp, vp, vp, vp, i32, vp], i32),
        "ckt_update_voltages_gpu": ([vp, vp, vp, vp, i32, vp], i32),
        "ckt_sanity_check_wires_gpu": ([vp, vp, u32, P(u64), vp], i32),
        "ckt_fold_init_gpu": ([vp, u32, u32, vp], i32),
        "ckt_fold_apply_gpu": ([vp, vp, u32, u32, vp], i32),
        "ckt_ctx_create": ([i32, P(vp)], i32),
        "ckt_ctx_destroy": ([vp], i32),
        "ckt_ctx_stream": ([vp], vp),
        "ckt_ctx_synchronize": ([vp], i32),
        "ckt_ctx_load": ([vp, vp, u32, vp, u32, vp, vp, u32], i32),
        "ckt_ctx_generate": ([vp, P(_GenParams)], i32),
        "ckt_ctx_generate_device": ([vp, P(_GenParams)], i32),
        "ckt_ctx_store_topology": ([vp, vp, vp, vp], i32),
        "ckt_ctx_store": ([vp, vp, vp], i32),
        "ckt_ctx_sizes": ([vp, P(u32), P(u32), P(u32)], i32),
        "ckt_ctx_pieces": ([vp, vp], i32),
        "ckt_calc_new_currents": ([vp, vp, u32, i32], i32),
        "ckt_distribute_charge": ([vp, vp, u32, i32], i32),
        "ckt_update_voltages": ([vp, vp, u32], i32),
        "ckt_sanity_check_wires": ([vp, vp, u32, P(u64)], i32),
        "ckt_execute_index_space": ([vp, i32, vp, u32, P(_TaskConfig)], i32),
        "ckt_run_loops": ([vp, i32, P(_LoopConfig), P(_LoopReport)], i32),
        "ckt_ctx_set_option": ([vp, i32, ctypes.c_int64], i32),
        "ckt_ctx_set_timing": ([vp, i32], i32),
        "ckt_ctx_kernel_time": ([vp, i32, P(ctypes.c_double), P(u64)], i32),
        "ckt_ctx_reset_timing": ([vp], i32),
        "ckt_batch_create": ([i32, u64, P(vp)], i32),
        "ckt_batch_destroy": ([vp], i32),
        "ckt_batch_set_option": ([vp, i32, ctypes.c_int64], i32),
        "ckt_batch_set_timing": ([vp, i32], i32),
        "ckt_batch_kernel_time": ([vp, i32, P(ctypes.c_double), P(u64)], i32),
        "ckt_batch_add_generated": ([vp, P(_GenParams), i32, P(u32)], i32),
        "ckt_batch_add_generated_device": ([vp, P(_GenParams), i32, P(u32)], i32),
        "ckt_batch_add_generated_many": ([vp, P(_GenParams), P(ctypes.c_int32), u32, i32,
                                          P(u32)], i32),
        "ckt_batch_add": ([vp, vp, u32, vp, u32, vp, vp, u32, i32, P(u32)], i32),
        "ckt_batch_remove": ([vp, u32], i32),
        "ckt_batch_count": ([vp, P(u32), P(u64)], i32),
        "ckt_footprint_bytes": ([u32, u32, u32, P(u64)], i32),
        "ckt_batch_run": ([vp, P(_BatchConfig), P(_BatchReport)], i32),
        "ckt_batch_step": ([vp, P(_BatchConfig), P(ctypes.c_int32)], i32),
        "ckt_batch_loops_left": ([vp, u32, P(ctypes.c_int32)], i32),
        "ckt_batch_store": ([vp, u32, vp, vp], i32),
        "ckt_batch_autotune": ([vp, i32, i32, P(_BatchConfig), P(ctypes.c_double)], i32),
        "ckt_batch_probe": ([vp, P(_BatchConfig), i32, P(ctypes.c_double)], i32),
        "ckt_server_create": ([i32, P(_ServerConfig), P(vp)], i32),
        "ckt_server_destroy": ([vp], i32),
        "ckt_server_run_trace": ([vp, P(_Request), u32, P(_RequestResult), P(_ServerReport)], i32),
        "ckt_server_set_keep_state": ([vp, i32], i32),
        "ckt_server_fetch": ([vp, u32, vp, vp], i32),
        "ckt_server_set_keep_limit": ([vp, u32], i32),
        "ckt_server_start": ([vp], i32),
        "ckt_server_stop": ([vp, i32], i32),
        "ckt_server_submit": ([vp, P(_GenParams), i32, ctypes.c_double, P(u32)], i32),
        "ckt_server_poll": ([vp, u32, P(i32), P(_RequestResult)], i32),
        "ckt_server_wait": ([vp, u32, ctypes.c_double, P(_RequestResult)], i32),
        "ckt_server_cancel": ([vp, u32], i32),
        "ckt_copy_kernel": ([vp, vp, sz, vp], i32),
        "ckt_debug_bounds_violations": ([i32, P(u64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def declared_symbols():
    """Names of the ABI functions this binding declares."""
    return sorted(k for k in dir(lib()) if k.startswith("ckt_"))


def _check(status):
    if status != 0:
        L = lib()
        raise CircuitError(status, f"{L.ckt_status_string(status).decode()}: "
                                   f"{L.ckt_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class GenParams:
    """Arguments of load_circuit (circuit.cc:1071-1073) and the CLI defaults
    of region_main (circuit.cc:128-133)."""
    num_pieces: int = 4
    nodes_per_piece: int = 2
    wires_per_piece: int = 4
    pct_wire_in_piece: int = 95
    random_seed: int = 12345
    profile: int = PROFILE_REFERENCE

    def c(self):
        return _GenParams(self.num_pieces, self.nodes_per_piece, self.wires_per_piece,
                          self.pct_wire_in_piece, self.random_seed, self.profile)


@dataclass
class HostCircuit:
    nodes: np.ndarray
    wires: np.ndarray
    locator: np.ndarray
    pieces: np.ndarray

    def copy(self):
        return HostCircuit(self.nodes.copy(), self.wires.copy(), self.locator.copy(),
                           self.pieces.copy())


def generate(params: GenParams) -> HostCircuit:
    """load_circuit on the host, returning the reference-layout records."""
    L = lib()
    n, w, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    cp = params.c()
    _check(L.ckt_generate_sizes(ctypes.byref(cp), ctypes.byref(n), ctypes.byref(w),
                                ctypes.byref(p)))
    nodes = np.zeros(n.value, NODE)
    wires = np.zeros(w.value, WIRE)
    loc = np.zeros(n.value, np.int32)
    pieces = np.zeros(p.value, PIECE)
    _check(L.ckt_generate(ctypes.byref(cp), _p(nodes), _p(wires), _p(loc), _p(pieces)))
    return HostCircuit(nodes, wires, loc, pieces)


def device_info(device=0):
    L = lib()
    g, cus = ctypes.c_int(), ctypes.c_int()
    lds, hbm = ctypes.c_size_t(), ctypes.c_size_t()
    _check(L.ckt_device_info(device, ctypes.byref(g), ctypes.byref(cus), ctypes.byref(lds),
                             ctypes.byref(hbm)))
    return {"gfx950": bool(g.value), "num_cus": cus.value, "lds_per_block": lds.value,
            "hbm_bytes": hbm.value}


class DeviceArray:
    """A device copy of a numpy array (for the accessor-form entry points)."""

    def __init__(self, host: np.ndarray, device=0):
        self.dtype, self.shape = host.dtype, host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        _check(lib().ckt_device_malloc(device, ctypes.byref(self.ptr), max(self.nbytes, 1)))
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        _check(lib().ckt_memcpy_htod(self.ptr, _p(host), self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().ckt_memcpy_dtoh(_p(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().ckt_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _piece_arg(piece):
    arr = np.ascontiguousarray(np.asarray(piece, dtype=PIECE).reshape(-1))
    return arr


# Accessor-form drop-ins (circuit.h:83-104).  Device pointers; synchronous.
def calc_new_currents_gpu(piece, wires: DeviceArray, pvt: DeviceArray, owned: DeviceArray,
                          ghost: DeviceArray, flag=POLICY_EXACT):
    pa = _piece_arg(piece)
    _check(lib().ckt_calc_new_currents_gpu(_p(pa), wires.ptr, pvt.ptr, owned.ptr, ghost.ptr, flag,
                                           None))


def distribute_charge_gpu(piece, wires: DeviceArray, pvt: DeviceArray, owned_fold: DeviceArray,
                          ghost_fold: DeviceArray):
    pa = _piece_arg(piece)
    _check(lib().ckt_distribute_charge_gpu(_p(pa), wires.ptr, pvt.ptr, owned_fold.ptr,
                                           ghost_fold.ptr, 0, None))


def update_voltages_gpu(piece, pvt: DeviceArray, owned: DeviceArray, locator: DeviceArray):
    pa = _piece_arg(piece)
    _check(lib().ckt_update_voltages_gpu(_p(pa), pvt.ptr, owned.ptr, locator.ptr, 0, None))


# The same over a table of pieces in one launch each (index-space form).
def calc_new_currents_index_gpu(pieces, wires: DeviceArray, pvt: DeviceArray, owned: DeviceArray,
                                ghost: DeviceArray, flag=POLICY_EXACT):
    pa = _piece_arg(pieces)
    _check(lib().ckt_calc_new_currents_index_gpu(_p(pa), len(pa), wires.ptr, pvt.ptr, owned.ptr,
                                                 ghost.ptr, flag, None))


def distribute_charge_index_gpu(pieces, wires: DeviceArray, pvt: DeviceArray,
                                owned_fold: DeviceArray, ghost_fold: DeviceArray):
    pa = _piece_arg(pieces)
    _check(lib().ckt_distribute_charge_index_gpu(_p(pa), len(pa), wires.ptr, pvt.ptr,
                                                 owned_fold.ptr, ghost_fold.ptr, 0, None))


def update_voltages_index_gpu(pieces, pvt: DeviceArray, owned: DeviceArray, locator: DeviceArray):
    pa = _piece_arg(pieces)
    _check(lib().ckt_update_voltages_index_gpu(_p(pa), len(pa), pvt.ptr, owned.ptr, locator.ptr, 0,
                                               None))


def sanity_check_wires_gpu(piece, wires: DeviceArray, num_nodes_total):
    pa = _piece_arg(piece)
    bad = ctypes.c_uint64()
    _check(lib().ckt_sanity_check_wires_gpu(_p(pa), wires.ptr, num_nodes_total, ctypes.byref(bad),
                                            None))
    return bad.value


def fold_init_gpu(fold: DeviceArray, first, count):
    _check(lib().ckt_fold_init_gpu(fold.ptr, first, count, None))


def fold_apply_gpu(owner: DeviceArray, fold: DeviceArray, first, count):
    _check(lib().ckt_fold_apply_gpu(owner.ptr, fold.ptr, first, count, None))


class Context:
    """One circuit resident in HBM (structure-of-arrays); the task variants."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        _check(lib().ckt_ctx_create(device, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().ckt_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load(self, c: HostCircuit):
        _check(lib().ckt_ctx_load(self.h, _p(c.nodes), len(c.nodes), _p(c.wires), len(c.wires),
                                  _p(c.locator), _p(c.pieces), len(c.pieces)))

    def generate(self, params: GenParams):
        cp = params.c()
        _check(lib().ckt_ctx_generate(self.h, ctypes.byref(cp)))

    def generate_device(self, params: GenParams):
        """The circuit generate() builds, bit for bit, generated in HBM."""
        cp = params.c()
        _check(lib().ckt_ctx_generate_device(self.h, ctypes.byref(cp)))

    def store_topology(self):
        """(locator, csr_row, csr_ent) of the loaded circuit."""
        n, w, _ = self.sizes()
        loc = np.zeros(n, np.int32)
        row = np.zeros(n + 1, np.uint32)
        ent = np.zeros(2 * w, np.uint32)
        _check(lib().ckt_ctx_store_topology(self.h, _p(loc), _p(row), _p(ent)))
        return loc, row, ent

    def sizes(self):
        n, w, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().ckt_ctx_sizes(self.h, ctypes.byref(n), ctypes.byref(w), ctypes.byref(p)))
        return n.value, w.value, p.value

    def pieces(self):
        _, _, p = self.sizes()
        out = np.zeros(p, PIECE)
        _check(lib().ckt_ctx_pieces(self.h, _p(out)))
        return out

    def store_nodes(self):
        """The node records only (no wire download)."""
        n, _, _ = self.sizes()
        nodes = np.zeros(n, NODE)
        _check(lib().ckt_ctx_store(self.h, _p(nodes), None))
        return nodes

    def store(self):
        n, w, _ = self.sizes()
        nodes = np.zeros(n, NODE)
        wires = np.zeros(w, WIRE)
        _check(lib().ckt_ctx_store(self.h, _p(nodes), _p(wires)))
        return nodes, wires

    def _pieces(self, pieces):
        if pieces is None:
            pieces = self.pieces()
        return _piece_arg(pieces)

    def calc_new_currents(self, pieces=None, policy=POLICY_EXACT):
        pa = self._pieces(pieces)
        _check(lib().ckt_calc_new_currents(self.h, _p(pa), len(pa), policy))

    def distribute_charge(self, pieces=None, mode=REDUCE_ATOMIC):
        pa = self._pieces(pieces)
        _check(lib().ckt_distribute_charge(self.h, _p(pa), len(pa), mode))

    def update_voltages(self, pieces=None):
        pa = self._pieces(pieces)
        _check(lib().ckt_update_voltages(self.h, _p(pa), len(pa)))

    def sanity_check_wires(self, pieces=None):
        pa = self._pieces(pieces)
        bad = ctypes.c_uint64()
        _check(lib().ckt_sanity_check_wires(self.h, _p(pa), len(pa), ctypes.byref(bad)))
        return bad.value

    def execute_index_space(self, task_id, pieces=None, policy=POLICY_EXACT, mode=REDUCE_ATOMIC):
        pa = self._pieces(pieces)
        cfg = _TaskConfig(policy, mode)
        _check(lib().ckt_execute_index_space(self.h, task_id, _p(pa), len(pa), ctypes.byref(cfg)))

    def run_loops(self, loops, policy=POLICY_EXACT, mode=REDUCE_ATOMIC, sync=0, use_graph=False,
                  per_piece=False):
        cfg = _LoopConfig(policy, mode, sync, int(use_graph), int(per_piece))
        rep = _LoopReport()
        _check(lib().ckt_run_loops(self.h, loops, ctypes.byref(cfg), ctypes.byref(rep)))
        return {"elapsed_s": rep.elapsed_s, "gflops": rep.gflops,
                "wire_updates_per_s": rep.wire_updates_per_s, "kernel_ms": list(rep.kernel_ms)}

    def synchronize(self):
        _check(lib().ckt_ctx_synchronize(self.h))

    def set_option(self, opt, value):
        _check(lib().ckt_ctx_set_option(self.h, opt, value))

    def set_timing(self, on=True):
        _check(lib().ckt_ctx_set_timing(self.h, int(on)))

    def kernel_time(self, task_id):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check(lib().ckt_ctx_kernel_time(self.h, task_id, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def reset_timing(self):
        _check(lib().ckt_ctx_reset_timing(self.h))

    @property
    def stream(self):
        return lib().ckt_ctx_stream(self.h)


class Batch:
    """Many independent circuits advanced by shared launches."""

    def __init__(self, device=0, budget_bytes=0):
        self.h = ctypes.c_void_p()
        _check(lib().ckt_batch_create(device, budget_bytes, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().ckt_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_generated(self, params: GenParams, loops):
        cid = ctypes.c_uint32()
        cp = params.c()
        _check(lib().ckt_batch_add_generated(self.h, ctypes.byref(cp), loops, ctypes.byref(cid)))
        return cid.value

    def add_generated_device(self, params: GenParams, loops):
        cid = ctypes.c_uint32()
        cp = params.c()
        _check(lib().ckt_batch_add_generated_device(self.h, ctypes.byref(cp), loops,
                                                    ctypes.byref(cid)))
        return cid.value

    def add_generated_many(self, params, loops, threads=8):
        n = len(params)
        arr = (_GenParams * n)(*[p.c() for p in params])
        lp = (ctypes.c_int32 * n)(*loops)
        ids = (ctypes.c_uint32 * n)()
        _check(lib().ckt_batch_add_generated_many(self.h, arr, lp, n, threads, ids))
        return list(ids)

    def add(self, c: HostCircuit, loops):
        cid = ctypes.c_uint32()
        _check(lib().ckt_batch_add(self.h, _p(c.nodes), len(c.nodes), _p(c.wires), len(c.wires),
                                   _p(c.locator), _p(c.pieces), len(c.pieces), loops,
                                   ctypes.byref(cid)))
        return cid.value

    def remove(self, cid):
        _check(lib().ckt_batch_remove(self.h, cid))

    def count(self):
        n, b = ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().ckt_batch_count(self.h, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def set_option(self, opt, value):
        _check(lib().ckt_batch_set_option(self.h, opt, value))

    def set_timing(self, on=True):
        _check(lib().ckt_batch_set_timing(self.h, int(on)))

    def kernel_time(self, task_id):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check(lib().ckt_batch_kernel_time(self.h, task_id, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    @staticmethod
    def config(policy=POLICY_FAST, mode=REDUCE_ATOMIC, use_graph=True, block_threads=0,
               persistent=0, waves_per_simd=0, wires_per_lane=0, tile_span=0, tile_order=0):
        return _BatchConfig(policy, mode, int(use_graph), block_threads, persistent,
                            waves_per_simd, wires_per_lane, tile_span, tile_order)

    def run(self, cfg=None):
        cfg = cfg or self.config()
        rep = _BatchReport()
        _check(lib().ckt_batch_run(self.h, ctypes.byref(cfg), ctypes.byref(rep)))
        return {"elapsed_s": rep.elapsed_s, "wire_updates": rep.wire_updates,
                "wire_updates_per_s": rep.wire_updates_per_s, "cnc_ms": rep.cnc_ms,
                "loops_run": rep.loops_run}

    def step(self, cfg=None):
        cfg = cfg or self.config()
        active = ctypes.c_int32()
        _check(lib().ckt_batch_step(self.h, ctypes.byref(cfg), ctypes.byref(active)))
        return active.value

    def loops_left(self, cid):
        v = ctypes.c_int32()
        _check(lib().ckt_batch_loops_left(self.h, cid, ctypes.byref(v)))
        return v.value

    def store(self, cid, num_nodes, num_wires):
        nodes = np.zeros(num_nodes, NODE)
        wires = np.zeros(num_wires, WIRE)
        _check(lib().ckt_batch_store(self.h, cid, _p(nodes), _p(wires)))
        return nodes, wires

    def autotune(self, policy=POLICY_FAST, probe_steps=500):
        best = _BatchConfig()
        ms = ctypes.c_double()
        _check(lib().ckt_batch_autotune(self.h, policy, probe_steps, ctypes.byref(best),
                                        ctypes.byref(ms)))
        return best, ms.value

    def probe(self, cfg, probe_steps=500):
        ms = ctypes.c_double()
        _check(lib().ckt_batch_probe(self.h, ctypes.byref(cfg), probe_steps, ctypes.byref(ms)))
        return ms.value


@dataclass
class Request:
    gen: GenParams
    num_loops: int
    arrival_s: float
    timeout_s: float = 0.0


class Server:
    """Continuous batching of circuit simulations: online (start, submit,
    poll, wait, cancel, stop) or replaying a request trace in real time
    (run_trace)."""

    def __init__(self, device=0, budget_bytes=0, max_active_wires=0, max_active=0,
                 policy=POLICY_FAST, mode=REDUCE_ATOMIC, gen_threads=8):
        self.h = ctypes.c_void_p()
        cfg = _ServerConfig(budget_bytes, max_active_wires, max_active, policy, mode, gen_threads)
        _check(lib().ckt_server_create(device, ctypes.byref(cfg), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().ckt_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_keep_state(self, keep=True, max_wires=None):
        _check(lib().ckt_server_set_keep_state(self.h, int(keep)))
        _check(lib().ckt_server_set_keep_limit(self.h, 0xFFFFFFFF if max_wires is None
                                               else int(max_wires)))

    # Online serving (ckt_server_start/submit/poll/wait/cancel/stop).
    def start(self):
        _check(lib().ckt_server_start(self.h))

    def stop(self, drain=True):
        _check(lib().ckt_server_stop(self.h, int(drain)))

    def submit(self, gen: GenParams, num_loops, timeout_s=0.0):
        rid = ctypes.c_uint32()
        g = gen.c()
        _check(lib().ckt_server_submit(self.h, ctypes.byref(g), num_loops, timeout_s,
                                       ctypes.byref(rid)))
        return rid.value

    @staticmethod
    def _result(x):
        return {"status": x.status, "num_wires": x.num_wires, "num_loops": x.num_loops,
                "arrival_s": x.arrival_s, "admit_s": x.admit_s, "finish_s": x.finish_s,
                "checksum": x.final_voltage_checksum}

    def poll(self, rid):
        done, res = ctypes.c_int(), _RequestResult()
        _check(lib().ckt_server_poll(self.h, rid, ctypes.byref(done), ctypes.byref(res)))
        return bool(done.value), self._result(res)

    def wait(self, rid, timeout_s=-1.0):
        res = _RequestResult()
        _check(lib().ckt_server_wait(self.h, rid, timeout_s, ctypes.byref(res)))
        return self._result(res)

    def cancel(self, rid):
        _check(lib().ckt_server_cancel(self.h, rid))

    def run_trace(self, requests):
        n = len(requests)
        reqs = (_Request * n)()
        for i, r in enumerate(requests):
            reqs[i] = _Request(r.gen.c(), r.num_loops, r.arrival_s, r.timeout_s)
        res = (_RequestResult * n)()
        rep = _ServerReport()
        _check(lib().ckt_server_run_trace(self.h, reqs, n, res, ctypes.byref(rep)))
        results = [self._result(x) for x in res]
        report = {k: getattr(rep, k) for k, _ in _ServerReport._fields_}
        return results, report

    def fetch(self, index, num_nodes, num_wires):
        nodes = np.zeros(num_nodes, NODE)
        wires = np.zeros(num_wires, WIRE)
        _check(lib().ckt_server_fetch(self.h, index, _p(nodes), _p(wires)))
        return nodes, wires


def poisson_trace(rate_per_s, duration_s, seed, wires_lo=2**12, wires_hi=2**20, loops_lo=1,
                  loops_hi=20, pct=95, profile=PROFILE_REFERENCE, base_seed=1000, timeout_s=0.0):
    """The served workload: Poisson arrivals; per request wires log-uniform in
    [wires_lo, wires_hi], nodes = wires/4, pieces = ceil(nodes/4096), uniform
    pieces (remainders dropped), loops uniform in [loops_lo, loops_hi];
    optional admission timeout per request."""
    rng = np.random.default_rng(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t > duration_s:
            break
        W = int(np.exp(rng.uniform(np.log(wires_lo), np.log(wires_hi))))
        out.append(Request(shape_for_wires(W, pct, base_seed + i, profile),
                           int(rng.integers(loops_lo, loops_hi + 1)), t, timeout_s))
        i += 1
    return out


def shape_for_wires(W, pct=95, seed=0, profile=PROFILE_REFERENCE, nodes_per_piece_max=4096):
    """Uniform-piece circuit of about W wires: nodes = W/4, pieces =
    ceil(nodes/4096), nodes/piece and wires/piece rounded down."""
    nodes = max(1, W // 4)
    pieces = -(-nodes // nodes_per_piece_max)
    return GenParams(pieces, max(1, nodes // pieces), max(1, W // pieces), pct, seed, profile)


def footprint_bytes(num_nodes, num_wires, num_pieces):
    b = ctypes.c_uint64()
    _check(lib().ckt_footprint_bytes(num_nodes, num_wires, num_pieces, ctypes.byref(b)))
    return b.value


def copy_kernel(dst: DeviceArray, src: DeviceArray, nbytes, stream=None):
    _check(lib().ckt_copy_kernel(dst.ptr, src.ptr, nbytes, stream))
