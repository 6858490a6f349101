"""Benchmark of the circuit hot path on one MI355X.

Workload (default): the full-size plain circuit -- 64 pieces x 16384 nodes x
65536 wires/piece (4,194,304 wires, 1,048,576 nodes), 95% of wires inside
their piece, seed 12345, WIRE_SEGMENTS=10, 10000 inner steps per loop, FP32,
structure-of-arrays in HBM.  One bench step is one loop of the three tasks
(calc_new_currents, distribute_charge, update_voltages) over all pieces; the
metric is wire-updates/s = wires x loops / wall time, inputs resident in HBM.
Policy (default): SCALED, with the ATOMIC reduction fused into
calc_new_currents; --policy selects EXACT (bit-identical to the oracle), RECIP
or FAST, and `detail.policies` times all four.

Printed: one JSON line (rank 0) with the metric, a `roofline` object for the
dominant kernel (calc_new_currents, FP32 VALU-bound) from HIP-event timing on
the launch stream, and a `cpu_baseline` object: the CPU oracle's restatement
of the reference task body timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector/matrix peak (public spec)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E (public spec)
FLOPS_PER_WIRE_STEP = 96   # (10*6 + 9*4), apps/circuit/circuit.cc:324
# FP32 flops each policy's step executes (an FMA counts 2), from its loop
# body's instruction mix, which the SQ_INSTS_VALU_FLOPS_FP32 counter confirms
# per launch (profiles/calc_new_currents_4M_wires_*_counters.md):
#   EXACT  the reference's 67 FP32 operations; its 29 divisions are FP64
#          multiplies by a per-wire reciprocal (29 FP64 flops, not counted here)
#   RECIP  10 x (subtract + 2 FMA) + 9 x (subtract + FMA) = 77
#   FAST   20 FMA + 8 add = 48
#   SCALED 10 FMA + 6 add, plus 10 multiplies per 4 steps = 28.5
EXECUTED_FLOPS = {"exact": 67, "recip": 77, "fast": 48, "scaled": 28.5}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed loops")
    ap.add_argument("--warmup", type=int, default=2, help="untimed loops")
    ap.add_argument("--policy", default="scaled", choices=["exact", "recip", "fast", "scaled"])
    ap.add_argument("--mode", default="atomic", choices=["atomic", "csr", "kahan64"])
    ap.add_argument("--workload", default="64x16384x65536",
                    choices=["4x1024x4096", "64x16384x65536"],
                    help="pieces x nodes/piece x wires/piece")
    ap.add_argument("--profile", default="reference", choices=["reference", "stable"],
                    help="generator profile of the timed circuit (stable: finite values)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the per-policy and per-kernel detail runs")
    ap.add_argument("--batched", action="store_true",
                    help="batched workload: 512 circuits in shared launches vs sequential")
    ap.add_argument("--served", action="store_true",
                    help="served workload: Poisson request stream with continuous batching")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-cap-wires", type=int, default=8 << 20)
    ap.add_argument("--serve-timeout", type=float, default=0.0,
                    help="served: admission deadline per request in seconds (0 = none)")
    ap.add_argument("--compare-seconds", type=float, default=2.0,
                    help="served: trace length of the batched-vs-sequential comparison")
    ap.add_argument("--upv-sweep", action="store_true",
                    help="update_voltages HBM bandwidth at growing node counts")
    ap.add_argument("--dsc-sweep", action="store_true",
                    help="distribute_charge bandwidth per reduction mode at growing sizes")
    ap.add_argument("--accessor", action="store_true",
                    help="the drop-in accessor form at 64x16384x65536: per-piece synchronous calls on "
                         "reference-layout instances")
    ap.add_argument("--gen", action="store_true",
                    help="circuit generation: host generator + upload vs the device generator")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run only the plain copy kernel over BYTES (profiler calibration)")
    return ap.parse_args()


WORKLOADS = {
    "4x1024x4096": dict(num_pieces=4, nodes_per_piece=1024, wires_per_piece=4096,
                        pct_wire_in_piece=95, random_seed=12345, loops=5),
    "64x16384x65536": dict(num_pieces=64, nodes_per_piece=16384, wires_per_piece=65536,
                           pct_wire_in_piece=95, random_seed=12345, loops=10),
}


def time_loops(C, ctx, loops, policy, mode):
    ctx.reset_timing()
    t0 = time.perf_counter()
    rep = ctx.run_loops(loops, policy=policy, mode=mode)
    wall = time.perf_counter() - t0
    k = {}
    for task in (C.CALC_NEW_CURRENTS, C.DISTRIBUTE_CHARGE, C.UPDATE_VOLTAGES):
        ms, n = ctx.kernel_time(task)
        k[task] = (ms / max(n, 1), n)
    return rep, wall, k


def cpu_baseline(C, params, seconds):
    """The oracle's calc_new_currents (restating circuit.cc:352-413) on a
    sample of the workload's wires, one wire-update per sampled wire."""
    import numpy as np
    from oracle import oracle as O
    host = C.generate(params)
    threads = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    # calibrate on a small sample, then size the timed sample to `seconds`
    probe = np.sort(rng.choice(len(host.wires), 64 * threads, replace=False)).astype(np.uint32)
    t = O.lib().orc_cnc_sample_f32(0, O.STEPS, threads, O._p(host.wires), O._p(host.nodes),
                                   O._p(probe), len(probe))
    per_wire = t / len(probe)
    count = int(max(len(probe), min(len(host.wires), seconds / per_wire)))
    idx = np.sort(rng.choice(len(host.wires), count, replace=False)).astype(np.uint32)
    t = O.lib().orc_cnc_sample_f32(0, O.STEPS, threads, O._p(host.wires), O._p(host.nodes),
                                   O._p(idx), len(idx))
    return {"value": count / t, "unit": "wire-updates/s", "cores": threads, "kind": "port",
            "sample": (f"calc_new_currents (10000 inner steps, FP32, reference operation order) "
                       f"of {count} randomly sampled wires of the workload, one loop, "
                       f"{threads} threads (one chunk of 16 wires per task); "
                       f"distribute_charge/update_voltages (<0.01% of the work) not timed; "
                       f"{t:.1f} s wall"),
            "seconds": t}


def batched_shapes(C, n=512, lo=2**10, hi=2**18, seed=4, profile=None):
    """The batched workload: wires log-uniform in [lo, hi], nodes = wires/4,
    pieces = ceil(nodes/4096), 95% in-piece, seeds 1000+i."""
    import numpy as np
    rng = np.random.default_rng(seed)
    prof = C.PROFILE_REFERENCE if profile is None else profile
    return [C.shape_for_wires(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))), 95, 1000 + i, prof)
            for i in range(n)]


def policy_of(C, name):
    return {"exact": C.POLICY_EXACT, "recip": C.POLICY_RECIP, "fast": C.POLICY_FAST,
            "scaled": C.POLICY_SCALED}[name]


def run_batched(C, args):
    loops = 5
    pol = policy_of(C, args.policy)
    shapes = batched_shapes(C)
    W_total = sum(s.num_pieces * s.wires_per_piece for s in shapes)
    b = C.Batch(0)
    t0 = time.perf_counter()
    b.add_generated_many(shapes, [loops] * len(shapes), threads=min(16, os.cpu_count() or 1))
    log(f"[batched] {len(shapes)} circuits, {W_total} wires, added in "
        f"{time.perf_counter() - t0:.1f} s")
    best, probe_ms = b.autotune(pol, probe_steps=500)
    log(f"[batched] autotune: block={best.block_threads} persistent={best.persistent} "
        f"waves/SIMD={best.waves_per_simd} ({probe_ms:.2f} ms per 500-step probe)")
    best.policy, best.reduce_mode, best.use_graph = pol, C.REDUCE_ATOMIC, 1
    b.set_timing(True)
    rep = b.run(best)
    b.close()
    log(f"[batched] {rep}")
    # Sequential launches, each circuit alone:
    #  (a) one launch per piece per task with a wait after each (how the
    #      reference runtime runs GPU point tasks);
    #  (b) plain launches: one per task over all pieces, no fusion, no graph;
    #  (c) this library's own loop driver: fused epilogue, graph-captured.
    seq_piece = seq_plain = seq_task = 0.0
    ctx = C.Context(0)
    for i, s in enumerate(shapes):
        ctx.generate(s)
        ctx.set_option(C.OPT_FUSE_DSC, 0)
        seq_piece += ctx.run_loops(loops, policy=pol, mode=C.REDUCE_ATOMIC,
                                   per_piece=True)["elapsed_s"]
        seq_plain += ctx.run_loops(loops, policy=pol,
                                   mode=C.REDUCE_ATOMIC)["elapsed_s"]
        ctx.set_option(C.OPT_FUSE_DSC, 1)
        seq_task += ctx.run_loops(loops, policy=pol, mode=C.REDUCE_ATOMIC,
                                  use_graph=True)["elapsed_s"]
        if i % 64 == 63:
            log(f"[batched] sequential: {i + 1} circuits")
    ctx.close()
    wu = W_total * loops
    out = {"metric": "batched wire-updates/s (512 circuits, one launch per task)",
           "value": rep["wire_updates_per_s"], "unit": "wire-updates/s", "n_gpus": 1,
           "dtype": "f32", "data": "synthetic (load_circuit generator, seeds 1000+i)",
           "config": {"workload": "512 circuits, wires log-uniform in [2^10, 2^18], nodes = wires/4, "
                                  "pieces = ceil(nodes/4096), 95% in-piece, 10000 inner steps, "
                                  f"5 loops each, {args.policy.upper()} policy, ATOMIC reduction, "
                                  "hipGraph per loop",
                      "total_wires": W_total,
                      "autotuned": {"block_threads": best.block_threads,
                                    "persistent": best.persistent,
                                    "waves_per_simd": best.waves_per_simd}},
           "batched_elapsed_s": rep["elapsed_s"],
           "sequential_per_piece": {"elapsed_s": seq_piece, "wire_updates_per_s": wu / seq_piece},
           "sequential_plain": {"elapsed_s": seq_plain, "wire_updates_per_s": wu / seq_plain},
           "sequential_fused_graph": {"elapsed_s": seq_task, "wire_updates_per_s": wu / seq_task},
           "speedup_vs_sequential_per_piece": seq_piece / rep["elapsed_s"],
           "speedup_vs_sequential_plain": seq_plain / rep["elapsed_s"],
           "speedup_vs_sequential_fused_graph": seq_task / rep["elapsed_s"]}
    print(json.dumps(out))


def run_served(C, args):
    budget = 200 * 10**9
    pol = policy_of(C, args.policy)
    trace = C.poisson_trace(args.serve_rate, args.serve_seconds, seed=5,
                            timeout_s=args.serve_timeout)
    offered = sum(r.gen.num_pieces * r.gen.wires_per_piece * r.num_loops for r in trace)
    log(f"[served] {len(trace)} requests over {args.serve_seconds} s, "
        f"{offered:.3g} wire-updates offered ({offered / args.serve_seconds:.3g}/s)")
    s = C.Server(0, budget_bytes=budget, max_active_wires=args.serve_cap_wires,
                 policy=pol, mode=C.REDUCE_ATOMIC, gen_threads=12)
    res, rep = s.run_trace(trace)
    s.close()
    log(f"[served] {rep}")
    # Continuous batching vs one request at a time on a short trace.
    short = C.poisson_trace(args.serve_rate, args.compare_seconds, seed=6)
    cmp = {}
    for name, max_active in (("continuous", 0), ("sequential", 1)):
        s = C.Server(0, budget_bytes=budget, max_active_wires=args.serve_cap_wires,
                     max_active=max_active, policy=pol, mode=C.REDUCE_ATOMIC,
                     gen_threads=12)
        _, r2 = s.run_trace(short)
        s.close()
        cmp[name] = {k: r2[k] for k in ("wire_updates_per_s", "elapsed_s", "latency_p50_s",
                                        "latency_p99_s", "completed")}
        log(f"[served] {name}: {cmp[name]}")
    # One request at a time in arrival order, in two launch styles, with
    # service times measured on the GPU and queueing computed from the
    # arrival times:
    #  - plain: one launch per task over all of the request's pieces, no
    #    fused epilogue, no graph;
    #  - reference style: one launch per piece per task and a wait after each
    #    (how the reference runtime executes GPU point tasks).
    import numpy as np
    for name, per_piece in (("sequential_plain", False), ("sequential_reference_style", True)):
        ctx = C.Context(0)
        ctx.set_option(C.OPT_FUSE_DSC, 0)
        free_at, lat, wu = 0.0, [], 0.0
        for r in short:
            ctx.generate(r.gen)
            svc = ctx.run_loops(r.num_loops, policy=pol, mode=C.REDUCE_ATOMIC,
                                per_piece=per_piece)["elapsed_s"]
            start = max(r.arrival_s, free_at)
            free_at = start + svc
            lat.append(free_at - r.arrival_s)
            wu += r.gen.num_pieces * r.gen.wires_per_piece * r.num_loops
        ctx.close()
        span = free_at - (short[0].arrival_s if short else 0.0)
        cmp[name] = {
            "wire_updates_per_s": wu / span if span > 0 else 0.0, "elapsed_s": span,
            "latency_p50_s": float(np.percentile(lat, 50)) if lat else 0.0,
            "latency_p99_s": float(np.percentile(lat, 99)) if lat else 0.0,
            "completed": len(lat)}
        log(f"[served] {name}: {cmp[name]}")
    out = {"metric": "served wire-updates/s with p50/p99 request latency",
           "value": rep["wire_updates_per_s"], "unit": "wire-updates/s", "n_gpus": 1,
           "dtype": "f32", "data": "synthetic (Poisson arrivals, load_circuit generator)",
           "config": {"workload": f"Poisson {args.serve_rate:g} req/s for {args.serve_seconds:g} s; "
                                  "wires log-uniform in [2^12, 2^20], nodes = wires/4, "
                                  "pieces = ceil(nodes/4096), 95% in-piece, loops U[1,20], "
                                  f"{args.policy.upper()} policy, ATOMIC reduction",
                      "hbm_budget_bytes": budget, "max_active_wires": args.serve_cap_wires},
           "latency_p50_s": rep["latency_p50_s"], "latency_p99_s": rep["latency_p99_s"],
           "latency_mean_s": rep["latency_mean_s"], "latency_max_s": rep["latency_max_s"],
           "offered_wire_updates_per_s": rep["offered_wire_updates_per_s"],
           "completed": rep["completed"], "rejected": rep["rejected"],
           "timed_out": rep["timed_out"], "admission_timeout_s": args.serve_timeout,
           "iterations": rep["iterations"], "max_resident": rep["max_resident"],
           "elapsed_s": rep["elapsed_s"],
           "compare_short_trace": {
               "seconds": args.compare_seconds, **cmp,
               "speedup": cmp["continuous"]["wire_updates_per_s"]
               / max(cmp["sequential"]["wire_updates_per_s"], 1e-30),
               "speedup_vs_plain": cmp["continuous"]["wire_updates_per_s"]
               / max(cmp["sequential_plain"]["wire_updates_per_s"], 1e-30),
               "speedup_vs_reference_style": cmp["continuous"]["wire_updates_per_s"]
               / max(cmp["sequential_reference_style"]["wire_updates_per_s"], 1e-30)}}
    print(json.dumps(out))


def run_upv_sweep(C, args):
    rows = []
    for P, npp in ((64, 16384), (256, 65536), (1024, 65536), (2048, 65536)):
        N = P * npp
        ctx = C.Context(0)
        ctx.generate_device(C.GenParams(P, npp, 0, 95, 1, C.PROFILE_REFERENCE))
        pieces = ctx.pieces()
        ctx.update_voltages(pieces)
        ctx.synchronize()
        ctx.set_timing(True)
        reps = 20
        for _ in range(reps):
            ctx.update_voltages(pieces)
        ms, n = ctx.kernel_time(C.UPDATE_VOLTAGES)
        avg = ms / n
        gbs = 24.0 * N / (avg * 1e-3) / 1e9
        rows.append({"nodes": N, "avg_launch_ms": avg, "achieved_GBs": gbs,
                     "frac_of_8TBs": gbs / PEAK_HBM_GBS, "bytes_per_launch": 24 * N})
        log(f"[upv] {rows[-1]}")
        ctx.close()
    # Reference point on the same GPU: the library's plain copy kernel
    # (128-bit ordinary loads and stores) over 1 GiB, 20 back-to-back launches
    # timed around one synchronisation.
    import numpy as np
    nbytes, reps = 1 << 30, 20
    src = C.DeviceArray(np.zeros(nbytes // 4, np.float32))
    dst = C.DeviceArray(np.zeros(nbytes // 4, np.float32))
    C.copy_kernel(dst, src, nbytes)
    C.lib().ckt_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(reps):
        C.copy_kernel(dst, src, nbytes)
    C.lib().ckt_device_synchronize(0)
    copy_ms = (time.perf_counter() - t0) / reps * 1e3
    copy_gbs = 2.0 * nbytes / (copy_ms * 1e-3) / 1e9
    copy = {"bytes_moved": 2 * nbytes, "avg_launch_ms": copy_ms, "achieved_GBs": copy_gbs,
            "frac_of_8TBs": copy_gbs / PEAK_HBM_GBS}
    log(f"[upv] plain copy: {copy}")
    print(json.dumps({"metric": "update_voltages HBM bandwidth", "unit": "GB/s",
                      "algorithmic_bytes_per_node": 24, "rows": rows, "plain_copy_1GiB": copy}))


def run_dsc_sweep(C, args):
    """distribute_charge alone, per reduction mode, on circuits of 64/256/1024
    pieces of 16384 nodes and 65536 wires (95% in-piece).  Algorithmic bytes:
    ATOMIC streams 16 B per wire (I[0], I[9], in_ptr, out_ptr) and reads and
    writes each node's charge once (8 B per node); CSR reads 8 B per
    contribution (index entry, gathered current; 2 per wire) and 12 B per node
    (row offset, charge read and written)."""
    rows = []
    modes = (("atomic_global", C.REDUCE_ATOMIC, 1), ("atomic_lds", C.REDUCE_ATOMIC, 2),
             ("csr", C.REDUCE_CSR, 0), ("kahan64", C.REDUCE_CSR_KAHAN64, 0))
    for P in (64, 256, 1024):
        npp, wpp = 16384, 65536
        N, W = P * npp, P * wpp
        ctx = C.Context(0)
        ctx.generate_device(C.GenParams(P, npp, wpp, 95, 12345, C.PROFILE_STABLE))
        pieces = ctx.pieces()
        ctx.set_timing(True)
        for name, mode, impl in modes:
            ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, impl)
            ctx.distribute_charge(pieces, mode)
            ctx.synchronize()
            ctx.reset_timing()
            reps = 20
            for _ in range(reps):
                ctx.distribute_charge(pieces, mode)
            ms, n = ctx.kernel_time(C.DISTRIBUTE_CHARGE)
            avg = ms / n
            nbytes = 16 * W + 8 * N if mode == C.REDUCE_ATOMIC else 16 * W + 12 * N
            gbs = nbytes / (avg * 1e-3) / 1e9
            rows.append({"wires": W, "nodes": N, "mode": name, "avg_launch_ms": avg,
                         "algorithmic_bytes": nbytes, "achieved_GBs": gbs,
                         "frac_of_8TBs": gbs / PEAK_HBM_GBS})
            log(f"[dsc] {rows[-1]}")
        ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, 0)
        ctx.close()
    print(json.dumps({"metric": "distribute_charge bandwidth per reduction mode", "unit": "GB/s",
                      "rows": rows}))


def run_accessor(C, args):
    """The three tasks through the accessor-form entry points -- what the
    reference's GPU task wrappers call (circuit.cc:475-550) -- on the 64-piece
    circuit (4.19 M wires) in the reference layouts (128-byte wires, 16-byte
    nodes) in device memory: one call per piece per task, each waiting for its
    kernel, as the reference runtime runs GPU point tasks; and the index-space
    accessor entry points, one call per task over all pieces.  On one GPU the
    private, shared and ghost node instances are one array; shared and ghost
    charges go to fold buffers that are applied once per loop (the runtime's
    reduction-instance apply).
    Beside it, the resident context in its per-piece mode and in its normal
    one-launch-per-task mode, on the same circuit."""
    import numpy as np
    wl = dict(WORKLOADS["64x16384x65536"])
    wl.pop("loops")
    policy = policy_of(C, args.policy)
    params = C.GenParams(**wl, profile=C.PROFILE_REFERENCE)
    host = C.generate(params)
    N, W = len(host.nodes), len(host.wires)
    dw = C.DeviceArray(host.wires)
    nodes = C.DeviceArray(host.nodes)
    loc = C.DeviceArray(host.locator)
    ofold = C.DeviceArray(np.zeros(N, np.float32))
    gfold = C.DeviceArray(np.zeros(N, np.float32))
    pieces = list(host.pieces)

    def loop():
        for p in pieces:
            C.calc_new_currents_gpu(p, dw, nodes, nodes, nodes, flag=policy)
        C.fold_init_gpu(ofold, 0, N)
        C.fold_init_gpu(gfold, 0, N)
        for p in pieces:
            C.distribute_charge_gpu(p, dw, nodes, ofold, gfold)
        C.fold_apply_gpu(nodes, ofold, 0, N)
        C.fold_apply_gpu(nodes, gfold, 0, N)
        for p in pieces:
            C.update_voltages_gpu(p, nodes, nodes, loc)

    def loop_index():
        C.calc_new_currents_index_gpu(pieces, dw, nodes, nodes, nodes, flag=policy)
        C.fold_init_gpu(ofold, 0, N)
        C.fold_init_gpu(gfold, 0, N)
        C.distribute_charge_index_gpu(pieces, dw, nodes, ofold, gfold)
        C.fold_apply_gpu(nodes, ofold, 0, N)
        C.fold_apply_gpu(nodes, gfold, 0, N)
        C.update_voltages_index_gpu(pieces, nodes, nodes, loc)

    loops = 3
    rows = {}
    for name, fn in (("accessor_form_per_piece_sync", loop),
                     ("accessor_form_index_space", loop_index)):
        fn()
        t0 = time.perf_counter()
        for _ in range(loops):
            fn()
        rows[name] = W * loops / (time.perf_counter() - t0)
        log(f"[accessor] {name}: {rows[name]:.4g} wire-updates/s")
    ctx = C.Context(0)
    ctx.generate(params)
    ctx.set_option(C.OPT_FUSE_DSC, 0)
    ctx.run_loops(1, policy=policy, mode=C.REDUCE_ATOMIC, per_piece=True)
    rows["context_per_piece_sync"] = ctx.run_loops(
        loops, policy=policy, mode=C.REDUCE_ATOMIC, per_piece=True)["wire_updates_per_s"]
    ctx.set_option(C.OPT_FUSE_DSC, 1)
    ctx.run_loops(1, policy=policy, mode=C.REDUCE_ATOMIC)
    rows["context_index_space"] = ctx.run_loops(
        loops, policy=policy, mode=C.REDUCE_ATOMIC)["wire_updates_per_s"]
    ctx.close()
    log(f"[accessor] {rows}")
    print(json.dumps({"metric": "wire-updates/s at 64x16384x65536 by calling form", "unit": "wire-updates/s",
                      "policy": args.policy, "loops": loops, "rows": rows}))


def run_gen(C, args):
    rows = []
    for P, npp, wpp in ((64, 16384, 65536), (1024, 16384, 65536)):
        gp = C.GenParams(P, npp, wpp, 95, 12345, C.PROFILE_REFERENCE)
        row = {"pieces": P, "nodes": P * npp, "wires": P * wpp}
        for name, fn in (("host", "generate"), ("device", "generate_device")):
            ctx = C.Context(0)
            getattr(ctx, fn)(gp)  # first call pays one-time costs
            t0 = time.perf_counter()
            getattr(ctx, fn)(gp)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            row[name + "_s"] = dt
            row[name + "_wires_per_s"] = P * wpp / dt
            ctx.close()
        row["speedup"] = row["host_s"] / row["device_s"]
        rows.append(row)
        log(f"[gen] {row}")
    print(json.dumps({"metric": "circuit generation (load_circuit) throughput", "unit": "wires/s",
                      "host": "generator on one host thread + slab upload",
                      "device": "start-state chain on the host, draws/classification/index on the GPU",
                      "rows": rows}))


def run_copy_calibration(C, nbytes):
    import numpy as np
    src = C.DeviceArray(np.zeros(nbytes // 4, np.float32))
    dst = C.DeviceArray(np.zeros(nbytes // 4, np.float32))
    for _ in range(5):
        C.copy_kernel(dst, src, nbytes)
    C.lib().ckt_device_synchronize(0)
    print(json.dumps({"copy_calibration_bytes": nbytes, "dispatches": 5}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from legionorigins_amd import circuit as C
    if args.copy_calibration:
        return run_copy_calibration(C, args.copy_calibration)
    if args.batched:
        return run_batched(C, args)
    if args.served:
        return run_served(C, args)
    if args.upv_sweep:
        return run_upv_sweep(C, args)
    if args.gen:
        return run_gen(C, args)
    if args.dsc_sweep:
        return run_dsc_sweep(C, args)
    if args.accessor:
        return run_accessor(C, args)

    policy = policy_of(C, args.policy)
    mode = {"atomic": C.REDUCE_ATOMIC, "csr": C.REDUCE_CSR,
            "kahan64": C.REDUCE_CSR_KAHAN64}[args.mode]
    wl = dict(WORKLOADS[args.workload])
    wl.pop("loops")
    params = C.GenParams(**wl, profile={"reference": C.PROFILE_REFERENCE,
                                        "stable": C.PROFILE_STABLE}[args.profile])
    info = C.device_info(0)
    log(f"[bench] device {info}")
    ctx = C.Context(0)
    t0 = time.perf_counter()
    ctx.generate(params)
    N, W, P = ctx.sizes()
    log(f"[bench] generated+uploaded {W} wires / {N} nodes / {P} pieces in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx.set_timing(True)
    if args.warmup:
        ctx.run_loops(args.warmup, policy=policy, mode=mode)
    log(f"[bench] warmup {args.warmup} loops done; timing {args.steps} loops")
    rep, wall, k = time_loops(C, ctx, args.steps, policy, mode)
    value = W * args.steps / wall
    cnc_ms, cnc_n = k[C.CALC_NEW_CURRENTS]
    ef = EXECUTED_FLOPS[args.policy]
    achieved = ef * C.STEPS * W / (cnc_ms * 1e-3) / 1e12
    ref_equiv = FLOPS_PER_WIRE_STEP * C.STEPS * W / (cnc_ms * 1e-3) / 1e12
    log(f"[bench] {value:.4g} wire-updates/s; calc_new_currents {cnc_ms:.3f} ms/launch "
        f"({achieved:.1f} TFLOP/s executed, {ref_equiv:.1f} reference-equivalent)")
    out = {
        "metric": "wire-updates/s at 1 GPU (calc_new_currents + distribute_charge + "
                  "update_voltages per loop)",
        "value": value,
        "unit": "wire-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (load_circuit generator, seed 12345, {args.profile} profile)",
        "config": {"workload": f"{P} pieces x {N // P} nodes x {W // P} wires/piece "
                               f"({W} wires, {N} nodes), 95% in-piece, seed 12345, "
                               f"WIRE_SEGMENTS=10, 10000 inner steps/loop, FP32, SoA",
                   "policy": args.policy, "reduction": args.mode,
                   "parallelism": "single GPU"},
        # calc_new_currents is FP32 vector-ALU bound.  `achieved` is the FP32
        # work the kernel executes (EXECUTED_FLOPS per wire-step x 10000 steps
        # x wires) over its mean launch time from HIP events on the launch
        # stream; `reference_equivalent` is the reference's own count (96 per
        # wire-step, circuit.cc:320-334) over the same time, which exceeds the
        # peak for policies that reach the same values with fewer operations.
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "calc_new_currents",
                     "executed_flops_per_wire_step": ef,
                     "flops_per_launch": ef * C.STEPS * W,
                     "avg_launch_ms": cnc_ms, "launches": cnc_n,
                     "reference_equivalent": {
                         "flops_per_wire_step": FLOPS_PER_WIRE_STEP, "tflops": ref_equiv,
                         "ratio_to_peak": ref_equiv / PEAK_FP32_TFLOPS}},
    }
    tp = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            t = json.load(f)
        key = f"calc_new_currents/{args.workload}/{args.policy}"
        if key in t:
            out["roofline"]["traffic"] = t[key]["bytes_per_launch"]
            out["roofline"]["traffic_source"] = ("profiles/traffic.json: rocprofv3 --pmc "
                                                 "FETCH_SIZE and WRITE_SIZE of this launch, "
                                                 "scaled by a plain copy's counters")
    detail = {"kernels_ms": {"calc_new_currents": cnc_ms,
                             "distribute_charge": k[C.DISTRIBUTE_CHARGE][0],
                             "update_voltages": k[C.UPDATE_VOLTAGES][0]},
              "gflops_reference_formula": rep["gflops"]}
    if not args.no_extra:
        pol = {}
        for name, pv in (("exact", C.POLICY_EXACT), ("recip", C.POLICY_RECIP),
                         ("fast", C.POLICY_FAST), ("scaled", C.POLICY_SCALED)):
            loops = 2 if pv == C.POLICY_EXACT else 4
            r2, w2, k2 = time_loops(C, ctx, loops, pv, mode)
            ms = k2[C.CALC_NEW_CURRENTS][0]
            ex = EXECUTED_FLOPS[name] * C.STEPS * W / ms / 1e9
            pol[name] = {"wire_updates_per_s": W * loops / w2, "cnc_ms": ms,
                         "executed_flops_per_wire_step": EXECUTED_FLOPS[name],
                         "tflops_executed": ex, "frac_of_fp32_peak": ex / PEAK_FP32_TFLOPS,
                         "tflops_reference_equivalent":
                             FLOPS_PER_WIRE_STEP * C.STEPS * W / ms / 1e9}
            log(f"[bench] policy {name}: {pol[name]}")
        detail["policies"] = pol
        red = {}
        for name, mv, impl in (("atomic_auto", C.REDUCE_ATOMIC, 0),
                               ("atomic_global", C.REDUCE_ATOMIC, 1),
                               ("atomic_lds", C.REDUCE_ATOMIC, 2), ("csr", C.REDUCE_CSR, 0),
                               ("kahan64", C.REDUCE_CSR_KAHAN64, 0)):
            ctx.set_option(C.OPT_STEPS, 1)
            ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, impl)
            ctx.set_option(C.OPT_FUSE_DSC, 0)
            r3, w3, k3 = time_loops(C, ctx, 5, C.POLICY_FAST, mv)
            ctx.set_option(C.OPT_STEPS, C.STEPS)
            ctx.set_option(C.OPT_DSC_ATOMIC_IMPL, 0)
            ctx.set_option(C.OPT_FUSE_DSC, 1)
            red[name] = {"distribute_charge_ms": k3[C.DISTRIBUTE_CHARGE][0]}
        # the full loop with distribute_charge as its own launch
        ctx.set_option(C.OPT_FUSE_DSC, 0)
        r4, w4, k4 = time_loops(C, ctx, 4, policy, mode)
        ctx.set_option(C.OPT_FUSE_DSC, 1)
        red["unfused_loop"] = {"wire_updates_per_s": W * 4 / w4,
                               "calc_new_currents_ms": k4[C.CALC_NEW_CURRENTS][0],
                               "distribute_charge_ms": k4[C.DISTRIBUTE_CHARGE][0]}
        # full loops with the deterministic reductions (their own launch)
        for name, mv in (("csr_loop", C.REDUCE_CSR), ("kahan64_loop", C.REDUCE_CSR_KAHAN64)):
            r6, w6, k6 = time_loops(C, ctx, 4, policy, mv)
            red[name] = {"wire_updates_per_s": W * 4 / w6,
                         "calc_new_currents_ms": k6[C.CALC_NEW_CURRENTS][0],
                         "distribute_charge_ms": k6[C.DISTRIBUTE_CHARGE][0]}
        detail["distribute_charge"] = red
        log(f"[bench] reductions: {red}")
        # The same workload in the stable profile: the reference parameters
        # overflow to inf/NaN within ~10 inner steps, the stable ones stay
        # finite, so this row shows the rate on finite data.
        ctx_s = C.Context(0)
        ctx_s.generate_device(C.GenParams(**wl, profile=C.PROFILE_STABLE))
        ctx_s.set_timing(True)
        ctx_s.run_loops(1, policy=policy, mode=mode)
        r5, w5, k5 = time_loops(C, ctx_s, 4, policy, mode)
        ms = k5[C.CALC_NEW_CURRENTS][0]
        detail["stable_profile"] = {
            "wire_updates_per_s": W * 4 / w5, "cnc_ms": ms,
            "tflops_executed": ef * C.STEPS * W / ms / 1e9,
            "tflops_reference_equivalent": FLOPS_PER_WIRE_STEP * C.STEPS * W / ms / 1e9}
        log(f"[bench] stable profile: {detail['stable_profile']}")
        ctx_s.close()
    out["detail"] = detail
    ctx.close()
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cb = cpu_baseline(C, params, args.cpu_seconds)
        out["cpu_baseline"] = {k: v for k, v in cb.items() if k != "seconds"}
        log(f"[bench] CPU baseline {cb['value']:.4g} wire-updates/s on {cb['cores']} threads")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
