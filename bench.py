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
dominant kernel (calc_new_currents, FP32 VALU-bound: executed flops over the
HIP-event time on the launch stream), a `cpu_baseline` object (the CPU
oracle's restatement of the reference task bodies, one loop over whole pieces,
one piece per host thread) and `detail`: per-policy and per-reduction timings,
the batched workload (512 circuits in shared launches against sequential
launches), the served workload (Poisson 200 requests/s for 60 s: p50/p99
latency and throughput against one request at a time with plain launches),
the FP64 error of each workload, and the rate of each calling form (the
accessor-form drop-ins per piece and per index space, the resident context).
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
# per launch (profiles/calc_new_currents_4M_wires_policy_counters.md):
#   EXACT  the reference's 67 FP32 operations (48 adds, 19 multiplies), and
#          its 29 correctly rounded divisions as two FMAs each (a split
#          reciprocal of the divisor): 67 + 29 x 4 = 183
#   RECIP  10 x (subtract + 2 FMA) + 9 x (subtract + FMA) = 77
#   FAST   20 FMA + 8 add = 48
#   SCALED 10 FMA + 6 add, plus 10 multiplies per 4 steps = 28.5
EXECUTED_FLOPS = {"exact": 183, "recip": 77, "fast": 48, "scaled": 28.5}
# baseline.cu on one Tesla M2090, 48 x 2500 x 16384, 10 loops, 181.033 s
# (apps/circuit/keeneland_results/baseline_48_1_1_1.stdio).
BASELINE_WIRE_UPDATES_PER_S = 48 * 16384 * 10 / 181.033


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed loops")
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed loops (the clock settles over the first few launches)")
    ap.add_argument("--policy", default="scaled", choices=["exact", "recip", "fast", "scaled"])
    ap.add_argument("--mode", default="atomic", choices=["atomic", "csr", "kahan64"])
    ap.add_argument("--workload", default="64x16384x65536",
                    choices=["4x1024x4096", "64x16384x65536"],
                    help="pieces x nodes/piece x wires/piece")
    ap.add_argument("--profile", default="reference", choices=["reference", "stable"],
                    help="generator profile of the timed circuit (stable: finite values)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip every detail run (policies, reductions, batched, served, FP64)")
    ap.add_argument("--skip-batched", action="store_true", help="detail without the batched run")
    ap.add_argument("--skip-served", action="store_true", help="detail without the served run")
    ap.add_argument("--skip-fp64", action="store_true", help="detail without the FP64 errors")
    ap.add_argument("--skip-accessor", action="store_true",
                    help="detail without the calling-form rates (accessor-form drop-ins)")
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
    ap.add_argument("--dsc-pieces", default="64,256,1024",
                    help="dsc sweep: circuit sizes in pieces of 16384 nodes x 65536 wires")
    ap.add_argument("--dsc-modes", default=None,
                    help="dsc sweep: comma-separated subset of atomic_global,atomic_lds,csr,kahan64")
    ap.add_argument("--dsc-reps", type=int, default=20, help="dsc sweep: timed launches per mode")
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


def host_threads():
    """Every core in this process's affinity mask, capped by OMP_NUM_THREADS
    where the environment sets it (the share of the host's CPUs given to
    this process)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def cpu_baseline(C, params):
    """One loop of the three task bodies, as the oracle restates them
    (circuit.cc:352-460; loop order :232-307), over whole pieces of the
    workload, one piece per thread (the reference's SharedMapper puts one
    point task on each CPU, circuit.cc:623-635): calc_new_currents on
    min(threads, pieces) pieces in parallel, then distribute_charge and
    update_voltages over those pieces."""
    import numpy as np
    from oracle import oracle as O
    host = C.generate(params)
    threads = host_threads()
    k = min(threads, len(host.pieces))
    pieces = np.ascontiguousarray(host.pieces[:k])
    W = int(pieces["num_wires"].sum())
    t0 = time.perf_counter()
    O.lib().orc_run_f32(0, O.STEPS, 1, 0, threads, 0, O._p(host.nodes), len(host.nodes),
                        O._p(host.wires), len(host.wires), O._p(host.locator), O._p(pieces), k)
    t = time.perf_counter() - t0
    return {"value": W / t, "unit": "wire-updates/s", "cores": threads, "kind": "port",
            "sample": (f"one loop of calc_new_currents (10000 inner steps, FP32, reference "
                       f"operation order), distribute_charge and update_voltages over {k} of "
                       f"the workload's {len(host.pieces)} pieces ({W} wires), one piece per "
                       f"thread on {threads} threads (affinity mask: "
                       f"{len(os.sched_getaffinity(0))} CPUs); {t:.1f} s wall"),
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


def batched_measure(C, policy_name, loops=5):
    """The batched workload in shared launches against the same circuits run
    one after another."""
    pol = policy_of(C, policy_name)
    shapes = batched_shapes(C)
    W_total = sum(s.num_pieces * s.wires_per_piece for s in shapes)
    b = C.Batch(0)
    t0 = time.perf_counter()
    b.add_generated_many(shapes, [loops] * len(shapes), threads=min(16, host_threads()))
    log(f"[batched] {len(shapes)} circuits, {W_total} wires, added in "
        f"{time.perf_counter() - t0:.1f} s")
    # The library's default schedule (256-thread persistent workgroups, 4
    # resident waves per SIMD, one tile per 256 wires of a piece, pieces in
    # order) against the autotuned one, as probes and as whole runs.
    default = C.Batch.config(policy=pol, mode=C.REDUCE_ATOMIC, block_threads=256, persistent=1,
                             waves_per_simd=4, wires_per_lane=1, tile_span=0, tile_order=0)
    default_probe_ms = b.probe(default, probe_steps=2000)
    best, _ = b.autotune(pol, probe_steps=2000)
    best_probe_ms = b.probe(best, probe_steps=2000)
    log(f"[batched] autotune: {best.describe()} ({best_probe_ms:.2f} ms per 2000-step probe, "
        f"default schedule {default_probe_ms:.2f} ms)")
    best.policy, best.reduce_mode, best.use_graph = pol, C.REDUCE_ATOMIC, 1
    b.set_timing(True)
    rep = b.run(best)
    b.close()
    log(f"[batched] {rep}")
    b = C.Batch(0)
    b.add_generated_many(shapes, [loops] * len(shapes), threads=min(16, host_threads()))
    rep_default = b.run(default)
    b.close()
    log(f"[batched] default schedule: {rep_default}")
    # Sequential launches, each circuit alone:
    #  (a) one launch per piece per task with a wait after each (how the
    #      reference runtime runs GPU point tasks);
    #  (b) plain launches: one per task over all pieces, no fusion, no graph;
    #  (c) this library's own loop driver: fused epilogue, graph-captured.
    seq_piece = seq_plain = seq_task = 0.0
    ctx = C.Context(0)
    for i, s in enumerate(shapes):
        ctx.generate_device(s)
        ctx.set_option(C.OPT_FUSE_DSC, 0)
        seq_piece += ctx.run_loops(loops, policy=pol, mode=C.REDUCE_ATOMIC,
                                   per_piece=True)["elapsed_s"]
        seq_plain += ctx.run_loops(loops, policy=pol,
                                   mode=C.REDUCE_ATOMIC)["elapsed_s"]
        ctx.set_option(C.OPT_FUSE_DSC, 1)
        seq_task += ctx.run_loops(loops, policy=pol, mode=C.REDUCE_ATOMIC,
                                  use_graph=True)["elapsed_s"]
        if i % 128 == 127:
            log(f"[batched] sequential: {i + 1} circuits")
    ctx.close()
    wu = W_total * loops
    return {"workload": "512 circuits, wires log-uniform in [2^10, 2^18], nodes = wires/4, "
                        "pieces = ceil(nodes/4096), 95% in-piece, 10000 inner steps, "
                        f"{loops} loops each, {policy_name.upper()} policy, ATOMIC reduction, "
                        "hipGraph per loop",
            "total_wires": W_total,
            "autotuned": best.describe(),
            "wire_updates_per_s": rep["wire_updates_per_s"], "elapsed_s": rep["elapsed_s"],
            "default_schedule": {"config": default.describe(),
                                 "wire_updates_per_s": rep_default["wire_updates_per_s"],
                                 "elapsed_s": rep_default["elapsed_s"],
                                 "probe_ms_2000_steps": default_probe_ms},
            "autotuned_probe_ms_2000_steps": best_probe_ms,
            "autotune_gain": rep_default["elapsed_s"] / rep["elapsed_s"],
            "sequential_per_piece": {"elapsed_s": seq_piece, "wire_updates_per_s": wu / seq_piece},
            "sequential_plain": {"elapsed_s": seq_plain, "wire_updates_per_s": wu / seq_plain},
            "sequential_fused_graph": {"elapsed_s": seq_task, "wire_updates_per_s": wu / seq_task},
            "speedup_vs_sequential_per_piece": seq_piece / rep["elapsed_s"],
            "speedup_vs_sequential_plain": seq_plain / rep["elapsed_s"],
            "speedup_vs_sequential_fused_graph": seq_task / rep["elapsed_s"]}


def run_batched(C, args):
    d = batched_measure(C, args.policy)
    out = {"metric": "batched wire-updates/s (512 circuits, one launch per task)",
           "value": d["wire_updates_per_s"], "unit": "wire-updates/s", "n_gpus": 1,
           "dtype": "f32", "data": "synthetic (load_circuit generator, seeds 1000+i)",
           "config": {"workload": d["workload"], "total_wires": d["total_wires"],
                      "autotuned": d["autotuned"]},
           **{k: v for k, v in d.items() if k not in ("workload", "total_wires", "autotuned")}}
    print(json.dumps(out))


def plain_loop_times(C, pol, sizes, profile):
    """Seconds per loop of one circuit alone with plain launches (one per task
    over all its pieces, no fused epilogue, no graph), per wire count."""
    out = []
    ctx = C.Context(0)
    ctx.set_option(C.OPT_FUSE_DSC, 0)
    for W in sizes:
        ctx.generate_device(C.shape_for_wires(int(W), 95, 77, profile))
        ctx.run_loops(1, policy=pol, mode=C.REDUCE_ATOMIC)
        out.append(ctx.run_loops(3, policy=pol, mode=C.REDUCE_ATOMIC)["elapsed_s"] / 3)
    ctx.close()
    return out


def fifo(arrivals, service):
    """One request at a time in arrival order: (latencies, end of the last)."""
    free, lat = 0.0, []
    for a, svc in zip(arrivals, service):
        free = max(a, free) + svc
        lat.append(free - a)
    return lat, free


def served_measure(C, policy_name, rate, seconds, cap_wires, timeout_s=0.0,
                   compare_seconds=2.0):
    """The served workload: a Poisson request stream served by continuous
    batching, with p50/p99 latency; against one request at a time with plain
    launches over the same trace."""
    import numpy as np
    budget = 200 * 10**9
    pol = policy_of(C, policy_name)
    trace = C.poisson_trace(rate, seconds, seed=5, timeout_s=timeout_s)
    work = [r.gen.num_pieces * r.gen.wires_per_piece * r.num_loops for r in trace]
    log(f"[served] {len(trace)} requests over {seconds} s, {sum(work):.3g} wire-updates "
        f"offered ({sum(work) / seconds:.3g}/s)")
    s = C.Server(0, budget_bytes=budget, max_active_wires=cap_wires, policy=pol,
                 mode=C.REDUCE_ATOMIC, gen_threads=12)
    _, rep = s.run_trace(trace)
    s.close()
    log(f"[served] {rep}")
    # One request at a time with plain launches, over the same trace: the
    # per-loop time of a circuit alone, measured at 17 wire counts from 2^12
    # to 2^20 and interpolated in log(wires), times each request's loops;
    # queueing in arrival order.  GPU time only (no generation).
    sizes = np.exp2(np.arange(12.0, 20.01, 0.5))
    t_loop = np.array(plain_loop_times(C, pol, sizes, C.PROFILE_REFERENCE))
    wires = np.array([r.gen.num_pieces * r.gen.wires_per_piece for r in trace], float)
    svc = np.interp(np.log2(wires), np.log2(sizes), t_loop) * np.array([r.num_loops for r in trace])
    lat, end = fifo([r.arrival_s for r in trace], svc)
    span = end - (trace[0].arrival_s if trace else 0.0)
    seq = {"wire_updates_per_s": sum(work) / span, "elapsed_s": span,
           "latency_p50_s": float(np.percentile(lat, 50)),
           "latency_p99_s": float(np.percentile(lat, 99)),
           "per_loop_s_by_wires": {str(int(w)): t for w, t in zip(sizes, t_loop)}}
    # The most any batching of these requests can gain over that: all of
    # their work at the rate one circuit that fills the GPU runs at in the
    # library's own loop (4,194,304 wires, fused epilogue, graph-captured).
    ctx = C.Context(0)
    ctx.generate_device(C.shape_for_wires(1 << 22, 95, 78, C.PROFILE_REFERENCE))
    ctx.run_loops(1, policy=pol, mode=C.REDUCE_ATOMIC, use_graph=True)
    full = ctx.run_loops(3, policy=pol, mode=C.REDUCE_ATOMIC, use_graph=True)
    ctx.close()
    r_full = full["wire_updates_per_s"]
    bound = float(svc.sum() / (sum(work) / r_full))
    log(f"[served] sequential plain (full trace): {seq['wire_updates_per_s']:.3g} wire-updates/s; "
        f"batching bound {bound:.2f}x")
    # Cross-check of that model on a short trace, measured directly.
    short = C.poisson_trace(rate, compare_seconds, seed=6)
    cmp = {}
    s = C.Server(0, budget_bytes=budget, max_active_wires=cap_wires, policy=pol,
                 mode=C.REDUCE_ATOMIC, gen_threads=12)
    _, r2 = s.run_trace(short)
    s.close()
    cmp["continuous"] = {k: r2[k] for k in ("wire_updates_per_s", "elapsed_s", "latency_p50_s",
                                            "latency_p99_s", "completed")}
    ctx = C.Context(0)
    ctx.set_option(C.OPT_FUSE_DSC, 0)
    svc_short = []
    for r in short:
        ctx.generate_device(r.gen)
        svc_short.append(ctx.run_loops(r.num_loops, policy=pol,
                                       mode=C.REDUCE_ATOMIC)["elapsed_s"])
    ctx.close()
    w_short = np.array([r.gen.num_pieces * r.gen.wires_per_piece for r in short], float)
    model_short = np.interp(np.log2(w_short), np.log2(sizes), t_loop) * \
        np.array([r.num_loops for r in short])
    lat2, end2 = fifo([r.arrival_s for r in short], svc_short)
    cmp["sequential_plain_measured"] = {
        "wire_updates_per_s": float((w_short * [r.num_loops for r in short]).sum()
                                    / (end2 - short[0].arrival_s)),
        "latency_p50_s": float(np.percentile(lat2, 50)),
        "latency_p99_s": float(np.percentile(lat2, 99)),
        "service_s": float(np.sum(svc_short)), "service_s_model": float(model_short.sum())}
    log(f"[served] short trace: {cmp}")
    return {"workload": f"Poisson {rate:g} req/s for {seconds:g} s; wires log-uniform in "
                        "[2^12, 2^20], nodes = wires/4, pieces = ceil(nodes/4096), 95% in-piece, "
                        f"loops U[1,20], {policy_name.upper()} policy, ATOMIC reduction",
            "hbm_budget_bytes": budget, "max_active_wires": cap_wires,
            "requests": len(trace),
            "wire_updates_per_s": rep["wire_updates_per_s"],
            "offered_wire_updates_per_s": rep["offered_wire_updates_per_s"],
            "latency_p50_s": rep["latency_p50_s"], "latency_p99_s": rep["latency_p99_s"],
            "latency_mean_s": rep["latency_mean_s"], "latency_max_s": rep["latency_max_s"],
            "completed": rep["completed"], "rejected": rep["rejected"],
            "timed_out": rep["timed_out"], "iterations": rep["iterations"],
            "max_resident": rep["max_resident"], "elapsed_s": rep["elapsed_s"],
            "sequential_plain": seq,
            "speedup_vs_sequential_plain": rep["wire_updates_per_s"] / seq["wire_updates_per_s"],
            "full_gpu_rate": r_full,
            "batching_bound_vs_sequential_plain": bound,
            "short_trace_check": {"seconds": compare_seconds, **cmp}}


def run_served(C, args):
    d = served_measure(C, args.policy, args.serve_rate, args.serve_seconds,
                       args.serve_cap_wires, args.serve_timeout, args.compare_seconds)
    out = {"metric": "served wire-updates/s with p50/p99 request latency",
           "value": d["wire_updates_per_s"], "unit": "wire-updates/s", "n_gpus": 1,
           "dtype": "f32", "data": "synthetic (Poisson arrivals, load_circuit generator)",
           "config": {"workload": d["workload"], "hbm_budget_bytes": d["hbm_budget_bytes"],
                      "max_active_wires": d["max_active_wires"]},
           **{k: v for k, v in d.items()
              if k not in ("workload", "hbm_budget_bytes", "max_active_wires")}}
    print(json.dumps(out))


def fp64_errors(C, policy_name):
    """Maximum relative node-voltage error against the FP64 oracle (the
    relaxation's FP64 limit, circuit.cc:352-460 in double) per workload,
    stable profile (in the reference profile the values are inf/NaN)."""
    import numpy as np
    from oracle import oracle as O
    pol = policy_of(C, policy_name)
    cases = [("4x1024x4096, 5 loops", (4, 1024, 4096, 95), 5, True),
             ("4x1024x4096, 200 loops, 95% in-piece", (4, 1024, 4096, 95), 200, True),
             ("4x1024x4096, 200 loops, 50% in-piece", (4, 1024, 4096, 50), 200, True),
             ("64x16384x65536, 10 loops", (64, 16384, 65536, 95), 10, False)]
    rows = {}
    for name, (P, npp, wpp, pct), loops, all_variants in cases:
        truth = O.run_f64_limit(O.generate(P, npp, wpp, pct, 12345, O.PROFILE_STABLE), loops,
                                nthreads=host_threads())[0]["voltage"]
        variants = [(f"{policy_name}+atomic", pol, C.REDUCE_ATOMIC),
                    ("exact+csr", C.POLICY_EXACT, C.REDUCE_CSR)]
        if all_variants:
            variants.append(("exact+csr_kahan64", C.POLICY_EXACT, C.REDUCE_CSR_KAHAN64))
        row = {}
        for vname, pv, mode in variants:
            with C.Context(0) as ctx:
                ctx.generate_device(C.GenParams(P, npp, wpp, pct, 12345, C.PROFILE_STABLE))
                ctx.run_loops(loops, policy=pv, mode=mode)
                v = ctx.store_nodes()["voltage"].astype(np.float64)
            fin = np.isfinite(v) & np.isfinite(truth)
            row[vname] = float(np.max(np.abs(v[fin] - truth[fin]) /
                                      np.maximum(np.abs(truth[fin]), 1e-30)))
            row[vname + "_nonfinite"] = int((~np.isfinite(v)).sum())
        rows[name] = row
        log(f"[fp64] {name}: {row}")
    return rows


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
    modes = tuple(m for m in modes if args.dsc_modes is None or m[0] in args.dsc_modes.split(","))
    for P in (int(x) for x in args.dsc_pieces.split(",")):
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
            for _ in range(args.dsc_reps):
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


def accessor_measure(C, policy_name, loops=3):
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
    policy = policy_of(C, policy_name)
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
    return {"unit": "wire-updates/s", "policy": policy_name, "loops": loops, "rows": rows}


def run_accessor(C, args):
    d = accessor_measure(C, args.policy)
    print(json.dumps({"metric": "wire-updates/s at 64x16384x65536 by calling form", **d}))


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
        # The reference's hand-coded single-GPU CUDA program (baseline.cu) on
        # one Tesla M2090: 48 pieces x 2500 nodes x 16384 wires, 10 loops in
        # 181.033 s (apps/circuit/keeneland_results/baseline_48_1_1_1.stdio),
        # 43,441 wire-updates/s; a different circuit size, so context only.
        "vs_baseline": value / BASELINE_WIRE_UPDATES_PER_S,
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
    # The rest of the metric: the batched and served workloads, and the FP64
    # error of each workload.  A failure in one of them is recorded in its
    # place and does not stop the line.
    extra = [] if args.no_extra else [
        ("batched", not args.skip_batched, lambda: batched_measure(C, args.policy)),
        ("served", not args.skip_served,
         lambda: served_measure(C, args.policy, args.serve_rate, args.serve_seconds,
                                args.serve_cap_wires, args.serve_timeout, args.compare_seconds)),
        ("fp64_error", not args.skip_fp64, lambda: fp64_errors(C, args.policy)),
        ("calling_forms", not args.skip_accessor, lambda: accessor_measure(C, args.policy))]
    for key, on, fn in extra:
        if not on:
            continue
        try:
            detail[key] = fn()
        except Exception as e:  # recorded, the line is still printed
            detail[key] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] {key} failed: {e}")
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cb = cpu_baseline(C, params)
        out["cpu_baseline"] = {k: v for k, v in cb.items() if k != "seconds"}
        log(f"[bench] CPU baseline {cb['value']:.4g} wire-updates/s on {cb['cores']} threads")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
