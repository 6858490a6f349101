"""Benchmark of the circuit hot path on one MI355X.

Workload (default): the full-size plain circuit -- 64 pieces x 16384 nodes x
65536 wires/piece (4,194,304 wires, 1,048,576 nodes), 95% of wires inside
their piece, seed 12345, WIRE_SEGMENTS=10, 10000 inner steps per loop, FP32,
structure-of-arrays in HBM.  One bench step is one loop of the three tasks
(calc_new_currents, distribute_charge, update_voltages) over all pieces; the
metric is wire-updates/s = wires x loops / wall time, inputs resident in HBM.

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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed loops")
    ap.add_argument("--warmup", type=int, default=2, help="untimed loops")
    ap.add_argument("--policy", default="fast", choices=["exact", "recip", "fast"])
    ap.add_argument("--mode", default="atomic", choices=["atomic", "csr", "kahan64"])
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the per-policy and per-kernel detail runs")
    return ap.parse_args()


WORKLOADS = {
    "c1": dict(num_pieces=4, nodes_per_piece=1024, wires_per_piece=4096, pct_wire_in_piece=95,
               random_seed=12345, loops=5),
    "c2": dict(num_pieces=64, nodes_per_piece=16384, wires_per_piece=65536, pct_wire_in_piece=95,
               random_seed=12345, loops=10),
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from legionorigins_amd import circuit as C

    policy = {"exact": C.POLICY_EXACT, "recip": C.POLICY_RECIP, "fast": C.POLICY_FAST}[args.policy]
    mode = {"atomic": C.REDUCE_ATOMIC, "csr": C.REDUCE_CSR,
            "kahan64": C.REDUCE_CSR_KAHAN64}[args.mode]
    wl = dict(WORKLOADS[args.workload])
    wl.pop("loops")
    params = C.GenParams(**wl, profile=C.PROFILE_REFERENCE)
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
    achieved = FLOPS_PER_WIRE_STEP * C.STEPS * W / (cnc_ms * 1e-3) / 1e12
    log(f"[bench] {value:.4g} wire-updates/s; calc_new_currents {cnc_ms:.3f} ms/launch "
        f"({achieved:.1f} TFLOP/s algorithmic)")
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
        "data": "synthetic (load_circuit generator, seed 12345)",
        "config": {"workload": f"{P} pieces x {N // P} nodes x {W // P} wires/piece "
                               f"({W} wires, {N} nodes), 95% in-piece, seed 12345, "
                               f"WIRE_SEGMENTS=10, 10000 inner steps/loop, FP32, SoA",
                   "policy": args.policy, "reduction": args.mode,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "calc_new_currents",
                     "flops_per_launch": FLOPS_PER_WIRE_STEP * C.STEPS * W,
                     "avg_launch_ms": cnc_ms, "launches": cnc_n},
    }
    tp = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            t = json.load(f)
        key = f"calc_new_currents/{args.workload}/{args.policy}"
        if key in t:
            out["roofline"]["traffic"] = t[key]["bytes_per_launch"]
    detail = {"kernels_ms": {"calc_new_currents": cnc_ms,
                             "distribute_charge": k[C.DISTRIBUTE_CHARGE][0],
                             "update_voltages": k[C.UPDATE_VOLTAGES][0]},
              "gflops_reference_formula": rep["gflops"]}
    if not args.no_extra:
        pol = {}
        for name, pv in (("exact", C.POLICY_EXACT), ("recip", C.POLICY_RECIP),
                         ("fast", C.POLICY_FAST)):
            loops = 2 if pv == C.POLICY_EXACT else 4
            r2, w2, k2 = time_loops(C, ctx, loops, pv, mode)
            ms = k2[C.CALC_NEW_CURRENTS][0]
            pol[name] = {"wire_updates_per_s": W * loops / w2, "cnc_ms": ms,
                         "tflops_algorithmic": FLOPS_PER_WIRE_STEP * C.STEPS * W / ms / 1e9}
            log(f"[bench] policy {name}: {pol[name]}")
        detail["policies"] = pol
        red = {}
        for name, mv in (("atomic", C.REDUCE_ATOMIC), ("csr", C.REDUCE_CSR),
                         ("kahan64", C.REDUCE_CSR_KAHAN64)):
            ctx.set_option(C.OPT_STEPS, 1)
            r3, w3, k3 = time_loops(C, ctx, 5, C.POLICY_FAST, mv)
            ctx.set_option(C.OPT_STEPS, C.STEPS)
            red[name] = {"distribute_charge_ms": k3[C.DISTRIBUTE_CHARGE][0]}
        detail["distribute_charge"] = red
        log(f"[bench] reductions: {red}")
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
