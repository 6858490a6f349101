"""Full-size parity check, too slow for the test suite: the full-size circuit
(64 pieces x 16384 nodes x 65536 wires, 95% in-piece, seed 12345) for its 10
loops on the GPU against the oracle's FP32 restatement of the reference tasks
(circuit.cc:352-460, sequential reduction order) on the host, every node and
wire compared: EXACT + CSR bit for bit; SCALED, FAST and RECIP with the fused ATOMIC
reduction (the benchmark's configurations) by non-finite class and within 1e-5
(node voltages elementwise, wire currents and internal voltages per wire,
relative to the wire's largest segment value).  Both profiles.  The oracle
needs about 1.7 core-hours per profile (~7 minutes on 16 threads), so this
script is run by hand and pytest does not collect it; the collected 10-loop
check at this size is test_gpu_fullsize.py's loop-by-loop comparison, which
runs the oracle's task bodies on a wire sample and every node each loop.

    python tests/full_size_exact_vs_oracle.py [--profile reference|stable] [--loops N]
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import (TOL, assert_bit_identical, assert_same_class, max_rel_err,  # noqa: E402
                     max_wire_rel_err)
from legionorigins_amd import circuit as C  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile", default="stable", choices=["reference", "stable"])
    ap.add_argument("--loops", type=int, default=10)
    a = ap.parse_args()
    prof = C.PROFILE_STABLE if a.profile == "stable" else C.PROFILE_REFERENCE
    P, npp, wpp, pct, seed = 64, 16384, 65536, 95, 12345
    t0 = time.time()
    with C.Context(0) as ctx:
        ctx.generate(C.GenParams(P, npp, wpp, pct, seed, prof))
        ctx.run_loops(a.loops, policy=C.POLICY_EXACT, mode=C.REDUCE_CSR)
        nodes, wires = ctx.store()
    print(f"[full size] GPU EXACT+CSR {a.loops} loops done in {time.time() - t0:.1f} s", flush=True)
    fast = {}
    for name, policy in (("scaled+atomic", C.POLICY_SCALED), ("fast+atomic", C.POLICY_FAST),
                         ("recip+atomic", C.POLICY_RECIP)):
        with C.Context(0) as ctx:
            ctx.generate(C.GenParams(P, npp, wpp, pct, seed, prof))
            ctx.run_loops(a.loops, policy=policy, mode=C.REDUCE_ATOMIC)
            fast[name] = ctx.store()
        print(f"[full size] GPU {name} {a.loops} loops done", flush=True)
    ref = O.generate(P, npp, wpp, pct, seed, prof)
    t0 = time.time()
    for loop in range(a.loops):
        O.run(ref, 1, reduce=0)
        print(f"[full size] oracle loop {loop + 1}/{a.loops} at {time.time() - t0:.0f} s", flush=True)
    assert_bit_identical(nodes["voltage"], ref.nodes["voltage"], "node voltage")
    assert_bit_identical(nodes["charge"], ref.nodes["charge"], "node charge")
    assert_bit_identical(wires["current"], ref.wires["current"], "wire current")
    assert_bit_identical(wires["voltage"], ref.wires["voltage"], "wire voltage")
    nan = float(np.isnan(nodes["voltage"]).mean())
    print(f"[full size] {a.profile} profile, {a.loops} loops: EXACT+CSR bit-identical to the oracle on "
          f"{len(nodes)} nodes and {len(wires)} wires (NaN fraction of node voltages {nan:.4f})")
    for name, (fn, fw) in fast.items():
        assert_same_class(fn["voltage"], ref.nodes["voltage"], f"{name} node voltage")
        assert_same_class(fw["current"], ref.wires["current"], f"{name} wire current")
        assert_same_class(fw["voltage"], ref.wires["voltage"], f"{name} wire voltage")
        fin = np.isfinite(ref.wires["current"]).all(axis=1) & np.isfinite(ref.wires["voltage"]).all(axis=1)
        e_v = max_rel_err(fn["voltage"], ref.nodes["voltage"])
        e_i = max_wire_rel_err(fw["current"][fin], ref.wires["current"][fin]) if fin.any() else 0.0
        e_w = max_wire_rel_err(fw["voltage"][fin], ref.wires["voltage"][fin]) if fin.any() else 0.0
        print(f"[full size] {name}: same non-finite classes as the oracle on every node and wire; "
              f"max relative error: node voltage {e_v:.3e}, wire current {e_i:.3e}, "
              f"internal voltage {e_w:.3e} ({int(fin.sum())} finite wires)", flush=True)
        assert max(e_v, e_i, e_w) <= TOL, name


if __name__ == "__main__":
    main()
