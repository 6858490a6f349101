"""Precision study (4 pieces x 1024 nodes x 4096 wires, 200 loops, stable profile,
pct 95 and 50).

For each policy (EXACT, FAST, SCALED) and reduction (ATOMIC, CSR, CSR_KAHAN64) the
GPU's node voltages after 200 loops are compared with the FP64 oracle (the
relaxation's FP64 limit, pinned against the step-by-step FP64 relaxation in
tests/test_oracle_physics.py).  The FP64 error is reported per configuration
(run pytest with -rP to see the table).  EXACT+CSR is bit-identical to the
FP32 CPU restatement (tests/test_gpu_parity.py), so it stands in for it here:
every other configuration must stay within 1e-5 of it after 200 loops.
"""
import numpy as np
import pytest

from helpers import TOL, max_rel_err

pytestmark = pytest.mark.gpu

LOOPS = 200


@pytest.mark.parametrize("pct", [95, 50])
def test_precision_study_200_loops(C, O, pct):
    P, npp, wpp, seed = 4, 1024, 4096, 12345
    truth_n, _ = O.run_f64_limit(O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE), LOOPS)
    tv = truth_n["voltage"]
    shared = O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE).locator == 0
    rows = {}
    ref32 = None
    for policy, pname in ((C.POLICY_EXACT, "exact"), (C.POLICY_FAST, "fast"),
                          (C.POLICY_SCALED, "scaled")):
        for mode, mname in ((C.REDUCE_CSR, "csr"), (C.REDUCE_ATOMIC, "atomic"),
                            (C.REDUCE_CSR_KAHAN64, "csr_kahan64")):
            with C.Context(0) as ctx:
                ctx.generate(C.GenParams(P, npp, wpp, pct, seed, C.PROFILE_STABLE))
                ctx.run_loops(LOOPS, policy=policy, mode=mode, use_graph=True)
                nodes, _ = ctx.store()
            v = nodes["voltage"]
            assert np.isfinite(v).all()
            if ref32 is None:
                ref32 = v.copy()  # EXACT + CSR
            rows[f"{pname}+{mname}"] = {
                "fp64_err_all": max_rel_err(v, tv),
                "fp64_err_shared": max_rel_err(v[shared], tv[shared]),
                "err_vs_fp32_reference": max_rel_err(v, ref32)}
    print(f"\nprecision study (pct={pct}, {LOOPS} loops): max relative node-voltage error")
    print(f"{'config':22s} {'vs FP64 (all)':>14s} {'vs FP64 (shared)':>17s} {'vs FP32 ref':>12s}")
    for k, r in rows.items():
        print(f"{k:22s} {r['fp64_err_all']:14.3e} {r['fp64_err_shared']:17.3e} "
              f"{r['err_vs_fp32_reference']:12.3e}")
    for k, r in rows.items():
        assert r["err_vs_fp32_reference"] <= TOL, k
        assert r["fp64_err_all"] <= 1e-3, k


def test_reduction_error_in_ulps(C):
    """The node voltages above are dominated by update_voltages' own
    rounding; this isolates the reductions: after one calc_new_currents
    (EXACT) and one distribute_charge, each node's charge is compared with the
    correctly rounded exact sum of the float contributions -dt*I[0], +dt*I[9]."""
    import math
    params = C.GenParams(4, 1024, 4096, 50, 12345, C.PROFILE_STABLE)
    dt = np.float32(1e-6)
    out = {}
    for mode, name in ((C.REDUCE_CSR, "csr"), (C.REDUCE_ATOMIC, "atomic"),
                       (C.REDUCE_CSR_KAHAN64, "csr_kahan64")):
        with C.Context(0) as ctx:
            ctx.generate(params)
            pieces = ctx.pieces()
            ctx.calc_new_currents(pieces, C.POLICY_EXACT)
            ctx.distribute_charge(pieces, mode)
            nodes, wires = ctx.store()
        out[name] = nodes["charge"]
    vals = [[] for _ in range(len(nodes))]
    q_in = (-dt * wires["current"][:, 0]).astype(np.float32)
    q_out = (dt * wires["current"][:, 9]).astype(np.float32)
    for k in range(len(wires)):
        vals[wires["in_ptr"][k]].append(float(q_in[k]))
        vals[wires["out_ptr"][k]].append(float(q_out[k]))
    exact = np.array([math.fsum(v) for v in vals])
    want = exact.astype(np.float32)
    ulp = np.spacing(np.maximum(np.abs(want), np.float32(1e-30)))
    host = C.generate(params)
    shared = host.locator == 0
    print(f"\nreduction error after one loop (pct 50, 4 pieces), max |charge - exact| in ulps of the exact sum")
    for name, q in out.items():
        e = np.abs(q.astype(np.float64) - exact) / ulp
        print(f"{name:12s} shared {e[shared].max():8.2f}   private {e[~shared].max():8.2f}")
        out[name] = e
    assert out["csr_kahan64"][shared].max() <= 0.5 + 1e-9
    assert np.array_equal(out["csr_kahan64"][~shared], out["csr"][~shared])
