"""Pins the oracle's task bodies against what the mathematics fixes.

calc_new_currents (circuit.cc:352-413, GPU seeding circuit_gpu.cu:107-112)
  * one step from rest is Ohm's law per segment, and the internal voltages
    then follow the capacitor law (exact in FP32);
  * the relaxation's fixed point is the backward-Euler step of the RLC
    ladder: KVL  L (I_i - I0_i)/dt + R I_i = v_{i+1} - v_i  and
    KCL  C (v_k - V0_k)/dt = I_{k-1} - I_k, solved here with numpy;
  * with the reference parameters the currents grow by -L/(dt R) per step
    and overflow FP32 after about 8-10 steps (divergence signature);
  * the CPU-literal seeding never reads the in-node voltage.
distribute_charge (circuit.cc:415-443)
  * charge conservation and a float64 scatter-add per node.
update_voltages (circuit.cc:1290-1307)
  * special cases q = 0 and leak = 0, charge cleared, locator dispatch.
"""
import math

import numpy as np
import pytest

f32 = np.float32
DT32 = f32(1e-6)


def _stable(O, P=2, npp=64, wpp=256, pct=80, seed=11):
    return O.generate(P, npp, wpp, pct, seed, O.PROFILE_STABLE)


def test_one_step_from_rest_is_ohms_law(O):
    c = _stable(O)
    O.calc_new_currents(c, steps=1)
    ref = _stable(O)
    w, n = ref.wires, ref.nodes
    vin = n["voltage"][w["in_ptr"]]
    vout = n["voltage"][w["out_ptr"]]
    R, C = w["resistance"], w["capacitance"]
    I = c.wires["current"]
    # segment 0 sees the in-node, segment 9 the out-node, the rest zero
    assert np.array_equal(I[:, 0], (f32(0) - vin) / R)
    assert np.array_equal(I[:, 9], (vout - f32(0)) / R)
    assert not I[:, 1:9].any()
    V = c.wires["voltage"]
    assert np.array_equal(V[:, 0], (DT32 * (I[:, 0] - I[:, 1])) / C)
    assert np.array_equal(V[:, 8], (DT32 * (I[:, 8] - I[:, 9])) / C)
    assert not V[:, 1:8].any()


def _steps_in_order(w, vin, vout, steps, inductive, capacitive):
    """The reference's step (circuit.cc:387-401) in numpy float32, one
    rounding per operation, with the inductive and capacitive terms formed
    by the given functions."""
    I0, V0 = w["current"].copy(), w["voltage"].copy()
    L, R, C = w["inductance"][:, None], w["resistance"][:, None], w["capacitance"][:, None]
    I = I0.copy()
    v = np.concatenate([vin[:, None], V0, vout[:, None]], axis=1)
    for _ in range(steps):
        I = ((v[:, 1:] - v[:, :-1]) - inductive(L, I, I0)) / R
        v = v.copy()
        v[:, 1:-1] = V0 + capacitive(C, I[:, :-1] - I[:, 1:])
    return I, v[:, 1:-1]


def test_inductive_and_capacitive_terms_in_reference_order(O):
    # From a state where the currents differ from the loop's starting
    # currents (so the inductive term is nonzero, unlike the one-step test
    # above), three steps of the oracle equal the reference's expressions
    # evaluated in their written order, L*(I - I0)/dt and dt*(I_i - I_i+1)/C,
    # bit for bit; each plausible reordering of either term gives different
    # bits on this state, so the comparison tells them apart.
    c = _stable(O, P=1, npp=256, wpp=4096, pct=100, seed=5)
    w, n = c.wires, c.nodes
    rng = np.random.default_rng(5)
    w["current"] = rng.uniform(-1, 1, w["current"].shape).astype(np.float32)
    w["voltage"] = rng.uniform(-1, 1, w["voltage"].shape).astype(np.float32)
    w["inductance"] = rng.uniform(1e-7, 1e-6, len(w)).astype(np.float32)
    before = w.copy()
    vin = n["voltage"][w["in_ptr"]].copy()
    vout = n["voltage"][w["out_ptr"]].copy()
    O.calc_new_currents(c, steps=3)
    ind = lambda L, I, I0: (L * (I - I0)) / DT32
    cap = lambda C, dI: (DT32 * dI) / C
    I, V = _steps_in_order(before, vin, vout, 3, ind, cap)
    assert np.array_equal(c.wires["current"], I)
    assert np.array_equal(c.wires["voltage"], V)
    others = [
        (lambda L, I, I0: L * ((I - I0) / DT32), cap),
        (lambda L, I, I0: (L / DT32) * (I - I0), cap),
        (lambda L, I, I0: (L * (I - I0)) * (f32(1) / DT32), cap),
        (ind, lambda C, dI: DT32 * (dI / C)),
        (ind, lambda C, dI: (DT32 / C) * dI),
    ]
    for k, (a, b) in enumerate(others):
        I2, V2 = _steps_in_order(before, vin, vout, 3, a, b)
        assert not (np.array_equal(I2, I) and np.array_equal(V2, V)), f"reordering {k} not told apart"


def test_cpu_literal_seeding_ignores_in_node(O):
    a = _stable(O)
    b = _stable(O)
    b.nodes["voltage"][:] += f32(0.5)  # changes every in/out node voltage
    b.nodes["voltage"][b.wires["out_ptr"]] = a.nodes["voltage"][a.wires["out_ptr"]]
    O.calc_new_currents(a, steps=3, seeding=O.SEED_CPU_LITERAL)
    O.calc_new_currents(b, steps=3, seeding=O.SEED_CPU_LITERAL)
    only_in = ~np.isin(a.wires["in_ptr"], a.wires["out_ptr"])
    assert np.array_equal(a.wires["current"][only_in], b.wires["current"][only_in])
    a2, b2 = _stable(O), _stable(O)
    b2.nodes["voltage"][:] += f32(0.5)
    O.calc_new_currents(a2, steps=3)
    O.calc_new_currents(b2, steps=3)
    assert not np.array_equal(a2.wires["current"][:, 0], b2.wires["current"][:, 0])


def _backward_euler(L, R, C, I0, V0, vin, vout, dt):
    """Solve KVL per segment and KCL per internal node for one time step."""
    n_i, n_v = 10, 9
    A = np.zeros((19, 19))
    b = np.zeros(19)
    for i in range(n_i):  # L(I_i - I0_i)/dt + R I_i - v_{i+1} + v_i = 0
        A[i, i] = L / dt + R
        b[i] = L / dt * I0[i]
        if i + 1 <= 9:
            A[i, n_i + i] -= 1.0  # v_{i+1}
        else:
            b[i] += vout
        if i >= 1:
            A[i, n_i + i - 1] += 1.0  # v_i
        else:
            b[i] -= vin
    for k in range(1, 10):  # C(v_k - V0_k)/dt - I_{k-1} + I_k = 0
        r = n_i + k - 1
        A[r, r] = C / dt
        A[r, k - 1] = -1.0
        A[r, k] = 1.0
        b[r] = C / dt * V0[k - 1]
    x = np.linalg.solve(A, b)
    return x[:n_i], x[n_i:]


def test_fixed_point_is_backward_euler_rlc_step(O):
    c0 = _stable(O, P=1, npp=64, wpp=96, pct=100, seed=5)
    rng = np.random.default_rng(1)
    c0.wires["current"][:] = rng.uniform(-0.5, 0.5, c0.wires["current"].shape).astype(f32)
    c0.wires["voltage"][:] = rng.uniform(-1, 1, c0.wires["voltage"].shape).astype(f32)
    c = c0.copy()
    n64, w64 = O.promote(c)
    O.calc_new_currents_f64(n64, w64, 0, len(w64))
    O.calc_new_currents(c)  # FP32, 10000 steps
    dt = float(DT32)
    worst64 = worst32 = 0.0
    for k in range(len(c0.wires)):
        w0 = c0.wires[k]
        vin = float(c0.nodes["voltage"][w0["in_ptr"]])
        vout = float(c0.nodes["voltage"][w0["out_ptr"]])
        I, V = _backward_euler(float(w0["inductance"]), float(w0["resistance"]),
                               float(w0["capacitance"]), w0["current"].astype(float),
                               w0["voltage"].astype(float), vin, vout, dt)
        scale = max(np.abs(I).max(), 1e-30)
        worst64 = max(worst64, np.abs(w64[k]["current"] - I).max() / scale)
        worst32 = max(worst32, np.abs(c.wires[k]["current"] - I).max() / scale)
        vs = max(np.abs(V).max(), 1e-30)
        worst64 = max(worst64, np.abs(w64[k]["voltage"] - V).max() / vs)
        worst32 = max(worst32, np.abs(c.wires[k]["voltage"] - V).max() / vs)
    assert worst64 < 1e-9, worst64
    assert worst32 < 1e-5, worst32


@pytest.mark.parametrize("profile", [0, 1], ids=["reference", "stable"])
def test_fp64_steps_match_exact_rational_arithmetic(O, profile):
    # The FP64 twin of calc_new_currents over k <= 3 steps against the same
    # recurrence (circuit.cc:387-401, GPU seeding circuit_gpu.cu:107-112) in
    # exact rational arithmetic from the same inputs, for 40 wires with
    # nonzero starting currents and voltages: per wire, within 1e-12 of the
    # wire's largest value.
    from fractions import Fraction as Fr
    c = O.generate(1, 64, 256, 100, 9, profile)
    rng = np.random.default_rng(9)
    c.wires["current"] = rng.uniform(-1, 1, c.wires["current"].shape).astype(np.float32)
    c.wires["voltage"] = rng.uniform(-1, 1, c.wires["voltage"].shape).astype(np.float32)
    for k in (1, 2, 3):
        n64, w64 = O.promote(c)
        O.calc_new_currents_f64(n64, w64, 0, 40, steps=k)
        dt = Fr(float(np.float32(1e-6)))
        for wi in range(40):
            w = c.wires[wi]
            L, R, C = Fr(float(w["inductance"])), Fr(float(w["resistance"])), Fr(float(w["capacitance"]))
            I0 = [Fr(float(x)) for x in w["current"]]
            V0 = [Fr(float(x)) for x in w["voltage"]]
            v = ([Fr(float(c.nodes["voltage"][w["in_ptr"]]))] + V0 +
                 [Fr(float(c.nodes["voltage"][w["out_ptr"]]))])
            I = list(I0)
            for _ in range(k):
                I = [((v[i + 1] - v[i]) - L * (I[i] - I0[i]) / dt) / R for i in range(10)]
                v = [v[0]] + [V0[i] + dt * (I[i] - I[i + 1]) / C for i in range(9)] + [v[10]]
            got_i, got_v = w64[wi]["current"], w64[wi]["voltage"]
            scale = max(max(abs(float(x)) for x in I), max(abs(float(x)) for x in v[1:10]))
            err = max(max(abs(float(I[i]) - got_i[i]) for i in range(10)),
                      max(abs(float(v[i + 1]) - got_v[i]) for i in range(9)))
            assert err <= 1e-12 * scale, (profile, k, wi, err, scale)


def test_fp64_limit_equals_stepwise_fp64(O):
    # The backward-Euler solve is the limit of the 10000-step relaxation in
    # the stable profile: the step-by-step FP64 oracle reaches it.
    c = O.generate(4, 256, 1024, 90, 3, O.PROFILE_STABLE)
    n1, w1 = O.run_f64(c, 3)
    n2, w2 = O.run_f64_limit(c, 3)
    assert np.abs(n1["voltage"] - n2["voltage"]).max() <= 1e-13 * np.abs(n1["voltage"]).max()
    assert np.abs(w1["current"] - w2["current"]).max() <= 1e-12 * np.abs(w1["current"]).max()
    assert np.abs(w1["voltage"] - w2["voltage"]).max() <= 1e-12 * np.abs(w1["voltage"]).max()


def test_reference_parameters_diverge_with_gain_L_over_dt_R(O):
    c = O.generate(1, 256, 200, 100, 12345, O.PROFILE_REFERENCE)
    n64, w64a = O.promote(c)
    _, w64b = O.promote(c)
    O.calc_new_currents_f64(n64, w64a, 0, len(w64a), steps=6)
    O.calc_new_currents_f64(n64, w64b, 0, len(w64b), steps=7)
    g = c.wires["inductance"].astype(float) / (float(DT32) * c.wires["resistance"].astype(float))
    ratio = w64b["current"] / w64a["current"]
    # From rest, step 1 excites only the end segments; afterwards each step
    # multiplies by T = -g + d(2 - shift).  A segment k places from the nearer
    # end therefore carries C(j-1, k) (-g)^(j-1-k) d^k, and from step j=6 to 7
    # grows by -g * 6/(6-k) (up to O(d/g) and O(1/g) terms).
    for k in (0, 1, 2):
        want = -g * 6.0 / (6 - k)
        assert np.allclose(ratio[:, k], want, rtol=1e-3)
        assert np.allclose(ratio[:, 9 - k], want, rtol=1e-3)
    # FP32: finite after 7 steps, overflowed and NaN well before 25
    a = c.copy()
    O.calc_new_currents(a, steps=7)
    assert np.isfinite(a.wires["current"]).all()
    b = c.copy()
    O.calc_new_currents(b, steps=25)
    assert np.isnan(b.wires["current"]).all()


def test_distribute_charge_conserves_charge(O):
    c = O.generate(4, 256, 1024, 70, 3, O.PROFILE_STABLE)
    rng = np.random.default_rng(2)
    c.wires["current"][:] = rng.uniform(-1, 1, c.wires["current"].shape).astype(f32)
    for p in c.pieces:
        O.distribute_charge(c, p)
    w = c.wires
    qin = (-DT32 * w["current"][:, 0]).astype(np.float64)
    qout = (DT32 * w["current"][:, 9]).astype(np.float64)
    exact = np.zeros(len(c.nodes))
    np.add.at(exact, w["in_ptr"], qin)
    np.add.at(exact, w["out_ptr"], qout)
    absum = np.zeros(len(c.nodes))
    np.add.at(absum, w["in_ptr"], np.abs(qin))
    np.add.at(absum, w["out_ptr"], np.abs(qout))
    err = np.abs(c.nodes["charge"].astype(np.float64) - exact)
    assert np.all(err <= 2 * np.finfo(f32).eps * absum * 8 + 1e-30)
    assert math.isclose(c.nodes["charge"].astype(np.float64).sum(), qin.sum() + qout.sum(),
                        rel_tol=1e-4, abs_tol=1e-9)


def test_kahan64_is_correctly_rounded_sum_on_shared_nodes(O):
    c = O.generate(4, 128, 1024, 50, 9, O.PROFILE_STABLE)
    rng = np.random.default_rng(3)
    c.wires["current"][:] = rng.uniform(-1, 1, c.wires["current"].shape).astype(f32)
    seq = c.copy()
    for p in seq.pieces:
        O.distribute_charge(seq, p)
    O.distribute_charge_kahan64(c)
    w = c.wires
    exact = np.zeros(len(c.nodes))
    vals = []
    for n in range(len(c.nodes)):
        vals.append([])
    for k in range(len(w)):
        vals[w["in_ptr"][k]].append(float(-DT32 * w["current"][k, 0]))
        vals[w["out_ptr"][k]].append(float(DT32 * w["current"][k, 9]))
    for n in range(len(c.nodes)):
        exact[n] = math.fsum(vals[n])
    shared = c.locator == 0
    got = c.nodes["charge"][shared]
    want = exact[shared].astype(f32)
    ulp = np.spacing(np.abs(want))
    assert np.all(np.abs(got - want) <= ulp)
    assert np.array_equal(c.nodes["charge"][~shared], seq.nodes["charge"][~shared])


def test_fold_buffers_match_direct_reduction(O):
    c = O.generate(4, 128, 512, 60, 4, O.PROFILE_STABLE)
    rng = np.random.default_rng(4)
    c.wires["current"][:] = rng.uniform(-1, 1, c.wires["current"].shape).astype(f32)
    direct = c.copy()
    for p in direct.pieces:
        O.distribute_charge(direct, p)
    owner = c.nodes.copy()
    N = len(owner)
    for p in c.pieces:
        shr = np.zeros(N, f32)
        ghost = np.zeros(N, f32)
        O.distribute_charge_fold(c.wires, owner, shr, ghost, int(p["first_wire"]),
                                 int(p["num_wires"]))
        O.fold_apply(owner, shr, 0, N)
        O.fold_apply(owner, ghost, 0, N)
    priv = c.locator == 1
    assert np.array_equal(owner["charge"][priv], direct.nodes["charge"][priv])
    assert np.allclose(owner["charge"], direct.nodes["charge"], rtol=1e-5, atol=1e-12)


def test_update_voltages_special_cases(O):
    c = O.generate(2, 512, 16, 90, 8)
    n = c.nodes
    rng = np.random.default_rng(5)
    n["charge"][:] = rng.uniform(-1e-3, 1e-3, len(n)).astype(f32)
    n["charge"][:100] = 0
    n["leakage"][100:200] = 0
    before = n.copy()
    pvt = c.nodes.copy()
    shr = c.nodes.copy()
    for p in c.pieces:
        O.update_voltages(c, p, pvt=pvt, shr=shr)
    priv = c.locator == 1
    got = np.where(priv, pvt["voltage"], shr["voltage"])
    # q = 0: pure leak
    assert np.array_equal(got[:100], before["voltage"][:100] * (f32(1) - before["leakage"][:100]))
    # leak = 0: pure charge-in
    i = slice(100, 200)
    assert np.array_equal(got[i], before["voltage"][i] + before["charge"][i] / before["capacitance"][i])
    # charge cleared in the instance that owns the node; the other untouched
    assert not np.where(priv, pvt["charge"], shr["charge"]).any()
    assert np.array_equal(np.where(priv, shr["voltage"], pvt["voltage"]),
                          np.where(priv, before["voltage"], before["voltage"]))


@pytest.mark.parametrize("grain,threads", [(0, 1), (0, 8), (64, 1), (64, 8), (7, 3)])
def test_loop_driver_equals_task_bodies_on_ragged_pieces(O, grain, threads):
    # The threaded loop driver against the three task bodies called one after
    # another (circuit.cc:232-307), on pieces whose wire counts are not
    # multiples of the work chunk (so chunks cross piece boundaries), for both
    # the FP32 run and its FP64 twin: every wire relaxed exactly once per loop.
    a = O.generate(3, 61, 237, 90, 41, O.PROFILE_STABLE)
    b = a.copy()
    c64 = a.copy()
    O.lib().orc_run_f32(0, 40, 3, 0, threads, grain, O._p(a.nodes), a.num_nodes, O._p(a.wires),
                        a.num_wires, O._p(a.locator), O._p(a.pieces), len(a.pieces))
    for _ in range(3):
        for p in b.pieces:
            O.calc_new_currents(b, p, steps=40)
        for p in b.pieces:
            O.distribute_charge(b, p)
        for p in b.pieces:
            O.update_voltages(b, p)
    assert a.wires.tobytes() == b.wires.tobytes()
    assert a.nodes.tobytes() == b.nodes.tobytes()
    n64, w64 = O.run_f64(c64, 3, steps=40, nthreads=threads)
    m64, x64 = O.promote(c64)
    for _ in range(3):
        O.calc_new_currents_f64(m64, x64, 0, c64.num_wires, steps=40)
        for p in c64.pieces:
            O.lib().orc_distribute_charge_f64(O._p(x64), O._p(m64), int(p["first_wire"]),
                                              int(p["num_wires"]))
        for p in c64.pieces:
            O.lib().orc_update_voltages_f64(O._p(m64), int(p["first_node"]), int(p["num_nodes"]))
    assert w64.tobytes() == x64.tobytes() and n64.tobytes() == m64.tobytes()
