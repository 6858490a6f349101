"""The bound behind the kernels' tail_stays_finite (ckt_kernels.hip): a wire
whose step contracts (infinity-norm of the substituted update at most 1/2,
negative capacitances included),
with 2^-20 <= R <= 2^20 and every starting current and voltage (node
voltages included) at most 2^40 in magnitude, keeps every intermediate value
of the reference's step (circuit.cc:387-401) below 2^104, so its last steps
cannot leave the finite range and the policies may skip the reference-order
tail for it.

Checked here on the CPU with numpy float32 (one rounding per operation, the
reference's order) on wires drawn at the edges of that region: parameters
spread over the allowed decades, starting values of magnitude up to 2^40
with random signs, and norms up to 1/2.  The guard is restated from the
kernel; a wire it rejects is not drawn.
"""
import numpy as np

f32 = np.float32
DT = f32(1e-6)


def guard(I, V, vin, vout, L, R, C):
    g = (L / DT) / R
    d = (f32(1) / R) * (DT / C)
    ad = np.abs(d)
    norm = np.maximum(np.abs(f32(2) * d - g) + f32(2) * ad, np.abs(d - g) + ad)
    ok = (norm <= 0.5) & (R >= 2.0**-20) & (R <= 2.0**20)
    ok &= (np.abs(vin) <= 2.0**40) & (np.abs(vout) <= 2.0**40)
    ok &= (np.abs(I) <= 2.0**40).all(1) & (np.abs(V) <= 2.0**40).all(1)
    return ok


def draw(rng, n):
    R = f32(2.0) ** rng.uniform(-20, 20, n).astype(f32)
    # norm <= 1/2 needs g <= 1 and d <= 1/4: draw g and d up to those and
    # derive L and C from them.
    g = rng.uniform(0, 1, n).astype(f32) * (rng.random(n) < 0.7)
    d = (f32(2.0) ** rng.uniform(-40, -2, n) * rng.choice([-1, 1], n, p=[0.1, 0.9])).astype(f32)
    L = (g * DT * R).astype(f32)
    C = (DT / (d * R)).astype(f32)
    mag = f32(2.0) ** rng.uniform(0, 40, (n, 1)).astype(f32)
    I = (rng.choice([-1, 1], (n, 10)) * rng.uniform(0.5, 1, (n, 10)) * mag).astype(f32)
    V = (rng.choice([-1, 1], (n, 9)) * rng.uniform(0.5, 1, (n, 9)) * mag).astype(f32)
    vin = (rng.choice([-1, 1], n) * mag[:, 0]).astype(f32)
    vout = (-vin).astype(f32)
    return I, V, vin, vout, L, R, C


def test_guarded_wires_keep_reference_intermediates_below_2_to_104():
    rng = np.random.default_rng(3)
    kept = 0
    worst = 0.0
    with np.errstate(over="raise", invalid="raise"):
        for _ in range(4):
            I0, V0, vin, vout, L, R, C = draw(rng, 20000)
            ok = guard(I0, V0, vin, vout, L, R, C)
            I0, V0, vin, vout = I0[ok], V0[ok], vin[ok], vout[ok]
            L, R, C = L[ok][:, None], R[ok][:, None], C[ok][:, None]
            kept += len(I0)
            I = I0.copy()
            v = np.concatenate([vin[:, None], V0, vout[:, None]], axis=1)
            for _ in range(200):
                dv = v[:, 1:] - v[:, :-1]
                ind = (L * (I - I0)) / DT
                num = dv - ind
                I = num / R
                q = (DT * (I[:, :-1] - I[:, 1:])) / C
                v = v.copy()
                v[:, 1:-1] = V0 + q
                for x in (dv, L * (I - I0), ind, num, I, q, v):
                    worst = max(worst, float(np.abs(x).max()))
    assert kept > 20000
    assert worst < 2.0**104, worst
    print(f"largest intermediate: 2^{np.log2(worst):.1f} over {kept} wires")


def test_guard_rejects_the_reference_profile_and_non_finite_values():
    rng = np.random.default_rng(4)
    n = 1000
    # the reference profile's parameters: L in [0.1, 0.11] H, R in [1, 11],
    # C in [0, 0.1]: g = L / (dt R) ~ 1e4..1e5, far from contracting
    L = rng.uniform(0.1, 0.11, n).astype(f32)
    R = rng.uniform(1, 11, n).astype(f32)
    C = rng.uniform(0, 0.1, n).astype(f32)
    z10, z9, z = np.zeros((n, 10), f32), np.zeros((n, 9), f32), np.zeros(n, f32)
    assert not guard(z10, z9, z, z, L, R, C).any()
    # stable-looking parameters pass until a value is NaN or too large
    L = np.full(n, 1e-7, f32)
    C = np.full(n, 0.005, f32)
    assert guard(z10, z9, z, z, L, R, C).all()
    I = z10.copy()
    I[::2, 3] = np.nan
    I[1::2, 3] = f32(2.0**41)
    assert not guard(I, z9, z, z, L, R, C).any()
