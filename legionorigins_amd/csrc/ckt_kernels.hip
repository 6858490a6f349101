// ckt_kernels.hip -- gfx950 kernels of the circuit hot path.
//
// calc_new_currents   apps/circuit/circuit_gpu.cu:61-149 (math :103-135)
// distribute_charge   apps/circuit/circuit_gpu.cu:219-286
// update_voltages     apps/circuit/circuit_gpu.cu:306-350
// sanity_check_wires  apps/circuit/circuit_gpu.cu:151-184
//
// Built with -ffp-contract=off: every a*b+c below is two rounded operations
// unless written as fmaf().  Division is IEEE-754 correctly rounded (the HIP
// default), FP32 denormals are preserved, packed-FP32 instructions are not
// generated (-packed-fp32-ops).
#include <hip/hip_runtime.h>

#include "ckt_divsplit.h"
#include "ckt_index.h"
#include "ckt_internal.h"

namespace ckt {

// float dt = DELTAT (circuit_gpu.cu:103, circuit.cc:376).
#define CKT_DT_F ((float)CKT_DELTAT)

// ---------------------------------------------------------------------------
// The relaxation of one wire (circuit_gpu.cu:102-135).
// On entry I[] and V[] hold wire.current and wire.voltage, vin/vout the
// voltages of the in and out nodes; on exit they hold the values written back.

template <int POLICY>
struct Relax;

// Correctly rounded a / b for a divisor that stays fixed across the step
// loop.  With rb = RN64(1/b) (one FP64 division per wire), RN32((double)a * rb)
// equals RN32(a / b) whenever the quotient is a normal float: a quotient of
// two floats is never exactly halfway between two normal floats and lies at
// least 2^-49 (relative) from every such midpoint, while the FP64 product is
// within 2^-52 of a / b, so both round the same way (the double-rounding
// argument of S. Figueroa, "When is double rounding innocuous?", 1995).  Zero
// and subnormal quotients take the IEEE division instead.  Infinite, NaN and
// zero operands give what the IEEE division gives.  One compare guards the
// rare path (|q| < 2^-126 holds for zero and subnormal q whether or not the
// compare treats subnormal inputs as zero); a zero dividend is sorted out
// inside it.
__device__ __forceinline__ float div_fixed(float a, double rb, float b) {
  float q = (float)((double)a * rb);
  if (__builtin_expect(fabsf(q) < 0x1p-126f, 0)) {
    asm volatile("");
    if (a != 0.f) q = a / b;
  }
  return q;
}

// div_fixed over N dividends by one divisor, with one guard for the group:
// the rare path runs when the smallest |quotient| is below 2^-126 and then
// takes the IEEE division for exactly the quotients div_fixed would (zero or
// subnormal, nonzero dividend), so the results are div_fixed's, bit for bit.
// fminf ignores NaN quotients, which div_fixed's compare also lets through.
// Four 3-input minimums (absolute values are free source modifiers) and one
// compare replace ten compares.
template <int N>
__device__ __forceinline__ void div_fixed_group(float (&q)[N], const float (&a)[N], double rb,
                                                float b) {
  float m = __builtin_huge_valf();
#pragma unroll
  for (int i = 0; i < N; i++) {
    q[i] = (float)((double)a[i] * rb);
    m = fminf(m, fabsf(q[i]));
  }
  if (__builtin_expect(m < 0x1p-126f, 0)) {
    asm volatile("");
#pragma unroll
    for (int i = 0; i < N; i++)
      if (fabsf(q[i]) < 0x1p-126f && a[i] != 0.f) q[i] = a[i] / b;
  }
}

// Whether `p` holds on every active lane of the wave.  Branches on it are
// uniform: EXACT picks its step variant by it, and the policies skip the tail
// only when the whole wave may (skipping changes the loop's trip count, and a
// trip count that differs between lanes turns the counter of the step loop
// into a per-lane register with an exec-mask update every iteration: +0.5
// VALU per step for SCALED, measured 2.6% more instructions and 6% more
// cycles; running the tail on a wire that could skip it is also correct).
__device__ __forceinline__ bool wave_all(bool p) { return __all(p); }

// A divisor with its split for div_split_group: (zh, zl) from make_div_split,
// zh = zl = 0 if it has none (every quotient then takes div_fixed).
struct SplitDivisor {
  float b, zh, zl;
  __device__ __forceinline__ explicit SplitDivisor(float divisor) : b(divisor) {
    make_div_split(divisor, &zh, &zl);
  }
};

// Correctly rounded a[i] / b for N dividends, FP32 only: q = RN(a zh - u),
// u = RN(-a zl + 0), two fmas each (div_split_q, ckt_divsplit.h); zero
// dividends get the zero of the right sign.  A minimum over |q| guards the
// group, and with kCheckU a maximum over |u| (fminf and fmaxf ignore NaN, and
// a NaN dividend gives a NaN quotient on every path); the minimum starts at
// |zh|, which is at least 2^-61 for a divisor with a split and 0 for one
// without.  The guard also fires on zero quotients, the common case once a
// wire has settled; a second test inside it looks for what the split does not
// cover, and then the whole group goes to div_fixed: a divisor without a
// split, an infinite u (infinite dividends, quotients beyond about 2^150), or
// a nonzero dividend below |b| 2^-98 (a quotient that may be below
// kDivSplitMin; the integer keys of div_split_key put zeros above every
// bound).  Without kCheckU the caller guarantees finite u (see split_step).
template <int N, bool kCheckU>
__device__ __forceinline__ void div_split_group(float (&q)[N], const float (&a)[N],
                                                const SplitDivisor& d) {
  float m = fabsf(d.zh), M = 0.f;
#pragma unroll
  for (int i = 0; i < N; i++) {
    float u;
    q[i] = div_split_q(a[i], d.zh, d.zl, &u);
    m = fminf(m, fabsf(q[i]));
    if (kCheckU) M = fmaxf(M, fabsf(u));
  }
  if (__builtin_expect(m < kDivSplitMin || (kCheckU && M > __FLT_MAX__), 0)) {
    asm volatile("");
    uint32_t lo = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < N; i++) lo = min(lo, div_split_key(a[i]));
    if (lo < div_split_key_bound(d.b) || (kCheckU && M > __FLT_MAX__) || d.zh == 0.f) {
      const double rb = 1.0 / (double)d.b;
#pragma unroll
      for (int i = 0; i < N; i++) q[i] = div_fixed(a[i], rb, d.b);
    }
  }
}

// One step of the reference (circuit_gpu.cu:114-129) with div_split_group.
// kCheckU = false needs every current and internal voltage entering the step,
// L, I0, V0 and the node voltages at most kStepBound in magnitude (or NaN),
// |R| >= 2^-20 and |C| >= 2^-40 (a group whose divisor has no split goes to
// div_fixed, whose quotients obey the same bounds): then |L (I - I0)| <=
// 2^81, its quotient by dt <= 2^101, the R dividends <= 2^102, their
// quotients <= 2^122, the C dividends <= 2^104, so no dividend is infinite and
// every u = RN(-a zl) is finite (|zl| <= 2^-24 |zh|: <= 2^-4 for dt, 2^-4 for
// R, 2^16 for C); the C quotients may overflow, which the fma rounds as IEEE
// division does.  (A smaller R lets the R quotients overflow, and the C
// dividends become infinite with them.)
template <bool kCheckU>
__device__ __forceinline__ void split_step(float (&ni)[kSeg], float (&nv)[kSeg + 1],
                                           const float (&I0)[kSeg], const float (&V0)[kSeg - 1],
                                           float L, const SplitDivisor& sdt,
                                           const SplitDivisor& sR, const SplitDivisor& sC) {
  const float dt = CKT_DT_F;
  float t[kSeg], q[kSeg], u[kSeg - 1], w[kSeg - 1];
#pragma unroll
  for (int i = 0; i < kSeg; i++) t[i] = L * (ni[i] - I0[i]);
  div_split_group<kSeg, kCheckU>(q, t, sdt);
#pragma unroll
  for (int i = 0; i < kSeg; i++) t[i] = (nv[i + 1] - nv[i]) - q[i];
  div_split_group<kSeg, kCheckU>(ni, t, sR);
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) u[i] = dt * (ni[i] - ni[i + 1]);
  div_split_group<kSeg - 1, kCheckU>(w, u, sC);
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = V0[i] + w[i];
}

// Largest |value| of the running currents and internal voltages (NaN ignored).
__device__ __forceinline__ float state_max(const float (&ni)[kSeg], const float (&nv)[kSeg + 1]) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kSeg; i++) s = fmaxf(s, fabsf(ni[i]));
#pragma unroll
  for (int i = 1; i < kSeg; i++) s = fmaxf(s, fabsf(nv[i]));
  return s;
}

// The reference's step in its order of operations (circuit_gpu.cu:114-129),
// `steps` times, on the running currents ni and voltages nv (nv[0], nv[kSeg]
// the node voltages), with I0 and V0 the wire's values at the start of the
// loop; writes the result to I0 and V0.  Every division is correctly rounded:
// the ten by dt, then the ten by R, then the nine by C, each group with one
// guard; by the FP32 split (split_step, after a per-wire check of the
// divisors) when kSplit, else by the FP64 reciprocal (div_fixed_group, no
// set-up: the short tails of the other policies).
template <bool kSplit>
__device__ __forceinline__ void reference_steps(float (&ni)[kSeg], float (&nv)[kSeg + 1],
                                                float (&I0)[kSeg], float (&V0)[kSeg - 1], float L,
                                                float R, float C, int steps) {
  const float dt = CKT_DT_F;
  if constexpr (kSplit) {
    const SplitDivisor sdt(dt), sR(R), sC(C);
    // The loop's constants against split_step's conditions, once; the running
    // state every K steps.  A wave whose lanes all meet them takes K steps
    // without the |u| guard.
    float cmax = fmaxf(fabsf(L), fmaxf(fabsf(nv[0]), fabsf(nv[kSeg])));
#pragma unroll
    for (int i = 0; i < kSeg; i++) cmax = fmaxf(cmax, fabsf(I0[i]));
#pragma unroll
    for (int i = 0; i < kSeg - 1; i++) cmax = fmaxf(cmax, fabsf(V0[i]));
    const bool consts_ok = cmax <= kStepBound && fabsf(R) >= 0x1p-20f && fabsf(C) >= 0x1p-40f;
    // Steps per state check: the largest of 8, 4, 3, 2 that every active lane
    // allows (votes, so inactive lanes take no part).
    const int kl = unchecked_steps(L, R, C);
    const int K = wave_all(kl >= 8) ? 8 : wave_all(kl >= 4) ? 4 : wave_all(kl >= 3) ? 3
                : wave_all(kl >= 2) ? 2 : 1;
    int j = 0;
    while (j < steps) {
      while (j < steps && wave_all(consts_ok && state_max(ni, nv) <= kStepBound)) {
        const int n = min(K, steps - j);
        for (int k = 0; k < n; k++) split_step<false>(ni, nv, I0, V0, L, sdt, sR, sC);
        j += n;
      }
      while (j < steps && !wave_all(consts_ok && state_max(ni, nv) <= kStepBound)) {
        split_step<true>(ni, nv, I0, V0, L, sdt, sR, sC);
        j++;
      }
    }
  } else {
    constexpr double rdt = 1.0 / (double)CKT_DT_F;
    const double rR = 1.0 / (double)R, rC = 1.0 / (double)C;
    for (int j = 0; j < steps; j++) {
      float t[kSeg], q[kSeg], u[kSeg - 1], w[kSeg - 1];
#pragma unroll
      for (int i = 0; i < kSeg; i++) t[i] = L * (ni[i] - I0[i]);
      div_fixed_group(q, t, rdt, dt);
#pragma unroll
      for (int i = 0; i < kSeg; i++) t[i] = (nv[i + 1] - nv[i]) - q[i];
      div_fixed_group(ni, t, rR, R);
#pragma unroll
      for (int i = 0; i < kSeg - 1; i++) u[i] = dt * (ni[i] - ni[i + 1]);
      div_fixed_group(w, u, rC, C);
#pragma unroll
      for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = V0[i] + w[i];
    }
  }
#pragma unroll
  for (int i = 0; i < kSeg; i++) I0[i] = ni[i];
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) V0[i] = nv[i + 1];
}

// A whole loop in the reference's order, seeded as the reference kernel
// seeds it (circuit_gpu.cu:107-112); kSplit as for reference_steps.
template <bool kSplit>
__device__ __forceinline__ void reference_loop(float (&I)[kSeg], float (&V)[kSeg - 1], float vin,
                                               float vout, float L, float R, float C, int steps) {
  float ni[kSeg], nv[kSeg + 1];
#pragma unroll
  for (int i = 0; i < kSeg; i++) ni[i] = I[i];
  nv[0] = vin;
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = V[i];
  nv[kSeg] = vout;
  reference_steps<kSplit>(ni, nv, I, V, L, R, C, steps);
}

// Reference order of operations; bit-identical to a strict FP32 CPU run.
template <>
struct Relax<CKT_POLICY_EXACT> {
  static __device__ __forceinline__ void run(float (&I)[kSeg], float (&V)[kSeg - 1], float vin,
                                             float vout, float L, float R, float C, int steps) {
    reference_loop<true>(I, V, vin, vout, L, R, C, steps);
  }
};

// The other policies run the last kTail steps of a loop in the reference's
// order (except for wires that provably stay finite through them; see
// tail_stays_finite).  Where a wire's values leave the finite range, the reference
// recurrence turns them to NaN one segment per step (an infinite current
// makes the voltage difference across its segment infinite, and the
// inductive term L*(I - I0)/dt the same infinity, so the next step takes
// inf - inf); which of a wire's values are +-inf, NaN or still finite at the
// end depends on the exact step at which each value overflowed, and that
// step depends on the reference's intermediate values (L*(I - I0)/dt is R
// times the current it yields).  A wire whose values are still finite kTail
// steps before the end therefore finishes in the reference's order and
// overflows where the reference does.  A wire that is already non-finite
// then overflowed within a step of the reference, whose NaN then reaches
// every current and internal voltage in at most 10 further steps (1.5 M
// wires with NaN, infinite or near-FLT_MAX node voltages in both generator
// profiles needed 10 at most), so it ends all-NaN, as it is written.
constexpr int kTail = 16;

// Wires with a mode whose gain per step is moderate in magnitude also run
// the whole loop in the reference's order.  Each step's rounding perturbs the
// state by a relative ~1e-7, and a mode with gain lambda carries the
// perturbation for about 1/(1 - |lambda|) steps, or, for |lambda| > 1, keeps
// amplifying it with the state: the reassociated steps then drift from the
// reference's by ~1e-7 per step the mode has grown through (1e-3 after 10000
// steps near |lambda| = 1, 1.4e-5 after 120 steps at |lambda| = 2.2), and
// where such a wire overflows near the end of the loop the drift moves the
// step it overflows at.  The step's modes are those of the tridiagonal update
// x <- K + A x with A = (2d - g) I - d B, B the path adjacency plus ones at
// both ends, whose eigenvalues are 2 cos(pi k / 10):
//   lambda_k = 2d - g - 2d cos(pi k / 10),  g = L / (dt R),  d = dt / (R C).
// Outside 0.95 <= |lambda_k| <= 1000 a perturbation decays by 5% per step
// (the accumulated drift stays below ~20 roundings) or grows over 1000x per
// step, so that a wire starting from values of at least 2^-96 (see
// tiny_input) overflows within 23 steps (1000^23 > FLT_MAX / 2^-96), within a
// step of the reference.  No wire of either generator profile comes near the
// band: the reference parameters have |lambda| ~ 1e4..1e5, the stable ones
// |lambda| < 0.12.
struct ModeGains {
  bool moderate;  // some |lambda_k| in [0.95, 1000]
  bool fast;      // some |lambda_k| above 1000
};

__device__ __forceinline__ ModeGains mode_gains(float L, float R, float C) {
  static_assert(kSeg == 10, "the table holds cos(pi k / 10)");
  constexpr float kCos[kSeg] = {1.0f,         0.95105652f,  0.80901699f,  0.58778525f,
                                0.30901699f,  0.0f,         -0.30901699f, -0.58778525f,
                                -0.80901699f, -0.95105652f};
  const float dt = CKT_DT_F;
  const float g = (L / dt) / R, d = (1.0f / R) * (dt / C);
  const float a = 2.0f * d - g, m2 = -2.0f * d;
  ModeGains mg{false, false};
#pragma unroll
  for (int k = 0; k < kSeg; k++) {
    const float lam = fabsf(fmaf(m2, kCos[k], a));
    mg.moderate |= lam >= 0.95f && lam <= 1000.0f;
    mg.fast |= lam > 1000.0f;
  }
  return mg;
}

// A wire with a fast-growing mode also runs in the reference's order when a
// value it starts from is nonzero but below 2^-96 in magnitude.  The
// reference's products L*(I - I0) and dt*(I_i - I_i+1) of such values fall
// below the normal range and lose bits or become 0 (a growing mode started
// from them then stays 0), where the reassociated steps scale them by other
// factors.  In a wire whose modes all decay, such values are rounding-level
// parts of a state its fixed point dominates, and they take the policy's own
// steps (values of 1e-30 occur among the internal voltages of the stable
// profile's early loops).
__device__ __forceinline__ bool tiny_input(const float (&I)[kSeg], const float (&V)[kSeg - 1],
                                           float vin, float vout) {
  bool tiny = (vin != 0.0f && fabsf(vin) < 0x1p-96f) || (vout != 0.0f && fabsf(vout) < 0x1p-96f);
#pragma unroll
  for (int i = 0; i < kSeg; i++) tiny |= I[i] != 0.0f && fabsf(I[i]) < 0x1p-96f;
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) tiny |= V[i] != 0.0f && fabsf(V[i]) < 0x1p-96f;
  return tiny;
}

// Whether a wire runs its whole loop in the reference's order (the cases
// above).
__device__ __forceinline__ bool whole_loop_in_reference_order(const float (&I)[kSeg],
                                                              const float (&V)[kSeg - 1],
                                                              float vin, float vout, float L,
                                                              float R, float C, int steps) {
  if (steps <= kTail) return true;
  const ModeGains mg = mode_gains(L, R, C);
  return mg.moderate || (mg.fast && tiny_input(I, V, vin, vout));
}

// A wire whose step contracts and whose parameters and starting values are
// bounded as below skips the tail: neither the reference's iteration nor the
// policy's can leave the finite range, so both end finite, and the policy's
// own steps run to the end.  With the step's infinity-norm at most 1/2 (the
// substituted update x <- K + A x of FAST below; |d| in the norm, so a
// negative capacitance is bounded too), every current stays within
// max(|x_1|, 2|K|) <= 2^63 when |I0|, |V0|, |vin|, |vout| <= 2^40 and
// 2^-20 <= R <= 2^20 (then |g| <= 1, |d| <= 1/4, |b| = |d| R <= 2^18,
// |1/C| <= 2^38, |L| <= R dt); every intermediate of the reference's step
// then stays below 2^104: voltages below 2^82, (v_{i+1} - v_i) - L (I - I0) /
// dt below 2^84, its quotient by R below 2^104, dt (I_i - I_{i+1}) / C below
// 2^82.  The stable profile's wires all pass; the reference profile's
// (g ~ 1e4..1e5) none.  A NaN anywhere fails a comparison and keeps the
// tail.
__device__ __forceinline__ bool tail_stays_finite(const float (&I)[kSeg], const float (&V)[kSeg - 1],
                                                  float vin, float vout, float L, float R,
                                                  float C) {
  const float dt = CKT_DT_F;
  const float g = (L / dt) / R, d = (1.0f / R) * (dt / C);
  const float ad = fabsf(d);
  const float norm = fmaxf(fabsf(2.0f * d - g) + 2.0f * ad, fabsf(d - g) + ad);
  bool ok = norm <= 0.5f && R >= 0x1p-20f && R <= 0x1p20f;
  ok &= fabsf(vin) <= 0x1p40f && fabsf(vout) <= 0x1p40f;
#pragma unroll
  for (int i = 0; i < kSeg; i++) ok &= fabsf(I[i]) <= 0x1p40f;
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) ok &= fabsf(V[i]) <= 0x1p40f;
  return ok;
}

// The internal voltages formed from the final currents x by the reference's
// voltage update (circuit.cc:398-401), V0_k + dt (x_{k-1} - x_k) / C, with
// x the currents: the end of a loop without the tail.
__device__ __forceinline__ void end_from_currents(float (&I)[kSeg], float (&V)[kSeg - 1],
                                                  const float (&x)[kSeg], float C) {
  const float dt = CKT_DT_F;
  const double rC = 1.0 / (double)C;
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) V[i] = V[i] + div_fixed(dt * (x[i] - x[i + 1]), rC, C);
#pragma unroll
  for (int i = 0; i < kSeg; i++) I[i] = x[i];
}

// The tail: the internal voltages formed from the body's currents x by the
// reference's voltage update, then the reference's steps.  A non-finite x
// gives an all-NaN wire.
__device__ __forceinline__ void finish_in_reference_order(float (&I)[kSeg], float (&V)[kSeg - 1],
                                                          float (&x)[kSeg], float vin, float vout,
                                                          float L, float R, float C) {
  bool finite = true;
#pragma unroll
  for (int i = 0; i < kSeg; i++) finite &= __builtin_isfinite(x[i]);
  if (!finite) {
    const float nan = __builtin_nanf("");
#pragma unroll
    for (int i = 0; i < kSeg; i++) I[i] = nan;
#pragma unroll
    for (int i = 0; i < kSeg - 1; i++) V[i] = nan;
    return;
  }
  const float dt = CKT_DT_F;
  const double rC = 1.0 / (double)C;
  float nv[kSeg + 1];
  nv[0] = vin;
#pragma unroll
  for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = V[i] + div_fixed(dt * (x[i] - x[i + 1]), rC, C);
  nv[kSeg] = vout;
  reference_steps<false>(x, nv, I, V, L, R, C, kTail);
}

// Per-wire constants hoisted, FMAs in the loop, voltages kept every step:
//   I_i <- h*(v_{i+1}-v_i) - g*(I_i - I0_i),  h = 1/R, g = (L/dt)/R
//   v_k <- V0_k + b*(I_{k-1} - I_k),          b = dt/C
template <>
struct Relax<CKT_POLICY_RECIP> {
  static __device__ __forceinline__ void run(float (&I)[kSeg], float (&V)[kSeg - 1], float vin,
                                             float vout, float L, float R, float C, int steps) {
    if (whole_loop_in_reference_order(I, V, vin, vout, L, R, C, steps)) {
      reference_loop<false>(I, V, vin, vout, L, R, C, steps);
      return;
    }
    const bool no_tail = wave_all(tail_stays_finite(I, V, vin, vout, L, R, C));
    const float dt = CKT_DT_F;
    const float h = 1.0f / R, g = (L / dt) / R, b = dt / C;
    float c0[kSeg], ni[kSeg], nv[kSeg + 1];
#pragma unroll
    for (int i = 0; i < kSeg; i++) {
      c0[i] = g * I[i];
      ni[i] = I[i];
    }
    nv[0] = vin;
#pragma unroll
    for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = V[i];
    nv[kSeg] = vout;
    // Step 1 from the seeded voltages: the running currents still equal I0,
    // so the inductive term is zero, as the reference computes it (its
    // L*(I - I0) is exactly 0 there; -g*I0 + g*I0 would leave the rounding of
    // g*I0, which steps with a large gain g amplify).
#pragma unroll
    for (int i = 0; i < kSeg; i++) ni[i] = h * (nv[i + 1] - nv[i]);
#pragma unroll
    for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = fmaf(b, ni[i] - ni[i + 1], V[i]);
#pragma unroll 2
    for (int j = no_tail ? 1 : kTail + 1; j < steps; j++) {
#pragma unroll
      for (int i = 0; i < kSeg; i++) ni[i] = fmaf(-g, ni[i], fmaf(h, nv[i + 1] - nv[i], c0[i]));
#pragma unroll
      for (int i = 0; i < kSeg - 1; i++) nv[i + 1] = fmaf(b, ni[i] - ni[i + 1], V[i]);
    }
    if (no_tail) {
#pragma unroll
      for (int i = 0; i < kSeg; i++) I[i] = ni[i];
#pragma unroll
      for (int i = 0; i < kSeg - 1; i++) V[i] = nv[i + 1];
      return;
    }
    finish_in_reference_order(I, V, ni, vin, vout, L, R, C);
  }
};

// After the first step every internal voltage is V0_k + b*(I_{k-1} - I_k),
// so substituting it into the next current update gives, per step,
//   I_i <- K_i + (2d - g) I_i - d (I_{i-1} + I_{i+1})   (i = 1..8)
//   I_0 <- K_0 + (d - g) I_0 - d I_1,  I_9 <- K_9 + (d - g) I_9 - d I_8
// with d = h*b and K_i = h*(V0 difference across segment i) + g*I0_i.
// The first step uses the seeded voltages (and its inductive term is zero
// because the running current still equals I0).  28 VALU per step: for the
// eight interior currents one add (I_{i-1} + I_{i+1}, two register operands)
// and two FMAs, for the two end currents two FMAs.  The add form issues
// faster than a third three-operand FMA.  The last kTail steps run in the
// reference's order.
template <>
struct Relax<CKT_POLICY_FAST> {
  static __device__ __forceinline__ void run(float (&I)[kSeg], float (&V)[kSeg - 1], float vin,
                                             float vout, float L, float R, float C, int steps) {
    if (whole_loop_in_reference_order(I, V, vin, vout, L, R, C, steps)) {
      reference_loop<false>(I, V, vin, vout, L, R, C, steps);
      return;
    }
    const bool no_tail = wave_all(tail_stays_finite(I, V, vin, vout, L, R, C));
    const int body = no_tail ? steps : steps - kTail;
    const float dt = CKT_DT_F;
    const float h = 1.0f / R, g = (L / dt) / R, b = dt / C;
    const float d = h * b;
    const float a_end = d - g, a_mid = 2.0f * d - g;
    // m gets a register of its own: folded into the FMAs as a negation
    // modifier it would force every outer FMA into the 8-byte three-source
    // encoding instead of the 4-byte accumulate form.
    float m = -d;
    asm volatile("" : "+v"(m));
    float K[kSeg], x[kSeg];
    {
      float dv[kSeg];
      dv[0] = V[0] - vin;
#pragma unroll
      for (int i = 1; i < kSeg - 1; i++) dv[i] = V[i] - V[i - 1];
      dv[kSeg - 1] = vout - V[kSeg - 2];
#pragma unroll
      for (int i = 0; i < kSeg; i++) {
        K[i] = fmaf(h, dv[i], g * I[i]);
        x[i] = h * dv[i];
      }
    }
#define CKT_FAST_STEP                                                              \
  {                                                                                \
    float y[kSeg];                                                                 \
    y[0] = fmaf(a_end, x[0], fmaf(m, x[1], K[0]));                                 \
    _Pragma("unroll") for (int i = 1; i < kSeg - 1; i++) y[i] =                    \
        fmaf(m, x[i - 1] + x[i + 1], fmaf(a_mid, x[i], K[i]));                     \
    y[kSeg - 1] = fmaf(a_end, x[kSeg - 1], fmaf(m, x[kSeg - 2], K[kSeg - 1]));     \
    _Pragma("unroll") for (int i = 0; i < kSeg; i++) x[i] = y[i];                  \
  }
    int j = 1;
    for (; j + 4 <= body; j += 4) {
      CKT_FAST_STEP CKT_FAST_STEP CKT_FAST_STEP CKT_FAST_STEP
    }
    for (; j < body; j++) CKT_FAST_STEP
#undef CKT_FAST_STEP
    // The substitution has no voltages in its step, so same-signed infinite
    // currents would stay infinite where the reference's turn to NaN; values
    // never become finite again in it, so a non-finite x here means the wire
    // left the finite range (finish_in_reference_order writes it all-NaN).
    if (no_tail)
      end_from_currents(I, V, x, C);
    else
      finish_in_reference_order(I, V, x, vin, vout, L, R, C);
  }
};


// FAST's setup, step and end state as a struct, for SCALED (the wires it
// starts from, and its fallback step).  FAST keeps its own copy above: built
// on this struct, with the same loop instructions, its kernel ran 11% slower
// (10 loops over 64 pieces x 65536 wires: 0.187-0.193 s against
// 0.168-0.174 s).
struct Tridiag {
  float b, d, g, a_end, a_mid, m;
  float K[kSeg], x[kSeg];  // x: the currents after step 1

  __device__ __forceinline__ Tridiag(const float (&I)[kSeg], const float (&V)[kSeg - 1], float vin,
                                     float vout, float L, float R, float C) {
    const float dt = CKT_DT_F;
    const float h = 1.0f / R;
    g = (L / dt) / R;
    b = dt / C;
    d = h * b;
    a_end = d - g;
    a_mid = 2.0f * d - g;
    // m gets a register of its own: folded into the FMAs as a negation
    // modifier it would force every outer FMA into the 8-byte three-source
    // encoding instead of the 4-byte accumulate form.
    m = -d;
    asm volatile("" : "+v"(m));
    float dv[kSeg];
    dv[0] = V[0] - vin;
#pragma unroll
    for (int i = 1; i < kSeg - 1; i++) dv[i] = V[i] - V[i - 1];
    dv[kSeg - 1] = vout - V[kSeg - 2];
#pragma unroll
    for (int i = 0; i < kSeg; i++) {
      K[i] = fmaf(h, dv[i], g * I[i]);
      x[i] = h * dv[i];
    }
  }

  // One step, as FAST's.
  __device__ __forceinline__ void step() {
    float y[kSeg];
    y[0] = fmaf(a_end, x[0], fmaf(m, x[1], K[0]));
#pragma unroll
    for (int i = 1; i < kSeg - 1; i++) y[i] = fmaf(m, x[i - 1] + x[i + 1], fmaf(a_mid, x[i], K[i]));
    y[kSeg - 1] = fmaf(a_end, x[kSeg - 1], fmaf(m, x[kSeg - 2], K[kSeg - 1]));
#pragma unroll
    for (int i = 0; i < kSeg; i++) x[i] = y[i];
  }

  __device__ __forceinline__ void steps4(int n) {
    for (int j = 0; j < n; j++) {
      step();
      step();
      step();
      step();
    }
  }
};

// The tridiagonal update x <- K + A x, iterated as the deviation from its
// fixed point: with x* = K + A x* (one tridiagonal solve per wire and loop)
// the deviation e = x - x* follows e <- A e.  A = a I + m (neighbours), so
// writing e = m^t w within a block of four steps leaves
//   w_i <- c w_i + (w_{i-1} + w_{i+1}),   c = a / m,
// one add and one accumulate-FMA per interior current and one FMA per end
// current, plus 10 multiplies by m^4 per block to bring w back to e.  The
// step commutes with the mirror i -> 9 - i (A's two ends carry the same
// coefficients), so the symmetric half s_i = w_i + w_{9-i} and the
// antisymmetric half a_i = w_i - w_{9-i} (i < 5) each follow a five-point
// stencil of the same form, closed at the middle by the mirror image
// (w_5 = +-w_4 there): 8 VALU per half, 16 per step, against 18 on the ten
// currents.  Every step updates all ten coupled values; the currents at any
// step are x* + m^t (s +- a) / 2.
//
// The rewrite is used for a wire when one of two conditions shows that it
// gives the reference's finite/non-finite outcome; other wires take the
// plain tridiagonal steps:
//   contracting: the step's infinity-norm |a_mid| + 2d (|a_end| + d at the
//     ends) is at most 1/2, so e shrinks at least 2x per step and the
//     halves stay below 2 (|c|+2)^4 max|e| (every row of either half's
//     stencil sums to at most |c|+2 in magnitude), checked to stay finite; m^4
//     is a normal float.  Rounding of m^4 scales e, which has decayed below
//     FP32 resolution of x* within a few dozen steps.
//   diverging: the eigenvalues of A lie in [-g, -g + 4d] (A = a_mid I + m B
//     with B the path adjacency plus ones at both ends, eigenvalues
//     2 cos(pi k / 10)); if -g + 4d <= -2 every mode grows at least 2x per
//     step, A is symmetric, so |e| >= 2^(steps) |e_0|, and if that passes
//     FLT_MAX by the end of the blocks, kTail steps before the end, the
//     reference's iteration has overflowed and spread NaN to every segment
//     by its last step.  Then this one has overflowed too (max(|s_i|, |a_i|)
//     >= m^t |e_i| elementwise, and an infinite half stays non-finite), and
//     both end all-NaN.
// Either way the last kTail steps run in the reference's order, unless the
// wire provably stays finite through them (tail_stays_finite).
// The reference parameters (g ~ 1e4..1e5) are diverging (and their inputs
// non-finite after the first loop); the stable ones (g <= 0.1, d <= 1e-3)
// contracting.
template <>
struct Relax<CKT_POLICY_SCALED> {
  static __device__ __forceinline__ void run(float (&I)[kSeg], float (&V)[kSeg - 1], float vin,
                                             float vout, float L, float R, float C, int steps) {
    if (whole_loop_in_reference_order(I, V, vin, vout, L, R, C, steps)) {
      reference_loop<false>(I, V, vin, vout, L, R, C, steps);
      return;
    }
    const bool no_tail = wave_all(tail_stays_finite(I, V, vin, vout, L, R, C));
    Tridiag t(I, V, vin, vout, L, R, C);
    const int rest = steps - 1 - (no_tail ? 0 : kTail);
    for (int j = 0; j < (rest & 3); j++) t.step();
    const int blocks = rest >> 2;
    if (blocks > 0) run_blocks(t, blocks);
    if (no_tail)
      end_from_currents(I, V, t.x, C);
    else
      finish_in_reference_order(I, V, t.x, vin, vout, L, R, C);
  }

 private:
  static constexpr int kHalf = kSeg / 2;
  static_assert(kSeg % 2 == 0 && kHalf >= 2, "the mirror split needs an even segment count");

  // `blocks` blocks of four steps from t.x, left in t.x: rewritten where one
  // of the two conditions holds, as plain tridiagonal steps otherwise.
  static __device__ __forceinline__ void run_blocks(Tridiag& t, int blocks) {
    // x*: (1 - a_i) x*_i + d (x*_{i-1} + x*_{i+1}) = K_i, by the Thomas
    // algorithm (diagonally dominant under either condition below).
    float xs[kSeg];
    {
      float cp[kSeg], kp[kSeg];
      const float De = 1.0f - t.a_end, Dm = 1.0f - t.a_mid;
      float r = 1.0f / De;
      cp[0] = t.d * r;
      kp[0] = t.K[0] * r;
#pragma unroll
      for (int i = 1; i < kSeg; i++) {
        r = 1.0f / fmaf(-t.d, cp[i - 1], i == kSeg - 1 ? De : Dm);
        cp[i] = t.d * r;
        kp[i] = fmaf(-t.d, kp[i - 1], t.K[i]) * r;
      }
      xs[kSeg - 1] = kp[kSeg - 1];
#pragma unroll
      for (int i = kSeg - 2; i >= 0; i--) xs[i] = fmaf(-cp[i], xs[i + 1], kp[i]);
    }
    float emax = 0.0f;
    bool e_finite = true;
#pragma unroll
    for (int i = 0; i < kSeg; i++) {
      const float di = t.x[i] - xs[i];
      emax = fmaxf(emax, fabsf(di));
      e_finite &= __builtin_isfinite(di);
    }
    // |w| grows at most by max(|c_mid| + 2, |c_end| + 1) = norm / d per step.
    const float norm = fmaxf(fabsf(t.a_mid) + 2.0f * t.d, fabsf(t.a_end) + t.d);
    const float grow = norm / t.d;
    const bool contracting =
        norm <= 0.5f && (!e_finite || emax * ((grow * grow) * (grow * grow)) <= 0x1p120f);
    // Steps of guaranteed growth: the blocks, which end kTail steps before
    // the end, time enough for the reference's NaN to reach every segment;
    // 2^132 max|e| is past FLT_MAX * sqrt(10).
    const int grown = min(4 * blocks, 256);
    const bool diverging = t.g - 4.0f * t.d >= 2.0f && grown > 0 &&
                           (!e_finite || (emax > 0.0f && emax >= ldexpf(1.0f, 132 - grown)));
    // d >= 2^-25 keeps m^4 a normal float (>= 2^-100).  Under either
    // condition the fixed-point system is diagonally dominant, so a
    // non-finite deviation means non-finite inputs (a NaN node voltage, say)
    // or a fixed point past FLT_MAX: the reference leaves the finite range
    // then, and so does either step here, the deviation staying non-finite.
    if (!(t.d >= 0x1p-25f && (contracting || diverging))) {
      t.steps4(blocks);
      return;
    }
    const float c_mid = t.a_mid / t.m, c_end = t.a_end / t.m;
    float m4 = (t.m * t.m) * (t.m * t.m);
    asm volatile("" : "+v"(m4));
    // The deviation is formed again here, from opaque copies of x*, so that
    // it is not kept live from the bound above across the fallback branch,
    // and split into its mirror-symmetric and antisymmetric halves.
    float s[kHalf], a[kHalf];
#pragma unroll
    for (int i = 0; i < kHalf; i++) {
      asm volatile("" : "+v"(xs[i]), "+v"(xs[kSeg - 1 - i]));
      const float lo = t.x[i] - xs[i], hi = t.x[kSeg - 1 - i] - xs[kSeg - 1 - i];
      s[i] = lo + hi;
      a[i] = lo - hi;
    }
    for (int j = 0; j < blocks; j++) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        half_step(s, c_end, c_mid, c_mid + 1.0f);
        half_step(a, c_end, c_mid, c_mid - 1.0f);
      }
#pragma unroll
      for (int i = 0; i < kHalf; i++) {
        s[i] *= m4;
        a[i] *= m4;
      }
    }
#pragma unroll
    for (int i = 0; i < kHalf; i++) {
      t.x[i] = fmaf(0.5f, s[i] + a[i], xs[i]);
      t.x[kSeg - 1 - i] = fmaf(0.5f, s[i] - a[i], xs[kSeg - 1 - i]);
    }
  }

  // One step of one half: w_0 <- c_end w_0 + w_1 at the wire end,
  // w_i <- c_mid w_i + (w_{i-1} + w_{i+1}) inside, and at the middle, where
  // the mirror image w_{9-i} = +-w_i stands in for the missing neighbour,
  // w_4 <- (c_mid +- 1) w_4 + w_3.
  static __device__ __forceinline__ void half_step(float (&w)[kHalf], float c_end, float c_mid,
                                                   float c_fold) {
    float y[kHalf];
    y[0] = fmaf(c_end, w[0], w[1]);
#pragma unroll
    for (int i = 1; i < kHalf - 1; i++) y[i] = fmaf(c_mid, w[i], w[i - 1] + w[i + 1]);
    y[kHalf - 1] = fmaf(c_fold, w[kHalf - 1], w[kHalf - 2]);
#pragma unroll
    for (int i = 0; i < kHalf; i++) w[i] = y[i];
  }
};

// ---------------------------------------------------------------------------
// calc_new_currents over SoA circuits.

// Global-address-space views of the SoA arrays: loads and stores use
// global_* instructions with a wave-uniform base, and the lane index is made
// opaque after the step loop so that the 19 per-lane store addresses are
// recomputed there instead of being held in VGPRs across the loop.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const uint32_t gcu32;

// FUSE: the wire's two charge contributions (distribute_charge's ATOMIC
// mode, circuit_gpu.cu:273-283) are added from registers right after the
// relaxation, so the loop drivers skip the separate launch.
template <int POLICY, bool FUSE>
__device__ __forceinline__ void cnc_one(const DevCircuit& c, uint32_t w, int steps) {
  const size_t ws = c.ws;
  gfloat* cur = (gfloat*)c.cur;
  gfloat* vol = (gfloat*)c.vol;
  if (!CKT_IN_BOUNDS(w < c.num_wires && c.in_ptr[w] < c.num_nodes && c.out_ptr[w] < c.num_nodes))
    return;
  float I[kSeg], V[kSeg - 1];
#pragma unroll
  for (int s = 0; s < kSeg; s++) I[s] = cur[s * ws + w];
#pragma unroll
  for (int s = 0; s < kSeg - 1; s++) V[s] = vol[s * ws + w];
  const float L = ((gfloat*)c.ind)[w], R = ((gfloat*)c.res)[w], C = ((gfloat*)c.wcap)[w];
  const float vin = ((gfloat*)c.voltage)[((gcu32*)c.in_ptr)[w]];
  const float vout = ((gfloat*)c.voltage)[((gcu32*)c.out_ptr)[w]];
  Relax<POLICY>::run(I, V, vin, vout, L, R, C, steps);
  asm volatile("" : "+v"(w));
#pragma unroll
  for (int s = 0; s < kSeg; s++) cur[s * ws + w] = I[s];
#pragma unroll
  for (int s = 0; s < kSeg - 1; s++) vol[s * ws + w] = V[s];
  if (FUSE) {
    const float dt = CKT_DT_F;
    unsafeAtomicAdd(&c.charge[c.in_ptr[w]], -dt * I[0]);
    unsafeAtomicAdd(&c.charge[c.out_ptr[w]], dt * I[kSeg - 1]);
  }
}

// One tile per workgroup.  A tile holds up to BLOCK wires per lane pass; lane
// l relaxes wires l, l + BLOCK, ... of it one after another.
template <int POLICY, int BLOCK, bool FUSE>
__global__ __launch_bounds__(BLOCK) void cnc_kernel(const DevCircuit* __restrict__ circuits,
                                                    const Tile* __restrict__ tiles, int steps) {
  const Tile t = tiles[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < t.count; i += BLOCK)
    cnc_one<POLICY, FUSE>(circuits[t.circuit], t.first + i, steps);
}

// Persistent form: a fixed grid pulls tiles from a queue counter.  The last
// block to finish resets the two counters, so consecutive launches on one
// stream reuse them without a separate clear.
template <int POLICY, int BLOCK, bool FUSE>
__global__ __launch_bounds__(BLOCK) void cnc_persistent_kernel(const DevCircuit* __restrict__ circuits,
                                                               const Tile* __restrict__ tiles,
                                                               uint32_t ntiles, uint32_t* counters,
                                                               int steps) {
  __shared__ uint32_t s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&counters[0], 1u);
    __syncthreads();
    const uint32_t ti = s_tile;
    __syncthreads();
    if (ti >= ntiles) break;
    const Tile t = tiles[ti];
    for (uint32_t i = threadIdx.x; i < t.count; i += BLOCK)
      cnc_one<POLICY, FUSE>(circuits[t.circuit], t.first + i, steps);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t done = atomicAdd(&counters[1], 1u);
    if (done == gridDim.x - 1) {
      atomicExch(&counters[0], 0u);
      atomicExch(&counters[1], 0u);
    }
  }
}

// ---------------------------------------------------------------------------
// distribute_charge.

// One lane per wire, float atomics into the node charges.
__global__ __launch_bounds__(256) void dsc_atomic_kernel(const DevCircuit* __restrict__ circuits,
                                                         const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (threadIdx.x >= t.count) return;
  const DevCircuit& c = circuits[t.circuit];
  const uint32_t w = t.first + threadIdx.x;
  if (!CKT_IN_BOUNDS(w < c.num_wires && c.in_ptr[w] < c.num_nodes && c.out_ptr[w] < c.num_nodes))
    return;
  const float dt = CKT_DT_F;
  const float q_in = -dt * c.cur[w];
  const float q_out = dt * c.cur[(size_t)(kSeg - 1) * c.ws + w];
  unsafeAtomicAdd(&c.charge[c.in_ptr[w]], q_in);
  unsafeAtomicAdd(&c.charge[c.out_ptr[w]], q_out);
}

#ifdef CKT_BOUNDS_CHECKS
static __device__ unsigned long long g_bounds_violations;
__device__ bool bounds_violation() {
  atomicAdd(&g_bounds_violations, 1ull);
  return false;
}
#endif

// Device-scope float add to global memory, result unused.
__device__ __forceinline__ void global_add(gfloat* p, float v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS-privatised reduction.  Workgroup b takes wire subrange b % splits of
// piece tile b / splits (the piece's wires cut into `splits` ranges of whole
// 1024-wire blocks).  Contributions to the piece's own node range are summed
// in LDS, everything else goes straight to HBM with float atomics.  Each
// workgroup writes its sums to its own row of `partial` (plain stores) and
// dsc_lds_sum_kernel adds the rows to the nodes after every atomic of this
// launch has landed.  Without a partial buffer (one subrange per piece) the
// workgroup adds its sums to the piece's nodes itself: private nodes receive
// contributions only from their own piece's wires, so they take a plain
// read-add-write, shared nodes (which other pieces' workgroups reach)
// atomics.  Each lane loads four wires' contributions before it adds any, so
// that 16 loads are in flight per lane.  (LDS float atomics, ds_add_f32, take
// about a lane per clock per CU and bound the kernel; integer LDS atomics
// and plain LDS accesses cost far less, but a counting sort of each chunk's
// contributions built on them, with its scan and barriers and one
// 160 KB workgroup per CU, ran at half this speed.)
template <bool SPLIT>
__global__ __launch_bounds__(1024) void dsc_lds_kernel(const DevCircuit* __restrict__ circuits,
                                                       const PieceTile* __restrict__ ptiles,
                                                       uint32_t splits, float* __restrict__ partial) {
  extern __shared__ float acc[];
  const uint32_t si = blockIdx.x % splits;
  const PieceTile pt = ptiles[blockIdx.x / splits];
  const ckt_piece_t p = circuits[pt.circuit].pieces[pt.piece];
  // The circuit's arrays, read once: the atomics below write memory the
  // compiler cannot tell apart from the descriptor table, so fields read
  // through it would be reloaded after every atomic.
  const uint32_t ws = circuits[pt.circuit].ws;
  [[maybe_unused]] const uint32_t num_wires = circuits[pt.circuit].num_wires;
  [[maybe_unused]] const uint32_t num_nodes = circuits[pt.circuit].num_nodes;
  if (!CKT_IN_BOUNDS((uint64_t)p.first_node + p.num_nodes <= num_nodes)) return;
  const gfloat* cur0 = (const gfloat*)circuits[pt.circuit].cur;
  const gfloat* cur9 = cur0 + (size_t)(kSeg - 1) * ws;
  const gcu32* in_ptr = (const gcu32*)circuits[pt.circuit].in_ptr;
  const gcu32* out_ptr = (const gcu32*)circuits[pt.circuit].out_ptr;
  gfloat* charge = (gfloat*)circuits[pt.circuit].charge;
  const uint8_t* locator = circuits[pt.circuit].locator;
  uint32_t k0, k1;
  lds_subrange(p.num_wires, splits, si, &k0, &k1);
  const uint32_t npp = p.num_nodes;
  for (uint32_t i = threadIdx.x; i < npp; i += blockDim.x) acc[i] = 0.f;
  __syncthreads();
  const float dt = CKT_DT_F;
  const uint32_t stride = blockDim.x;
  constexpr int kU = 4;
  for (uint32_t base = k0 + threadIdx.x; base < k1; base += kU * stride) {
    float q[2 * kU];
    uint32_t node[2 * kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t k = base + u * stride;
      uint32_t w = p.first_wire + (k < k1 ? k : k0);
      if (!CKT_IN_BOUNDS(w < num_wires)) w = 0;
      q[2 * u] = -dt * cur0[w];
      q[2 * u + 1] = dt * cur9[w];
      node[2 * u] = in_ptr[w];
      node[2 * u + 1] = out_ptr[w];
    }
#pragma unroll
    for (int j = 0; j < 2 * kU; j++) {
      if (base + (j >> 1) * stride >= k1) break;
      if (!CKT_IN_BOUNDS(node[j] < num_nodes)) continue;
      const uint32_t l = node[j] - p.first_node;
      if (l < npp) atomicAdd(&acc[l], q[j]);
      else global_add(&charge[node[j]], q[j]);
    }
  }
  __syncthreads();
  if constexpr (SPLIT) {
    float* row = partial + pt.partial + (size_t)si * npp;
    for (uint32_t i = threadIdx.x; i < npp; i += blockDim.x) row[i] = acc[i];
  } else {
    for (uint32_t i = threadIdx.x; i < npp; i += blockDim.x) {
      const uint32_t n = p.first_node + i;
      const float v = acc[i];
      if (locator[n]) charge[n] = charge[n] + v;
      else if (v != 0.f) global_add(&charge[n], v);
    }
  }
}

// The rows of the split reduction added to the nodes: node i of piece tile
// blockIdx.y, in subrange order.  Every atomic of the first kernel (other
// pieces' contributions to this piece's shared nodes) has landed, and piece
// node ranges are disjoint, so each node has one writer here.
__global__ __launch_bounds__(256) void dsc_lds_sum_kernel(const DevCircuit* __restrict__ circuits,
                                                          const PieceTile* __restrict__ ptiles,
                                                          uint32_t splits,
                                                          const float* __restrict__ partial) {
  const PieceTile pt = ptiles[blockIdx.y];
  const DevCircuit& c = circuits[pt.circuit];
  const ckt_piece_t p = c.pieces[pt.piece];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= p.num_nodes) return;
  if (!CKT_IN_BOUNDS(p.first_node + i < c.num_nodes)) return;
  const float* row = partial + pt.partial + i;
  float v = 0.f;
  for (uint32_t s = 0; s < splits; s++) v += row[(size_t)s * p.num_nodes];
  const uint32_t n = p.first_node + i;
  c.charge[n] = c.charge[n] + v;
}

// One lane per node: the node's contributions in (wire, in-before-out) order,
// added one at a time to its charge -- the order in which the reference's
// CPU tasks apply them when run one piece after another.
// Node tiles in XCD-contiguous order (xcd_tile): neighbouring node tiles --
// whose wires, and so whose gathered currents, lie in the same piece -- share
// an L2.
template <bool KAHAN>
__global__ __launch_bounds__(256) void dsc_csr_kernel(const DevCircuit* __restrict__ circuits,
                                                      const Tile* __restrict__ tiles) {
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (threadIdx.x >= t.count) return;
  const DevCircuit& c = circuits[t.circuit];
  const uint32_t n = t.first + threadIdx.x;
  if (!CKT_IN_BOUNDS(n < c.num_nodes)) return;
  const float dt = CKT_DT_F;
  const float* cur0 = c.cur;
  const float* cur9 = c.cur + (size_t)(kSeg - 1) * c.ws;
  const uint32_t k0 = c.csr_row[n], k1 = c.csr_row[n + 1];
  if (!CKT_IN_BOUNDS(k0 <= k1 && k1 <= 2ull * c.num_wires)) return;
  if (!KAHAN || c.locator[n]) {
    float acc = c.charge[n];
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t e = c.csr_ent[k];
      const uint32_t w = e >> 1;
      if (!CKT_IN_BOUNDS(w < c.num_wires)) continue;
      const float v = (e & 1u) ? dt * cur9[w] : -dt * cur0[w];
      acc += v;
    }
    c.charge[n] = acc;
  } else {
    double s = (double)c.charge[n], comp = 0.0;
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t e = c.csr_ent[k];
      const uint32_t w = e >> 1;
      if (!CKT_IN_BOUNDS(w < c.num_wires)) continue;
      const float v = (e & 1u) ? dt * cur9[w] : -dt * cur0[w];
      const double y = (double)v - comp;
      const double tt = s + y;
      comp = (tt - s) - y;
      s = tt;
    }
    c.charge[n] = (float)s;
  }
}

// ---------------------------------------------------------------------------
// update_voltages (circuit_gpu.cu:330-345; circuit.cc:1299-1302).

__device__ __forceinline__ float upv_one(float v, float q, float C, float lk) {
  v = v + q / C;
  return v * (1.f - lk);
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4f gv4f;

template <bool NT>
__device__ __forceinline__ v4f load4(const gv4f* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <bool NT>
__device__ __forceinline__ void store4(v4f v, gv4f* p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Tiles of up to 2048 nodes: each lane updates two groups of 4 consecutive
// nodes, the groups interleaved across the wave so that every 128-bit access
// of a wave is contiguous.  Aligned groups (tile start on a 16-byte boundary;
// node arrays start 256-byte aligned) use 128-bit loads and stores, ragged
// ones go one node at a time.  NT: nontemporal accesses, for launches whose
// node arrays are larger than the last-level cache (they are streamed once
// per loop); smaller ones keep ordinary accesses so that a cache-resident
// working set stays resident.
template <bool NT>
__global__ __launch_bounds__(256) void upv_kernel(const DevCircuit* __restrict__ circuits,
                                                  const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const DevCircuit& c = circuits[t.circuit];
  gfloat* charge = (gfloat*)c.charge;
  gfloat* voltage = (gfloat*)c.voltage;
  gfloat* ncap = (gfloat*)c.ncap;
  gfloat* leak = (gfloat*)c.leak;
  const bool aligned = (t.first & 3u) == 0;
  if (!CKT_IN_BOUNDS((uint64_t)t.first + t.count <= c.num_nodes && t.count <= 2048u)) return;
#pragma unroll
  for (uint32_t g = 0; g < 2; g++) {
    const uint32_t k0 = upv_group_first(g, threadIdx.x);
    if (k0 >= t.count) continue;
    const uint32_t n = t.first + k0;
    if (aligned && k0 + 4 <= t.count) {
      const v4f q = load4<NT>((const gv4f*)(charge + n));
      const v4f C = load4<NT>((const gv4f*)(ncap + n));
      const v4f lk = load4<NT>((const gv4f*)(leak + n));
      v4f v = load4<NT>((const gv4f*)(voltage + n));
      v.x = upv_one(v.x, q.x, C.x, lk.x);
      v.y = upv_one(v.y, q.y, C.y, lk.y);
      v.z = upv_one(v.z, q.z, C.z, lk.z);
      v.w = upv_one(v.w, q.w, C.w, lk.w);
      store4<NT>(v, (gv4f*)(voltage + n));
      store4<NT>((v4f){0.f, 0.f, 0.f, 0.f}, (gv4f*)(charge + n));
    } else {
      const uint32_t end = t.count - k0 < 4u ? t.count - k0 : 4u;
      for (uint32_t k = 0; k < end; k++) {
        voltage[n + k] = upv_one(voltage[n + k], charge[n + k], ncap[n + k], leak[n + k]);
        charge[n + k] = 0.f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Wire invariant validator (generalises sanity_check_positions,
// circuit_gpu.cu:151-184): location codes valid, in-node inside the piece,
// out-node inside the circuit, and location consistent with the locator and
// the piece's node range.

__global__ __launch_bounds__(256) void sanity_kernel(const DevCircuit* __restrict__ circuits,
                                                     const Tile* __restrict__ tiles,
                                                     unsigned long long* bad) {
  const Tile t = tiles[blockIdx.x];
  bool ok = true;
  if (threadIdx.x < t.count) {
    const DevCircuit& c = circuits[t.circuit];
    const ckt_piece_t p = c.pieces[t.piece];
    const uint32_t w = t.first + threadIdx.x;
    const uint32_t a = c.in_ptr[w], b = c.out_ptr[w];
    const uint32_t la = c.in_loc[w], lb = c.out_loc[w];
    ok = la <= 2u && lb <= 2u && (a - p.first_node) < p.num_nodes && b < c.num_nodes;
    if (ok) {
      const uint32_t ends[2] = {a, b}, locs[2] = {la, lb};
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const bool own = (ends[e] - p.first_node) < p.num_nodes;
        const bool priv = c.locator[ends[e]] != 0;
        if (locs[e] == CKT_PRIVATE_PTR) ok = ok && own && priv;
        else if (locs[e] == CKT_SHARED_PTR) ok = ok && own && !priv;
        else ok = ok && !own && !priv;
      }
    }
  }
  const unsigned long long m = __ballot(!ok);
  if ((threadIdx.x & 63u) == 0 && m) atomicAdd(bad, (unsigned long long)__popcll(m));
}

__global__ void nop_kernel() {}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Accessor-form kernels over the reference AoS layouts (absolute indexing,
// GPU_Accessor::ref / AccessorGPUReductionFold, lowlevel.h:749-881).

__device__ __forceinline__ const ckt_node_t* pick_instance(int32_t loc, const ckt_node_t* pvt,
                                                           const ckt_node_t* owned,
                                                           const ckt_node_t* ghost) {
  switch (loc) { // get_node, circuit_gpu.cu:40-59
    case CKT_PRIVATE_PTR: return pvt;
    case CKT_SHARED_PTR: return owned;
    case CKT_GHOST_PTR: return ghost;
    default: return nullptr;
  }
}

// The accessor-form kernels run over a piece table (an index-space launch;
// a per-piece call is a table of one).  Workgroups are numbered across the
// pieces: piece p owns workgroups [first_block[p], first_block[p + 1]), each
// covering 256 consecutive elements (wires or nodes) of it.
__device__ __forceinline__ uint32_t piece_of_block(const PieceIndex& ix, uint32_t b,
                                                   uint32_t* elem) {
  return piece_of_block(ix.first_block, ix.num_pieces, b, threadIdx.x, elem);
}

template <int POLICY>
__global__ __launch_bounds__(256) void cnc_aos_kernel(PieceIndex ix, ckt_wire_t* __restrict__ wires,
                                                      const ckt_node_t* pvt, const ckt_node_t* owned,
                                                      const ckt_node_t* ghost, int steps,
                                                      unsigned long long* bad) {
  uint32_t k;
  const ckt_piece_t p = ix.pieces[piece_of_block(ix, blockIdx.x, &k)];
  if (k >= p.num_wires) return;
  ckt_wire_t& wire = wires[p.first_wire + k];
  const ckt_node_t* in_base = pick_instance(wire.in_loc, pvt, owned, ghost);
  const ckt_node_t* out_base = pick_instance(wire.out_loc, pvt, owned, ghost);
  if (!in_base || !out_base) {
    atomicAdd(bad, 1ull);
    return;
  }
  float I[kSeg], V[kSeg - 1];
#pragma unroll
  for (int s = 0; s < kSeg; s++) I[s] = wire.current[s];
#pragma unroll
  for (int s = 0; s < kSeg - 1; s++) V[s] = wire.voltage[s];
  const float vin = in_base[wire.in_ptr].voltage, vout = out_base[wire.out_ptr].voltage;
  Relax<POLICY>::run(I, V, vin, vout, wire.inductance, wire.resistance, wire.capacitance, steps);
#pragma unroll
  for (int s = 0; s < kSeg; s++) wire.current[s] = I[s];
#pragma unroll
  for (int s = 0; s < kSeg - 1; s++) wire.voltage[s] = V[s];
}

__device__ __forceinline__ bool reduce_local(int32_t loc, uint32_t ptr, float v, ckt_node_t* pvt,
                                             float* owned_fold, float* ghost_fold) {
  switch (loc) { // reduce_local, circuit_gpu.cu:219-243
    case CKT_PRIVATE_PTR: unsafeAtomicAdd(&pvt[ptr].charge, v); return true;
    case CKT_SHARED_PTR: unsafeAtomicAdd(&owned_fold[ptr], v); return true;
    case CKT_GHOST_PTR: unsafeAtomicAdd(&ghost_fold[ptr], v); return true;
    default: return false;
  }
}

__global__ __launch_bounds__(256) void dsc_aos_kernel(PieceIndex ix,
                                                      const ckt_wire_t* __restrict__ wires,
                                                      ckt_node_t* pvt, float* owned_fold,
                                                      float* ghost_fold, unsigned long long* bad) {
  uint32_t k;
  const ckt_piece_t p = ix.pieces[piece_of_block(ix, blockIdx.x, &k)];
  if (k >= p.num_wires) return;
  const ckt_wire_t& wire = wires[p.first_wire + k];
  const float dt = CKT_DT_F;
  bool ok = reduce_local(wire.in_loc, wire.in_ptr, -dt * wire.current[0], pvt, owned_fold, ghost_fold);
  ok = reduce_local(wire.out_loc, wire.out_ptr, dt * wire.current[kSeg - 1], pvt, owned_fold,
                    ghost_fold) && ok;
  if (!ok) atomicAdd(bad, 1ull);
}

__global__ __launch_bounds__(256) void upv_aos_kernel(PieceIndex ix, ckt_node_t* pvt,
                                                      ckt_node_t* owned,
                                                      const int32_t* __restrict__ locator) {
  uint32_t k;
  const ckt_piece_t p = ix.pieces[piece_of_block(ix, blockIdx.x, &k)];
  if (k >= p.num_nodes) return;
  const uint32_t i = p.first_node + k;
  ckt_node_t& nd = locator[i] ? pvt[i] : owned[i];
  float v = nd.voltage;
  v = v + nd.charge / nd.capacitance;
  v = v * (1.f - nd.leakage);
  nd.voltage = v;
  nd.charge = 0.f;
}

__global__ __launch_bounds__(256) void sanity_aos_kernel(ckt_piece_t p, const ckt_wire_t* wires,
                                                         uint32_t num_nodes_total,
                                                         unsigned long long* bad) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = true;
  if (tid < p.num_wires) {
    const ckt_wire_t& wire = wires[p.first_wire + tid];
    ok = (uint32_t)wire.in_loc <= 2u && (uint32_t)wire.out_loc <= 2u &&
         (wire.in_ptr - p.first_node) < p.num_nodes && wire.out_ptr < num_nodes_total;
  }
  const unsigned long long m = __ballot(!ok);
  if ((threadIdx.x & 63u) == 0 && m) atomicAdd(bad, (unsigned long long)__popcll(m));
}

__global__ __launch_bounds__(256) void fold_init_kernel(float* fold, uint32_t first, uint32_t count) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < count) fold[first + tid] = 0.0f; // AccumulateCharge::identity, circuit.cc:30
}

__global__ __launch_bounds__(256) void fold_apply_kernel(ckt_node_t* owner, const float* fold,
                                                         uint32_t first, uint32_t count) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < count) owner[first + tid].charge += fold[first + tid]; // apply<true>, circuit.cc:32-36
}

// One 16-byte element per lane, one workgroup per 4 KiB: many independent
// loads in flight (a grid-stride loop over a capped grid keeps one per lane
// and reached ~60% of HBM; this form ~75%).
__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src, size_t n16) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers.

static inline uint32_t blocks_for(size_t n, int b) { return (uint32_t)((n + b - 1) / b); }

// Plain grid (one tile per workgroup) or persistent grid (`grid` workgroups
// pulling tiles), for every policy x workgroup size x fused epilogue.
template <int POLICY, int BLOCK, bool FUSE>
static void cnc_go(bool persistent, const DevCircuit* c, const Tile* t, uint32_t n, int grid,
                   uint32_t* ctr, int steps, hipStream_t s) {
  if (persistent)
    hipLaunchKernelGGL((cnc_persistent_kernel<POLICY, BLOCK, FUSE>), dim3(grid), dim3(BLOCK), 0, s,
                       c, t, n, ctr, steps);
  else
    hipLaunchKernelGGL((cnc_kernel<POLICY, BLOCK, FUSE>), dim3(n), dim3(BLOCK), 0, s, c, t, steps);
}

template <int POLICY, bool FUSE>
static hipError_t cnc_by_block(bool persistent, const DevCircuit* c, const Tile* t, uint32_t n,
                               int block, int grid, uint32_t* ctr, int steps, hipStream_t s) {
  switch (block) {
    case 64: cnc_go<POLICY, 64, FUSE>(persistent, c, t, n, grid, ctr, steps, s); break;
    case 128: cnc_go<POLICY, 128, FUSE>(persistent, c, t, n, grid, ctr, steps, s); break;
    case 256: cnc_go<POLICY, 256, FUSE>(persistent, c, t, n, grid, ctr, steps, s); break;
    case 512: cnc_go<POLICY, 512, FUSE>(persistent, c, t, n, grid, ctr, steps, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int POLICY>
static hipError_t cnc_by_fuse(bool fuse, bool persistent, const DevCircuit* c, const Tile* t,
                              uint32_t n, int block, int grid, uint32_t* ctr, int steps,
                              hipStream_t s) {
  return fuse ? cnc_by_block<POLICY, true>(persistent, c, t, n, block, grid, ctr, steps, s)
              : cnc_by_block<POLICY, false>(persistent, c, t, n, block, grid, ctr, steps, s);
}

static hipError_t cnc_launch(int policy, bool fuse, bool persistent, const DevCircuit* c,
                             const Tile* t, uint32_t n, int block, int grid, uint32_t* ctr,
                             int steps, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (persistent && grid < 1) return hipErrorInvalidValue;
  switch (policy) {
    case CKT_POLICY_EXACT:
      return cnc_by_fuse<CKT_POLICY_EXACT>(fuse, persistent, c, t, n, block, grid, ctr, steps, s);
    case CKT_POLICY_RECIP:
      return cnc_by_fuse<CKT_POLICY_RECIP>(fuse, persistent, c, t, n, block, grid, ctr, steps, s);
    case CKT_POLICY_FAST:
      return cnc_by_fuse<CKT_POLICY_FAST>(fuse, persistent, c, t, n, block, grid, ctr, steps, s);
    case CKT_POLICY_SCALED:
      return cnc_by_fuse<CKT_POLICY_SCALED>(fuse, persistent, c, t, n, block, grid, ctr, steps, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_cnc(int policy, bool fuse_dsc, const DevCircuit* c, const Tile* t, uint32_t n,
                      int block, int steps, hipStream_t s) {
  return cnc_launch(policy, fuse_dsc, false, c, t, n, block, 0, nullptr, steps, s);
}

hipError_t launch_cnc_persistent(int policy, bool fuse_dsc, const DevCircuit* c, const Tile* t,
                                 uint32_t n, int block, int grid, uint32_t* ctr, int steps,
                                 hipStream_t s) {
  return cnc_launch(policy, fuse_dsc, true, c, t, n, block, grid, ctr, steps, s);
}

hipError_t launch_dsc_atomic(const DevCircuit* c, const Tile* t, uint32_t n, int block,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (block != 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dsc_atomic_kernel, dim3(n), dim3(256), 0, s, c, t);
  return hipGetLastError();
}

// Bytes of dynamic LDS the reduction needs for pieces of up to max_nodes
// nodes.
size_t dsc_lds_bytes(uint32_t max_nodes) {
  const size_t b = (size_t)max_nodes * 4;
  return b < 256 ? 256 : b;
}

template <bool SPLIT>
static hipError_t dsc_lds_go(const DevCircuit* c, const PieceTile* pt, uint32_t grid,
                             uint32_t splits, float* partial, size_t lds_bytes, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    // Allows dynamic LDS beyond the default limit where the runtime needs it.
    (void)hipFuncSetAttribute((const void*)dsc_lds_kernel<SPLIT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    attr_set = true;
  }
  hipLaunchKernelGGL(dsc_lds_kernel<SPLIT>, dim3(grid), dim3(1024), lds_bytes, s, c, pt, splits,
                     partial);
  return hipGetLastError();
}

hipError_t launch_dsc_lds(const DevCircuit* c, const PieceTile* pt, uint32_t n, uint32_t splits,
                          uint32_t max_nodes, float* partial, size_t lds_bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (splits < 1 || (splits > 1 && !partial)) return hipErrorInvalidValue;
  if (lds_bytes < dsc_lds_bytes(max_nodes)) return hipErrorInvalidValue;
  const uint64_t grid = (uint64_t)n * splits;
  if (grid > 0x7fffffffu) return hipErrorInvalidValue;
  hipError_t e = partial ? dsc_lds_go<true>(c, pt, (uint32_t)grid, splits, partial, lds_bytes, s)
                         : dsc_lds_go<false>(c, pt, (uint32_t)grid, 1u, nullptr, lds_bytes, s);
  if (e != hipSuccess || !partial || max_nodes == 0) return e;
  hipLaunchKernelGGL(dsc_lds_sum_kernel, dim3(blocks_for(max_nodes, 256), n), dim3(256), 0, s, c,
                     pt, splits, (const float*)partial);
  return hipGetLastError();
}

hipError_t launch_dsc_csr(int kahan, const DevCircuit* c, const Tile* t, uint32_t n, int block,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (block != 256) return hipErrorInvalidValue;
  if (kahan) hipLaunchKernelGGL(dsc_csr_kernel<true>, dim3(n), dim3(256), 0, s, c, t);
  else hipLaunchKernelGGL(dsc_csr_kernel<false>, dim3(n), dim3(256), 0, s, c, t);
  return hipGetLastError();
}

hipError_t launch_upv(const DevCircuit* c, const Tile* t, uint32_t n, int block, bool nontemporal,
                      hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (block != 256) return hipErrorInvalidValue;
  if (nontemporal) hipLaunchKernelGGL(upv_kernel<true>, dim3(n), dim3(256), 0, s, c, t);
  else hipLaunchKernelGGL(upv_kernel<false>, dim3(n), dim3(256), 0, s, c, t);
  return hipGetLastError();
}

hipError_t launch_sanity(const DevCircuit* c, const Tile* t, uint32_t n, int block,
                         unsigned long long* bad, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (block != 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sanity_kernel, dim3(n), dim3(256), 0, s, c, t, bad);
  return hipGetLastError();
}

hipError_t launch_nop(hipStream_t s) {
  hipLaunchKernelGGL(nop_kernel, dim3(1), dim3(64), 0, s);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_cnc_aos(int policy, const PieceIndex& ix, uint32_t blocks, ckt_wire_t* wires,
                          const ckt_node_t* pvt, const ckt_node_t* owned, const ckt_node_t* ghost,
                          int steps, unsigned long long* bad, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  const dim3 g(blocks), b(256);
  switch (policy) {
    case CKT_POLICY_EXACT:
      hipLaunchKernelGGL(cnc_aos_kernel<CKT_POLICY_EXACT>, g, b, 0, s, ix, wires, pvt, owned, ghost, steps, bad);
      break;
    case CKT_POLICY_RECIP:
      hipLaunchKernelGGL(cnc_aos_kernel<CKT_POLICY_RECIP>, g, b, 0, s, ix, wires, pvt, owned, ghost, steps, bad);
      break;
    case CKT_POLICY_FAST:
      hipLaunchKernelGGL(cnc_aos_kernel<CKT_POLICY_FAST>, g, b, 0, s, ix, wires, pvt, owned, ghost, steps, bad);
      break;
    case CKT_POLICY_SCALED:
      hipLaunchKernelGGL(cnc_aos_kernel<CKT_POLICY_SCALED>, g, b, 0, s, ix, wires, pvt, owned, ghost, steps, bad);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dsc_aos(const PieceIndex& ix, uint32_t blocks, const ckt_wire_t* wires,
                          ckt_node_t* pvt, float* owned_fold, float* ghost_fold,
                          unsigned long long* bad, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(dsc_aos_kernel, dim3(blocks), dim3(256), 0, s, ix, wires, pvt, owned_fold,
                     ghost_fold, bad);
  return hipGetLastError();
}

hipError_t launch_upv_aos(const PieceIndex& ix, uint32_t blocks, ckt_node_t* pvt,
                          ckt_node_t* owned, const int32_t* locator, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(upv_aos_kernel, dim3(blocks), dim3(256), 0, s, ix, pvt, owned, locator);
  return hipGetLastError();
}

hipError_t launch_sanity_aos(const ckt_piece_t p, const ckt_wire_t* wires, uint32_t num_nodes_total,
                             unsigned long long* bad, hipStream_t s) {
  if (p.num_wires == 0) return hipSuccess;
  hipLaunchKernelGGL(sanity_aos_kernel, dim3(blocks_for(p.num_wires, 256)), dim3(256), 0, s, p,
                     wires, num_nodes_total, bad);
  return hipGetLastError();
}

hipError_t launch_fold_init(float* fold, uint32_t first, uint32_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(fold_init_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, fold, first,
                     count);
  return hipGetLastError();
}

hipError_t launch_fold_apply(ckt_node_t* owner, const float* fold, uint32_t first, uint32_t count,
                             hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(fold_apply_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, s, owner, fold,
                     first, count);
  return hipGetLastError();
}

// The bounds-check counter of this device (read and cleared), or
// hipErrorNotSupported in builds without CKT_BOUNDS_CHECKS.
hipError_t read_bounds_violations(uint64_t* count) {
#ifdef CKT_BOUNDS_CHECKS
  unsigned long long h = 0;
  hipError_t e = hipMemcpyFromSymbol(&h, HIP_SYMBOL(g_bounds_violations), sizeof(h));
  if (e != hipSuccess) return e;
  const unsigned long long zero = 0;
  e = hipMemcpyToSymbol(HIP_SYMBOL(g_bounds_violations), &zero, sizeof(zero));
  *count = h;
  return e;
#else
  *count = 0;
  return hipErrorNotSupported;
#endif
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return hipErrorInvalidValue;
  const size_t n16 = bytes / 16;
  if (n16 > (size_t)0x7fffffff * 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(n16, 256)), dim3(256), 0, s, (uint4*)dst,
                     (const uint4*)src, n16);
  return hipGetLastError();
}

}  // namespace ckt
