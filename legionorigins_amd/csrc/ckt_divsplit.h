// ckt_divsplit.h -- correctly rounded FP32 division by a divisor that stays
// fixed over many dividends, with FP32 operations only: two fmas per quotient
// (the first one a multiply),
//
//     q = RN(x * zh + RN(x * zl)),     zh + zl ~ 1/y,
//
// the method of N. Brisebarre, J.-M. Muller and S. K. Raina, "Accelerating
// correctly rounded floating-point division when the divisor is known in
// advance", IEEE Transactions on Computers 53(8), 2004.  It returns RN(x / y)
// for every x for most divisors, and a divisor can be checked once, against a
// short list of dividends, before it is used.  Host-and-device functions, so
// that host code can check them exhaustively.
//
// Why the list is complete.  Scale x and y to significands X, Y in [1, 2)
// with integer significands Xi = X 2^23, Yi = Y 2^23 (while every value
// below stays normal, scaling by powers of two changes nothing), and let the
// split satisfy |zh + zl - 1/Y| + 2^-24 |zl| <= E.  The fma rounds
//     V = X zh + RN(X zl) = X/Y + X (zh + zl - 1/Y) + (RN(X zl) - X zl),
// so |V - Q| <= X E < 2E for Q = X/Y.  RN(V) differs from RN(Q) only if a
// rounding midpoint M lies between them or V = M, so |Q - M| <= X E.
//   * X >= Y: Q is in [1, 2), M = (2k+1) 2^-24, and
//     Q - M = (Xi 2^24 - Yi (2k+1)) / (Y 2^47), so D = |Xi 2^24 - Yi (2k+1)|
//     satisfies D <= X Y E 2^47 < E 2^49.
//   * X < Y: Q is in [1/2, 1), M = (2k+1) 2^-25, D = |Xi 2^25 - Yi (2k+1)|
//     and D < E 2^50.
// D is a nonzero integer.  With Yi = Y' 2^s, Y' odd, D = 2^s d with
// d = |Xi 2^(24-s) - Y' (2k+1)| (2^(25-s) when X < Y) odd, so
// Xi = +-d 2^-(24-s) (mod Y').  For each odd d up to the bound these are one
// or two residues, and every Xi of the case's range with one of them is a
// candidate; if the split gives RN(X / Y) for every candidate, it gives it for
// every X.  zh = RN(1/Y) and zl = RN(1/Y - zh) give E <= 1.5 2^-49; 98.7% of
// divisors pass with it, 99.9% with zl moved one ulp either way, which is
// also tried.  The rest are refused (make_div_split returns false) and their
// quotients are computed another way.
//
// When the result holds for the actual operands.  The split of y is the split
// of Y scaled by 2^-e (exact: |y| in [2^-40, 2^60] keeps zh and zl normal,
// since 2^-48 |zh| <= |zl| <= 2^-24 |zh|; a power of two gets zl = zh 2^-60,
// too small to move any normal quotient, so that zl is never zero).  Then
// q = RN(x/y) whenever t = RN(x zl) is finite and |q| >= 2^-100, and an
// overflowing quotient rounds to the infinity of its sign inside the fma.
// t can be subnormal there, with an absolute error up to 2^-150 on top of the
// relative one; with s = 2^(ex - ey) >= 2^-101 that adds less than 1/2 to the
// bound on D when X >= Y and less than 1 when X < Y (D < E 2^49 + 2^-102 / s,
// D < E 2^50 + 2^-101 / s), which the candidate lists include.  t enters the
// fma negated, as u = -t = fma(-x, zl, +0), which is +0 for a zero x of
// either sign: the fma then adds -0 to x zh, and a zero dividend gets the zero
// of IEEE division's sign.  The caller detects the rest: an infinite dividend,
// and a quotient beyond about 2^150, make u infinite (q can then be NaN or an
// infinity of the wrong sign); a quotient below 2^-100 in magnitude needs a
// nonzero dividend below |y| 2^-98 (div_split_key below div_split_key_bound).
// NaN dividends give NaN quotients and NaN u.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ckt {

__host__ __device__ inline uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float f32_from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

// The smallest |quotient| for which div_split_q is guaranteed (see above).
constexpr float kDivSplitMin = 0x1p-100f;

// One quotient, RN(x * zh - u) with u = RN(-x * zl + 0) (= -RN(x * zl), and
// +0 for a zero x), and u for the caller's test.
__host__ __device__ inline float div_split_q(float x, float zh, float zl, float* u_out) {
  const float u = fmaf(-x, zl, 0.f);
  *u_out = u;
  return fmaf(x, zh, -u);
}
__host__ __device__ inline float div_split_q(float x, float zh, float zl) {
  float u;
  return div_split_q(x, zh, zl, &u);
}

// A dividend's key for the small-dividend test: nonzero |x| in increasing
// order (the bits without the sign, doubled, minus 2), zeros of either sign
// wrapped to the top, NaN above every finite bound.
__host__ __device__ inline uint32_t div_split_key(float x) { return (f32_bits(x) << 1) - 2u; }
// Dividends with a key below this are nonzero and smaller than |y| 2^-98
// (rounded: at least |y| 2^-98 (1 - 2^-10) for |y| >= 2^-40), so every other
// dividend has a quotient of at least 2^-99 in magnitude, or zero.
__host__ __device__ inline uint32_t div_split_key_bound(float y) {
  return div_split_key(fabsf(y) * 0x1p-98f);
}

// 2^-k modulo an odd m < 2^24: k halvings modulo m.
__host__ __device__ inline uint32_t inv_pow2_mod(uint32_t m, int k) {
  uint32_t h = 1;
  for (int i = 0; i < k; i++) h = (h & 1u) ? (h + m) >> 1 : h >> 1;
  return h;
}

// Xi in [2^23, 2^24) as the float Xi 2^-23 in [1, 2).
__host__ __device__ inline float significand_float(uint32_t Xi) {
  return f32_from_bits((127u << 23) | (Xi & 0x7fffffu));
}

// True if div_split_q(X, zh, zl) == RN(X / Y) for every candidate X (see the
// top of this file) of a divisor Y = Yi 2^-23 with D below dA (X >= Y) and dB
// (X < Y).  Yi is not a power of two.
__host__ __device__ inline bool div_split_candidates_pass(uint32_t Yi, float zh, float zl, int dA,
                                                          int dB) {
  const float Y = significand_float(Yi);
  const int s = __builtin_ctz(Yi);
  const uint32_t Yp = Yi >> s;
  for (int c = 0; c < 2; c++) {
    const int dmax = c ? dB : dA;
    const uint32_t lo = c ? (1u << 23) : Yi, hi = c ? Yi : (1u << 24);
    if ((1 << s) > dmax) continue;  // also keeps Yp >= 2^20 below
    const uint32_t inv = inv_pow2_mod(Yp, (c ? 25 : 24) - s);
    for (int d = 1; (d << s) <= dmax; d += 2) {
      const uint32_t a = (uint32_t)(((uint64_t)d * inv) % Yp);
      for (int sg = 0; sg < 2; sg++) {
        const uint32_t r = sg ? Yp - a : a;  // a != 0: d < Yp and inv is a unit
        uint32_t X = r;
        if (X < lo) X += ((lo - X + Yp - 1) / Yp) * Yp;
        for (; X < hi; X += Yp) {
          const float x = significand_float(X);
          if (f32_bits(div_split_q(x, zh, zl)) != f32_bits(x / Y)) return false;
        }
      }
    }
  }
  return true;
}

// Chooses a split (zh, zl) of 1/y for which div_split_q gives RN(x / y)
// whenever x is zero, or u is finite and |result| >= kDivSplitMin.  Returns
// false, with zh = zl = 0, for divisors outside [2^-40, 2^60] in magnitude and
// for those none of the tried splits is verified for.
__host__ __device__ inline bool make_div_split(float y, float* zh_out, float* zl_out) {
  *zh_out = 0.f;
  *zl_out = 0.f;
  const float ay = fabsf(y);
  if (!(ay >= 0x1p-40f && ay <= 0x1p60f)) return false;
  const uint32_t by = f32_bits(ay);
  const int e = (int)(by >> 23) - 127;
  const uint32_t Yi = (by & 0x7fffffu) | 0x800000u;
  const float sign = y < 0.f ? -1.f : 1.f;
  if (Yi == 0x800000u) {  // a power of two: 1/y is exact
    *zh_out = sign * ldexpf(1.f, -e);
    *zl_out = sign * ldexpf(1.f, -e - 60);
    return true;
  }
  const float Y = significand_float(Yi);
  const float zh = 1.0f / Y;
  const double rem = 1.0 - (double)Y * (double)zh;  // exact: a multiple of 2^-47 below 2^-23
  const float zl0 = (float)(rem / (double)Y);       // RN(1/Y - zh) (53 >= 2*24 + 2)
  for (int m = 0; m < 3; m++) {
    const uint32_t b0 = f32_bits(zl0);
    const float zl = m == 0 ? zl0 : f32_from_bits(m == 1 ? b0 + 1u : b0 - 1u);  // one ulp out, in
    const float big = fmaxf(fabsf(zl0), fabsf(zl));
    const double ulp = ldexp(1.0, (int)((f32_bits(big) >> 23) & 0xffu) - 127 - 23);
    const double E = (m ? 1.5 : 0.5) * ulp + ldexp((double)big, -24);
    const int dA = (int)(E * 0x1p49 + 0.51), dB = (int)(E * 0x1p50 + 1.01);
    if (div_split_candidates_pass(Yi, zh, zl, dA, dB)) {
      *zh_out = sign * ldexpf(zh, -e);
      *zl_out = sign * ldexpf(zl, -e);
      return true;
    }
  }
  return false;
}

constexpr float kStepBound = 0x1p40f;

// The EXACT step (circuit_gpu.cu:114-129) without its guard over |u| =
// |RN(a zl)|: a wave may take it while every lane's running currents and
// internal voltages and loop constants are within kStepBound, its |R| is at
// least 2^-20 and its |C| at least 2^-40 (then no dividend or u is infinite;
// see split_step), and for this many steps after one check of that state, for
// one wire.  From a state and loop constants of magnitude at most S (NaN aside),
// one step (circuit_gpu.cu:114-129, 1/dt < 2^20) gives
//   |L (I - I0)| / dt <= 2^21 |L| S,  R dividends <= (2 + 2^21 |L|) S,
//   new currents <= g2 S with g2 = (2 + 2^21 |L|) / |R|,
//   C dividends <= 2 dt g2 S <= 2^-18 g2 S,  their quotients <= g3 S with
//   g3 = 2^-18 g2 / |C|,  new internal voltages <= (1 + g3) S,
// so the state grows by at most G = max(g2, 1 + g3) per step and every value
// of a step stays within H S, H = max(2 + 2^21 |L|, g2, 1 + g3); every u is at
// most 2^16 times its dividend (|zl| <= 2^-24 |zh| <= 2^16 for |R|, |C| >=
// 2^-40).  From S <= kStepBound = 2^40, K steps keep every value below
// 2^(40 + lg H + (K - 1) lg G) and every u below 2^16 times that, finite
// while lg H + (K - 1) lg G < 71 (lg v = ilogb(v) + 1 >= log2 v, rounding
// included); the last step's C quotients may overflow (the next step is
// checked again).  1 for constants that allow no more (or are not finite),
// at most 8; one step is always allowed under the conditions above.
__host__ __device__ inline int unchecked_steps(float L, float R, float C) {
  const float a = fmaf(fabsf(L), 0x1p21f, 2.f);
  const float g2 = a / fabsf(R);
  const float g3 = g2 * 0x1p-18f / fabsf(C);
  const float G = fmaxf(g2, 1.f + g3), H = fmaxf(fmaxf(a, g2), 1.f + g3);
  if (!(G <= 0x1p60f && H <= 0x1p60f)) return 1;
  const int lgG = ilogbf(fmaxf(G, 1.f)) + 1, lgH = ilogbf(fmaxf(H, 1.f)) + 1;
  const int k = 1 + (70 - lgH) / lgG;
  return k < 1 ? 1 : (k > 8 ? 8 : k);
}

}  // namespace ckt
