// Host checks of the split division (ckt_divsplit.h), the division of the
// EXACT calc_new_currents step, and of the bound on its unguarded steps
// (unchecked_steps): for every divisor make_div_split accepts,
// div_split_q(x) == x / y (IEEE, correctly rounded) for every dividend
// significand, for random finite operands outside the kernel's exact-path
// tests, and for zero dividends of either sign.  Built and run by tests/test_division_method.py;
// exits non-zero after printing the first failures.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "ckt_divsplit.h"

using namespace ckt;

static int fails = 0;
#define CHECK(cond, ...)                                \
  do {                                                  \
    if (!(cond)) {                                      \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                         \
      std::printf("\n");                                \
      if (++fails > 10) std::exit(1);                   \
    }                                                   \
  } while (0)

// Every dividend significand, scaled to y's binade times 2^k (quotients in
// (1/2, 2) 2^k), against IEEE division, on 8 threads.
static long exhaustive_mismatches(float y, float zh, float zl, int k) {
  const float scale = ldexpf(1.f, ilogbf(fabsf(y)) + k);
  std::vector<long> bad(8, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < 8; t++)
    th.emplace_back([&, t] {
      for (uint32_t Xi = (1u << 23) + t; Xi < (1u << 24); Xi += 8) {
        const float x = significand_float(Xi) * scale;
        if (f32_bits(div_split_q(x, zh, zl)) != f32_bits(x / y)) bad[t]++;
      }
    });
  for (auto& x : th) x.join();
  long n = 0;
  for (long b : bad) n += b;
  return n;
}

static void check_divisor_exhaustive(float y, const char* what) {
  float zh, zl;
  if (!make_div_split(y, &zh, &zl)) {
    std::printf("  %s (%a): refused\n", what, y);
    return;
  }
  // quotients near 1, and just above kDivSplitMin, where RN(x zl) is subnormal
  for (int k : {0, -99}) {
    const long bad = exhaustive_mismatches(y, zh, zl, k);
    CHECK(bad == 0, "%s (%a), quotients near 2^%d: %ld of 2^23 dividends differ from IEEE division",
          what, y, k, bad);
  }
}

// A divisor whose first split (zl = RN(1/Y - zh)) fails on a candidate.
static bool base_split_fails(uint32_t Yi) {
  const float Y = significand_float(Yi);
  const float zh = 1.0f / Y;
  const float zl0 = (float)((1.0 - (double)Y * (double)zh) / (double)Y);
  return !div_split_candidates_pass(Yi, zh, zl0, 5, 10);
}

// With arguments: prints "<divisor> base|moved|refused" for each (strtof
// syntax, hexadecimal floats included) and exits.
static int classify(int argc, char** argv) {
  for (int i = 1; i < argc; i++) {
    const float y = std::strtof(argv[i], nullptr);
    float zh, zl;
    const char* kind = "refused";
    if (make_div_split(y, &zh, &zl)) {
      const uint32_t Yi = (f32_bits(fabsf(y)) & 0x7fffffu) | 0x800000u;
      const float Y = significand_float(Yi);
      const float zh0 = 1.0f / Y;
      const float zl0 = (float)((1.0 - (double)Y * (double)zh0) / (double)Y);
      const int e = ilogbf(fabsf(y));
      const float base = Yi == 0x800000u ? zl : (y < 0.f ? -1.f : 1.f) * ldexpf(zl0, -e);
      kind = f32_bits(zl) == f32_bits(base) ? "base" : "moved";
    }
    std::printf("%s %s\n", argv[i], kind);
  }
  return 0;
}

// The EXACT step's unguarded runs: from a state and loop constants within
// kStepBound, unchecked_steps(L, R, C) steps of the reference recurrence
// (circuit_gpu.cu:114-129, IEEE arithmetic) give no infinite dividend,
// quotient or u = RN(-a zl) (for a divisor with a split), except that the
// last step's capacitive quotients may overflow.  States near the
// bound with random signs; constants over the whole allowed range and the
// generator's.  Returns a histogram of K.
static void check_unchecked_steps(std::mt19937_64& rng, long hist[9]) {
  const float dt = (float)1e-6;
  std::uniform_real_distribution<double> u01(0, 1);
  auto logu = [&](double lo, double hi) { return (float)std::exp2(lo + (hi - lo) * u01(rng)); };
  auto sgn = [&]() { return (rng() & 1) ? 1.f : -1.f; };
  float zhT, zlT;
  make_div_split(dt, &zhT, &zlT);
  for (int it = 0; it < 300000; it++) {
    float L, R, C;
    if (it % 3 == 0) {
      L = logu(-30, 40); R = logu(-40, 60); C = logu(-40, 60);
    } else if (it % 3 == 1) {
      L = (float)(u01(rng) * 0.01 + 0.1); R = (float)(u01(rng) * 10 + 1); C = (float)(u01(rng) * 0.1);
    } else {
      L = (float)((u01(rng) * 0.01 + 0.1) * 1e-6); R = (float)(u01(rng) * 10 + 1);
      C = (float)(u01(rng) * 0.01 + 0.001);
    }
    L *= sgn(); R *= sgn(); C *= sgn();
    if (!(fabsf(R) >= 0x1p-20f && fabsf(C) >= 0x1p-40f)) continue;  // the kernel's conditions
    const int K = unchecked_steps(L, R, C);
    CHECK(K >= 1 && K <= 8, "K = %d", K);
    hist[K]++;
    if (it % 3 == 1) CHECK(K >= 3 || fabsf(C) < 1e-5f, "reference-profile constants give K = %d", K);
    if (it % 3 == 2) CHECK(K == 8, "stable-profile constants give K = %d", K);
    float zhR, zlR, zhC, zlC;
    const bool okR = make_div_split(R, &zhR, &zlR), okC = make_div_split(C, &zhC, &zlC);
    const float S = it % 7 ? kStepBound : logu(0, 40);
    float I0[10], V0[9], ni[10], nv[11];
    for (int i = 0; i < 10; i++) { ni[i] = S * sgn(); I0[i] = S * sgn(); }
    for (int i = 0; i < 9; i++) V0[i] = S * sgn();
    for (int i = 0; i < 11; i++) nv[i] = (i == 0 || i == 10 || it % 2) ? S * sgn() : V0[i - 1];
    bool bad = false;
    for (int k = 0; k < K && !bad; k++) {
      float q[10], w[9];
      for (int i = 0; i < 10; i++) {
        const float t = L * (ni[i] - I0[i]);
        q[i] = t / dt;
        bad |= !std::isfinite(t) || !std::isfinite(q[i]) || !std::isfinite(fmaf(-t, zlT, 0.f));
      }
      for (int i = 0; i < 10; i++) {
        const float a = (nv[i + 1] - nv[i]) - q[i];
        ni[i] = a / R;
        bad |= !std::isfinite(a) || !std::isfinite(ni[i]) || (okR && !std::isfinite(fmaf(-a, zlR, 0.f)));
      }
      for (int i = 0; i < 9; i++) {
        const float a = dt * (ni[i] - ni[i + 1]);
        w[i] = a / C;
        // the last step's capacitive quotients may overflow (the fma rounds
        // them as IEEE division does); the next step is checked again
        bad |= !std::isfinite(a) || (k + 1 < K && !std::isfinite(w[i])) ||
               (okC && !std::isfinite(fmaf(-a, zlC, 0.f)));
      }
      for (int i = 0; i < 9; i++) nv[i + 1] = V0[i] + w[i];
    }
    CHECK(!bad, "L=%a R=%a C=%a S=%a: a value left the finite range within K=%d steps", L, R, C, S, K);
  }
}

int main(int argc, char** argv) {
  if (argc > 1) return classify(argc, argv);
  std::mt19937_64 rng(2004);
  // 1. Dividend significands, exhaustively, for the step's divisors and edges.
  check_divisor_exhaustive((float)1e-6, "dt");
  check_divisor_exhaustive(1.f, "1");
  check_divisor_exhaustive(1.5f, "1.5");
  check_divisor_exhaustive(0x1.fffffep0f, "largest significand");
  check_divisor_exhaustive(0x1.000002p0f, "smallest significand above 1");
  check_divisor_exhaustive(-3.25f, "negative");
  int accepted_moved = 0;
  for (int k = 0; k < 40; k++) {
    uint32_t Yi = (1u << 23) | (uint32_t)(rng() & 0x7fffffu);
    if (k % 2) {  // divisors that need the moved split (or are refused)
      do Yi = (1u << 23) | (uint32_t)(rng() & 0x7fffffu);
      while (!base_split_fails(Yi));
      float zh, zl;
      accepted_moved += make_div_split(significand_float(Yi), &zh, &zl);
    }
    if (k % 5 == 2) Yi = (Yi & ~3u) | 2u;  // significand = 2 (mod 4)
    check_divisor_exhaustive(significand_float(Yi) * ldexpf(1.f, (int)(rng() % 41) - 20), "random");
  }
  CHECK(accepted_moved >= 15, "only %d of 20 divisors needing a moved split got one", accepted_moved);

  // 2. Random finite operands at every exponent the method allows, both signs.
  long tested = 0, overflowed = 0;
  for (int i = 0; i < 4000000; i++) {
    const float y = f32_from_bits((uint32_t)rng());
    const float x = f32_from_bits((uint32_t)rng());
    float zh, zl, u;
    if (!make_div_split(y, &zh, &zl)) continue;
    const float q = div_split_q(x, zh, zl, &u), want = x / y;
    if (std::isnan(x)) {
      CHECK(std::isnan(q) && std::isnan(u), "NaN / %a", y);
      continue;
    }
    // the kernel's tests for its exact path: u infinite, or a nonzero
    // dividend below |y| 2^-98
    const bool small = div_split_key(x) < div_split_key_bound(y);
    CHECK(small == (x != 0.f && fabsf(x) < fabsf(y) * 0x1p-98f), "key of %a against %a", x, y);
    if (!std::isfinite(u) || small) {
      CHECK(std::isfinite(x) || !std::isfinite(u), "%a / %a: finite u", x, y);
      continue;
    }
    CHECK(x == 0.f || fabsf(q) >= kDivSplitMin, "%a / %a = %a below the minimum, not caught", x, y, q);
    tested++;
    if (std::isinf(want)) overflowed++;
    CHECK(f32_bits(q) == f32_bits(want), "x=%a y=%a: %a vs %a", x, y, q, want);
  }
  CHECK(tested > 500000 && overflowed > 1000, "only %ld random operand pairs in range (%ld overflowing)",
        tested, overflowed);

  // 3. Special dividends: zeros give the zero of IEEE division's sign for
  // every split (zl of either sign), infinities an infinite u, NaN a NaN q;
  // powers of two and negative numbers as divisors too.
  int opposite = 0;
  for (int i = 0; i < 100000; i++) {
    float y = i % 10 ? (float)(1 + (rng() % 1000000)) * 1e-3f : ldexpf(1.f, (int)(rng() % 101) - 40);
    if (i % 3 == 0) y = -y;
    float zh, zl, u;
    if (!make_div_split(y, &zh, &zl)) continue;
    CHECK(zl != 0.f, "zl = 0 for %a", y);
    opposite += (zl < 0.f) != (zh < 0.f);
    CHECK(std::isnan(div_split_q(NAN, zh, zl)), "NaN / %a", y);
    CHECK(f32_bits(div_split_q(0.f, zh, zl)) == f32_bits(0.f / y), "+0 / %a", y);
    CHECK(f32_bits(div_split_q(-0.f, zh, zl)) == f32_bits(-0.f / y), "-0 / %a", y);
    div_split_q(INFINITY, zh, zl, &u);
    CHECK(std::isinf(u), "inf / %a: finite u", y);
    div_split_q(-INFINITY, zh, zl, &u);
    CHECK(std::isinf(u), "-inf / %a: finite u", y);
  }
  CHECK(opposite > 10000, "only %d splits with zl of the opposite sign", opposite);

  // 4. Share of divisors with a verified split: random significands, and the
  // generator's resistances and capacitances.
  long acc = 0, accR = 0, accC = 0, accC2 = 0;
  const long N = 400000;
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (long i = 0; i < N; i++) {
    float zh, zl;
    acc += make_div_split(significand_float((1u << 23) | (uint32_t)(rng() & 0x7fffffu)), &zh, &zl);
    accR += make_div_split((float)(u(rng) * 10 + 1), &zh, &zl);
    accC += make_div_split((float)(u(rng) * 0.01 + 0.001), &zh, &zl);
    accC2 += make_div_split((float)(u(rng) * 0.1), &zh, &zl);
  }
  std::printf("  accepted: random significands %.5f, R %.5f, C (stable) %.5f, C (reference) %.5f\n",
              (double)acc / N, (double)accR / N, (double)accC / N, (double)accC2 / N);
  CHECK(acc > 0.998 * N && accR > 0.998 * N && accC > 0.998 * N && accC2 > 0.998 * N,
        "fewer than 99.8%% of divisors get a split");
  long hist[9] = {0};
  check_unchecked_steps(rng, hist);
  std::printf("  unguarded steps per state check:");
  for (int k = 1; k <= 8; k++) std::printf(" K=%d: %ld", k, hist[k]);
  std::printf("\n");
  if (fails) return 1;
  std::printf("split division: ok\n");
  return 0;
}
