// Exhaustive host checks of the kernels' index arithmetic (ckt_index.h):
// every element of a launch is covered by exactly one workgroup lane.
// Built and run by tests/test_debug_build.py; exits non-zero on the first
// failure, after printing it.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ckt_index.h"

using namespace ckt;

static int fails = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);               \
      std::printf("\n");                      \
      if (++fails > 10) std::exit(1);         \
    }                                         \
  } while (0)

// xcd_tile is a bijection on [0, n) that gives XCD x (workgroups b with
// b % 8 == x) a contiguous range of tiles.
static void check_xcd_tile() {
  for (uint32_t n = 1; n <= 4100; n++) {
    std::vector<int> seen(n, 0);
    for (uint32_t b = 0; b < n; b++) {
      const uint32_t t = xcd_tile(b, n);
      CHECK(t < n, "xcd_tile(%u, %u) = %u out of range", b, n, t);
      if (t < n) seen[t]++;
      if (b >= kXcds) CHECK(t == xcd_tile(b - kXcds, n) + 1, "XCD %u not contiguous at b=%u n=%u", b % kXcds, b, n);
    }
    for (uint32_t t = 0; t < n; t++) CHECK(seen[t] == 1, "tile %u of %u covered %d times", t, n, seen[t]);
  }
}

// piece_of_block against a linear scan, with empty pieces anywhere.
static void check_piece_of_block() {
  std::mt19937 rng(7);
  for (int trial = 0; trial < 2000; trial++) {
    const uint32_t P = 1 + rng() % 40;
    std::vector<uint32_t> elems(P), first(P + 1, 0);
    for (uint32_t p = 0; p < P; p++) elems[p] = (rng() % 3 == 0) ? 0 : rng() % 2000;
    for (uint32_t p = 0; p < P; p++) first[p + 1] = first[p] + (elems[p] + 255) / 256;
    std::vector<uint32_t> covered_p, covered_e;
    for (uint32_t b = 0; b < first[P]; b++) {
      uint32_t want = 0;
      for (uint32_t p = 0; p < P; p++)
        if (first[p] <= b && b < first[p + 1]) want = p;
      for (uint32_t lane = 0; lane < 256; lane += 37) {
        uint32_t e;
        const uint32_t p = piece_of_block(first.data(), P, b, lane, &e);
        CHECK(p == want, "block %u: piece %u, want %u", b, p, want);
        CHECK(e == (b - first[want]) * 256 + lane, "block %u lane %u: element %u", b, lane, e);
      }
    }
  }
}

// LDS reduction subranges partition [0, num_wires) in order, in whole
// 1024-wire blocks except at the end.
static void check_lds_subrange() {
  for (uint32_t nw = 0; nw <= 70000; nw += (nw < 5000 ? 1 : 97)) {
    for (uint32_t splits = 1; splits <= 16; splits++) {
      uint32_t next = 0;
      for (uint32_t si = 0; si < splits; si++) {
        uint32_t k0, k1;
        lds_subrange(nw, splits, si, &k0, &k1);
        CHECK(k0 == next || (k0 == nw && k1 == nw), "nw=%u splits=%u si=%u: [%u,%u) after %u", nw, splits, si, k0, k1, next);
        CHECK(k0 <= k1 && k1 <= nw, "nw=%u splits=%u si=%u: [%u,%u)", nw, splits, si, k0, k1);
        CHECK(k1 == nw || (k1 - k0) % 1024 == 0, "nw=%u splits=%u si=%u: partial block", nw, splits, si);
        next = k1;
      }
      CHECK(next == nw, "nw=%u splits=%u: covered up to %u", nw, splits, next);
    }
  }
}

// update_voltages: lanes 0..255 x groups 0..1 x 4 consecutive nodes cover a
// tile of up to 2048 nodes exactly once.
static void check_upv_groups() {
  for (uint32_t count = 0; count <= 2048; count++) {
    std::vector<int> seen(count, 0);
    for (uint32_t g = 0; g < 2; g++)
      for (uint32_t lane = 0; lane < 256; lane++) {
        const uint32_t k0 = upv_group_first(g, lane);
        if (k0 >= count) continue;
        const uint32_t end = count - k0 < 4 ? count - k0 : 4;
        for (uint32_t k = 0; k < end; k++) seen[k0 + k]++;
      }
    for (uint32_t k = 0; k < count; k++) CHECK(seen[k] == 1, "count %u: node %u covered %d times", count, k, seen[k]);
  }
}

// calc_new_currents' tile schedule, over random sets of pieces (circuits of
// 1-12 contiguous pieces of 0-3000 wires, some pieces left out, some lists
// out of wire order):
//  - every wire of every listed piece is in exactly one tile, and each tile
//    holds 1..tile wires of one circuit;
//  - without span, tiles stay inside one piece and number the sum of
//    ceil(piece / tile); longest-first lists the tiles' pieces by
//    non-increasing size; interleaved lists them by non-decreasing tile
//    index within their piece;
//  - with span, a circuit whose pieces are all listed in wire order gets
//    ceil(wires / tile) tiles, so at most one partial.
struct Run {
  uint32_t circuit, first, count, piece;
};
static void check_cnc_schedule() {
  std::mt19937 rng(7);
  for (int trial = 0; trial < 3000; trial++) {
    std::vector<Run> pieces;
    std::vector<uint32_t> wires;   // per circuit
    std::vector<bool> complete;    // every piece listed, in wire order
    const uint32_t circuits = 1 + rng() % 5;
    for (uint32_t c = 0; c < circuits; c++) {
      const uint32_t np = 1 + rng() % 12;
      uint32_t first = 0;
      bool all = true;
      std::vector<Run> mine;
      for (uint32_t p = 0; p < np; p++) {
        const uint32_t n = rng() % 7 == 0 ? 0 : rng() % 3000;
        if (rng() % 6 != 0) mine.push_back(Run{c, first, n, p});
        else all = false;
        first += n;
      }
      if (rng() % 4 == 0 && mine.size() > 1 && mine[0].count && mine.back().count) {
        std::swap(mine[0], mine.back());
        all = false;
      }
      pieces.insert(pieces.end(), mine.begin(), mine.end());
      wires.push_back(first);
      complete.push_back(all);
    }
    // the listed piece holding wire w of circuit c
    auto piece_of = [&](uint32_t c, uint32_t w) -> const Run* {
      for (const Run& p : pieces)
        if (p.circuit == c && w >= p.first && w - p.first < p.count) return &p;
      return nullptr;
    };
    const uint32_t sizes[4] = {64, 256, 512, 2048};
    const uint32_t tile = sizes[rng() % 4];
    for (int span = 0; span <= 1; span++)
      for (int order = 0; order <= 2; order++) {
        const std::vector<Run> t = cnc_schedule<Run>(pieces, tile, span, order);
        std::vector<std::vector<int>> seen(circuits);
        for (uint32_t c = 0; c < circuits; c++) seen[c].assign(wires[c], 0);
        std::vector<size_t> per_circuit(circuits, 0);
        bool ok = true;
        for (const Run& x : t) {
          const bool in = x.circuit < circuits && x.count >= 1 && x.count <= tile &&
                          (uint64_t)x.first + x.count <= wires[x.circuit];
          CHECK(in, "trial %d: tile {%u, %u, %u} malformed", trial, x.circuit, x.first, x.count);
          if (!in) { ok = false; continue; }
          per_circuit[x.circuit]++;
          for (uint32_t k = 0; k < x.count; k++) seen[x.circuit][x.first + k]++;
        }
        if (!ok) continue;
        for (const Run& p : pieces)
          for (uint32_t k = 0; k < p.count; k++)
            CHECK(seen[p.circuit][p.first + k] == 1, "trial %d span %d order %d: wire %u covered %d times",
                  trial, span, order, p.first + k, seen[p.circuit][p.first + k]);
        if (!span) {
          size_t want = 0;
          for (const Run& p : pieces) want += (p.count + tile - 1) / tile;
          CHECK(t.size() == want, "trial %d order %d: %zu tiles, expected %zu", trial, order, t.size(), want);
          uint32_t prev_size = UINT32_MAX, prev_k = 0;
          for (const Run& x : t) {
            const Run* p = piece_of(x.circuit, x.first);
            CHECK(p && x.first + x.count <= p->first + p->count, "trial %d: tile crosses a piece end", trial);
            if (!p) continue;
            const uint32_t k = (x.first - p->first) / tile;
            CHECK((x.first - p->first) % tile == 0, "trial %d: tile not at a multiple of the tile size", trial);
            if (order == 1) CHECK(p->count <= prev_size, "trial %d: piece sizes not non-increasing", trial);
            if (order == 2) CHECK(k >= prev_k, "trial %d: interleaved tile index decreases", trial);
            prev_size = p->count;
            prev_k = k;
          }
        } else {
          for (uint32_t c = 0; c < circuits; c++)
            if (complete[c])
              CHECK(per_circuit[c] == (wires[c] + tile - 1) / tile,
                    "trial %d order %d: circuit %u of %u wires in %zu tiles", trial, order, c, wires[c],
                    per_circuit[c]);
        }
      }
  }
}

int main() {
  check_cnc_schedule();
  check_xcd_tile();
  check_piece_of_block();
  check_lds_subrange();
  check_upv_groups();
  if (fails) return 1;
  std::printf("index arithmetic: ok\n");
  return 0;
}
