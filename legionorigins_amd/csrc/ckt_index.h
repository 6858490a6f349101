// ckt_index.h -- the index arithmetic of the kernels' launches (which
// workgroup covers which elements), as host-and-device functions so that
// host code can check it exhaustively, and the optional bounds checks of the
// kernels' derived indices (the counterpart of the reference's
// POINTER_CHECKS accessors, runtime/lowlevel.h:818-832).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace ckt {

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  xcd_tile gives the workgroups of
// one XCD a contiguous range of the n tiles, so neighbouring tiles share an
// L2 instead of being spread over all eight.  A bijection on [0, n).
constexpr uint32_t kXcds = 8;
__host__ __device__ inline uint32_t xcd_tile(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds;
  const uint32_t x = b % kXcds, k = b / kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Piece-table launches: piece p owns workgroups [first_block[p],
// first_block[p + 1]) (non-decreasing), each covering 256 consecutive
// elements of it.  Returns the piece of workgroup b -- the last p with
// first_block[p] <= b, so pieces without elements own no workgroup -- and
// the element of lane `lane` in it.
__host__ __device__ inline uint32_t piece_of_block(const uint32_t* first_block, uint32_t num_pieces,
                                                   uint32_t b, uint32_t lane, uint32_t* elem) {
  uint32_t lo = 0, hi = num_pieces;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first_block[mid] <= b) lo = mid;
    else hi = mid;
  }
  *elem = (b - first_block[lo]) * 256u + lane;
  return lo;
}

// Wire subrange si of `splits` of a piece's num_wires wires in the LDS
// reduction: whole 1024-wire blocks, [*k0, *k1); trailing subranges may be
// empty.
__host__ __device__ inline void lds_subrange(uint32_t num_wires, uint32_t splits, uint32_t si,
                                             uint32_t* k0, uint32_t* k1) {
  const uint32_t chunk = ((num_wires + splits - 1) / splits + 1023u) & ~1023u;
  const uint64_t a = (uint64_t)si * chunk;
  *k0 = a < num_wires ? (uint32_t)a : num_wires;
  *k1 = (uint64_t)*k0 + chunk < num_wires ? *k0 + chunk : num_wires;
}

// update_voltages: lane `lane` of a tile of up to 2048 nodes updates nodes
// [4 (g 256 + lane), +4) for g = 0, 1 (two lane-interleaved groups of 4).
__host__ __device__ inline uint32_t upv_group_first(uint32_t g, uint32_t lane) {
  return 4u * (g * 256u + lane);
}

// calc_new_currents' tile list (host side).  `pieces` holds one element per
// piece, {circuit, first wire, wires, piece}.  A run is a range of
// consecutive wires of one circuit: a piece, or with `span` adjacent pieces
// of one circuit merged, so that only a run's last tile is partial.  Tiles
// hold at most `tile` wires and are listed run by run in the order given
// (order 0), longest run first (1, stable), or interleaved (2: tile k of
// every run before tile k + 1 of any).  T is any struct with those four
// uint32_t members.
template <class T>
std::vector<T> cnc_schedule(const std::vector<T>& pieces, uint32_t tile, int span, int order) {
  std::vector<T> runs;
  for (const T& p : pieces) {
    if (!p.count) continue;
    if (span && !runs.empty() && runs.back().circuit == p.circuit &&
        (uint64_t)runs.back().first + runs.back().count == p.first)
      runs.back().count += p.count;
    else
      runs.push_back(p);
  }
  if (order == 1)
    std::stable_sort(runs.begin(), runs.end(), [](const T& a, const T& b) { return a.count > b.count; });
  std::vector<T> t;
  auto tile_of = [&](const T& r, uint32_t off) {
    T x = r;
    x.first = r.first + off;
    x.count = std::min(tile, r.count - off);
    return x;
  };
  if (order == 2) {
    for (uint64_t off = 0;; off += tile) {
      const size_t before = t.size();
      for (const T& r : runs)
        if (off < r.count) t.push_back(tile_of(r, (uint32_t)off));
      if (t.size() == before) break;
    }
  } else {
    for (const T& r : runs)
      for (uint64_t off = 0; off < r.count; off += tile) t.push_back(tile_of(r, (uint32_t)off));
  }
  return t;
}

}  // namespace ckt

// CKT_BOUNDS_CHECKS (a build option, off by default): every index a kernel
// derives before it touches a circuit array is checked against the array's
// size; an out-of-range index is counted in a device counter
// (ckt_debug_bounds_violations reads it) and the access is skipped instead
// of made.  Without the option the checks compile to nothing.
#ifdef CKT_BOUNDS_CHECKS
namespace ckt {
__device__ bool bounds_violation();
}
#define CKT_IN_BOUNDS(cond) (__builtin_expect(!!(cond), 1) || ::ckt::bounds_violation())
#else
#define CKT_IN_BOUNDS(cond) true
#endif
