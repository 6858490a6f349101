// ckt_generator.cpp -- load_circuit (apps/circuit/circuit.cc:1071-1288).
//
// The reference draws from glibc srand48/drand48 in a fixed order: three
// draws per node (voltage, capacitance, leakage), then per wire resistance,
// inductance, capacitance, the in-node index, the in-piece test and either
// one index (in-piece) or a piece and an index (cross-piece).  This file
// implements the POSIX 48-bit LCG itself (re-entrant state, so circuits can
// be generated concurrently) and replaces random_element's walk over
// std::set (circuit.cc:1042-1049) by the equivalent O(1) index arithmetic:
// while wires are drawn, piece n's private set is every node of piece n in
// pointer order, so the k-th element is first_node + k.
#include <cmath>
#include <cstring>
#include <vector>

#include "ckt_gen.h"
#include "ckt_internal.h"

namespace ckt {

namespace {

// POSIX drand48: X <- (a X + c) mod 2^48, a = 0x5DEECE66D, c = 0xB; srand48
// puts the seed in the high 32 bits and 0x330E in the low 16; the result is
// X / 2^48, exact in a double.
struct Lcg48 {
  uint64_t x;
  explicit Lcg48(int32_t seed) : x(((uint64_t)(uint32_t)seed << 16) | 0x330Eull) {}
  double next() {
    x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
    return (double)x * 0x1p-48;  // exact: x < 2^48 and the scale is a power of two
  }
};

}  // namespace

int gen_validate(const ckt_gen_params_t* p, uint32_t* N, uint32_t* W, uint32_t* P) {
  if (!p) return fail(CKT_ERR_INVALID_ARG, "generator: null params");
  if (p->num_pieces < 1 || p->nodes_per_piece < 1 || p->wires_per_piece < 0)
    return fail(CKT_ERR_INVALID_ARG, "generator: need pieces >= 1, nodes/piece >= 1, wires/piece >= 0");
  if (p->profile != CKT_PROFILE_REFERENCE && p->profile != CKT_PROFILE_STABLE)
    return fail(CKT_ERR_INVALID_ARG, "generator: unknown profile");
  const uint64_t n = (uint64_t)p->num_pieces * (uint64_t)p->nodes_per_piece;
  const uint64_t w = (uint64_t)p->num_pieces * (uint64_t)p->wires_per_piece;
  if (n >= kMaxElements || w >= kMaxElements)
    return fail(CKT_ERR_OUT_OF_RANGE, "generator: more than 2^30 nodes or wires");
  if (N) *N = (uint32_t)n;
  if (W) *W = (uint32_t)w;
  if (P) *P = (uint32_t)p->num_pieces;
  return CKT_OK;
}

int gen_run(const ckt_gen_params_t* p, GenSink& out) {
  uint32_t N, W, P;
  int st = gen_validate(p, &N, &W, &P);
  if (st) return st;
  const uint32_t npp = (uint32_t)p->nodes_per_piece, wpp = (uint32_t)p->wires_per_piece;
  const bool stable = p->profile == CKT_PROFILE_STABLE;
  Lcg48 rng(p->random_seed);

  // Nodes (circuit.cc:1109-1136).  The double expressions are rounded to
  // float on assignment, as in the reference.
  for (uint32_t i = 0; i < N; i++) {
    const float voltage = (float)(2 * rng.next() - 1);
    const float capacitance = (float)(rng.next() + 1);
    const float leakage = (float)((double)0.1f * rng.next());
    out.node(i, voltage, capacitance, leakage);
  }

  // privacy_map[1] membership (circuit.cc:1183-1184) as a byte per node.
  std::vector<uint8_t> shared(N, 0);
  // Wires (circuit.cc:1147-1193).
  for (uint32_t n = 0; n < P; n++) {
    const uint32_t base = n * npp;
    for (uint32_t i = 0; i < wpp; i++) {
      const uint32_t w = n * wpp + i;
      float R, L, C;
      if (stable) {
        R = (float)(rng.next() * 10 + 1);
        L = (float)((rng.next() * 0.01 + 0.1) * 1e-6);
        C = (float)(rng.next() * 0.01 + 0.001);
      } else {
        R = (float)(rng.next() * 10 + 1);
        L = (float)(rng.next() * 0.01 + 0.1);
        C = (float)(rng.next() * 0.1);
      }
      const uint32_t in_ptr = base + (uint32_t)(int)(rng.next() * (double)npp);
      uint32_t out_ptr;
      if ((100 * rng.next()) < p->pct_wire_in_piece) {
        out_ptr = base + (uint32_t)(int)(rng.next() * (double)npp);
      } else {
        int nn = (int)(rng.next() * (p->num_pieces - 1));
        if (nn >= (int)n) nn++;
        // One piece: the reference would index piece 1, which does not
        // exist; the pick wraps to piece 0.
        if (nn >= p->num_pieces) nn -= p->num_pieces;
        out_ptr = (uint32_t)nn * npp + (uint32_t)(int)(rng.next() * (double)npp);
        shared[out_ptr] = 1;
      }
      out.wire(w, in_ptr, out_ptr, L, R, C);
    }
  }

  // Second pass: locator (circuit.cc:1210-1220) and locations
  // (find_location, circuit.cc:1051-1069): an end inside the wire's own
  // piece is private or shared by the node's flag; an end in another piece
  // can only be the out-end of one of this piece's cross wires, i.e. a ghost.
  for (uint32_t i = 0; i < N; i++) out.locator(i, shared[i] ? 0 : 1);
  for (uint32_t n = 0; n < P; n++) {
    const uint32_t lo = n * npp;
    for (uint32_t i = 0; i < wpp; i++) {
      const uint32_t w = n * wpp + i;
      uint32_t ends[2];
      out.ends(w, &ends[0], &ends[1]);
      int32_t loc[2];
      for (int e = 0; e < 2; e++) {
        const bool own = ends[e] - lo < npp;
        loc[e] = own ? (shared[ends[e]] ? CKT_SHARED_PTR : CKT_PRIVATE_PTR) : CKT_GHOST_PTR;
      }
      out.locations(w, loc[0], loc[1]);
    }
    out.piece(n, ckt_piece_t{n * wpp, wpp, n * npp, npp});
  }
  return CKT_OK;
}

// Sink writing the reference AoS layouts.
struct AosSink final : GenSink {
  ckt_node_t* nodes;
  ckt_wire_t* wires;
  int32_t* loc;
  ckt_piece_t* pieces;
  void node(uint32_t i, float v, float c, float l) override {
    nodes[i] = ckt_node_t{0.f, v, c, l};
  }
  void wire(uint32_t w, uint32_t in, uint32_t out, float L, float R, float C) override {
    ckt_wire_t& x = wires[w];
    std::memset(&x, 0, sizeof(x));
    x.in_ptr = in;
    x.out_ptr = out;
    x.inductance = L;
    x.resistance = R;
    x.capacitance = C;
  }
  void ends(uint32_t w, uint32_t* in, uint32_t* out) override {
    *in = wires[w].in_ptr;
    *out = wires[w].out_ptr;
  }
  void locations(uint32_t w, int32_t in_loc, int32_t out_loc) override {
    wires[w].in_loc = in_loc;
    wires[w].out_loc = out_loc;
  }
  void locator(uint32_t i, int32_t v) override {
    if (loc) loc[i] = v;
  }
  void piece(uint32_t n, const ckt_piece_t& p) override {
    if (pieces) pieces[n] = p;
  }
};

}  // namespace ckt

using namespace ckt;

extern "C" {

int ckt_generate_sizes(const ckt_gen_params_t* params, uint32_t* num_nodes, uint32_t* num_wires,
                       uint32_t* num_pieces) {
  return gen_validate(params, num_nodes, num_wires, num_pieces);
}

int ckt_generate(const ckt_gen_params_t* params, ckt_node_t* nodes, ckt_wire_t* wires,
                 int32_t* locator, ckt_piece_t* pieces) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  if (!nodes || (!wires && W)) return fail(CKT_ERR_INVALID_ARG, "ckt_generate: null output array");
  AosSink sink;
  sink.nodes = nodes;
  sink.wires = wires;
  sink.loc = locator;
  sink.pieces = pieces;
  return gen_run(params, sink);
}

}  // extern "C"
