// ckt_gen.h -- generator interface (implementation in ckt_generator.cpp).
#pragma once

#include <stdint.h>

#include "ckt_internal.h"

namespace ckt {

// Node and wire counts are kept below 2^30 so that the distribute_charge
// index can pack a wire index and an end bit into 32 bits.
constexpr uint64_t kMaxElements = 1ull << 30;

// Receives the generated circuit.  Calls arrive in this order: node() for
// every node, wire() for every wire, locator() for every node, then per piece
// ends()/locations() for its wires followed by piece().
struct GenSink {
  virtual ~GenSink() = default;
  virtual void node(uint32_t i, float voltage, float capacitance, float leakage) = 0;
  virtual void wire(uint32_t w, uint32_t in_ptr, uint32_t out_ptr, float inductance,
                    float resistance, float capacitance) = 0;
  virtual void ends(uint32_t w, uint32_t* in_ptr, uint32_t* out_ptr) = 0;
  virtual void locations(uint32_t w, int32_t in_loc, int32_t out_loc) = 0;
  virtual void locator(uint32_t i, int32_t v) = 0;
  virtual void piece(uint32_t n, const ckt_piece_t& p) = 0;
};

int gen_validate(const ckt_gen_params_t* p, uint32_t* N, uint32_t* W, uint32_t* P);
int gen_run(const ckt_gen_params_t* p, GenSink& out);

}  // namespace ckt
