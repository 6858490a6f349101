// ckt_internal.h -- declarations shared by the HIP kernels and the host side
// of the circuit library.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>

#include "../../include/circuit_abi.h"

namespace ckt {

constexpr int kSeg = CKT_WIRE_SEGMENTS;

// Device view of one circuit held as structure-of-arrays in HBM.
//   wire arrays are segment-major: cur[s * ws + w], vol[s * ws + w];
//   ws (wire stride) is num_wires rounded up to 64 so that every segment row
//   starts on a 256-byte boundary.
struct DevCircuit {
  float* cur;            // [kSeg][ws]
  float* vol;            // [kSeg-1][ws]
  float* ind;            // [ws]  inductance
  float* res;            // [ws]  resistance
  float* wcap;           // [ws]  wire capacitance
  uint32_t* in_ptr;      // [ws]
  uint32_t* out_ptr;     // [ws]
  uint8_t* in_loc;       // [ws]
  uint8_t* out_loc;      // [ws]
  float* charge;         // [N]
  float* voltage;        // [N]
  float* ncap;           // [N]  node capacitance
  float* leak;           // [N]
  uint8_t* locator;      // [N]  1 = private, 0 = shared
  uint32_t* csr_row;     // [N+1] node -> first contribution
  uint32_t* csr_ent;     // [2W]  contribution = wire << 1 | side (0 in, 1 out)
  ckt_piece_t* pieces;   // [P]
  uint32_t num_wires, ws, num_nodes, num_pieces;
};

// One unit of work of a launch: `count` consecutive elements (wires or nodes)
// of circuit `circuit`, starting at `first`, belonging to piece `piece`.
struct Tile {
  uint32_t circuit;
  uint32_t first;
  uint32_t count;
  uint32_t piece;
};

// A whole piece as one unit (LDS-privatised charge reduction).
struct PieceTile {
  uint32_t circuit;
  uint32_t piece;
  uint32_t partial;  // split LDS reduction: first float of the piece's rows in `partial`
  uint32_t pad;
};

// Host launchers (defined in ckt_kernels.hip).  All are asynchronous on `s`.
// fuse_dsc adds each wire's two charge contributions with float atomics at the
// end of the relaxation (the ATOMIC distribute_charge, fused).
hipError_t launch_cnc(int policy, bool fuse_dsc, const DevCircuit* circuits, const Tile* tiles,
                      uint32_t ntiles, int block, int steps, hipStream_t s);
hipError_t launch_cnc_persistent(int policy, bool fuse_dsc, const DevCircuit* circuits,
                                 const Tile* tiles, uint32_t ntiles, int block, int grid,
                                 uint32_t* counters, int steps, hipStream_t s);
hipError_t launch_dsc_atomic(const DevCircuit* circuits, const Tile* tiles, uint32_t ntiles,
                             int block, hipStream_t s);
// Dynamic LDS bytes of the LDS-privatised reduction for pieces of up to
// max_nodes nodes.
size_t dsc_lds_bytes(uint32_t max_nodes);
// LDS-privatised ATOMIC reduction over n piece tiles, each piece's wires in
// `splits` subranges (splits > 1 needs `partial`, splits x the tiles' node
// counts floats at each tile's `partial` offset, and max_nodes, the largest
// tile's node count).
hipError_t launch_dsc_lds(const DevCircuit* circuits, const PieceTile* ptiles, uint32_t n,
                          uint32_t splits, uint32_t max_nodes, float* partial, size_t lds_bytes,
                          hipStream_t s);
hipError_t launch_dsc_csr(int kahan, const DevCircuit* circuits, const Tile* node_tiles,
                          uint32_t ntiles, int block, hipStream_t s);
// nontemporal: streaming accesses for node arrays larger than the last-level
// cache.
hipError_t launch_upv(const DevCircuit* circuits, const Tile* node_tiles, uint32_t ntiles,
                      int block, bool nontemporal, hipStream_t s);
hipError_t launch_sanity(const DevCircuit* circuits, const Tile* tiles, uint32_t ntiles,
                         int block, unsigned long long* bad, hipStream_t s);
hipError_t launch_nop(hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);

// The piece table of an accessor-form launch, in device memory: pieces[p]
// and first_block[p], the number of 256-element workgroups of the pieces
// before p (first_block[num_pieces] = the launch's workgroup count), counted
// over wires (calc_new_currents, distribute_charge) or nodes
// (update_voltages).
struct PieceIndex {
  const ckt_piece_t* pieces;
  const uint32_t* first_block;
  uint32_t num_pieces;
};

// Accessor-form (reference AoS layout) launchers over a piece table;
// `blocks` = first_block[num_pieces].
hipError_t launch_cnc_aos(int policy, const PieceIndex& ix, uint32_t blocks, ckt_wire_t* wires,
                          const ckt_node_t* pvt, const ckt_node_t* owned, const ckt_node_t* ghost,
                          int steps, unsigned long long* bad, hipStream_t s);
hipError_t launch_dsc_aos(const PieceIndex& ix, uint32_t blocks, const ckt_wire_t* wires,
                          ckt_node_t* pvt, float* owned_fold, float* ghost_fold,
                          unsigned long long* bad, hipStream_t s);
hipError_t launch_upv_aos(const PieceIndex& ix, uint32_t blocks, ckt_node_t* pvt,
                          ckt_node_t* owned, const int32_t* locator, hipStream_t s);
hipError_t launch_sanity_aos(const ckt_piece_t piece, const ckt_wire_t* wires,
                             uint32_t num_nodes_total, unsigned long long* bad, hipStream_t s);
hipError_t launch_fold_init(float* fold, uint32_t first, uint32_t count, hipStream_t s);
hipError_t launch_fold_apply(ckt_node_t* owner, const float* fold, uint32_t first, uint32_t count,
                             hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t read_bounds_violations(uint64_t* count);

// Error reporting (thread-local last error).
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
void clear_error();

}  // namespace ckt

#define CKT_HIP(expr)                                   \
  do {                                                  \
    hipError_t ckt_e_ = (expr);                         \
    if (ckt_e_ != hipSuccess) return ckt::hip_fail(ckt_e_, #expr); \
  } while (0)
