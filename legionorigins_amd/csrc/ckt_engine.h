// ckt_engine.h -- device-resident circuits and the launches over them.
//
// An Engine owns one HIP stream on one device and any number of circuits,
// each held in one HBM slab as structure-of-arrays.  A launch covers a list
// of (circuit, piece) references -- one piece is one point of the
// reference's index-space launch (circuit.cc:250-251) -- and is cut into
// tiles of consecutive wires or nodes.  The single-circuit context and the
// batch are both Engines.
#pragma once

#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "ckt_gen.h"
#include "ckt_internal.h"

namespace ckt {

struct SlabLayout {
  size_t cur, vol, ind, res, wcap, in_ptr, out_ptr, in_loc, out_loc;
  size_t charge, voltage, ncap, leak, locator, csr_row, csr_ent, pieces;
  size_t total;
  uint32_t ws;
  static SlabLayout make(uint32_t N, uint32_t W, uint32_t P);
};

// Host image of a slab, filled on the host and copied to HBM in one transfer.
struct SlabImage {
  SlabLayout lay{};
  uint32_t N = 0, W = 0, P = 0;
  std::vector<uint8_t> bytes;
  std::vector<ckt_piece_t> pieces;
  template <class T>
  T* at(size_t off) {
    return reinterpret_cast<T*>(bytes.data() + off);
  }
  void reset(uint32_t n, uint32_t w, uint32_t p);
};

int image_from_aos(SlabImage& img, const ckt_node_t* nodes, uint32_t N, const ckt_wire_t* wires,
                   uint32_t W, const int32_t* locator, const ckt_piece_t* pieces, uint32_t P);
int image_from_gen(SlabImage& img, const ckt_gen_params_t* params);

// Generates the circuit `params` describes straight into a device slab of
// layout `L` (ckt_devgen.hip), bit-identical to image_from_gen; fills
// `pieces`.  Synchronises `s` before returning.
int devgen_build(const ckt_gen_params_t* params, const SlabLayout& L, void* slab, hipStream_t s,
                 std::vector<ckt_piece_t>& pieces);

// The host part of the device generator, done ahead of time (thread-safe):
// every wire's drand48 state at its first draw.
struct DevgenPlan {
  ckt_gen_params_t params{};
  uint32_t N = 0, W = 0, P = 0;
  std::vector<uint64_t> wire_state;
};
int devgen_plan(const ckt_gen_params_t* params, DevgenPlan& plan);
// Device scratch a planned build of N nodes and W wires needs.
size_t devgen_scratch_bytes(uint32_t N, uint32_t W);
// Enqueues the planned build on `s` without waiting: `staged_state` (pinned
// host memory holding plan.wire_state) and `scratch` (devgen_scratch_bytes)
// must stay untouched until `s` has run the enqueued work.  Fills `pieces`.
int devgen_enqueue(const DevgenPlan& plan, const uint64_t* staged_state, const SlabLayout& L,
                   void* slab, void* scratch, hipStream_t s, std::vector<ckt_piece_t>& pieces);

struct Circuit {
  uint32_t id = 0;
  uint32_t num_nodes = 0, num_wires = 0, num_pieces = 0;
  SlabLayout lay{};
  void* slab = nullptr;
  DevCircuit dev{};
  std::vector<ckt_piece_t> pieces;
  std::unordered_map<uint64_t, uint32_t> piece_index;  // descriptor -> index
  bool csr_ok = false;   // the pieces partition the wires
  bool lds_ok = false;   // node ranges of the pieces are disjoint
  bool stream_ordered = false;  // slab from the engine stream's memory pool
  int32_t loops_left = 0;
};

struct PieceRef {
  uint32_t slot;   // index into the engine's circuit table
  uint32_t piece;  // index into the circuit's pieces
};

enum EngineOption {
  OPT_STEPS = 0,
  OPT_CNC_BLOCK = 1,
  OPT_PERSISTENT = 2,
  OPT_WAVES_PER_SIMD = 3,
  OPT_DSC_ATOMIC_IMPL = 4,  // 0 auto, 1 global atomics, 2 LDS per piece
  OPT_FUSE_DSC = 5,         // loop drivers: ATOMIC distribute_charge fused into calc_new_currents
  OPT_WIRES_PER_LANE = 6,   // calc_new_currents tile = workgroup size x this (1, 2, 4) wires
  OPT_TILE_SPAN = 7,        // 0: tiles within a piece; 1: within runs of adjacent pieces
  OPT_TILE_ORDER = 8,       // 0: as given; 1: longest runs first; 2: runs interleaved
};

class Engine {
 public:
  Engine() = default;
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  int init(int device, uint64_t budget_bytes);
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  int num_cus() const { return num_cus_; }

  // Circuits.
  int add(SlabImage& img, int32_t loops, uint32_t* id);
  // Generates the circuit on the device (no host image, no upload).
  int add_generated_on_device(const ckt_gen_params_t* params, int32_t loops, uint32_t* id);
  // Same circuit from a plan made earlier: the generation is only enqueued
  // on the engine stream (after whatever runs there), nothing waits.  The
  // plan's start states are copied into pinned staging memory that sync()
  // recycles, so the plan may be dropped on return.
  int add_planned(const DevgenPlan& plan, int32_t loops, uint32_t* id);
  int remove(uint32_t id);
  int remove_many(const uint32_t* ids, size_t n);
  Circuit* find(uint32_t id);
  int store(uint32_t id, ckt_node_t* nodes, ckt_wire_t* wires);
  int store_topology(uint32_t id, int32_t* locator, uint32_t* csr_row, uint32_t* csr_ent);
  int download_voltages(uint32_t id, std::vector<float>& v);
  // Enqueues a copy of the node voltages into `dst` (pinned host memory, at
  // least num_nodes floats) on the engine stream; valid after sync().
  int copy_voltages_async(uint32_t id, float* dst);
  size_t count() const { return slots_.size(); }
  uint64_t bytes_in_use() const { return in_use_; }
  uint64_t budget() const { return budget_; }
  const std::vector<Circuit*>& slots() const { return slots_; }
  int slot_of(uint32_t id) const;
  // Reference to the loaded piece equal to `p` in circuit `slot`.
  int ref_for(uint32_t slot, const ckt_piece_t& p, PieceRef* out) const;
  void all_refs(std::vector<PieceRef>& refs, bool only_active) const;

  // Task launches (asynchronous on the engine stream).
  int cnc(const std::vector<PieceRef>& refs, int policy, bool fuse_dsc = false);
  int dsc(const std::vector<PieceRef>& refs, int mode);
  int upv(const std::vector<PieceRef>& refs);
  int sanity(const std::vector<PieceRef>& refs, uint64_t* bad);
  int nop();
  // Waits for the engine stream; recycles the staging memory of planned adds.
  int sync();

  // One loop (cnc, dsc, upv) of `refs` captured in a graph and replayed
  // `times` times.
  int run_graph(const std::vector<PieceRef>& refs, int policy, int mode, int times);
  // One loop of `refs`: calc_new_currents, distribute_charge, update_voltages
  // (the first two as one fused launch when OPT_FUSE_DSC is on and the mode
  // is ATOMIC).
  int loop_once(const std::vector<PieceRef>& refs, int policy, int mode);

  int set_option(int opt, int64_t value);
  int64_t option(int opt) const;

  // Event timing per task id.
  void set_timing(bool on) { timing_ = on; }
  int kernel_time(int task, double* ms, uint64_t* launches);
  void reset_timing();

 private:
  struct DevBuf {
    void* p = nullptr;
    uint32_t n = 0;
    bool pooled = false;  // from the stream's memory pool (freed stream-ordered)
  };
  struct Pending {
    int task;
    hipEvent_t a, b;
  };
  void attach(void* slab, const SlabLayout& L, uint32_t N, uint32_t W, uint32_t P,
              const std::vector<ckt_piece_t>& pieces, int32_t loops, uint32_t* id,
              bool stream_ordered = false);
  int ensure_table();
  int tiles_for(int kind, const std::vector<PieceRef>& refs, int block, DevBuf* out);
  int upload_buf(uint64_t key, const void* data, size_t bytes, uint32_t n, DevBuf* out);
  int prefetch_tiles(const std::vector<PieceRef>& refs, int policy, int mode);
  // calc_new_currents' workgroup size: OPT_CNC_BLOCK, or with -1 (default) 64
  // for EXACT and 256 for the other policies.
  int cnc_block_for(int policy) const {
    return cnc_block_ > 0 ? cnc_block_ : (policy == CKT_POLICY_EXACT ? 64 : 256);
  }
  // loop_once's launches; every tile list must already be cached.
  int launch_loop(const std::vector<PieceRef>& refs, int policy, int mode);
  void drop_caches();
  int stage(const void* src, size_t bytes, void** out);
  int timed_begin(int task, hipEvent_t* a);
  int timed_end(int task, hipEvent_t a);
  int csr_nodes_ok(const std::vector<PieceRef>& refs) const;
  // The LDS-privatised ATOMIC reduction over `refs`: whether it applies, its
  // LDS bytes, the subranges per piece and the floats of partial sums it
  // needs (allocated here, before any launch or capture).
  struct LdsPlan {
    bool lds = false;
    size_t lds_bytes = 0;
    uint32_t splits = 1, max_nodes = 0;
    uint64_t partial_floats = 0;
  };
  int lds_plan(const std::vector<PieceRef>& refs, LdsPlan* plan);

  int device_ = -1;
  hipStream_t stream_ = nullptr;
  int num_cus_ = 0;
  size_t lds_limit_ = 0;
  uint64_t budget_ = 0, in_use_ = 0;
  uint32_t next_id_ = 1;
  std::vector<Circuit*> slots_;
  std::map<uint32_t, Circuit*> by_id_;
  DevCircuit* d_table_ = nullptr;
  size_t table_cap_ = 0;
  bool table_dirty_ = true;
  uint64_t generation_ = 1;
  std::unordered_map<uint64_t, DevBuf> tile_cache_;
  std::unordered_map<uint64_t, hipGraphExec_t> graphs_;
  uint32_t* d_counters_ = nullptr;
  unsigned long long* d_bad_ = nullptr;
  float* d_partial_ = nullptr;  // split LDS reduction's partial sums (grow-only)
  uint64_t partial_cap_ = 0;    // floats
  int steps_ = CKT_STEPS;
  int cnc_block_ = -1;       // -1: chosen per policy (cnc_block_for)
  int persistent_ = -1;      // -1: chosen per policy (launch_shape)
  int waves_per_simd_ = -1;  // -1: chosen per policy
  int dsc_atomic_impl_ = 0;
  int fuse_dsc_ = 1;
  int wires_per_lane_ = 1;
  int tile_span_ = 0;
  int tile_order_ = 0;
  bool timing_ = false;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> event_pool_;
  // Planned device generation: one device scratch (grow-only; the builds on
  // the stream run one after another) and pinned staging chunks for the
  // start states, recycled at sync().
  bool pool_ok_ = false;  // the device's default memory pool is usable
  void* devgen_scratch_ = nullptr;
  size_t devgen_scratch_cap_ = 0;
  struct PinnedChunk {
    void* p;
    size_t cap;
  };
  std::vector<PinnedChunk> staging_;
  size_t staging_chunk_ = 0, staging_off_ = 0;
  double acc_ms_[CKT_NUM_TASK_IDS] = {0};
  uint64_t acc_n_[CKT_NUM_TASK_IDS] = {0};
};

}  // namespace ckt
