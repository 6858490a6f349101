// ckt_devgen.hip -- load_circuit (apps/circuit/circuit.cc:1071-1288) built
// directly in HBM.
//
// The reference draws from one drand48 stream in a fixed order: three draws
// per node, then per wire six (in-piece) or seven (cross-piece) draws.  A
// node's first draw is a fixed jump from the seed, so every node is drawn
// independently on the device.  A wire's first draw depends on how many
// earlier wires were cross-piece, which is only known by walking the chain:
// the host does that walk (two 48-bit multiply-adds per wire: jump to the
// in-piece test, then jump to the next wire).  The synchronous build streams
// the start states to the device in chunks, overlapped with the device
// drawing the previous chunk; the planned build takes a walk done earlier
// (on any thread) and only enqueues device work, so a server can issue it
// behind a running iteration.  The device then marks shared nodes, derives
// the locations, and builds the node -> (wire, end) index with a stable
// radix sort.  The slab is bit-identical to the host generator's image
// (ckt_generator.cpp + finalize_image).  The library is compiled with
// -ffp-contract=off, so the double expressions round exactly as on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ckt_engine.h"

namespace ckt {

namespace {

constexpr uint64_t kMask48 = (1ull << 48) - 1;
constexpr uint64_t kA = 0x5DEECE66Dull, kC = 0xBull;
constexpr uint32_t kChunk = 1u << 22;  // wires per host -> device start-state chunk

// s -> a*s + c (mod 2^48): a number of steps of the drand48 recurrence.
struct Jump {
  uint64_t a, c;
};

// Jumps of 3 * 2^k draws: node i starts 3*i draws after the seed.
struct NodeJumps {
  Jump j[32];
};

__host__ __device__ __forceinline__ uint64_t apply(const Jump& j, uint64_t s) {
  return (j.a * s + j.c) & kMask48;
}

// `first` then `second`.
Jump then(const Jump& first, const Jump& second) {
  return Jump{(second.a * first.a) & kMask48, (second.a * first.c + second.c) & kMask48};
}

Jump jump(uint64_t n) {
  Jump result{1, 0}, step{kA, kC};
  while (n) {
    if (n & 1) result = then(result, step);
    step = then(step, step);
    n >>= 1;
  }
  return result;
}

// One drand48 draw: advance, then return the new state / 2^48 (exact).
__device__ __forceinline__ double draw(uint64_t& s) {
  s = (kA * s + kC) & kMask48;
  return (double)s * 0x1p-48;
}

// Nodes (circuit.cc:1109-1136); the doubles round to float on assignment.
__global__ __launch_bounds__(256) void gen_nodes(uint32_t N, uint64_t seed_state, NodeJumps t,
                                                 float* __restrict__ voltage,
                                                 float* __restrict__ ncap,
                                                 float* __restrict__ leak) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  uint64_t s = seed_state;
  for (uint32_t k = 0, m = i; m; k++, m >>= 1)
    if (m & 1) s = apply(t.j[k], s);
  voltage[i] = (float)(2 * draw(s) - 1);
  ncap[i] = (float)(draw(s) + 1);
  leak[i] = (float)((double)0.1f * draw(s));
}

// Wires first .. first+count-1 (circuit.cc:1147-1193); state[k] is wire
// first+k's start state.  A cross-piece wire's out-node becomes shared.
__global__ __launch_bounds__(256) void gen_wires(uint32_t first, uint32_t count, uint32_t P,
                                                 uint32_t npp, uint32_t wpp, int32_t pct,
                                                 int32_t stable, const uint64_t* __restrict__ state,
                                                 float* __restrict__ ind, float* __restrict__ res,
                                                 float* __restrict__ wcap,
                                                 uint32_t* __restrict__ in_ptr,
                                                 uint32_t* __restrict__ out_ptr, uint8_t* locator) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= count) return;
  const uint32_t w = first + k;
  const uint32_t n = w / wpp;
  const uint32_t base = n * npp;
  uint64_t s = state[k];
  float R, L, C;
  if (stable) {  // the same draws; inductance x 1e-6, capacitance U(0.001, 0.011)
    R = (float)(draw(s) * 10 + 1);
    L = (float)((draw(s) * 0.01 + 0.1) * 1e-6);
    C = (float)(draw(s) * 0.01 + 0.001);
  } else {
    R = (float)(draw(s) * 10 + 1);
    L = (float)(draw(s) * 0.01 + 0.1);
    C = (float)(draw(s) * 0.1);
  }
  const uint32_t in = base + (uint32_t)(int)(draw(s) * (double)npp);
  uint32_t out;
  if ((100 * draw(s)) < pct) {
    out = base + (uint32_t)(int)(draw(s) * (double)npp);
  } else {
    int nn = (int)(draw(s) * (double)((int)P - 1));
    if (nn >= (int)n) nn++;
    if (nn >= (int)P) nn -= (int)P;  // one piece: the pick wraps to piece 0
    out = (uint32_t)nn * npp + (uint32_t)(int)(draw(s) * (double)npp);
    locator[out] = 0;  // every writer stores the same value
  }
  ind[w] = L;
  res[w] = R;
  wcap[w] = C;
  in_ptr[w] = in;
  out_ptr[w] = out;
}

// Locations (find_location, circuit.cc:1051-1069) and the index entries:
// entry 2w+e (end e of wire w) is keyed by that end's node.
__global__ __launch_bounds__(256) void gen_locations(uint32_t W, uint32_t npp, uint32_t wpp,
                                                     const uint32_t* __restrict__ in_ptr,
                                                     const uint32_t* __restrict__ out_ptr,
                                                     const uint8_t* __restrict__ locator,
                                                     uint8_t* __restrict__ in_loc,
                                                     uint8_t* __restrict__ out_loc,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals,
                                                     uint32_t* counts) {
  const uint32_t w = blockIdx.x * 256u + threadIdx.x;
  if (w >= W) return;
  const uint32_t lo = (w / wpp) * npp;
  const uint32_t a = in_ptr[w], b = out_ptr[w];
  in_loc[w] = a - lo < npp ? (locator[a] ? CKT_PRIVATE_PTR : CKT_SHARED_PTR) : CKT_GHOST_PTR;
  out_loc[w] = b - lo < npp ? (locator[b] ? CKT_PRIVATE_PTR : CKT_SHARED_PTR) : CKT_GHOST_PTR;
  keys[2 * w] = a;
  keys[2 * w + 1] = b;
  vals[2 * w] = w << 1;
  vals[2 * w + 1] = (w << 1) | 1u;
  atomicAdd(&counts[a], 1u);
  atomicAdd(&counts[b], 1u);
}

__global__ __launch_bounds__(256) void fill_bytes(uint8_t* p, uint8_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = v;
}

// Piece n owns wires [n*wpp, (n+1)*wpp) and nodes [n*npp, (n+1)*npp)
// (circuit.cc:1092-1104,1140-1157).
__global__ __launch_bounds__(256) void gen_pieces(uint32_t P, uint32_t npp, uint32_t wpp,
                                                  ckt_piece_t* pieces) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n < P) pieces[n] = ckt_piece_t{n * wpp, wpp, n * npp, npp};
}

uint32_t grid_for(size_t n) { return (uint32_t)((n + 255) / 256); }

uint64_t seed_state_of(const ckt_gen_params_t* p) {
  return ((uint64_t)(uint32_t)p->random_seed << 16) | 0x330Eull;
}

// Start states of wires [0, W): the only sequential part of load_circuit.
// Per chunk of up to `chunk` wires, fn.begin(first, n) returns where to write
// the n states and fn.end(first, n) is called once they are written.
template <class Sink>
void walk_wire_states(const ckt_gen_params_t* p, uint32_t N, uint32_t W, size_t chunk, Sink&& fn) {
  const Jump j5 = jump(5), j6 = jump(6), j7 = jump(7);
  const int32_t pct = p->pct_wire_in_piece;
  uint64_t x = apply(jump(3ull * N), seed_state_of(p));
  for (uint32_t first = 0; first < W; first += (uint32_t)chunk) {
    const uint32_t n = (uint32_t)std::min<size_t>(chunk, W - first);
    uint64_t* h = fn.begin(first, n);
    for (uint32_t i = 0; i < n; i++) {
      h[i] = x;
      // The wire's fifth draw is the in-piece test (circuit.cc:1167).
      const double u = (double)apply(j5, x) * 0x1p-48;
      x = ((100 * u) < pct) ? apply(j6, x) : apply(j7, x);
    }
    fn.end(first, n);
  }
}

// Slab arrays the kernels write.
struct SlabView {
  uint8_t* b;
  const SlabLayout& L;
  float* f(size_t off) const { return (float*)(b + off); }
  uint32_t* u(size_t off) const { return (uint32_t*)(b + off); }
};

// Zero the slab, mark every node private, draw the nodes, write the pieces.
int enqueue_nodes(const ckt_gen_params_t* p, uint32_t N, uint32_t P, const SlabView& v,
                  hipStream_t s) {
  CKT_HIP(launch_fill_u32((uint32_t*)v.b, 0u, v.L.total / 4, s));
  hipLaunchKernelGGL(fill_bytes, dim3(grid_for(N)), dim3(256), 0, s, v.b + v.L.locator,
                     (uint8_t)1, N);
  CKT_HIP(hipGetLastError());
  NodeJumps t{};
  for (int k = 0; k < 32; k++) t.j[k] = jump(3ull << k);
  hipLaunchKernelGGL(gen_nodes, dim3(grid_for(N)), dim3(256), 0, s, N, seed_state_of(p), t,
                     v.f(v.L.voltage), v.f(v.L.ncap), v.f(v.L.leak));
  CKT_HIP(hipGetLastError());
  hipLaunchKernelGGL(gen_pieces, dim3(grid_for(P)), dim3(256), 0, s, P,
                     (uint32_t)p->nodes_per_piece, (uint32_t)p->wires_per_piece,
                     (ckt_piece_t*)(v.b + v.L.pieces));
  CKT_HIP(hipGetLastError());
  return CKT_OK;
}

int enqueue_wires(const ckt_gen_params_t* p, uint32_t P, uint32_t first, uint32_t n,
                  const uint64_t* d_state, const SlabView& v, hipStream_t s) {
  if (n == 0) return CKT_OK;
  hipLaunchKernelGGL(gen_wires, dim3(grid_for(n)), dim3(256), 0, s, first, n, P,
                     (uint32_t)p->nodes_per_piece, (uint32_t)p->wires_per_piece,
                     p->pct_wire_in_piece, (int32_t)(p->profile == CKT_PROFILE_STABLE), d_state,
                     v.f(v.L.ind), v.f(v.L.res), v.f(v.L.wcap), v.u(v.L.in_ptr), v.u(v.L.out_ptr),
                     v.b + v.L.locator);
  CKT_HIP(hipGetLastError());
  return CKT_OK;
}

// Scratch of the index build, carved from one buffer.
struct IndexScratch {
  uint32_t *keys, *vals, *keys_sorted, *counts;
  void* work;
  size_t scan_bytes, sort_bytes;
};

unsigned key_bits(uint32_t N) {
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < N) bits++;
  return bits;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Bytes of scratch for N nodes and W wires; with `base`, carves it.
int index_scratch(uint32_t N, uint32_t W, void* base, IndexScratch* sc, size_t* bytes) {
  size_t scan_bytes = 0, sort_bytes = 0;
  CKT_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (size_t)N, rocprim::plus<uint32_t>(), (hipStream_t)0));
  if (W)
    CKT_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                      (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)2 * W, 0,
                                      key_bits(N), (hipStream_t)0));
  const size_t e = align256((size_t)2 * W * sizeof(uint32_t));
  const size_t c = align256((size_t)N * sizeof(uint32_t));
  const size_t w = align256(std::max<size_t>(std::max(scan_bytes, sort_bytes), 16));
  *bytes = 3 * e + c + w;
  if (base && sc) {
    uint8_t* b = (uint8_t*)base;
    sc->keys = (uint32_t*)b;
    sc->vals = (uint32_t*)(b + e);
    sc->keys_sorted = (uint32_t*)(b + 2 * e);
    sc->counts = (uint32_t*)(b + 3 * e);
    sc->work = b + 3 * e + c;
    sc->scan_bytes = scan_bytes;
    sc->sort_bytes = sort_bytes;
  }
  return CKT_OK;
}

// Locations, then the node -> (wire, end) index: counts -> row offsets; a
// stable sort of the entries by node keeps each node's entries in (wire,
// in-before-out) order.
int enqueue_index(const ckt_gen_params_t* p, uint32_t N, uint32_t W, const SlabView& v,
                  const IndexScratch& sc, hipStream_t s) {
  uint32_t* csr_row = v.u(v.L.csr_row);
  CKT_HIP(launch_fill_u32(sc.counts, 0u, N, s));
  if (W) {
    hipLaunchKernelGGL(gen_locations, dim3(grid_for(W)), dim3(256), 0, s, W,
                       (uint32_t)p->nodes_per_piece, (uint32_t)p->wires_per_piece,
                       v.u(v.L.in_ptr), v.u(v.L.out_ptr), v.b + v.L.locator, v.b + v.L.in_loc,
                       v.b + v.L.out_loc, sc.keys, sc.vals, sc.counts);
    CKT_HIP(hipGetLastError());
  }
  size_t scan_bytes = sc.scan_bytes, sort_bytes = sc.sort_bytes;
  CKT_HIP(rocprim::inclusive_scan(sc.work, scan_bytes, sc.counts, csr_row + 1, (size_t)N,
                                  rocprim::plus<uint32_t>(), s));
  if (W)
    CKT_HIP(rocprim::radix_sort_pairs(sc.work, sort_bytes, sc.keys, sc.keys_sorted, sc.vals,
                                      v.u(v.L.csr_ent), (size_t)2 * W, 0, key_bits(N), s));
  return CKT_OK;
}

// Temporary device and pinned buffers and events of one synchronous build;
// released after the stream has drained.
struct Temps {
  hipStream_t s;
  std::vector<void*> dev, host;
  std::vector<hipEvent_t> events;
  explicit Temps(hipStream_t st) : s(st) {}
  ~Temps() {
    (void)hipStreamSynchronize(s);
    for (void* p : dev) (void)hipFree(p);
    for (void* p : host) (void)hipHostFree(p);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
  hipError_t device(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e == hipSuccess) dev.push_back(*p);
    return e;
  }
  hipError_t pinned(void** p, size_t bytes) {
    hipError_t e = hipHostMalloc(p, std::max<size_t>(bytes, 16), hipHostMallocDefault);
    if (e == hipSuccess) host.push_back(*p);
    return e;
  }
  hipError_t event(hipEvent_t* ev) {
    hipError_t e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e == hipSuccess) events.push_back(*ev);
    return e;
  }
};

}  // namespace

// Synchronous build: the host walks the start-state chain in chunks of 2^22
// wires, two chunks in flight, so copying and drawing chunk c overlaps the
// walk over chunk c+1.
int devgen_build(const ckt_gen_params_t* params, const SlabLayout& L, void* slab, hipStream_t s,
                 std::vector<ckt_piece_t>& pieces) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  const SlabView v{(uint8_t*)slab, L};
  Temps tmp(s);
  if ((st = enqueue_nodes(params, N, P, v, s))) return st;

  uint64_t* h_ring[2] = {nullptr, nullptr};
  uint64_t* d_ring[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};
  const size_t chunk = std::min<size_t>(kChunk, std::max<uint32_t>(W, 1));
  for (int k = 0; k < 2; k++) {
    CKT_HIP(tmp.pinned((void**)&h_ring[k], chunk * sizeof(uint64_t)));
    CKT_HIP(tmp.device((void**)&d_ring[k], chunk * sizeof(uint64_t)));
    CKT_HIP(tmp.event(&copied[k]));
  }
  struct RingSink {
    uint64_t** h;
    uint64_t** d;
    hipEvent_t* ev;
    hipStream_t s;
    const ckt_gen_params_t* p;
    uint32_t P;
    const SlabView& v;
    uint32_t c = 0;
    int st = CKT_OK;
    uint64_t* begin(uint32_t, uint32_t) {
      // chunk c-2's copy must have left this buffer
      if (c >= 2 && st == CKT_OK) {
        hipError_t e = hipEventSynchronize(ev[c & 1]);
        if (e != hipSuccess) st = hip_fail(e, "hipEventSynchronize");
      }
      return h[c & 1];
    }
    void end(uint32_t first, uint32_t n) {
      const int k = (int)(c & 1);
      c++;
      if (st) return;
      hipError_t e = hipMemcpyAsync(d[k], h[k], n * sizeof(uint64_t), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipEventRecord(ev[k], s);
      if (e != hipSuccess) {
        st = hip_fail(e, "wire start states upload");
        return;
      }
      st = enqueue_wires(p, P, first, n, d[k], v, s);
    }
  } sink{h_ring, d_ring, copied, s, params, P, v};
  walk_wire_states(params, N, W, chunk, sink);
  if (sink.st) return sink.st;

  size_t bytes = 0;
  IndexScratch sc{};
  if ((st = index_scratch(N, W, nullptr, nullptr, &bytes))) return st;
  void* base = nullptr;
  CKT_HIP(tmp.device(&base, bytes));
  if ((st = index_scratch(N, W, base, &sc, &bytes))) return st;
  if ((st = enqueue_index(params, N, W, v, sc, s))) return st;
  pieces.resize(P);
  for (uint32_t n = 0; n < P; n++)
    pieces[n] = ckt_piece_t{n * (uint32_t)params->wires_per_piece,
                            (uint32_t)params->wires_per_piece, n * (uint32_t)params->nodes_per_piece,
                            (uint32_t)params->nodes_per_piece};
  CKT_HIP(hipStreamSynchronize(s));
  return CKT_OK;
}

int devgen_plan(const ckt_gen_params_t* params, DevgenPlan& plan) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  plan.params = *params;
  plan.N = N;
  plan.W = W;
  plan.P = P;
  plan.wire_state.resize(W);
  struct VecSink {
    uint64_t* base;
    uint64_t* begin(uint32_t first, uint32_t) { return base + first; }
    void end(uint32_t, uint32_t) {}
  } sink{plan.wire_state.data()};
  walk_wire_states(params, N, W, std::max<uint32_t>(W, 1), sink);
  return CKT_OK;
}

size_t devgen_scratch_bytes(uint32_t N, uint32_t W) {
  size_t bytes = 0;
  if (index_scratch(N, W, nullptr, nullptr, &bytes)) return 0;
  return align256((size_t)W * sizeof(uint64_t)) + bytes;
}

int devgen_enqueue(const DevgenPlan& plan, const uint64_t* staged_state, const SlabLayout& L,
                   void* slab, void* scratch, hipStream_t s, std::vector<ckt_piece_t>& pieces) {
  const ckt_gen_params_t* p = &plan.params;
  const SlabView v{(uint8_t*)slab, L};
  int st;
  if ((st = enqueue_nodes(p, plan.N, plan.P, v, s))) return st;
  uint64_t* d_state = (uint64_t*)scratch;
  if (plan.W) {
    CKT_HIP(hipMemcpyAsync(d_state, staged_state, (size_t)plan.W * sizeof(uint64_t),
                           hipMemcpyHostToDevice, s));
    if ((st = enqueue_wires(p, plan.P, 0, plan.W, d_state, v, s))) return st;
  }
  size_t bytes = 0;
  IndexScratch sc{};
  if ((st = index_scratch(plan.N, plan.W, (uint8_t*)scratch +
                                               align256((size_t)plan.W * sizeof(uint64_t)),
                          &sc, &bytes)))
    return st;
  if ((st = enqueue_index(p, plan.N, plan.W, v, sc, s))) return st;
  pieces.resize(plan.P);
  for (uint32_t n = 0; n < plan.P; n++)
    pieces[n] = ckt_piece_t{n * (uint32_t)p->wires_per_piece, (uint32_t)p->wires_per_piece,
                            n * (uint32_t)p->nodes_per_piece, (uint32_t)p->nodes_per_piece};
  return CKT_OK;
}

}  // namespace ckt
