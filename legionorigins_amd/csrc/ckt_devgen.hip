// ckt_devgen.hip -- load_circuit (apps/circuit/circuit.cc:1071-1288) built
// directly in HBM.
//
// The reference draws from one drand48 stream in a fixed order: three draws
// per node, then per wire six (in-piece) or seven (cross-piece) draws.  A
// node's first draw is a fixed jump from the seed, so every node is drawn
// independently on the device.  A wire's first draw depends on how many
// earlier wires were cross-piece, which is only known by walking the chain:
// the host does that walk (two 48-bit multiply-adds per wire: jump to the
// in-piece test, then jump to the next wire) and streams each wire's start
// state to the device in chunks, overlapped with the device drawing the
// previous chunk.  The device then marks shared nodes, derives the
// locations, and builds the node -> (wire, end) index with a stable radix
// sort.  The slab is bit-identical to the host generator's image
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

uint32_t grid_for(size_t n) { return (uint32_t)((n + 255) / 256); }

// Temporary device and pinned buffers and events of one build; released
// after the stream has drained.
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

int devgen_build(const ckt_gen_params_t* params, const SlabLayout& L, void* slab, hipStream_t s,
                 std::vector<ckt_piece_t>& pieces) {
  uint32_t N, W, P;
  int st = gen_validate(params, &N, &W, &P);
  if (st) return st;
  const uint32_t npp = (uint32_t)params->nodes_per_piece, wpp = (uint32_t)params->wires_per_piece;
  const int32_t pct = params->pct_wire_in_piece;
  const int32_t stable = params->profile == CKT_PROFILE_STABLE;
  const uint64_t seed_state = ((uint64_t)(uint32_t)params->random_seed << 16) | 0x330Eull;

  uint8_t* b = (uint8_t*)slab;
  float* ind = (float*)(b + L.ind);
  float* res = (float*)(b + L.res);
  float* wcap = (float*)(b + L.wcap);
  uint32_t* in_ptr = (uint32_t*)(b + L.in_ptr);
  uint32_t* out_ptr = (uint32_t*)(b + L.out_ptr);
  uint8_t* locator = b + L.locator;
  uint32_t* csr_row = (uint32_t*)(b + L.csr_row);

  Temps tmp(s);
  // Zero everything (currents, internal voltages, charges, padding), then
  // mark every node private until a cross-piece wire reaches it.
  CKT_HIP(launch_fill_u32((uint32_t*)slab, 0u, L.total / 4, s));
  hipLaunchKernelGGL(fill_bytes, dim3(grid_for(N)), dim3(256), 0, s, locator, (uint8_t)1, N);
  CKT_HIP(hipGetLastError());

  NodeJumps t{};
  for (int k = 0; k < 32; k++) t.j[k] = jump(3ull << k);
  hipLaunchKernelGGL(gen_nodes, dim3(grid_for(N)), dim3(256), 0, s, N, seed_state, t,
                     (float*)(b + L.voltage), (float*)(b + L.ncap), (float*)(b + L.leak));
  CKT_HIP(hipGetLastError());

  // Wire start states, two chunks in flight.
  uint64_t* h_ring[2] = {nullptr, nullptr};
  uint64_t* d_ring[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};
  const size_t chunk = std::min<size_t>(kChunk, std::max<uint32_t>(W, 1));
  for (int k = 0; k < 2; k++) {
    CKT_HIP(tmp.pinned((void**)&h_ring[k], chunk * sizeof(uint64_t)));
    CKT_HIP(tmp.device((void**)&d_ring[k], chunk * sizeof(uint64_t)));
    CKT_HIP(tmp.event(&copied[k]));
  }
  const Jump j5 = jump(5), j6 = jump(6), j7 = jump(7);
  uint64_t x = apply(jump(3ull * N), seed_state);
  for (uint32_t first = 0, c = 0; first < W; first += (uint32_t)chunk, c++) {
    const int k = (int)(c & 1);
    const uint32_t n = (uint32_t)std::min<size_t>(chunk, W - first);
    if (c >= 2) CKT_HIP(hipEventSynchronize(copied[k]));  // chunk c-2's copy has left the buffer
    uint64_t* h = h_ring[k];
    for (uint32_t i = 0; i < n; i++) {
      h[i] = x;
      // The wire's fifth draw is the in-piece test (circuit.cc:1167).
      const double u = (double)apply(j5, x) * 0x1p-48;
      x = ((100 * u) < pct) ? apply(j6, x) : apply(j7, x);
    }
    CKT_HIP(hipMemcpyAsync(d_ring[k], h, n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    CKT_HIP(hipEventRecord(copied[k], s));
    hipLaunchKernelGGL(gen_wires, dim3(grid_for(n)), dim3(256), 0, s, first, n, P, npp, wpp, pct,
                       stable, d_ring[k], ind, res, wcap, in_ptr, out_ptr, locator);
    CKT_HIP(hipGetLastError());
  }

  // Index: counts -> row offsets; a stable sort of the entries by node keeps
  // each node's entries in (wire, in-before-out) order.
  uint32_t *keys = nullptr, *vals = nullptr, *keys_sorted = nullptr, *counts = nullptr;
  CKT_HIP(tmp.device((void**)&keys, (size_t)2 * W * sizeof(uint32_t)));
  CKT_HIP(tmp.device((void**)&vals, (size_t)2 * W * sizeof(uint32_t)));
  CKT_HIP(tmp.device((void**)&keys_sorted, (size_t)2 * W * sizeof(uint32_t)));
  CKT_HIP(tmp.device((void**)&counts, (size_t)N * sizeof(uint32_t)));
  CKT_HIP(launch_fill_u32(counts, 0u, N, s));
  if (W) {
    hipLaunchKernelGGL(gen_locations, dim3(grid_for(W)), dim3(256), 0, s, W, npp, wpp, in_ptr,
                       out_ptr, locator, b + L.in_loc, b + L.out_loc, keys, vals, counts);
    CKT_HIP(hipGetLastError());
  }
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < N) bits++;
  size_t scan_bytes = 0, sort_bytes = 0;
  CKT_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, counts, csr_row + 1, (size_t)N,
                                  rocprim::plus<uint32_t>(), s));
  if (W)
    CKT_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys, keys_sorted, vals,
                                      (uint32_t*)(b + L.csr_ent), (size_t)2 * W, 0, bits, s));
  void* work = nullptr;
  CKT_HIP(tmp.device(&work, std::max(scan_bytes, sort_bytes)));
  CKT_HIP(rocprim::inclusive_scan(work, scan_bytes, counts, csr_row + 1, (size_t)N,
                                  rocprim::plus<uint32_t>(), s));
  if (W)
    CKT_HIP(rocprim::radix_sort_pairs(work, sort_bytes, keys, keys_sorted, vals,
                                      (uint32_t*)(b + L.csr_ent), (size_t)2 * W, 0, bits, s));

  pieces.resize(P);
  for (uint32_t n = 0; n < P; n++) pieces[n] = ckt_piece_t{n * wpp, wpp, n * npp, npp};
  CKT_HIP(hipMemcpyAsync(b + L.pieces, pieces.data(), P * sizeof(ckt_piece_t),
                         hipMemcpyHostToDevice, s));
  CKT_HIP(hipStreamSynchronize(s));
  return CKT_OK;
}

}  // namespace ckt
