// ckt_common.cpp -- status strings, device utilities and the accessor-form
// drop-ins (the entry points of apps/circuit/circuit.h:83-104).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "ckt_internal.h"

namespace ckt {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CKT_ERR_NO_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return CKT_ERR_NO_DEVICE;
  return CKT_ERR_HIP;
}

void clear_error() { g_last_error.clear(); }

// Runs a synchronous accessor-form call: launch on `s`, then wait, as the
// reference's host launchers do (cudaDeviceSynchronize, circuit_gpu.cu:204).
static int finish(hipError_t launched, hipStream_t s, const char* what) {
  if (launched != hipSuccess) return hip_fail(launched, what);
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, what);
  return CKT_OK;
}

// Device scratch of an accessor-form call over n pieces: the bad-record
// counter (8 bytes, then 8 of padding), first_block[n + 1] (PieceIndex), and
// the piece table on a 16-byte boundary.
constexpr size_t index_pieces_offset(uint32_t n) {
  return (16 + 4 * ((size_t)n + 1) + 15) & ~(size_t)15;
}
constexpr size_t index_bytes(uint32_t n) {
  return index_pieces_offset(n) + sizeof(ckt_piece_t) * (size_t)n;
}
struct IndexLayout {
  size_t first_block, pieces, bytes;
  explicit IndexLayout(uint32_t n)
      : first_block(16), pieces(index_pieces_offset(n)), bytes(index_bytes(n)) {}
};

// The pieces of one call must not share wires (or nodes, for a task over
// nodes): the reference's pieces partition the circuit, and a wire in two
// pieces of one launch would be relaxed twice at once.  `order` is scratch
// for n indices (the caller's, so the allocation-free entry point stays
// allocation-free).
static int check_disjoint(const ckt_piece_t* pieces, uint32_t n, bool over_nodes, uint32_t* order,
                          const char* what) {
  auto first = [&](uint32_t i) -> uint64_t {
    return over_nodes ? pieces[i].first_node : pieces[i].first_wire;
  };
  auto count = [&](uint32_t i) -> uint64_t {
    return over_nodes ? pieces[i].num_nodes : pieces[i].num_wires;
  };
  uint32_t m = 0;
  for (uint32_t i = 0; i < n; i++)
    if (count(i)) order[m++] = i;
  std::sort(order, order + m, [&](uint32_t a, uint32_t b) { return first(a) < first(b); });
  for (uint32_t k = 1; k < m; k++) {
    const uint32_t a = order[k - 1], b = order[k];
    if (first(a) + count(a) > first(b))
      return fail(CKT_ERR_INVALID_ARG,
                  std::string(what) + ": pieces " + std::to_string(a) + " and " +
                      std::to_string(b) + " share " + (over_nodes ? "nodes" : "wires"));
  }
  return CKT_OK;
}

// Host image of the scratch's head: the zeroed counter (image[0..3]) and
// first_block[] over each piece's wires or nodes (image[4..4+n]); n + 5 words.
static int stage_index(const ckt_piece_t* pieces, uint32_t n, bool over_nodes, uint32_t* image,
                       uint32_t* blocks) {
  image[0] = image[1] = image[2] = image[3] = 0;
  uint64_t acc = 0;
  for (uint32_t p = 0; p < n; p++) {
    image[4 + p] = (uint32_t)acc;
    const uint32_t count = over_nodes ? pieces[p].num_nodes : pieces[p].num_wires;
    acc += (count + 255u) / 256u;
    if (acc > 0x7fffffffu)
      return fail(CKT_ERR_OUT_OF_RANGE, "accessor-form call: more than 2^31 workgroups");
  }
  image[4 + n] = (uint32_t)acc;
  *blocks = (uint32_t)acc;
  return CKT_OK;
}

// One accessor-form task over a piece table, with its scratch at `dev`
// (IndexLayout(n).bytes of device memory) and `image` (n + 5 words of host
// memory) to stage it: copies the table in, launches, reads the bad-record
// counter into *nbad (if not NULL) and waits, as the reference launchers do
// (cudaDeviceSynchronize, circuit_gpu.cu:204).
template <class Launch>
static int run_indexed(const ckt_piece_t* pieces, uint32_t n, bool over_nodes, unsigned char* dev,
                       uint32_t* image, hipStream_t s, const char* what, uint64_t* nbad,
                       Launch launch) {
  const IndexLayout lay(n);
  uint32_t blocks = 0;
  int st = stage_index(pieces, n, over_nodes, image, &blocks);
  if (st) return st;
  const PieceIndex ix{(const ckt_piece_t*)(dev + lay.pieces),
                      (const uint32_t*)(dev + lay.first_block), n};
  unsigned long long* d_bad = (unsigned long long*)dev;
  unsigned long long h_bad = 0;
  hipError_t e = hipMemcpyAsync(dev, image, 4 * ((size_t)n + 5), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev + lay.pieces, pieces, sizeof(ckt_piece_t) * n, hipMemcpyHostToDevice,
                       s);
  if (e == hipSuccess) e = launch(ix, blocks, d_bad);
  if (e == hipSuccess && nbad)
    e = hipMemcpyAsync(&h_bad, d_bad, sizeof(h_bad), hipMemcpyDeviceToHost, s);
  // The copies read and write host memory the caller owns: wait in every case.
  const hipError_t w = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, what);
  if (w != hipSuccess) return hip_fail(w, what);
  if (nbad) *nbad = h_bad;
  return CKT_OK;
}

// run_indexed with its scratch from the stream-ordered allocator, released
// behind the call.
template <class Launch>
static int run_indexed_alloc(const ckt_piece_t* pieces, uint32_t n, bool over_nodes, hipStream_t s,
                             const char* what, uint64_t* nbad, Launch launch) {
  std::vector<uint32_t> image((size_t)n + 5);
  if (int st = check_disjoint(pieces, n, over_nodes, image.data(), what)) return st;
  unsigned char* dev = nullptr;
  CKT_HIP(hipMallocAsync((void**)&dev, IndexLayout(n).bytes, s));
  const int st = run_indexed(pieces, n, over_nodes, dev, image.data(), s, what, nbad, launch);
  const hipError_t e = hipFreeAsync(dev, s);
  if (st) return st;
  if (e != hipSuccess) return hip_fail(e, what);
  return CKT_OK;
}

// Device scratch of ckt_calc_new_currents_index_gpu_noalloc: the layout
// above for CKT_NOALLOC_MAX_PIECES pieces, one copy per device (each GPU the
// library runs on holds it until the process ends).  In host code `scratch`
// names no device memory; kernels and copies get the per-device address that
// hipGetSymbolAddress returns, looked up once per device on the first call
// made with that device current.
static_assert(index_bytes(CKT_NOALLOC_MAX_PIECES) == CKT_NOALLOC_SCRATCH_BYTES,
              "CKT_NOALLOC_SCRATCH_BYTES is the layout's size for CKT_NOALLOC_MAX_PIECES");
static __device__ __align__(256) unsigned char scratch[CKT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_status[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];

static int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  CKT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(CKT_ERR_NO_DEVICE, "allocation-free call: device index beyond 63");
  std::call_once(g_scratch_once[dev], [dev] {
    g_scratch_status[dev] = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
  });
  if (g_scratch_status[dev] != hipSuccess)
    return hip_fail(g_scratch_status[dev], "hipGetSymbolAddress(scratch)");
  *out = (unsigned char*)g_scratch_ptr[dev];
  return CKT_OK;
}

static int check_pieces(const ckt_piece_t* pieces, uint32_t n, const char* what) {
  if (!pieces && n) return fail(CKT_ERR_INVALID_ARG, std::string(what) + ": null piece table");
  return CKT_OK;
}

static int bad_locations(uint64_t nbad, const char* what) {
  if (!nbad) return CKT_OK;
  return fail(CKT_ERR_INVALID_ARG, std::string(what) + ": " + std::to_string(nbad) +
                                       " wires with an invalid PointerLocation");
}

}  // namespace ckt

using namespace ckt;

extern "C" {

int ckt_abi_version(void) { return CKT_ABI_VERSION; }

const char* ckt_status_string(int status) {
  switch (status) {
    case CKT_OK: return "ok";
    case CKT_ERR_INVALID_ARG: return "invalid argument";
    case CKT_ERR_OUT_OF_RANGE: return "out of range";
    case CKT_ERR_NO_MEMORY: return "out of memory";
    case CKT_ERR_BUDGET: return "over HBM budget";
    case CKT_ERR_HIP: return "HIP runtime error";
    case CKT_ERR_NO_DEVICE: return "no usable device";
    case CKT_ERR_STATE: return "invalid state";
    case CKT_ERR_UNSUPPORTED: return "unsupported";
    case CKT_ERR_NOT_FOUND: return "not found";
    case CKT_ERR_CANCELLED: return "cancelled";
    case CKT_ERR_TIMEOUT: return "timeout";
    default: return "unknown status";
  }
}

const char* ckt_last_error(void) { return g_last_error.c_str(); }

int ckt_device_info(int device, int* is_gfx950, int* num_cus, size_t* lds_per_block,
                    size_t* hbm_bytes) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(CKT_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count) return fail(CKT_ERR_NO_DEVICE, "device index out of range");
  hipDeviceProp_t prop;
  CKT_HIP(hipGetDeviceProperties(&prop, device));
  if (is_gfx950) *is_gfx950 = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
  if (num_cus) *num_cus = prop.multiProcessorCount;
  if (lds_per_block) *lds_per_block = prop.sharedMemPerBlock;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  return CKT_OK;
}

int ckt_device_malloc(int device, void** ptr, size_t bytes) {
  if (!ptr) return fail(CKT_ERR_INVALID_ARG, "ckt_device_malloc: null ptr");
  *ptr = nullptr;
  CKT_HIP(hipSetDevice(device));
  CKT_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return CKT_OK;
}

int ckt_device_free(void* ptr) {
  if (ptr) CKT_HIP(hipFree(ptr));
  return CKT_OK;
}

int ckt_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return fail(CKT_ERR_INVALID_ARG, "ckt_memcpy_htod: null pointer");
  if (bytes) CKT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CKT_OK;
}

int ckt_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return fail(CKT_ERR_INVALID_ARG, "ckt_memcpy_dtoh: null pointer");
  if (bytes) CKT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CKT_OK;
}

int ckt_device_synchronize(int device) {
  CKT_HIP(hipSetDevice(device));
  CKT_HIP(hipDeviceSynchronize());
  return CKT_OK;
}

// ---------------------------------------------------------------------------
// Accessor-form drop-ins.

int ckt_calc_new_currents_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces,
                                    ckt_wire_t* wires, const ckt_node_t* pvt,
                                    const ckt_node_t* owned, const ckt_node_t* ghost, int flag,
                                    void* stream) {
  const char* what = "calc_new_currents";
  if (!wires || !pvt || !owned || !ghost)
    return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: null instance");
  if (int st = check_pieces(pieces, num_pieces, what)) return st;
  if (flag < CKT_POLICY_EXACT || flag > CKT_POLICY_SCALED)
    return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: flag is not a CKT_POLICY_*");
  hipStream_t s = (hipStream_t)stream;
  uint64_t nbad = 0;
  int st = run_indexed_alloc(pieces, num_pieces, false, s, what, &nbad,
                             [&](const PieceIndex& ix, uint32_t blocks, unsigned long long* bad) {
                               return launch_cnc_aos(flag, ix, blocks, wires, pvt, owned, ghost,
                                                     CKT_STEPS, bad, s);
                             });
  return st ? st : bad_locations(nbad, what);
}

int ckt_calc_new_currents_index_gpu_noalloc(const ckt_piece_t* pieces, uint32_t num_pieces,
                                            ckt_wire_t* wires, const ckt_node_t* pvt,
                                            const ckt_node_t* owned, const ckt_node_t* ghost,
                                            int flag, void* stream) {
  const char* what = "calc_new_currents";
  if (!wires || !pvt || !owned || !ghost)
    return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: null instance");
  if (int st = check_pieces(pieces, num_pieces, what)) return st;
  if (flag < CKT_POLICY_EXACT || flag > CKT_POLICY_SCALED)
    return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: flag is not a CKT_POLICY_*");
  if (num_pieces > CKT_NOALLOC_MAX_PIECES)
    return fail(CKT_ERR_OUT_OF_RANGE,
                "ckt_calc_new_currents_index_gpu_noalloc: " + std::to_string(num_pieces) +
                    " pieces, more than CKT_NOALLOC_MAX_PIECES (" +
                    std::to_string(CKT_NOALLOC_MAX_PIECES) + ")");
  uint32_t image[CKT_NOALLOC_MAX_PIECES + 5];
  if (int st = check_disjoint(pieces, num_pieces, false, image, what)) return st;
  unsigned char* dev = nullptr;
  if (int st = noalloc_scratch(&dev)) return st;
  hipStream_t s = (hipStream_t)stream;
  uint64_t nbad = 0;
  int st = run_indexed(pieces, num_pieces, false, dev, image, s, what, &nbad,
                       [&](const PieceIndex& ix, uint32_t blocks, unsigned long long* bad) {
                         return launch_cnc_aos(flag, ix, blocks, wires, pvt, owned, ghost,
                                               CKT_STEPS, bad, s);
                       });
  return st ? st : bad_locations(nbad, what);
}

int ckt_calc_new_currents_gpu(const ckt_piece_t* piece, ckt_wire_t* wires, const ckt_node_t* pvt,
                              const ckt_node_t* owned, const ckt_node_t* ghost, int flag,
                              void* stream) {
  if (!piece) return fail(CKT_ERR_INVALID_ARG, "ckt_calc_new_currents_gpu: null argument");
  return ckt_calc_new_currents_index_gpu(piece, 1, wires, pvt, owned, ghost, flag, stream);
}

int ckt_distribute_charge_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces,
                                    const ckt_wire_t* wires, ckt_node_t* pvt, float* owned_fold,
                                    float* ghost_fold, int flag, void* stream) {
  const char* what = "distribute_charge";
  if (!wires || !pvt || !owned_fold || !ghost_fold)
    return fail(CKT_ERR_INVALID_ARG, "distribute_charge: null instance");
  if (int st = check_pieces(pieces, num_pieces, what)) return st;
  if (flag != 0) return fail(CKT_ERR_INVALID_ARG, "distribute_charge: flag must be 0");
  hipStream_t s = (hipStream_t)stream;
  uint64_t nbad = 0;
  int st = run_indexed_alloc(pieces, num_pieces, false, s, what, &nbad,
                             [&](const PieceIndex& ix, uint32_t blocks, unsigned long long* bad) {
                               return launch_dsc_aos(ix, blocks, wires, pvt, owned_fold,
                                                     ghost_fold, bad, s);
                             });
  return st ? st : bad_locations(nbad, what);
}

int ckt_distribute_charge_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires, ckt_node_t* pvt,
                              float* owned_fold, float* ghost_fold, int flag, void* stream) {
  if (!piece) return fail(CKT_ERR_INVALID_ARG, "ckt_distribute_charge_gpu: null argument");
  return ckt_distribute_charge_index_gpu(piece, 1, wires, pvt, owned_fold, ghost_fold, flag,
                                         stream);
}

int ckt_update_voltages_index_gpu(const ckt_piece_t* pieces, uint32_t num_pieces, ckt_node_t* pvt,
                                  ckt_node_t* owned, const int32_t* locator, int flag,
                                  void* stream) {
  const char* what = "update_voltages";
  if (!pvt || !owned || !locator)
    return fail(CKT_ERR_INVALID_ARG, "update_voltages: null instance");
  if (int st = check_pieces(pieces, num_pieces, what)) return st;
  if (flag != 0) return fail(CKT_ERR_INVALID_ARG, "update_voltages: flag must be 0");
  hipStream_t s = (hipStream_t)stream;
  return run_indexed_alloc(pieces, num_pieces, true, s, what, nullptr,
                           [&](const PieceIndex& ix, uint32_t blocks, unsigned long long*) {
                             return launch_upv_aos(ix, blocks, pvt, owned, locator, s);
                           });
}

int ckt_update_voltages_gpu(const ckt_piece_t* piece, ckt_node_t* pvt, ckt_node_t* owned,
                            const int32_t* locator, int flag, void* stream) {
  if (!piece) return fail(CKT_ERR_INVALID_ARG, "ckt_update_voltages_gpu: null argument");
  return ckt_update_voltages_index_gpu(piece, 1, pvt, owned, locator, flag, stream);
}

int ckt_sanity_check_wires_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires,
                               uint32_t num_nodes_total, uint64_t* num_bad, void* stream) {
  if (!piece || !wires || !num_bad)
    return fail(CKT_ERR_INVALID_ARG, "ckt_sanity_check_wires_gpu: null argument");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* bad = nullptr;
  CKT_HIP(hipMallocAsync((void**)&bad, sizeof(*bad), s));
  unsigned long long h = 0;
  hipError_t e = hipMemsetAsync(bad, 0, sizeof(*bad), s);
  if (e == hipSuccess) e = launch_sanity_aos(*piece, wires, num_nodes_total, bad, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, s);
  const hipError_t f = hipFreeAsync(bad, s);
  const hipError_t w = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "sanity_check_wires");
  if (f != hipSuccess) return hip_fail(f, "sanity_check_wires");
  if (w != hipSuccess) return hip_fail(w, "sanity_check_wires");
  *num_bad = h;
  return CKT_OK;
}

int ckt_fold_init_gpu(float* fold, uint32_t first, uint32_t count, void* stream) {
  if (!fold) return fail(CKT_ERR_INVALID_ARG, "ckt_fold_init_gpu: null fold");
  hipStream_t s = (hipStream_t)stream;
  return finish(launch_fold_init(fold, first, count, s), s, "fold_init");
}

int ckt_fold_apply_gpu(ckt_node_t* owner, const float* fold, uint32_t first, uint32_t count,
                       void* stream) {
  if (!owner || !fold) return fail(CKT_ERR_INVALID_ARG, "ckt_fold_apply_gpu: null argument");
  hipStream_t s = (hipStream_t)stream;
  return finish(launch_fold_apply(owner, fold, first, count, s), s, "fold_apply");
}

int ckt_debug_bounds_violations(int device, uint64_t* count) {
  if (!count) return fail(CKT_ERR_INVALID_ARG, "ckt_debug_bounds_violations: null count");
  *count = 0;
  CKT_HIP(hipSetDevice(device));
  CKT_HIP(hipDeviceSynchronize());
  const hipError_t e = read_bounds_violations(count);
  if (e == hipErrorNotSupported)
    return fail(CKT_ERR_UNSUPPORTED, "the library was built without bounds checks (CKT_BOUNDS_CHECKS)");
  if (e != hipSuccess) return hip_fail(e, "bounds-check counter");
  return CKT_OK;
}

int ckt_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
  if (!dst || !src) return fail(CKT_ERR_INVALID_ARG, "ckt_copy_kernel: null pointer");
  if (bytes % 16) return fail(CKT_ERR_INVALID_ARG, "ckt_copy_kernel: size must be a multiple of 16");
  hipError_t e = launch_copy(dst, src, bytes, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "copy_kernel");
  return CKT_OK;
}

}  // extern "C"
