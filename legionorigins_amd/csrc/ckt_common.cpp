// ckt_common.cpp -- status strings, device utilities and the accessor-form
// drop-ins (the entry points of apps/circuit/circuit.h:83-104).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <string>

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

// Device counter for bad-record counts of the accessor-form kernels; one per
// host thread and device, allocated on first use (so concurrent callers on
// different threads never share one).
static unsigned long long* bad_counter(int* status) {
  static thread_local unsigned long long* ptrs[64] = {nullptr};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= 64) {
    *status = hip_fail(e == hipSuccess ? hipErrorInvalidDevice : e, "hipGetDevice");
    return nullptr;
  }
  if (!ptrs[dev]) {
    e = hipMalloc((void**)&ptrs[dev], sizeof(unsigned long long));
    if (e != hipSuccess) {
      *status = hip_fail(e, "hipMalloc(bad counter)");
      ptrs[dev] = nullptr;
      return nullptr;
    }
  }
  *status = CKT_OK;
  return ptrs[dev];
}

static int read_bad(unsigned long long* d_bad, hipStream_t s, uint64_t* out) {
  unsigned long long h = 0;
  CKT_HIP(hipMemcpyAsync(&h, d_bad, sizeof(h), hipMemcpyDeviceToHost, s));
  CKT_HIP(hipStreamSynchronize(s));
  *out = h;
  return CKT_OK;
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

int ckt_calc_new_currents_gpu(const ckt_piece_t* piece, ckt_wire_t* wires, const ckt_node_t* pvt,
                              const ckt_node_t* owned, const ckt_node_t* ghost, int flag,
                              void* stream) {
  if (!piece || !wires || !pvt || !owned || !ghost)
    return fail(CKT_ERR_INVALID_ARG, "ckt_calc_new_currents_gpu: null argument");
  if (flag < CKT_POLICY_EXACT || flag > CKT_POLICY_SCALED)
    return fail(CKT_ERR_INVALID_ARG, "ckt_calc_new_currents_gpu: flag is not a CKT_POLICY_*");
  int st;
  unsigned long long* bad = bad_counter(&st);
  if (!bad) return st;
  hipStream_t s = (hipStream_t)stream;
  CKT_HIP(hipMemsetAsync(bad, 0, sizeof(*bad), s));
  st = finish(launch_cnc_aos(flag, piece->first_wire, piece->num_wires, wires, pvt, owned, ghost,
                             CKT_STEPS, bad, s),
              s, "calc_new_currents");
  if (st) return st;
  uint64_t nbad = 0;
  if ((st = read_bad(bad, s, &nbad))) return st;
  if (nbad) return fail(CKT_ERR_INVALID_ARG, "calc_new_currents: " + std::to_string(nbad) +
                                                 " wires with an invalid PointerLocation");
  return CKT_OK;
}

int ckt_distribute_charge_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires, ckt_node_t* pvt,
                              float* owned_fold, float* ghost_fold, int flag, void* stream) {
  if (!piece || !wires || !pvt || !owned_fold || !ghost_fold)
    return fail(CKT_ERR_INVALID_ARG, "ckt_distribute_charge_gpu: null argument");
  if (flag != 0) return fail(CKT_ERR_INVALID_ARG, "ckt_distribute_charge_gpu: flag must be 0");
  int st;
  unsigned long long* bad = bad_counter(&st);
  if (!bad) return st;
  hipStream_t s = (hipStream_t)stream;
  CKT_HIP(hipMemsetAsync(bad, 0, sizeof(*bad), s));
  st = finish(launch_dsc_aos(piece->first_wire, piece->num_wires, wires, pvt, owned_fold, ghost_fold,
                             bad, s),
              s, "distribute_charge");
  if (st) return st;
  uint64_t nbad = 0;
  if ((st = read_bad(bad, s, &nbad))) return st;
  if (nbad) return fail(CKT_ERR_INVALID_ARG, "distribute_charge: " + std::to_string(nbad) +
                                                 " wires with an invalid PointerLocation");
  return CKT_OK;
}

int ckt_update_voltages_gpu(const ckt_piece_t* piece, ckt_node_t* pvt, ckt_node_t* owned,
                            const int32_t* locator, int flag, void* stream) {
  if (!piece || !pvt || !owned || !locator)
    return fail(CKT_ERR_INVALID_ARG, "ckt_update_voltages_gpu: null argument");
  if (flag != 0) return fail(CKT_ERR_INVALID_ARG, "ckt_update_voltages_gpu: flag must be 0");
  hipStream_t s = (hipStream_t)stream;
  return finish(launch_upv_aos(piece->first_node, piece->num_nodes, pvt, owned, locator, s), s,
                "update_voltages");
}

int ckt_sanity_check_wires_gpu(const ckt_piece_t* piece, const ckt_wire_t* wires,
                               uint32_t num_nodes_total, uint64_t* num_bad, void* stream) {
  if (!piece || !wires || !num_bad)
    return fail(CKT_ERR_INVALID_ARG, "ckt_sanity_check_wires_gpu: null argument");
  int st;
  unsigned long long* bad = bad_counter(&st);
  if (!bad) return st;
  hipStream_t s = (hipStream_t)stream;
  CKT_HIP(hipMemsetAsync(bad, 0, sizeof(*bad), s));
  st = finish(launch_sanity_aos(*piece, wires, num_nodes_total, bad, s), s, "sanity_check_wires");
  if (st) return st;
  return read_bad(bad, s, num_bad);
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

int ckt_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
  if (!dst || !src) return fail(CKT_ERR_INVALID_ARG, "ckt_copy_kernel: null pointer");
  if (bytes % 16) return fail(CKT_ERR_INVALID_ARG, "ckt_copy_kernel: size must be a multiple of 16");
  hipError_t e = launch_copy(dst, src, bytes, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "copy_kernel");
  return CKT_OK;
}

}  // extern "C"
