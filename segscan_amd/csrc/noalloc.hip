// Allocation-free SpMM entry point: segscan_spmm_csr with its scratch (ebsr carries,
// ebpr tile rows) in a static device array instead of a caller-provided workspace,
// for callers that cannot call hipMalloc (code under graph capture, serving loops).
//
// `scratch` is a device symbol: in host code its name is a host-side stand-in, not a
// device address. Its address on each GPU is taken once per device, with that device
// current, by hipGetSymbolAddress; only that address is passed to the kernels.
#include <cstdio>
#include <mutex>

#include <hip/hip_runtime.h>

#include "../../include/segscan.h"
#include "abi_internal.hpp"

namespace {

constexpr size_t kScratchBytes = SEGSCAN_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on first use).
segscan_status_t scratch_on_current_device(void* stream, void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return segscan::set_error(SEGSCAN_EHIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return segscan::set_error(SEGSCAN_EUNSUPPORTED, "device ordinal beyond 63");
  if (stream) {
    int sdev = -1;
    e = hipStreamGetDevice(static_cast<hipStream_t>(stream), &sdev);
    if (e != hipSuccess) return segscan::set_error(SEGSCAN_EHIP, hipGetErrorString(e));
    if (sdev != dev)
      return segscan::set_error(SEGSCAN_EINVAL, "stream belongs to another device than the current one");
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return segscan::set_error(SEGSCAN_EHIP, hipGetErrorString(e));
    if (!p) return segscan::set_error(SEGSCAN_EHIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SEGSCAN_OK;
}

}  // namespace

extern "C" segscan_status_t segscan_spmm_csr_noalloc(const segscan_csr_t* A, int64_t N, const void* B, int64_t ldb,
                                                     segscan_dtype_t b_dtype, float* C, int64_t ldc,
                                                     segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                                     uint32_t flags, void* stream) {
  // validation and planning first, with no HIP call
  size_t need = 0;
  segscan_status_t st = segscan::spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, nullptr, 0,
                                               stream, &need);
  if (st != SEGSCAN_OK) return st;
  if (need > kScratchBytes) {
    char msg[160];
    snprintf(msg, sizeof(msg), "plan needs %zu bytes of scratch; the allocation-free entry point holds %zu",
             need, kScratchBytes);
    return segscan::set_error(SEGSCAN_EWORKSPACE, msg);
  }
  void* ws = nullptr;
  if (need > 0) {
    st = scratch_on_current_device(stream, &ws);
    if (st != SEGSCAN_OK) return st;
  }
  return segscan::spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, ws, ws ? kScratchBytes : 0,
                                stream, nullptr);
}
