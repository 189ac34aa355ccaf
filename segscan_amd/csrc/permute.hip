// Fixed permutation of per-entry values in two streaming passes.
//
// out[i] = in[src[i]] for a permutation src of [0, n) that stays the same from call to
// call (values kept in the transpose's order, read in A's order: src = the inverse of
// the transpose's perm). A direct gather makes n random 4-byte reads, each a full line
// from outside L2 (4.19M values: 66 us). Here the plan sorts the work once on the host
// and each run streams:
//   pass 1: input tile t (kTile values, read coalesced) is staged in LDS in order of its
//           values' destination buckets (kBucket consecutive outputs each) and written to
//           `mid` bucket by bucket, tile runs one after another inside each bucket;
//   pass 2: bucket b (kBucket values, contiguous in `mid`) is loaded into LDS and written
//           to out[b * kBucket ...] in final order, coalesced.
// Per value the plan holds a 16-bit rank inside its tile (pass 1's LDS slot), the `mid`
// position of each staged slot, and a 16-bit slot inside its bucket (pass 2's LDS read).
#include <algorithm>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/segscan.h"
#include "abi_internal.hpp"
#include "kernel_params.hpp"

namespace segscan {
namespace {

constexpr int kTile = 16384;    // values per pass-1 tile (64 KB of LDS)
constexpr int kBucket = kPermBucket;  // outputs per pass-2 bucket (8192: 32 KB of LDS)
constexpr int kPermTpb = 1024;

__global__ void __launch_bounds__(kPermTpb)
permute_pass1_kernel(const float* __restrict__ in, const uint16_t* __restrict__ rank1,
                     const int32_t* __restrict__ mid_pos, float* __restrict__ mid, int64_t n) {
  __shared__ float s[kTile];
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kTile;
  const int len = static_cast<int>(min<int64_t>(kTile, n - t0));
  // every stream here is read once: non-temporal loads leave L2 to the kernels around the
  // plan (fwd + bwd step 0.3518 -> 0.3494 ms over five pairs)
  for (int k = threadIdx.x; k < len; k += kPermTpb)
    s[__builtin_nontemporal_load(rank1 + t0 + k)] = __builtin_nontemporal_load(in + t0 + k);
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += kPermTpb) mid[__builtin_nontemporal_load(mid_pos + t0 + k)] = s[k];
}

__global__ void __launch_bounds__(kPermTpb)
permute_pass2_kernel(const float* __restrict__ mid, const uint16_t* __restrict__ slot2, float* __restrict__ out,
                     int64_t n) {
  __shared__ float s[kBucket];
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kBucket;
  const int len = static_cast<int>(min<int64_t>(kBucket, n - b0));
  for (int k = threadIdx.x; k < len; k += kPermTpb) s[k] = mid[b0 + k];
  __syncthreads();
  for (int k = threadIdx.x; k < len; k += kPermTpb) out[b0 + k] = s[slot2[b0 + k]];
}

}  // namespace
}  // namespace segscan

struct segscan_permute_plan {
  int device = 0;
  int64_t n = 0;
  void* mem = nullptr;  // rank1, slot2 (uint16), mid_pos (int32), mid, aux (float)
  uint16_t* rank1 = nullptr;
  uint16_t* slot2 = nullptr;
  int32_t* mid_pos = nullptr;
  float* mid = nullptr;
  float* aux = nullptr;  // A-order values for segscan_spmm_csr_permuted's unfused launches
};

namespace segscan {

PermuteBuffers permute_buffers(const segscan_permute_plan* plan) {
  PermuteBuffers b;
  b.device = plan->device;
  b.n = plan->n;
  b.mid = plan->mid;
  b.slot2 = plan->slot2;
  b.aux = plan->aux;
  return b;
}

hipError_t permute_pass1(const segscan_permute_plan* plan, const float* in, hipStream_t s) {
  const unsigned tiles = static_cast<unsigned>((plan->n + kTile - 1) / kTile);
  permute_pass1_kernel<<<tiles, kPermTpb, 0, s>>>(in, plan->rank1, plan->mid_pos, plan->mid, plan->n);
  return hipGetLastError();
}

hipError_t permute_pass2(const segscan_permute_plan* plan, float* out, hipStream_t s) {
  const unsigned buckets = static_cast<unsigned>((plan->n + kBucket - 1) / kBucket);
  permute_pass2_kernel<<<buckets, kPermTpb, 0, s>>>(plan->mid, plan->slot2, out, plan->n);
  return hipGetLastError();
}

}  // namespace segscan

namespace {

constexpr int64_t kMaxI32 = 2147483647LL;

size_t a256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

segscan_status_t hip_fail(hipError_t e, const char* what) {
  char msg[256];
  std::snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
  return segscan::set_error(SEGSCAN_EHIP, msg);
}

segscan_status_t einval(const char* msg) { return segscan::set_error(SEGSCAN_EINVAL, msg); }

}  // namespace

extern "C" {

segscan_status_t segscan_permute_plan_create(const int32_t* src, int64_t n, void* stream,
                                             segscan_permute_plan_t* plan) {
  using segscan::kBucket;
  using segscan::kTile;
  if (!plan) return einval("plan is NULL");
  *plan = nullptr;
  if (n < 1 || n > kMaxI32) return einval("n out of [1, 2^31)");
  if (!src) return einval("src is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  std::vector<int32_t> h_src(static_cast<size_t>(n));
  if (e == hipSuccess) e = hipMemcpyAsync(h_src.data(), src, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "permute plan: read src");
  // dst[q]: the output position that takes in[q]
  std::vector<int32_t> dst(static_cast<size_t>(n), -1);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t q = h_src[i];
    if (q < 0 || q >= n || dst[q] >= 0) return einval("src is not a permutation of [0, n)");
    dst[q] = static_cast<int32_t>(i);
  }
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int64_t nbuckets = (n + kBucket - 1) / kBucket;
  std::vector<uint16_t> rank1(static_cast<size_t>(n)), slot2(static_cast<size_t>(n));
  std::vector<int32_t> mid_pos(static_cast<size_t>(n));
  // filled[b]: values of bucket b placed in mid so far (tiles in order), i.e. the offset of
  // the current tile's run inside bucket b
  std::vector<int32_t> filled(static_cast<size_t>(nbuckets), 0);
  std::vector<uint64_t> key(kTile);
  for (int64_t t = 0; t < ntiles; ++t) {
    const int64_t q0 = t * kTile;
    const int len = static_cast<int>(std::min<int64_t>(kTile, n - q0));
    // the tile's values sorted by destination bucket, tile order inside a bucket
    for (int i = 0; i < len; ++i)
      key[i] = (static_cast<uint64_t>(dst[q0 + i] / kBucket) << 32) | static_cast<uint32_t>(i);
    std::sort(key.begin(), key.begin() + len);
    for (int k = 0; k < len; ++k) {
      const int64_t q = q0 + static_cast<int64_t>(key[k] & 0xffffffffu);
      const int64_t b = static_cast<int64_t>(key[k] >> 32);
      const int32_t m = filled[b]++;  // slot inside bucket b of mid
      rank1[q] = static_cast<uint16_t>(k);
      mid_pos[q0 + k] = static_cast<int32_t>(b * kBucket + m);
      slot2[dst[q]] = static_cast<uint16_t>(m);
    }
  }
  auto* pl = new segscan_permute_plan();
  pl->device = device;
  pl->n = n;
  const size_t b16 = a256(static_cast<size_t>(n) * 2), b32 = a256(static_cast<size_t>(n) * 4);
  e = hipMalloc(&pl->mem, 2 * b16 + 3 * b32);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "permute plan: allocation");
  }
  char* base = static_cast<char*>(pl->mem);
  pl->rank1 = reinterpret_cast<uint16_t*>(base);
  pl->slot2 = reinterpret_cast<uint16_t*>(base + b16);
  pl->mid_pos = reinterpret_cast<int32_t*>(base + 2 * b16);
  pl->mid = reinterpret_cast<float*>(base + 2 * b16 + b32);
  pl->aux = reinterpret_cast<float*>(base + 2 * b16 + 2 * b32);
  e = hipMemcpyAsync(pl->rank1, rank1.data(), static_cast<size_t>(n) * 2, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(pl->slot2, slot2.data(), static_cast<size_t>(n) * 2, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pl->mid_pos, mid_pos.data(), static_cast<size_t>(n) * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host tables go out of scope
  if (e != hipSuccess) {
    (void)hipFree(pl->mem);
    delete pl;
    return hip_fail(e, "permute plan: upload");
  }
  *plan = pl;
  segscan::set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_permute_run(segscan_permute_plan_t plan, const float* in, float* out, void* stream) {
  if (!plan) return einval("plan is NULL");
  if (!in || !out) return einval("in or out is NULL");
  if (in == out) return einval("in and out must not alias");
  int device = 0;
  hipError_t de = hipGetDevice(&device);
  if (de != hipSuccess) return hip_fail(de, "permute run: current device");
  if (device != plan->device) return einval("the current device is not the plan's");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = segscan::permute_pass1(plan, in, s);
  if (e == hipSuccess) e = segscan::permute_pass2(plan, out, s);
  if (e != hipSuccess) return hip_fail(e, "permute launch");
  segscan::set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_permute_plan_destroy(segscan_permute_plan_t plan) {
  if (!plan) return SEGSCAN_OK;
  hipError_t e = hipSuccess;
  if (plan->mem) e = hipFree(plan->mem);
  delete plan;
  if (e != hipSuccess) return hip_fail(e, "permute plan: free");
  return SEGSCAN_OK;
}

}  // extern "C"
