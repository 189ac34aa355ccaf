// Library-internal entry points shared between the C ABI's translation units
// (not exported from the shared library's public header).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "../../include/segscan.h"

namespace segscan {

// Body of segscan_spmm_csr. With plan_bytes non-NULL it validates and plans only (no
// HIP call) and stores the workspace the launch needs (0 if it launches nothing).
// Values an ebsr SpMM reads from a permutation plan's staged buckets instead of A.vals.
// The planning run (plan_bytes non-NULL) sets `fused` when the resolved launch can read
// them: ebsr with groups of 8 to 64 lanes, each workgroup round covering a part of one
// bucket (at least 2048 entries). A launch with a
// StagedValues whose `fused` is false is refused.
struct StagedValues {
  const float* mid = nullptr;       // bucket-major staging buffer, kPermBucket values a bucket
  const uint16_t* slot = nullptr;   // [nnz] entry e's slot inside its bucket
  bool fused = false;
};

// Body of segscan_spmm_csr. With plan_bytes non-NULL it validates and plans only (no
// HIP call) and stores the workspace the launch needs (0 if it launches nothing).
segscan_status_t spmm_csr_impl(const segscan_csr_t* A, int64_t N, const void* B, int64_t ldb,
                               segscan_dtype_t b_dtype, float* C, int64_t ldc, segscan_spmm_variant_t variant,
                               const segscan_knobs_t* knobs, uint32_t flags, void* workspace, size_t workspace_bytes,
                               void* stream, size_t* plan_bytes, StagedValues* staged = nullptr);

// A permutation plan's device buffers and passes (permute.hip).
struct PermuteBuffers {
  int device = 0;                   // the device whose memory holds the plan
  int64_t n = 0;
  const float* mid = nullptr;       // pass 1's output, bucket-major
  const uint16_t* slot2 = nullptr;  // [n] output position's slot inside its bucket
  float* aux = nullptr;             // [n] an output buffer the plan owns
};
PermuteBuffers permute_buffers(const segscan_permute_plan* plan);
hipError_t permute_pass1(const segscan_permute_plan* plan, const float* in, hipStream_t s);
hipError_t permute_pass2(const segscan_permute_plan* plan, float* out, hipStream_t s);

// Sets the calling thread's segscan_last_error() text and returns st.
segscan_status_t set_error(segscan_status_t st, const char* msg);

}  // namespace segscan
