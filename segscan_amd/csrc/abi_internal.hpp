// Library-internal entry points shared between the C ABI's translation units
// (not exported from the shared library's public header).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/segscan.h"

namespace segscan {

// Body of segscan_spmm_csr. With plan_bytes non-NULL it validates and plans only (no
// HIP call) and stores the workspace the launch needs (0 if it launches nothing).
segscan_status_t spmm_csr_impl(const segscan_csr_t* A, int64_t N, const void* B, int64_t ldb,
                               segscan_dtype_t b_dtype, float* C, int64_t ldc, segscan_spmm_variant_t variant,
                               const segscan_knobs_t* knobs, uint32_t flags, void* workspace, size_t workspace_bytes,
                               void* stream, size_t* plan_bytes);

// Sets the calling thread's segscan_last_error() text and returns st.
segscan_status_t set_error(segscan_status_t st, const char* msg);

}  // namespace segscan
