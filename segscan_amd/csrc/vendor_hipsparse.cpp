// hipSPARSE SpMM comparator: the MI355X counterpart of the reference's
// csrspmm_cusparse (test-cg/include/spmm/spmm.cuh:640-694), used only to time
// the vendor library on the same inputs ("best vs vendor", README §5). It is
// built into its own library (libsegscan_vendor.so) so that the product
// library does not depend on hipSPARSE, and it only calls the public API.
#include <hip/hip_runtime.h>
#include <hipsparse/hipsparse.h>

#include <cstdint>
#include <cstdio>

namespace {

struct Plan {
  hipsparseHandle_t handle = nullptr;
  hipsparseSpMatDescr_t A = nullptr;
  hipsparseDnMatDescr_t B = nullptr;
  hipsparseDnMatDescr_t C = nullptr;
  hipsparseSpMMAlg_t alg = HIPSPARSE_SPMM_ALG_DEFAULT;
  void* buffer = nullptr;
  float alpha = 1.0f;
  float beta = 0.0f;
};

void destroy(Plan* p) {
  if (!p) return;
  if (p->buffer) (void)hipFree(p->buffer);
  if (p->A) hipsparseDestroySpMat(p->A);
  if (p->B) hipsparseDestroyDnMat(p->B);
  if (p->C) hipsparseDestroyDnMat(p->C);
  if (p->handle) hipsparseDestroy(p->handle);
  delete p;
}

}  // namespace

extern "C" {

// b_dtype: 0 fp32 only (the vendor SpMM has no fp32-A / 16-bit-B / fp32-C form; it
// throws for that combination instead of returning a status, so it is refused here,
// returning -1000). alg: 0 default, 1/2/3 CSR_ALG1/2/3. Returns 0 on success, else a
// hipSPARSE / HIP error code (negative for HIP errors). Descriptors, buffer size query, buffer allocation
// and preprocessing happen here, outside any timed region.
int vendor_spmm_create(int64_t M, int64_t K, int64_t nnz, const int32_t* rowptr, const int32_t* colind,
                       const float* vals, int64_t N, const void* B, int b_dtype, float* C, int alg, void* stream,
                       void** plan_out) {
  *plan_out = nullptr;
  if (b_dtype != 0) return -1000;
  Plan* p = new Plan();
  const hipsparseSpMMAlg_t algs[4] = {HIPSPARSE_SPMM_ALG_DEFAULT, HIPSPARSE_SPMM_CSR_ALG1, HIPSPARSE_SPMM_CSR_ALG2,
                                      HIPSPARSE_SPMM_CSR_ALG3};
  p->alg = algs[(alg >= 0 && alg < 4) ? alg : 0];
  const hipDataType bt = HIP_R_32F;
  hipsparseStatus_t st;
#define VCHECK(x)                  \
  do {                             \
    st = (x);                      \
    if (st != HIPSPARSE_STATUS_SUCCESS) { \
      destroy(p);                  \
      return static_cast<int>(st); \
    }                              \
  } while (0)
  VCHECK(hipsparseCreate(&p->handle));
  VCHECK(hipsparseSetStream(p->handle, static_cast<hipStream_t>(stream)));
  VCHECK(hipsparseCreateCsr(&p->A, M, K, nnz, const_cast<int32_t*>(rowptr), const_cast<int32_t*>(colind),
                            const_cast<float*>(vals), HIPSPARSE_INDEX_32I, HIPSPARSE_INDEX_32I,
                            HIPSPARSE_INDEX_BASE_ZERO, HIP_R_32F));
  VCHECK(hipsparseCreateDnMat(&p->B, K, N, N, const_cast<void*>(B), bt, HIPSPARSE_ORDER_ROW));
  VCHECK(hipsparseCreateDnMat(&p->C, M, N, N, C, HIP_R_32F, HIPSPARSE_ORDER_ROW));
  size_t bytes = 0;
  VCHECK(hipsparseSpMM_bufferSize(p->handle, HIPSPARSE_OPERATION_NON_TRANSPOSE, HIPSPARSE_OPERATION_NON_TRANSPOSE,
                                  &p->alpha, p->A, p->B, &p->beta, p->C, HIP_R_32F, p->alg, &bytes));
  if (bytes) {
    hipError_t e = hipMalloc(&p->buffer, bytes);
    if (e != hipSuccess) {
      destroy(p);
      return -static_cast<int>(e);
    }
  }
  VCHECK(hipsparseSpMM_preprocess(p->handle, HIPSPARSE_OPERATION_NON_TRANSPOSE, HIPSPARSE_OPERATION_NON_TRANSPOSE,
                                  &p->alpha, p->A, p->B, &p->beta, p->C, HIP_R_32F, p->alg, p->buffer));
#undef VCHECK
  *plan_out = p;
  return 0;
}

int vendor_spmm_run(void* plan) {
  Plan* p = static_cast<Plan*>(plan);
  return static_cast<int>(hipsparseSpMM(p->handle, HIPSPARSE_OPERATION_NON_TRANSPOSE,
                                        HIPSPARSE_OPERATION_NON_TRANSPOSE, &p->alpha, p->A, p->B, &p->beta, p->C,
                                        HIP_R_32F, p->alg, p->buffer));
}

void vendor_spmm_destroy(void* plan) { destroy(static_cast<Plan*>(plan)); }

}  // extern "C"
