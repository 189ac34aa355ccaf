// CPU oracle for the SpMM / segment_coo path — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this
// library, and only as the checker (or as the timed CPU baseline); the product
// path (segscan_amd, libsegscan_hip.so) never calls it.
//
// Restated from the reference (zhang677/segScan); every function names the
// lines it follows. The reference path itself is CUDA-only (check.cuh includes
// cuda_runtime.h and cusparse.h; the segment_coo data loader needs Boost), so
// it cannot be compiled here and oracle/_ref stays empty.
//
// How each function is pinned (tests/test_oracle.py):
//   oracle_row_of_nnz        bisect.bisect_right brute force over random rowptrs,
//                            including empty rows and e >= nnz
//   oracle_spmm_f32          dense float64 GEMM on tiny inputs (numpy), the
//                            all-ones known answer C[i][j] = deg(i), integer-valued
//                            inputs (exact in any order), linearity in B, and the
//                            committed fixtures in tests/golden/
//   oracle_spmm_f64 / _abs   same checks at float64; |A|.|B| against numpy
//   oracle_segment_coo_f32   np.add.at brute force, src = 1 -> counts, src[e] = e
//   oracle_transpose         numpy stable argsort by column, CSC round trip
//   oracle_generate_index    sortedness, run-length clamps, total count, last-run
//                            remainder, early stop (properties of dataloader.h:33-67)
//   oracle_check_result      tolerance edge cases (1e-2 relative)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

extern "C" {

int oracle_version(void) { return 1; }

// test-cg/include/util/check.cuh:88-102 (binary_search_segment_number):
// first j in [1, M] with rowptr[j] > e, minus one; M - 1 when there is none.
int32_t oracle_row_of_nnz(const int32_t* rowptr, int32_t M, int64_t e) {
  int32_t lo = 1, hi = M, mid;
  while (lo < hi) {
    mid = (lo + hi) >> 1;
    if (rowptr[mid] <= e) lo = mid + 1;
    else hi = mid;
  }
  return hi - 1;
}

// test-cg/include/util/check.cuh:139-158 (util::spmm_reference_host): loop order
// i -> p -> j, fp32 "C += val * B" per element, accumulating into C (callers
// zero it). vals == NULL means every value is 1 (check.cuh:104-111).
void oracle_spmm_f32(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const float* vals,
                     const float* B, int64_t ldb, float* C, int64_t ldc) {
  for (int64_t i = 0; i < M; i++) {
    const int32_t begin = rowptr[i];
    const int32_t end = rowptr[i + 1];
    for (int32_t p = begin; p < end; p++) {
      const int64_t k = colind[p];
      const float val = vals ? vals[p] : 1.0f;
      for (int64_t j = 0; j < N; j++) {
        C[i * ldc + j] += val * B[k * ldb + j];
      }
    }
  }
}

// Same loops in float64 (error reference). B is given in fp32 (the values the
// GPU reads after widening bf16/f16).
void oracle_spmm_f64(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const float* vals,
                     const float* B, int64_t ldb, double* C, int64_t ldc) {
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t k = colind[p];
      const double val = vals ? static_cast<double>(vals[p]) : 1.0;
      for (int64_t j = 0; j < N; j++) C[i * ldc + j] += val * static_cast<double>(B[k * ldb + j]);
    }
  }
}

// Same loops with float64 operands (finite-difference gradient checks perturb the
// inputs below fp32 resolution).
void oracle_spmm_f64d(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const double* vals,
                      const double* B, int64_t ldb, double* C, int64_t ldc) {
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t k = colind[p];
      const double val = vals ? vals[p] : 1.0;
      for (int64_t j = 0; j < N; j++) C[i * ldc + j] += val * B[k * ldb + j];
    }
  }
}

void oracle_sddmm_f64d(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const double* X,
                       int64_t ld_x, const double* Y, int64_t ld_y, double* out) {
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t k = colind[p];
      double s = 0.0;
      for (int64_t j = 0; j < N; j++) s += X[i * ld_x + j] * Y[k * ld_y + j];
      out[p] = s;
    }
  }
}

// D[i][j] = sum_p |a_ip| |b_{k(p) j}|: the denominator of the normalized error
// max |C - C64| / (|A| |B|).
void oracle_spmm_abs(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const float* vals,
                     const float* B, int64_t ldb, double* D, int64_t ldd) {
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t k = colind[p];
      const double val = vals ? std::fabs(static_cast<double>(vals[p])) : 1.0;
      for (int64_t j = 0; j < N; j++) D[i * ldd + j] += val * std::fabs(static_cast<double>(B[k * ldb + j]));
    }
  }
}

// new-cg/include/utils/check.cuh:44-52 (segment_coo_sequencial): for every
// element i, for every column j, dst[index[i] * N + j] += src[i * N + j].
void oracle_segment_coo_f32(const float* src, int64_t ld_src, const int32_t* index, int64_t nnz, int64_t N,
                            float* dst, int64_t ld_dst) {
  for (int64_t i = 0; i < nnz; i++) {
    for (int64_t j = 0; j < N; j++) {
      dst[static_cast<int64_t>(index[i]) * ld_dst + j] += src[i * ld_src + j];
    }
  }
}

void oracle_segment_coo_f64(const float* src, int64_t ld_src, const int32_t* index, int64_t nnz, int64_t N,
                            double* dst, int64_t ld_dst, double* absd) {
  for (int64_t i = 0; i < nnz; i++) {
    for (int64_t j = 0; j < N; j++) {
      const double v = static_cast<double>(src[i * ld_src + j]);
      dst[static_cast<int64_t>(index[i]) * ld_dst + j] += v;
      if (absd) absd[static_cast<int64_t>(index[i]) * ld_dst + j] += std::fabs(v);
    }
  }
}

// test-cg/include/util/check.cuh:116-136 (util::check_result): fails when any
// |c - c_ref| > 1e-2 * |c_ref|; returns the number of failing rows (the
// reference's break leaves the inner loop after the first failure of a row).
int64_t oracle_check_result(int64_t M, int64_t N, const float* C, const float* C_ref) {
  int64_t bad_rows = 0;
  for (int64_t i = 0; i < M; i++) {
    for (int64_t j = 0; j < N; j++) {
      const float c = C[i * N + j];
      const float c_ref = C_ref[i * N + j];
      if (std::fabs(c - c_ref) > 1e-2 * std::fabs(c_ref)) {
        bad_rows++;
        break;
      }
    }
  }
  return bad_rows;
}

// test-cg/include/dataloader/dataloader.hpp:123-144 (transpose) and :109-121
// (compressedRow): counting sort of the nnz by column in CSR order, giving for
// every CSC entry q its row and the CSR position it came from.
void oracle_transpose(int64_t M, int64_t K, const int32_t* rowptr, const int32_t* colind, int32_t* colptr,
                      int32_t* rowind, int32_t* perm) {
  std::vector<int64_t> hist(static_cast<size_t>(K) + 1, 0);
  const int64_t nnz = rowptr[M];
  for (int64_t t = 0; t < nnz; t++) hist[colind[t]]++;
  colptr[0] = 0;
  for (int64_t c = 0; c < K; c++) colptr[c + 1] = static_cast<int32_t>(colptr[c] + hist[c]);
  std::vector<int64_t> cursor(colptr, colptr + K);
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t q = cursor[colind[p]]++;
      rowind[q] = static_cast<int32_t>(i);
      perm[q] = p;
    }
  }
}

// out[p] = sum_j X[row(p)][j] * Y[colind[p]][j] in float64 (SDDMM reference).
void oracle_sddmm_f64(int64_t M, int64_t N, const int32_t* rowptr, const int32_t* colind, const float* X,
                      int64_t ld_x, const float* Y, int64_t ld_y, double* out, double* absd) {
  for (int64_t i = 0; i < M; i++) {
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int64_t k = colind[p];
      double s = 0.0, a = 0.0;
      for (int64_t j = 0; j < N; j++) {
        const double t = static_cast<double>(X[i * ld_x + j]) * static_cast<double>(Y[k * ld_y + j]);
        s += t;
        a += std::fabs(t);
      }
      out[p] = s;
      if (absd) absd[p] = a;
    }
  }
}

// new-cg/include/dataloader/dataloader.h:33-67 (generateIndex): segment i gets
// mi = clamp(round(Normal(avg, avg * cv)), min_seg, max_seg) copies of i, the
// last segment of the range takes the remainder, generation stops once the
// total is reached. The reference writes past the buffer when a segment
// overshoots (dataloader.h:56-58); here the run is cut at `total`. Uses the
// reference's engine and distribution types, seeded with `seed`
// (std::default_random_engine's default seed is 1). Returns dst_len, the
// number of non-empty segments.
int32_t oracle_generate_index(int32_t range, int32_t min_seg, int32_t max_seg, int32_t total, double cv,
                              uint32_t seed, int32_t* result) {
  const double avg = static_cast<double>(total) / range;
  std::default_random_engine generator(seed);
  std::normal_distribution<double> distribution(avg, avg * cv);
  int64_t current = 0;
  int32_t dst_len = 0;
  for (int32_t i = 0; i < range; ++i) {
    int64_t mi = static_cast<int64_t>(std::round(distribution(generator)));
    mi = std::max<int64_t>(mi, min_seg);
    mi = std::min<int64_t>(mi, max_seg);
    if (i == range - 1) mi = total - current;
    for (int64_t j = 0; j < mi && current + j < total; ++j) result[current + j] = i;
    current += mi;
    if (mi > 0) dst_len += 1;
    if (current >= total) break;
  }
  return dst_len;
}

}  // extern "C"
