// segscan.hpp — C++ host interface with the reference's launcher names and
// template knob lists, implemented on the C ABI (segscan.h) and the gfx950 HIP
// kernels of libsegscan_hip.so.
//
// Drop-in for the reference's headers:
//   test-cg/include/spmm/spmm.cuh      csrspmm_parreduce_nnzbalance (:167), csrspmm_parreduce_rowbalance (:310),
//                                      csrspmm_rowcaching_nnzbalance (:522), csrspmm_seqreduce_rowbalance (:621)
//   test-cg/include/spmm/spmm_cg.cuh   csrspmm_parreduce_nnzbalance_cg (:176), csrspmm_parreduce_rowbalance_cg (:333),
//                                      csrspmm_rowcaching_nnzbalance_cg (:553), csrspmm_seqreduce_rowbalance_cg (:664)
//   new-cg/include/segscan/segscan.cuh segment_coo (:150), segment_coo_new (:327), segment_coo_sr (:817),
//                                      segment_coo_noshmem_sr (:868), segment_coo_noshmem_lessatom_sr (:918),
//                                      segment_coo_sorted_sr (:974)
//   test-cg/include/dataloader/dataloader.hpp SpMatCsrDescr_t (:146-168), read_mtx_file (:22),
//                                      compressedRow (:109), transpose (:123), DataLoader (:170),
//                                      SingleDataLoader (:185)
//
// Semantics kept: atomic variants add into C / dst (callers zero them first, as
// the reference harness does), store variants overwrite; the launchers are
// asynchronous on the default stream, except that the segment_coo family reads
// the last index value back to size dst (the reference's segment_coo launchers
// synchronise too, segscan.cuh:179); an optional trailing dst_len argument (the
// row count of dst) skips that read, for hipGraph capture. Differences: failures throw
// std::runtime_error (the reference's SpMM launchers check nothing; its
// segment_coo launchers exit()), values are read as float in every variant, and
// scratch space (carries, ebpr tile rows) is cached in the descriptor.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../segscan.h"

namespace segscan {

inline void check_status(segscan_status_t st, const char* what) {
  if (st != SEGSCAN_OK) throw std::runtime_error(std::string(what) + ": " + segscan_last_error());
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device array with a host mirror, like util::RamArray (test-cg/include/util/ramArray.cuh:21-98).
template <class T>
struct DeviceArray {
  std::vector<T> h;
  T* d = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t len) { create(len); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : h(std::move(o.h)), d(o.d), n(o.n) { o.d = nullptr; o.n = 0; }
  ~DeviceArray() { if (d) (void)hipFree(d); }
  void create(size_t len) {
    h.assign(len, T());
    n = len;
    if (d) check_hip(hipFree(d), "hipFree");
    d = nullptr;
    if (len) check_hip(hipMalloc(&d, len * sizeof(T)), "hipMalloc");
  }
  void create(size_t len, const std::vector<T>& src) { create(len); std::copy(src.begin(), src.begin() + len, h.begin()); }
  void fill_default_one() { std::fill(h.begin(), h.end(), T(1)); }
  void fill_zero_h() { std::fill(h.begin(), h.end(), T(0)); }
  void upload() { if (n) check_hip(hipMemcpy(d, h.data(), n * sizeof(T), hipMemcpyHostToDevice), "upload"); }
  void download() { if (n) check_hip(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "download"); }
  void reset() { fill_zero_h(); if (n) check_hip(hipMemset(d, 0, n * sizeof(T)), "reset"); }
};

// CSR descriptor; values default to 1 as in the reference (dataloader.hpp:154-155).
template <class Index, class DType>
struct SpMatCsrDescr_t {
  static_assert(sizeof(Index) == 4, "int32 indices");
  static_assert(sizeof(DType) == 4, "fp32 values");
  SpMatCsrDescr_t(int ncol_, std::vector<Index>& indptr, std::vector<Index>& indices) {
    nrow = static_cast<int>(indptr.size()) - 1;
    ncol = ncol_;
    nnz = static_cast<int>(indices.size());
    sp_csrptr.create(nrow + 1, indptr);
    sp_csrind.create(nnz, indices);
    sp_data.create(nnz);
    sp_data.fill_default_one();
    for (int r = 0; r < nrow; ++r) max_row = std::max<int>(max_row, indptr[r + 1] - indptr[r]);
  }
  void upload() {
    sp_csrptr.upload();
    sp_csrind.upload();
    sp_data.upload();
  }
  segscan_csr_t desc() const {
    segscan_csr_t d{};
    d.nrow = nrow;
    d.ncol = ncol;
    d.nnz = nnz;
    d.rowptr = reinterpret_cast<const int32_t*>(sp_csrptr.d);
    d.colind = reinterpret_cast<const int32_t*>(sp_csrind.d);
    d.vals = reinterpret_cast<const float*>(sp_data.d);
    d.max_row_nnz = max_row;
    return d;
  }
  // Grow-only scratch for the carry and tile-row buffers.
  void* workspace(size_t bytes) {
    if (bytes > ws_bytes) {
      if (ws) check_hip(hipFree(ws), "hipFree");
      ws = nullptr;
      check_hip(hipMalloc(&ws, bytes), "hipMalloc workspace");
      ws_bytes = bytes;
    }
    return ws;
  }
  SpMatCsrDescr_t(SpMatCsrDescr_t&& o) noexcept
      : nrow(o.nrow), ncol(o.ncol), nnz(o.nnz), max_row(o.max_row), sp_csrptr(std::move(o.sp_csrptr)),
        sp_csrind(std::move(o.sp_csrind)), sp_data(std::move(o.sp_data)), ws(o.ws), ws_bytes(o.ws_bytes) {
    o.ws = nullptr;
    o.ws_bytes = 0;
  }
  SpMatCsrDescr_t(const SpMatCsrDescr_t&) = delete;
  SpMatCsrDescr_t& operator=(const SpMatCsrDescr_t&) = delete;
  ~SpMatCsrDescr_t() { if (ws) (void)hipFree(ws); }
  int nrow = 0;
  int ncol = 0;
  int nnz = 0;
  int max_row = 0;  // longest row; guides SEGSCAN_SPMM_AUTO
  DeviceArray<Index> sp_csrptr;
  DeviceArray<Index> sp_csrind;
  DeviceArray<DType> sp_data;
  void* ws = nullptr;
  size_t ws_bytes = 0;
};

namespace detail {

inline segscan_knobs_t knobs(int gf, int tf, int tpb, int bn, int bd, int cf = 0, int tnz = 0, int policy = 0) {
  segscan_knobs_t k{};
  k.group_factor = gf == 0 ? -1 : gf;
  k.tile_factor = tf == 0 ? -1 : tf;
  k.thread_per_block = tpb;
  k.block_numer = bn;
  k.block_denom = bd;
  k.coarsen_factor = cf;
  k.thread_nz = tnz;
  k.write_policy = policy;
  return k;
}

template <class Index, class DType>
void spmm(SpMatCsrDescr_t<Index, DType>& A, int N, const DType* B, DType* C, segscan_spmm_variant_t v,
          const segscan_knobs_t* k, uint32_t flags) {
  segscan_csr_t d = A.desc();
  size_t bytes = 0;
  check_status(segscan_spmm_workspace_size(&d, N, SEGSCAN_F32, v, k, &bytes), "spmm workspace");
  void* ws = bytes ? A.workspace(bytes) : nullptr;
  check_status(segscan_spmm_csr(&d, N, B, N, SEGSCAN_F32, reinterpret_cast<float*>(C), N, v, k, flags, ws, bytes,
                                nullptr),
               "spmm");
}

// Cached workspace for the segment_coo launchers (they take no descriptor).
inline void* seg_workspace(size_t bytes) {
  static thread_local void* ws = nullptr;
  static thread_local size_t cap = 0;
  if (bytes > cap) {
    if (ws) check_hip(hipFree(ws), "hipFree");
    ws = nullptr;
    check_hip(hipMalloc(&ws, bytes), "hipMalloc workspace");
    cap = bytes;
  }
  return ws;
}

// dst_len < 0: read the last index back (a blocking copy, as the reference's launchers take
// no output length); dst_len >= 0: the caller's row count of dst, no host round trip (the
// form to use under hipGraph capture).
template <class ValueType, class IndexType>
void segment(const ValueType* src, const IndexType* index, int nnz, int N, ValueType* dst, segscan_seg_variant_t v,
             const segscan_knobs_t* k, int dst_rows) {
  static_assert(sizeof(ValueType) == 4 && sizeof(IndexType) == 4, "float values, int32 index");
  if (nnz <= 0 || N <= 0) return;
  int64_t dst_len = dst_rows;
  if (dst_rows < 0) {
    IndexType last = 0;
    check_hip(hipMemcpy(&last, index + nnz - 1, sizeof(IndexType), hipMemcpyDeviceToHost), "read last index");
    dst_len = static_cast<int64_t>(last) + 1;
  }
  size_t bytes = 0;
  check_status(segscan_segment_coo_workspace_size(nnz, N, dst_len, SEGSCAN_F32, v, k, &bytes), "segment workspace");
  void* ws = bytes ? seg_workspace(bytes) : nullptr;
  check_status(segscan_segment_coo(src, N, SEGSCAN_F32, reinterpret_cast<const int32_t*>(index), nnz, N,
                                   reinterpret_cast<float*>(dst), N, dst_len, v, k, SEGSCAN_ACCUMULATE, ws, bytes,
                                   nullptr),
               "segment_coo");
}

}  // namespace detail

// ---- plain SpMM (spmm.cuh) ----

template <typename Index, typename DType>
void csrspmm_parreduce_nnzbalance(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_PLAIN_EBPR, nullptr, SEGSCAN_ACCUMULATE);
}
template <typename Index, typename DType>
void csrspmm_parreduce_rowbalance(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_PLAIN_RBPR, nullptr, 0);
}
template <typename Index, typename DType>
void csrspmm_rowcaching_nnzbalance(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_PLAIN_EBSR, nullptr, SEGSCAN_ACCUMULATE);
}
template <typename Index, typename DType>
void csrspmm_seqreduce_rowbalance(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_PLAIN_RBSR, nullptr, 0);
}

// ---- tunable SpMM (spmm_cg.cuh), same template parameter lists ----

template <typename Index, typename DType, int group_factor, int thread_per_block, int tile_factor, int block_numer,
          int block_denom>
void csrspmm_parreduce_nnzbalance_cg(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  const segscan_knobs_t k = detail::knobs(group_factor, tile_factor, thread_per_block, block_numer, block_denom);
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_EBPR, &k, SEGSCAN_ACCUMULATE);
}
template <typename Index, typename DType, int group_factor, int thread_per_block, int tile_factor, int block_numer,
          int block_denom>
void csrspmm_parreduce_rowbalance_cg(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  const segscan_knobs_t k = detail::knobs(group_factor, tile_factor, thread_per_block, block_numer, block_denom);
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_RBPR, &k, SEGSCAN_ACCUMULATE);
}
template <typename Index, typename DType, int group_factor, int thread_per_block, int CoarsenFactor, int ThreadNz,
          int block_numer, int block_denom>
void csrspmm_rowcaching_nnzbalance_cg(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  const segscan_knobs_t k = detail::knobs(group_factor, 0, thread_per_block, block_numer, block_denom, CoarsenFactor,
                                          ThreadNz, SEGSCAN_WRITE_ATOMIC);
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_EBSR, &k, SEGSCAN_ACCUMULATE);
}
template <typename Index, typename DType, int coarsen_factor, int thread_per_block, int tile_factor, int block_numer,
          int block_denom>
void csrspmm_seqreduce_rowbalance_cg(SpMatCsrDescr_t<Index, DType>& A, const int N, const DType* B, DType* C) {
  const segscan_knobs_t k =
      detail::knobs(0, tile_factor, thread_per_block, block_numer, block_denom, coarsen_factor);
  detail::spmm(A, N, B, C, SEGSCAN_SPMM_RBSR, &k, 0);
}

// ---- segment_coo (segscan.cuh), same template parameter lists ----

template <typename ValueType, typename IndexType, int group_factor, int thread_per_block, int tile_factor,
          int block_numer, int block_denom>
void segment_coo(const ValueType* src, const IndexType* index, const int nnz, const int N, ValueType* dst, const int dst_len = -1) {
  const segscan_knobs_t k = detail::knobs(group_factor, tile_factor, thread_per_block, block_numer, block_denom);
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_PR, &k, dst_len);
}
template <typename ValueType, typename IndexType, int group_factor, int thread_per_block, int tile_factor,
          int blockDimY>
void segment_coo_new(const ValueType* src, const IndexType* index, const int nnz, const int N, ValueType* dst, const int dst_len = -1) {
  segscan_knobs_t k = detail::knobs(group_factor, tile_factor, thread_per_block, 1, 1);
  k.block_dim_y = blockDimY;
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_PR_NEW, &k, dst_len);
}
template <typename ValueType, typename IndexType, int group_factor, int thread_per_block, int CoarsenFactor,
          int ThreadNz, int block_numer, int block_denom>
void segment_coo_sr(const ValueType* src, const IndexType* index, const int nnz, const int N, ValueType* dst, const int dst_len = -1) {
  const segscan_knobs_t k = detail::knobs(group_factor, 0, thread_per_block, block_numer, block_denom, CoarsenFactor,
                                          ThreadNz, SEGSCAN_WRITE_ATOMIC);
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_SR, &k, dst_len);
}
template <typename ValueType, typename IndexType, int group_factor, int thread_per_block, int CoarsenFactor,
          int ThreadNz, int block_numer, int block_denom>
void segment_coo_noshmem_sr(const ValueType* src, const IndexType* index, const int nnz, const int N,
                            ValueType* dst, const int dst_len = -1) {
  const segscan_knobs_t k = detail::knobs(group_factor, 0, thread_per_block, block_numer, block_denom, CoarsenFactor,
                                          ThreadNz, SEGSCAN_WRITE_ATOMIC);
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_NOSHMEM_SR, &k, dst_len);
}
template <typename ValueType, typename IndexType, int group_factor, int thread_per_block, int CoarsenFactor,
          int ThreadNz, int block_numer, int block_denom>
void segment_coo_noshmem_lessatom_sr(const ValueType* src, const IndexType* index, const int nnz, const int N,
                                     ValueType* dst, const int dst_len = -1) {
  const segscan_knobs_t k = detail::knobs(group_factor, 0, thread_per_block, block_numer, block_denom, CoarsenFactor,
                                          ThreadNz, SEGSCAN_WRITE_LESSATOM);
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_LESSATOM_SR, &k, dst_len);
}
template <typename ValueType, typename IndexType, int CoarsenFactor, int ThreadNz, int DtileSize, int DtileNum,
          int BlockDimY>
void segment_coo_sorted_sr(const ValueType* src, const IndexType* index, const int nnz, const int N, ValueType* dst, const int dst_len = -1) {
  segscan_knobs_t k = detail::knobs(0, 0, 0, 1, 1, CoarsenFactor, ThreadNz, SEGSCAN_WRITE_LESSATOM);
  k.group_factor = 0;
  k.tile_factor = 0;
  k.thread_per_block = 0;
  k.dtile_size = DtileSize;
  k.dtile_num = DtileNum;
  k.block_dim_y = BlockDimY;
  detail::segment(src, index, nnz, N, dst, SEGSCAN_SEG_SORTED_SR, &k, dst_len);
}

// ---- Matrix Market input (test-cg/include/dataloader/dataloader.hpp:22-192) ----------------
// flags: SEGSCAN_MTX_REFERENCE_QUIRKS reproduces read_mtx_file's mirroring and truncation;
// the default reads the full matrix. Values are dropped, as in the reference.

// read_mtx_file (:22-107): same outputs, coo_rowind_buffer = the row of each entry.
inline void read_mtx_file(const char* filename, int& nrow, int& ncol, int& nnz, std::vector<int>& csr_indptr_buffer,
                          std::vector<int>& csr_indices_buffer, std::vector<int>& coo_rowind_buffer,
                          uint32_t flags = 0) {
  segscan_host_csr_t m;
  if (segscan_mtx_read(filename, flags & SEGSCAN_MTX_REFERENCE_QUIRKS, &m) != SEGSCAN_OK)
    throw std::runtime_error(std::string("read_mtx_file: ") + segscan_mtx_last_error());
  nrow = static_cast<int>(m.nrow);
  ncol = static_cast<int>(m.ncol);
  nnz = static_cast<int>(m.nnz);
  csr_indptr_buffer.assign(m.rowptr, m.rowptr + m.nrow + 1);
  csr_indices_buffer.assign(m.colind, m.colind + m.nnz);
  coo_rowind_buffer.resize(static_cast<size_t>(m.nnz));
  for (int r = 0; r < nrow; ++r)
    std::fill(coo_rowind_buffer.begin() + m.rowptr[r], coo_rowind_buffer.begin() + m.rowptr[r + 1], r);
  segscan_host_csr_free(&m);
}

// compressedRow (:109-121): row pointer of a row-sorted row-index list.
template <class Index>
void compressedRow(Index nrow, const std::vector<Index>& rowind, std::vector<Index>& rowptr) {
  rowptr.assign(static_cast<size_t>(nrow) + 1, 0);
  for (Index r : rowind) ++rowptr[static_cast<size_t>(r) + 1];
  for (Index r = 0; r < nrow; ++r) rowptr[r + 1] += rowptr[r];
}

// transpose (:123-144): stable counting sort of the entries by column; row_t gets the
// column ids (sorted), col_t the row ids (ascending within a column).
template <class Index>
void transpose(Index ncol, const std::vector<Index>& row, const std::vector<Index>& col, std::vector<Index>& row_t,
               std::vector<Index>& col_t) {
  std::vector<Index> next(static_cast<size_t>(ncol) + 1, 0);
  for (Index c : col) ++next[static_cast<size_t>(c) + 1];
  for (Index c = 0; c < ncol; ++c) next[c + 1] += next[c];
  row_t.resize(col.size());
  col_t.resize(col.size());
  for (size_t e = 0; e < col.size(); ++e) {
    const Index q = next[col[e]]++;
    row_t[q] = col[e];
    col_t[q] = row[e];
  }
}

// SingleDataLoader (:185-192).
template <class Index, class DType>
SpMatCsrDescr_t<Index, DType> SingleDataLoader(const char* filename, uint32_t flags = 0) {
  int nrow, ncol, nnz;
  std::vector<Index> indptr, indices, coorow;
  read_mtx_file(filename, nrow, ncol, nnz, indptr, indices, coorow, flags);
  return SpMatCsrDescr_t<Index, DType>(ncol, indptr, indices);
}

// DataLoader (:170-183): the matrix and its transpose, both built on the host.
template <class Index, class DType>
std::tuple<SpMatCsrDescr_t<Index, DType>, SpMatCsrDescr_t<Index, DType>> DataLoader(const char* filename,
                                                                                   uint32_t flags = 0) {
  int nrow, ncol, nnz;
  std::vector<Index> indptr, indices, coorow, t_indptr, t_coorow, t_indices;
  read_mtx_file(filename, nrow, ncol, nnz, indptr, indices, coorow, flags);
  transpose<Index>(ncol, coorow, indices, t_coorow, t_indices);
  compressedRow<Index>(ncol, t_coorow, t_indptr);
  return std::make_tuple(SpMatCsrDescr_t<Index, DType>(ncol, indptr, indices),
                         SpMatCsrDescr_t<Index, DType>(nrow, t_indptr, t_indices));
}

}  // namespace segscan
