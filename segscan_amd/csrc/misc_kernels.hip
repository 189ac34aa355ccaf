// Support kernels: zero fill, nnz-chunk partition, carry fix-up, CSR->CSC
// transpose, SDDMM, segment gather / counts and the device-side input checks.
#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

namespace {

constexpr int kBlock = 256;

inline dim3 grid_for(int64_t work, int per_block, int64_t cap = 65536) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return dim3(static_cast<unsigned>(g));
}

}  // namespace

// ---- zero fill (a kernel, not hipMemsetAsync, so that it is safe inside captured graphs) ----

__global__ void fill_zero_contig_kernel(float* __restrict__ C, int64_t n) {
  const int64_t n4 = n / 4;
  float4* C4 = reinterpret_cast<float4*>(C);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += stride)
    C4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t i = n4 * 4 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) C[i] = 0.f;
}

__global__ void fill_zero_strided_kernel(float* __restrict__ C, int64_t rows, int64_t N, int64_t ldc) {
  const int64_t n = rows * N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = i / N;
    C[r * ldc + (i - r * N)] = 0.f;
  }
}

__global__ void fill_i32_kernel(int32_t* __restrict__ a, int64_t n, int32_t value) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = value;
}

hipError_t launch_fill_zero(float* C, int64_t rows, int64_t N, int64_t ldc, hipStream_t s) {
  if (rows <= 0 || N <= 0) return hipSuccess;
  if (ldc == N && (reinterpret_cast<uintptr_t>(C) & 15) == 0) {
    fill_zero_contig_kernel<<<grid_for(rows * N / 4 + 1, kBlock, 16384), kBlock, 0, s>>>(C, rows * N);
  } else {
    fill_zero_strided_kernel<<<grid_for(rows * N, kBlock, 16384), kBlock, 0, s>>>(C, rows, N, ldc);
  }
  return hipGetLastError();
}

// ---- partition: row of the first nnz of every chunk (and of the nnz just before it) ----

__global__ void partition_rows_kernel(const int32_t* __restrict__ rowptr, int M, int64_t nnz, int64_t chunk,
                                      int64_t nchunks, int32_t* __restrict__ chunk_row,
                                      int32_t* __restrict__ chunk_prev) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < nchunks; c += stride) {
    const int64_t e0 = c * chunk;
    chunk_row[c] = row_of_nnz(rowptr, M, e0 < nnz ? e0 : nnz - 1);
    if (chunk_prev) chunk_prev[c] = (c == 0) ? -1 : row_of_nnz(rowptr, M, e0 - 1);
  }
}

hipError_t launch_partition_rows(const int32_t* rowptr, int M, int64_t nnz, int64_t chunk, int64_t nchunks,
                                 int32_t* chunk_row, int32_t* chunk_prev, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  partition_rows_kernel<<<grid_for(nchunks, kBlock), kBlock, 0, s>>>(rowptr, M, nnz, chunk, nchunks, chunk_row,
                                                                     chunk_prev);
  return hipGetLastError();
}

// ---- carry fix-up of the owner write policy ----

__global__ void carry_fixup_kernel(const int32_t* __restrict__ carry_row, const float* __restrict__ carry_val,
                                   int64_t nchunks, int N, float* __restrict__ C, int64_t ldc) {
  const int64_t n = nchunks * N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t c = i / N;
    const int32_t r = carry_row[c];
    if (r >= 0) {
      const int64_t j = i - c * N;
      unsafeAtomicAdd(C + static_cast<int64_t>(r) * ldc + j, carry_val[i]);
    }
  }
}

hipError_t launch_carry_fixup(const int32_t* carry_row, const float* carry_val, int64_t nchunks, int N, float* C,
                              int64_t ldc, hipStream_t s) {
  if (nchunks <= 0 || N <= 0) return hipSuccess;
  carry_fixup_kernel<<<grid_for(nchunks * N, kBlock), kBlock, 0, s>>>(carry_row, carry_val, nchunks, N, C, ldc);
  return hipGetLastError();
}

// ---- CSR -> CSC transpose ---------------------------------------------------------------
// 1. column histogram; 2. exclusive scan -> colptr; 3. scatter every nnz position
// to its column with an atomic cursor (order within a column arbitrary);
// 4. rank each position among its column's positions (count of smaller ones)
// and place it there, which restores CSR order within every column — the same
// result as a stable sort by column; 5. row index and value of every entry.

__global__ void col_histogram_kernel(const int32_t* __restrict__ colind, int64_t nnz, int32_t* __restrict__ counts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz; p += stride)
    atomicAdd(counts + colind[p], 1);
}

// Block-local inclusive scan of 1024 entries per block (4 per thread); writes
// exclusive prefixes to out[i+1] (out[0] written by the caller's finish pass)
// and the block total to block_sums.
__global__ void scan_block_kernel(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out,
                                  int32_t* __restrict__ block_sums) {
  __shared__ int32_t s[1024];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 1024;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) s[i] = (base + i < n) ? in[base + i] : 0;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int32_t tmp[4];
    for (int q = 0; q < 4; ++q) {
      const int i = threadIdx.x + q * 256;
      tmp[q] = (i >= off) ? s[i - off] : 0;
    }
    __syncthreads();
    for (int q = 0; q < 4; ++q) s[threadIdx.x + q * 256] += tmp[q];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < 1024; i += blockDim.x)
    if (base + i < n) out[base + i + 1] = s[i];  // inclusive within block
  if (threadIdx.x == 0) block_sums[blockIdx.x] = s[1023];
}

// Serial scan of the block totals (one thread; nblocks = ceil(K / 1024) is small).
__global__ void scan_block_sums_kernel(int32_t* __restrict__ block_sums, int64_t nblocks) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int32_t run = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
      const int32_t v = block_sums[b];
      block_sums[b] = run;
      run += v;
    }
  }
}

__global__ void scan_finish_kernel(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ block_sums) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i + 1] += block_sums[i / 1024];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

__global__ void col_scatter_kernel(const int32_t* __restrict__ colind, int64_t nnz, const int32_t* __restrict__ colptr,
                                   int32_t* __restrict__ cursor, int32_t* __restrict__ tmp) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < nnz; p += stride) {
    const int c = colind[p];
    const int q = colptr[c] + atomicAdd(cursor + c, 1);
    tmp[q] = static_cast<int32_t>(p);
  }
}

__global__ void col_rank_kernel(const int32_t* __restrict__ colind, int64_t nnz, const int32_t* __restrict__ colptr,
                                const int32_t* __restrict__ tmp, int32_t* __restrict__ perm) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz; q += stride) {
    const int32_t pos = tmp[q];
    const int c = colind[pos];
    const int lo = colptr[c], hi = colptr[c + 1];
    int rank = 0;
    for (int i = lo; i < hi; ++i) rank += (tmp[i] < pos) ? 1 : 0;
    perm[lo + rank] = pos;
  }
}

__global__ void col_finish_kernel(const int32_t* __restrict__ rowptr, int M, const float* __restrict__ vals,
                                  int64_t nnz, const int32_t* __restrict__ perm, int32_t* __restrict__ rowind,
                                  float* __restrict__ vals_t) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < nnz; q += stride) {
    const int32_t pos = perm[q];
    rowind[q] = row_of_nnz(rowptr, M, pos);
    if (vals_t) vals_t[q] = vals ? vals[pos] : 1.0f;
  }
}

hipError_t launch_transpose(const int32_t* rowptr, const int32_t* colind, const float* vals, int M, int K,
                            int64_t nnz, int32_t* colptr, int32_t* rowind, int32_t* perm, float* vals_t,
                            int32_t* counts, int32_t* block_sums, int64_t nblocks, hipStream_t s) {
  if (K == 0) {  // no columns: colptr = {0}, and validation guarantees nnz == 0
    fill_i32_kernel<<<1, 64, 0, s>>>(colptr, 1, 0);
    return hipGetLastError();
  }
  // counts doubles as the scatter cursor; tmp lives in rowind until the finish pass.
  fill_i32_kernel<<<grid_for(K, kBlock), kBlock, 0, s>>>(counts, K, 0);
  if (nnz > 0) col_histogram_kernel<<<grid_for(nnz, kBlock), kBlock, 0, s>>>(colind, nnz, counts);
  scan_block_kernel<<<dim3(static_cast<unsigned>(nblocks)), 256, 0, s>>>(counts, K, colptr, block_sums);
  scan_block_sums_kernel<<<1, 64, 0, s>>>(block_sums, nblocks);
  scan_finish_kernel<<<grid_for(K, kBlock), kBlock, 0, s>>>(colptr, K, block_sums);
  if (nnz > 0) {
    fill_i32_kernel<<<grid_for(K, kBlock), kBlock, 0, s>>>(counts, K, 0);
    int32_t* tmp = rowind;
    col_scatter_kernel<<<grid_for(nnz, kBlock), kBlock, 0, s>>>(colind, nnz, colptr, counts, tmp);
    col_rank_kernel<<<grid_for(nnz, kBlock), kBlock, 0, s>>>(colind, nnz, colptr, tmp, perm);
    col_finish_kernel<<<grid_for(nnz, kBlock), kBlock, 0, s>>>(rowptr, M, vals, nnz, perm, rowind, vals_t);
  }
  return hipGetLastError();
}

// ---- SDDMM on A's pattern: out[p] = <X[row(p), :], Y[colind[p], :]> ----------------------
// The kernel (sddmm_sr_kernel) lives with the sequential-walk kernels; here the
// shape: CF columns per lane at 128-bit width where N and alignment allow, a
// group of G >= ceil(N / CF) lanes (at most 64, wider N loops), 256-nnz chunks.

hipError_t launch_sddmm(VType vt, const int32_t* rowptr, const int32_t* colind, int M, int64_t nnz, int N,
                        const void* X, int64_t ld_x, const void* Y, int64_t ld_y, float* out, hipStream_t s) {
  if (nnz <= 0) return hipSuccess;
  const int es = (vt == VType::F32) ? 4 : 2;
  const int vec = 16 / es;
  auto aligned = [&](int cf) {
    const size_t vb = static_cast<size_t>(es) * cf;
    const size_t a = vb < 16 ? vb : 16;
    return (reinterpret_cast<uintptr_t>(X) % a == 0) && (reinterpret_cast<uintptr_t>(Y) % a == 0) &&
           (ld_x % cf == 0) && (ld_y % cf == 0);
  };
  int cf = 1;
  for (int c = 2; c <= vec; c <<= 1)
    if (N % c == 0 && aligned(c)) cf = c;
  int g = 1;
  while (g < 64 && g * cf < N) g <<= 1;
  SddmmParams p{};
  p.M = M;
  p.N = N;
  p.nnz = nnz;
  p.rowptr = rowptr;
  p.colind = colind;
  p.X = X;
  p.ld_x = ld_x;
  p.Y = Y;
  p.ld_y = ld_y;
  p.out = out;
  p.chunk = (g >= 256) ? g : 256;
  p.nchunks = (nnz + p.chunk - 1) / p.chunk;
  const int tpb = 256;
  const int gpb = tpb / g;
  int64_t blocks = (p.nchunks + gpb - 1) / gpb;
  if (blocks > 4096) blocks = 4096;  // ~16 waves per CU on 256 CUs, the rest grid-strided
  p.total_groups = blocks * gpb;
  Launch l;
  l.grid = dim3(static_cast<unsigned>(blocks));
  l.block = dim3(tpb);
  l.shmem = 0;
  l.stream = s;
  bool ok = true;
  hipError_t e = launch_sddmm_sr(vt, cf, g, p, l, ok);
  return ok ? e : hipErrorInvalidValue;
}

// ---- segment gather / counts ----

__global__ void segment_gather_kernel(const float* __restrict__ ddst, int64_t ld_ddst, const int32_t* __restrict__ index,
                                      int64_t E, int N, const int32_t* __restrict__ counts, float* __restrict__ dsrc,
                                      int64_t ld_dsrc) {
  const int64_t n = E * N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t e = i / N;
    const int64_t j = i - e * N;
    const int s = index[e];
    float scale = 1.0f;
    if (counts) scale = 1.0f / static_cast<float>(max(counts[s], 1));
    dsrc[e * ld_dsrc + j] = ddst[static_cast<int64_t>(s) * ld_ddst + j] * scale;
  }
}

hipError_t launch_segment_gather(const float* ddst, int64_t ld_ddst, const int32_t* index, int64_t E, int N,
                                 const int32_t* counts, float* dsrc, int64_t ld_dsrc, hipStream_t s) {
  if (E <= 0 || N <= 0) return hipSuccess;
  segment_gather_kernel<<<grid_for(E * N, kBlock, 65536), kBlock, 0, s>>>(ddst, ld_ddst, index, E, N, counts, dsrc,
                                                                         ld_dsrc);
  return hipGetLastError();
}

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t n, int32_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void segment_counts_kernel(const int32_t* __restrict__ index, int64_t E, int64_t dst_len,
                                      int32_t* __restrict__ counts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < dst_len; s += stride) {
    const int64_t lo = lower_bound_i32(index, E, static_cast<int32_t>(s));
    const int64_t hi = lower_bound_i32(index, E, static_cast<int32_t>(s + 1));
    counts[s] = static_cast<int32_t>(hi - lo);
  }
}

hipError_t launch_segment_counts(const int32_t* index, int64_t E, int64_t dst_len, int32_t* counts, hipStream_t s) {
  if (dst_len <= 0) return hipSuccess;
  segment_counts_kernel<<<grid_for(dst_len, kBlock), kBlock, 0, s>>>(index, E, dst_len, counts);
  return hipGetLastError();
}

// ---- device-side input checks ----

__global__ void check_csr_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind, int64_t M,
                                 int64_t K, int64_t nnz, int32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < M; r += stride) {
    int32_t f = 0;
    const int64_t a = rowptr[r], b = rowptr[r + 1];
    if (r == 0 && a != 0) f |= 1;
    if (r == M - 1 && b != nnz) f |= 1;
    if (b < a || a < 0 || b > nnz) {
      f |= 1;
    } else {
      for (int64_t q = a; q < b; ++q) {
        const int c = colind[q];
        if (c < 0 || c >= K) f |= 2;
        if (q > a && colind[q - 1] >= c) f |= 4;
      }
    }
    if (f) atomicOr(flags, f);
  }
}

hipError_t launch_check_csr(const int32_t* rowptr, const int32_t* colind, int64_t M, int64_t K, int64_t nnz,
                            int32_t* flags, hipStream_t s) {
  fill_i32_kernel<<<1, 64, 0, s>>>(flags, 1, 0);
  if (M > 0) check_csr_kernel<<<grid_for(M, kBlock), kBlock, 0, s>>>(rowptr, colind, M, K, nnz, flags);
  return hipGetLastError();
}

__global__ void check_index_kernel(const int32_t* __restrict__ index, int64_t E, int64_t dst_len,
                                   int32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += stride) {
    int32_t f = 0;
    const int32_t v = index[e];
    if (v < 0 || v >= dst_len) f |= 2;
    if (e > 0 && index[e - 1] > v) f |= 1;
    if (f) atomicOr(flags, f);
  }
}

hipError_t launch_check_index(const int32_t* index, int64_t E, int64_t dst_len, int32_t* flags, hipStream_t s) {
  fill_i32_kernel<<<1, 64, 0, s>>>(flags, 1, 0);
  if (E > 0) check_index_kernel<<<grid_for(E, kBlock), kBlock, 0, s>>>(index, E, dst_len, flags);
  return hipGetLastError();
}

}  // namespace segscan
