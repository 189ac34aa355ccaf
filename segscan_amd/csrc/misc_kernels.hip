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

// Carries: one lane per (chunk, column). Entry-free blocks: each wave tests 64 blocks at
// once and zeroes the rows of those that hold no entry (rows x N flattened over its
// lanes); consecutive windows of 64 blocks go to consecutive waves, so one long empty
// run is zeroed by many waves.
__global__ void carry_fixup_kernel(const int32_t* __restrict__ carry_row, const float* __restrict__ carry_val,
                                   int64_t nchunks, int N, float* __restrict__ C, int64_t ldc,
                                   const uint8_t* __restrict__ zflag, const int32_t* __restrict__ rowptr,
                                   int64_t M) {
  const int64_t n = nchunks * N;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t i = tid; i < n; i += stride) {
    const int64_t c = i / N;
    const int32_t r = carry_row[c];
    if (r >= 0) {
      const int64_t j = i - c * N;
      unsafeAtomicAdd(C + static_cast<int64_t>(r) * ldc + j, carry_val[i]);
    }
  }
  if (zflag == nullptr && rowptr == nullptr) return;
  const int64_t nblk = zero_blocks(M);
  const int lane = threadIdx.x & 63;
  // Row stores: 16-byte units when rows allow them. A row of U units takes U lanes;
  // 64 / U rows go at once when U <= 64, else the lanes loop over the row's units.
  const bool vec = N % 4 == 0 && ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
  const int U = vec ? N / 4 : N;
  const int rows_at_once = U <= 64 ? 64 / U : 1;
  const int lane_row = U <= 64 ? lane / U : 0;
  const int lane_unit = U <= 64 ? lane - lane_row * U : lane;
  for (int64_t w = tid >> 6; w * 64 < nblk; w += stride >> 6) {
    const int64_t j = w * 64 + lane;
    bool empty = false;
    if (j < nblk) {
      if (zflag) empty = zflag[j] != 0;
      else empty = rowptr[j << kZeroRowsLog] == rowptr[min((j + 1) << kZeroRowsLog, M)];
    }
    uint64_t m = __ballot(empty);
    while (m) {
      const int64_t blk = w * 64 + (__ffsll(static_cast<unsigned long long>(m)) - 1);
      m &= m - 1;
      const int64_t r0 = blk << kZeroRowsLog;
      const int nrows = static_cast<int>(min(r0 + kZeroRows, M) - r0);
      if (lane_row < rows_at_once) {
        for (int rr = lane_row; rr < nrows; rr += rows_at_once) {
          float* row = C + (r0 + rr) * ldc;
          for (int u = lane_unit; u < U; u += (U <= 64 ? U : 64)) {
            if (vec) *reinterpret_cast<float4*>(row + 4 * static_cast<int64_t>(u)) = make_float4(0.f, 0.f, 0.f, 0.f);
            else row[u] = 0.0f;
          }
        }
      }
    }
  }
}

hipError_t launch_carry_fixup(const int32_t* carry_row, const float* carry_val, int64_t nchunks, int N, float* C,
                              int64_t ldc, const uint8_t* zflag, const int32_t* rowptr, int64_t M, hipStream_t s) {
  if (nchunks <= 0 || N <= 0) return hipSuccess;
  // at least one wave per 64 blocks of flags, up to the carry work's grid or more
  const int64_t waves = (zflag || rowptr) ? (zero_blocks(M) + 63) / 64 : 0;
  const int64_t work = std::max<int64_t>(nchunks * N, waves * 64);
  carry_fixup_kernel<<<grid_for(work, kBlock), kBlock, 0, s>>>(carry_row, carry_val, nchunks, N, C, ldc, zflag, rowptr, M);
  return hipGetLastError();
}

// ---- CSR -> CSC transpose ---------------------------------------------------------------
// A stable LSD radix sort of the column indices. Every entry carries
// (row << 32 | position); CSR order is row-major, so a stable sort by column
// leaves the rows of each column ascending — the reference's counting-sort
// transpose (dataloader.hpp:123-144). No global atomics: tile histograms in LDS,
// a device-wide scan of the (digit-major) tile histograms, then a stable scatter
// in which equal digits are ranked inside a wave by ballots and across the
// block's waves and rounds through LDS counters. The last pass writes perm,
// rowind and the permuted values directly; colptr is a lower bound per column
// over the sorted keys.

// One 64-ary search per wave (the row of its first entry), then the wave's
// rowptr window resolves every lane (wave_row_lookup).
__global__ void transpose_init_kernel(const int32_t* __restrict__ rowptr, int M, int64_t nnz,
                                      uint64_t* __restrict__ pay) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t wbase = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63); wbase < nnz;
       wbase += stride) {
    const int64_t p = wbase + lane;
    const int e = static_cast<int>(p < nnz ? p : nnz - 1);
    const int r0 = static_cast<int>(wave_search_gt(rowptr, 1, M, wbase)) - 1;  // rowptr[M] = nnz > wbase
    const uint32_t r = static_cast<uint32_t>(wave_row_lookup(rowptr, M, r0, e));
    if (p < nnz) pay[p] = (static_cast<uint64_t>(r) << 32) | static_cast<uint32_t>(p);
  }
}

constexpr int kRadixThreads = 256;
constexpr int kRadixItems = static_cast<int>(kTransposeTile / kRadixThreads);  // 16
constexpr int kRadixMaxBuckets = 512;

__global__ void __launch_bounds__(kRadixThreads)
radix_upsweep_kernel(const int32_t* __restrict__ keys, int64_t nnz, int shift, int mask, int buckets, int64_t ntiles,
                     int32_t* __restrict__ hist) {
  __shared__ int32_t cnt[kRadixMaxBuckets];
  for (int b = threadIdx.x; b < buckets; b += kRadixThreads) cnt[b] = 0;
  __syncthreads();
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTransposeTile;
  for (int j = 0; j < kRadixItems; ++j) {
    const int64_t idx = base + static_cast<int64_t>(j) * kRadixThreads + threadIdx.x;
    if (idx < nnz) atomicAdd(&cnt[(keys[idx] >> shift) & mask], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < buckets; b += kRadixThreads) hist[static_cast<int64_t>(b) * ntiles + tile] = cnt[b];
}

// Stable scatter of one tile. Wave w holds the tile's entries [w * 1024, (w + 1) * 1024),
// 16 per lane, loaded up front (64 consecutive entries per round, coalesced). Each wave
// ranks its entries among equal digits by itself: ballots over the digit bits give the
// lanes of a round that share a digit, and a wave-private LDS count per digit carries the
// rank across rounds (no block barrier per round). One barrier later, a per-digit prefix
// over the waves places each wave's run after the earlier waves' runs of the same digit,
// which keeps the scatter stable (entries of one digit stay in tile order). The tile is
// then placed in LDS in digit order and consecutive lanes copy consecutive LDS slots out,
// so each digit's run of the tile is written as one contiguous stretch.
// LAST: write perm and rowind instead of the next pass's payload. VALS: the
// entries' values travel with them (val_in: the CSR values in the first pass,
// or NULL for all ones), ending in val_out (vals_t in the last pass).
template <bool LAST, bool VALS>
__global__ void __launch_bounds__(kRadixThreads)
radix_downsweep_kernel(const int32_t* __restrict__ keys_in, const uint64_t* __restrict__ pay_in, int64_t nnz,
                       int shift, int mask, int dbits, int buckets, int64_t ntiles, const int32_t* __restrict__ hist,
                       const int32_t* __restrict__ offs, int32_t* __restrict__ keys_out, uint64_t* __restrict__ pay_out,
                       int32_t* __restrict__ perm, int32_t* __restrict__ rowind, const float* __restrict__ val_in,
                       float* __restrict__ val_out) {
  constexpr int kWaves = kRadixThreads / 64;
  constexpr int kPerWave = kTransposeTile / kWaves;  // 1024 = kRadixItems rounds of 64
  __shared__ int32_t s_base[kRadixMaxBuckets];    // global start of each digit's run of this tile
  __shared__ int32_t s_lstart[kRadixMaxBuckets];  // tile-local start of each digit's run
  __shared__ int32_t s_wc[kWaves][kRadixMaxBuckets];  // per-wave digit counts, then their prefixes
  __shared__ int32_t s_key[kTransposeTile];
  __shared__ uint64_t s_pay[kTransposeTile];
  __shared__ float s_val[VALS ? kTransposeTile : 1];
  const int64_t tile = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int b = threadIdx.x; b < buckets; b += kRadixThreads) {
    s_base[b] = offs[static_cast<int64_t>(b) * ntiles + tile];
    for (int w = 0; w < kWaves; ++w) s_wc[w][b] = 0;
  }
  if (wave == 0) {  // exclusive scan of this tile's digit counts
    const int per = (buckets + 63) / 64;
    const int b0 = lane * per;
    int sum = 0;
    for (int b = b0; b < b0 + per && b < buckets; ++b) sum += hist[static_cast<int64_t>(b) * ntiles + tile];
    int incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off);
      if (lane >= off) incl += t;
    }
    int run = incl - sum;
    for (int b = b0; b < b0 + per && b < buckets; ++b) {
      s_lstart[b] = run;
      run += hist[static_cast<int64_t>(b) * ntiles + tile];
    }
  }
  const uint64_t below = (1ull << lane) - 1ull;
  const int64_t base = tile * kTransposeTile;
  const int tile_n = static_cast<int>(nnz - base < kTransposeTile ? nnz - base : kTransposeTile);
  const int wbase = wave * kPerWave;
  int32_t key[kRadixItems];
  uint64_t pay[kRadixItems];
  float val[kRadixItems];
  int32_t lpos[kRadixItems];
#pragma unroll
  for (int j = 0; j < kRadixItems; ++j) {
    const int i = wbase + j * 64 + lane;
    key[j] = i < tile_n ? keys_in[base + i] : 0;
    pay[j] = i < tile_n ? pay_in[base + i] : 0ull;
    if constexpr (VALS) val[j] = (i < tile_n && val_in) ? val_in[base + i] : 1.0f;
  }
  __syncthreads();  // s_wc zeroed, s_lstart and s_base written
#pragma unroll
  for (int j = 0; j < kRadixItems; ++j) {
    const bool valid = wbase + j * 64 + lane < tile_n;
    const int d = (key[j] >> shift) & mask;
    // lanes of this round holding the same digit
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank_w = __popcll(peers & below);
    // the wave's own count of digit d so far (only this wave writes s_wc[wave])
    const int c = valid ? s_wc[wave][d] : 0;
    lpos[j] = c + rank_w;
    if (valid && rank_w == 0) s_wc[wave][d] = c + __popcll(peers);
  }
  __syncthreads();  // every wave's counts final
  for (int b = threadIdx.x; b < buckets; b += kRadixThreads) {
    int run = 0;
    for (int w = 0; w < kWaves; ++w) {
      const int cw = s_wc[w][b];
      s_wc[w][b] = run;
      run += cw;
    }
  }
  __syncthreads();  // per-wave prefixes in s_wc
#pragma unroll
  for (int j = 0; j < kRadixItems; ++j) {
    if (wbase + j * 64 + lane < tile_n) {
      const int d = (key[j] >> shift) & mask;
      const int slot = s_lstart[d] + s_wc[wave][d] + lpos[j];
      s_key[slot] = key[j];
      s_pay[slot] = pay[j];
      if constexpr (VALS) s_val[slot] = val[j];
    }
  }
  __syncthreads();  // the tile sits in LDS in digit order
  for (int slot = threadIdx.x; slot < tile_n; slot += kRadixThreads) {
    const int32_t k = s_key[slot];
    const uint64_t v = s_pay[slot];
    const int d = (k >> shift) & mask;
    const int64_t pos = s_base[d] + (slot - s_lstart[d]);
    keys_out[pos] = k;
    if constexpr (LAST) {
      perm[pos] = static_cast<int32_t>(v & 0xffffffffull);
      rowind[pos] = static_cast<int32_t>(v >> 32);
    } else {
      pay_out[pos] = v;
    }
    if constexpr (VALS) val_out[pos] = s_val[slot];
  }
}

// Block-local inclusive scan of 1024 entries per block (4 per thread); writes
// exclusive prefixes to out[i+1] (out[0] written by the finish pass) and the
// block total to block_sums.
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

// Exclusive scan of the block totals in place, one workgroup of 1024 lanes
// walking them 1024 at a time with a running carry.
__global__ void __launch_bounds__(1024) scan_block_sums_kernel(int32_t* __restrict__ block_sums, int64_t nblocks) {
  __shared__ int32_t s[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  for (int64_t base = 0; base < nblocks; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int32_t v = i < nblocks ? block_sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblocks) block_sums[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
}

__global__ void scan_finish_kernel(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ block_sums) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i + 1] += block_sums[i / 1024];
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

// colptr[c] = first sorted position whose key is >= c (c <= K; keys < K).
__global__ void colptr_from_keys_kernel(const int32_t* __restrict__ keys, int64_t nnz, int K,
                                        int32_t* __restrict__ colptr) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c <= K; c += stride) {
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (keys[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    colptr[c] = static_cast<int32_t>(lo);
  }
}

static hipError_t launch_exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int32_t* bsums, int64_t nblocks,
                                 hipStream_t s) {
  scan_block_kernel<<<dim3(static_cast<unsigned>(nblocks)), 256, 0, s>>>(in, n, out, bsums);
  scan_block_sums_kernel<<<1, 1024, 0, s>>>(bsums, nblocks);
  scan_finish_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(out, n, bsums);
  return hipGetLastError();
}

hipError_t launch_transpose(const int32_t* rowptr, const int32_t* colind, const float* vals, int M, int K,
                            int64_t nnz, int32_t* colptr, int32_t* rowind, int32_t* perm, float* vals_t,
                            void* workspace, const TransposePlan& tp, hipStream_t s) {
  if (nnz == 0 || K == 0) {  // every column empty (K == 0 implies nnz == 0 by validation)
    fill_i32_kernel<<<grid_for(static_cast<int64_t>(K) + 1, kBlock), kBlock, 0, s>>>(colptr, K + 1, 0);
    return hipGetLastError();
  }
  char* ws = static_cast<char*>(workspace);
  int32_t* keys_a = reinterpret_cast<int32_t*>(ws + tp.off_keys_a);
  int32_t* keys_b = reinterpret_cast<int32_t*>(ws + tp.off_keys_b);
  uint64_t* pay_a = reinterpret_cast<uint64_t*>(ws + tp.off_pay_a);
  uint64_t* pay_b = reinterpret_cast<uint64_t*>(ws + tp.off_pay_b);
  int32_t* hist = reinterpret_cast<int32_t*>(ws + tp.off_hist);
  int32_t* offs = reinterpret_cast<int32_t*>(ws + tp.off_offs);
  int32_t* bsums = reinterpret_cast<int32_t*>(ws + tp.off_bsums);
  transpose_init_kernel<<<grid_for(nnz, kBlock), kBlock, 0, s>>>(rowptr, M, nnz, pay_a);
  const int mask = tp.buckets - 1;
  const dim3 tiles(static_cast<unsigned>(tp.ntiles));
  const bool carry_vals = vals_t != nullptr;
  float* val_a = reinterpret_cast<float*>(ws + tp.off_val_a);
  float* val_b = reinterpret_cast<float*>(ws + tp.off_val_b);
  const int32_t* kin = colind;
  const uint64_t* pin = pay_a;
  const float* vin = vals;  // NULL: all ones
  int32_t* kout = keys_b;
  uint64_t* pout = pay_b;
  float* vout = val_b;
  for (int pass = 0; pass < tp.passes; ++pass) {
    const int shift = pass * tp.dbits;
    radix_upsweep_kernel<<<tiles, kRadixThreads, 0, s>>>(kin, nnz, shift, mask, tp.buckets, tp.ntiles, hist);
    hipError_t e = launch_exclusive_scan(hist, tp.hist_n, offs, bsums, tp.scan_blocks, s);
    if (e != hipSuccess) return e;
    const bool last = pass + 1 == tp.passes;
    float* vdst = last ? vals_t : vout;
#define SEGSCAN_DOWNSWEEP(L, V)                                                                                  \
  radix_downsweep_kernel<L, V><<<tiles, kRadixThreads, 0, s>>>(kin, pin, nnz, shift, mask, tp.dbits, tp.buckets, \
                                                               tp.ntiles, hist, offs, kout, L ? nullptr : pout,   \
                                                               L ? perm : nullptr, L ? rowind : nullptr, vin, vdst)
    if (last && carry_vals) SEGSCAN_DOWNSWEEP(true, true);
    else if (last) SEGSCAN_DOWNSWEEP(true, false);
    else if (carry_vals) SEGSCAN_DOWNSWEEP(false, true);
    else SEGSCAN_DOWNSWEEP(false, false);
#undef SEGSCAN_DOWNSWEEP
    if (!last) {
      kin = kout;
      pin = pout;
      if (carry_vals) vin = vout;
      kout = (kout == keys_b) ? keys_a : keys_b;
      pout = (pout == pay_b) ? pay_a : pay_b;
      vout = (vout == val_b) ? val_a : val_b;
    }
  }
  colptr_from_keys_kernel<<<grid_for(static_cast<int64_t>(K) + 1, kBlock), kBlock, 0, s>>>(kout, nnz, K, colptr);
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

// dsrc[e, :] = ddst[index[e], :] * scale. Each lane keeps one column position of
// the row it is on (vpr lanes per row, rows_per_block rows per workgroup pass) and
// walks rows grid-strided: no per-element division, and 16-byte loads and stores
// when N and the leading dimensions allow (V = 4). The 16-byte path keeps four rows
// in flight per lane (four index loads, then four row loads) and writes dsrc, which
// nothing rereads soon, with non-temporal stores: 2.39M x 64 rows over 256 segments
// 157 -> 106 us, 4.19M x 64 over 262144 segments 311 -> 193 us.
template <int V>
__global__ void segment_gather_kernel(const float* __restrict__ ddst, int64_t ld_ddst, const int32_t* __restrict__ index,
                                      int64_t E, int N, const int32_t* __restrict__ counts, float* __restrict__ dsrc,
                                      int64_t ld_dsrc, int vpr, int rows_per_block) {
  const int t = threadIdx.x;
  const int ry = t / vpr;
  const int col = (t - ry * vpr) * V;
  if (ry >= rows_per_block) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * rows_per_block;
  if constexpr (V == 4) {
    constexpr int R = 4;
    for (int64_t e0 = static_cast<int64_t>(blockIdx.x) * rows_per_block + ry; e0 < E; e0 += R * stride) {
      int sr[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int64_t e = e0 + i * stride;
        sr[i] = e < E ? index[e] : 0;  // past E: row 0 (E > 0 entries name valid rows), not stored
      }
      float4 xr[R];
      float sc[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        xr[i] = *reinterpret_cast<const float4*>(ddst + static_cast<int64_t>(sr[i]) * ld_ddst + col);
        sc[i] = counts ? 1.0f / static_cast<float>(max(counts[sr[i]], 1)) : 1.0f;
      }
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int64_t e = e0 + i * stride;
        if (e < E) {
          float* dst = dsrc + e * ld_dsrc + col;
          __builtin_nontemporal_store(xr[i].x * sc[i], dst);
          __builtin_nontemporal_store(xr[i].y * sc[i], dst + 1);
          __builtin_nontemporal_store(xr[i].z * sc[i], dst + 2);
          __builtin_nontemporal_store(xr[i].w * sc[i], dst + 3);
        }
      }
    }
  } else {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * rows_per_block + ry; e < E; e += stride) {
      const int s = index[e];
      const float scale = counts ? 1.0f / static_cast<float>(max(counts[s], 1)) : 1.0f;
      const float* src = ddst + static_cast<int64_t>(s) * ld_ddst;
      float* dst = dsrc + e * ld_dsrc;
      for (int c = col; c < N; c += vpr) dst[c] = src[c] * scale;
    }
  }
}

hipError_t launch_segment_gather(const float* ddst, int64_t ld_ddst, const int32_t* index, int64_t E, int N,
                                 const int32_t* counts, float* dsrc, int64_t ld_dsrc, hipStream_t s) {
  if (E <= 0 || N <= 0) return hipSuccess;
  const bool vec = N % 4 == 0 && N <= 4 * kBlock && ld_ddst % 4 == 0 && ld_dsrc % 4 == 0 &&
                   reinterpret_cast<uintptr_t>(ddst) % 16 == 0 && reinterpret_cast<uintptr_t>(dsrc) % 16 == 0;
  const int vpr = vec ? N / 4 : (N < kBlock ? N : kBlock);
  const int rows_per_block = kBlock / vpr;
  const dim3 grid = grid_for(E, vec ? rows_per_block * 4 : rows_per_block, 65536);  // 4 rows per lane pass (V = 4)
  if (vec)
    segment_gather_kernel<4><<<grid, kBlock, 0, s>>>(ddst, ld_ddst, index, E, N, counts, dsrc, ld_dsrc, vpr,
                                                     rows_per_block);
  else
    segment_gather_kernel<1><<<grid, kBlock, 0, s>>>(ddst, ld_ddst, index, E, N, counts, dsrc, ld_dsrc, vpr,
                                                     rows_per_block);
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

// ---- gather probe: the fabric rate of random dense-row reads ----
// Each group of G lanes reads rows idx[i] of the table (16 bytes per lane, 16 loads in
// flight per group) over its chunk of indices and keeps a running sum, written once per
// lane at the end so that no load is dropped. It reads as the SpMM's dense-row gather does,
// without the CSR streams or the output.

template <int G>
__global__ void __launch_bounds__(256) gather_probe_kernel(const uint4* __restrict__ table, int64_t row_vecs,
                                                           const int32_t* __restrict__ idx, int64_t n, int64_t chunk,
                                                           float* __restrict__ out) {
  constexpr int U = 16;
  const int gl = threadIdx.x % G;
  const int64_t group = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
  const int64_t i0 = group * chunk;
  const int64_t i1 = min(n, i0 + chunk);
  uint32_t acc = 0;
  for (int64_t i = i0; i < i1; i += U) {
    uint4 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = (i + u < i1) ? idx[i + u] : idx[i];
      r[u] = table[k * row_vecs + gl];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u < i1) acc += r[u].x ^ r[u].y ^ r[u].z ^ r[u].w;
    }
  }
  out[static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x] = static_cast<float>(acc & 0xffff);
}

hipError_t launch_gather_probe(const void* table, int64_t row_bytes, const int32_t* idx, int64_t n, int64_t chunk,
                               float* out, int64_t out_len, hipStream_t s) {
  const int G = static_cast<int>(row_bytes / 16);
  const int64_t groups = (n + chunk - 1) / chunk;
  const int64_t blocks = (groups * G + 255) / 256;
  if (blocks * 256 > out_len) return hipErrorInvalidValue;
  const uint4* t = static_cast<const uint4*>(table);
  switch (G) {
    case 8: gather_probe_kernel<8><<<static_cast<unsigned>(blocks), 256, 0, s>>>(t, G, idx, n, chunk, out); break;
    case 16: gather_probe_kernel<16><<<static_cast<unsigned>(blocks), 256, 0, s>>>(t, G, idx, n, chunk, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- inverse permutation: inv[perm[i]] = i ----

__global__ void invert_permutation_kernel(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t j = perm[i];
    if (j >= 0 && j < n) inv[j] = static_cast<int32_t>(i);
  }
}

hipError_t launch_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  invert_permutation_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(perm, n, inv);
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
