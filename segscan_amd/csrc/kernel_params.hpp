// Launch parameter blocks passed by value to the kernels, and the declarations
// of the per-family launch functions (defined in the .hip files, called by the
// host dispatcher in segscan_abi.cpp).
#pragma once

#include <stdint.h>
#include <stddef.h>

#include <hip/hip_runtime.h>

namespace segscan {

enum class VType : int { F32 = 0, BF16 = 1, F16 = 2 };

// XCDs of an MI355X (gfx950), each with its own L2. Workgroups are dispatched to them round
// robin: workgroup w runs on XCD w % kXcds.
constexpr int kXcds = 8;

// Permutation plans (permute.hip) stage values in buckets of kPermBucket consecutive
// output positions; an ebsr SpMM whose workgroups each cover a part of one bucket reads
// its values from that bucket (SrParams::vslot).
constexpr int kPermBucket = 8192;

// Owner write policy: empty output rows are tracked in blocks of kZeroRows rows, one
// flag byte per block (owner_empty_run, carry_fixup_kernel).
constexpr int kZeroRowsLog = 4;
constexpr int kZeroRows = 1 << kZeroRowsLog;
__host__ __device__ constexpr int64_t zero_blocks(int64_t rows) { return (rows + kZeroRows - 1) >> kZeroRowsLog; }

// Parallel-reduction SpMM / segment_coo (ebpr, rbpr, segment_coo, segment_coo_new).
// Lanes of a workgroup are laid out [Y][P][X]: X lanes over consecutive nnz
// (the tile), P column panels of CF columns, Y tiles per workgroup.
struct PrParams {
  int M;              // rows of A (SpMM) / dst_len (segment)
  int N;              // dense columns
  int64_t nnz;        // nnz (SpMM) / E (segment)
  const int32_t* rowptr;   // SpMM only
  const int32_t* colind;   // SpMM only
  const float* vals;       // SpMM only, nullable
  const int32_t* vmap;     // nullable: entry e's value is vals[vmap[e]]
  const int32_t* index;    // segment only
  const void* B;           // dense operand (B or src)
  int64_t ldb;
  float* C;                // output (C or dst)
  int64_t ldc;
  int X, P, Y;        // lane layout
  int G;              // group size (template; kept for checks)
  int slice;          // columns per blockIdx.y
  int64_t ntiles;     // tiles of X nnz
  const int32_t* tile_row;  // row of the first nnz of every tile (ebpr)
  int row_store;      // rbpr: 1 = one partial per row -> plain store, 0 = atomic partials
  int accumulate;
};

// Sequential-reduction SpMM / segment_coo (ebsr, segment_coo_sr family).
// Groups of G lanes cover G*CF columns; each group walks chunks of `chunk`
// consecutive nnz in order.
struct SrParams {
  int M;
  int N;
  int64_t nnz;
  const int32_t* rowptr;
  const int32_t* colind;
  const float* vals;
  const int32_t* vmap;        // nullable: entry e's value is vals[vmap[e]]
  // nullable (SpMM): values staged by a permutation plan's first pass. Entry e's value is
  // vmid[(e / kPermBucket) * kPermBucket + vslot[e]]; chunk x groups per workgroup divides
  // kPermBucket, so each workgroup round reads one bucket, loaded into LDS.
  const float* vmid;
  const uint16_t* vslot;
  const int32_t* index;
  const void* B;
  int64_t ldb;
  float* C;
  int64_t ldc;
  int tpb;            // lanes per workgroup (multiple of 64)
  int64_t chunk;      // nnz per chunk
  int64_t nchunks;
  int64_t total_groups;
  const int32_t* chunk_row;   // row of the first nnz of each chunk (SpMM)
  const int32_t* chunk_prev;  // row of the nnz just before each chunk, -1 for chunk 0 (SpMM)
  int32_t* carry_row;         // [nchunks] owner policy: row of the chunk's carry or -1
  float* carry_val;           // [nchunks x N] owner policy
  uint8_t* zflag;             // [zero_blocks(M)] segments, owner policy, not accumulating: entry-free blocks
  int policy;                 // 1 atomic, 2 owner store + carry, 3 lessatom
  int accumulate;
};

// SDDMM on A's pattern (sequential walk over nnz chunks).
struct SddmmParams {
  int M;
  int N;
  int64_t nnz;
  const int32_t* rowptr;
  const int32_t* colind;
  const void* X;
  int64_t ld_x;
  const void* Y;
  int64_t ld_y;
  float* out;
  int64_t chunk;
  int64_t nchunks;
  int64_t total_groups;
};

// Fused SpMM backward over At = A^T (dB = At.dC and dvals = SDDMM(dC, B) on A's
// pattern from one gather of dC per entry). Groups of G lanes cover all N
// columns (G * CF >= N) and walk chunks of `chunk` consecutive At entries.
struct BwdParams {
  int K;                   // rows of At (columns of A)
  int N;
  int64_t nnz;
  const int32_t* rowptr;   // At: K + 1 offsets
  const int32_t* colind;   // At: row of A of every entry
  const int32_t* perm;     // At position -> A position
  const float* vals_t;     // values in At order, nullable (then vals through perm)
  const float* vals;       // A's values in A's order, nullable (all ones when vals_t is null too)
  const void* dC;          // gathered rows, M x N (or indexed through dc_map)
  int64_t ld_dc;
  const int32_t* dc_map;   // nullable: A's row r reads dC row dc_map[r]
  const void* B;           // row j of At reads B[j], K x N
  int64_t ld_b;
  float* dB;               // K x N, nullable (not computed)
  int64_t ld_db;
  float* dvals;            // nnz, A's order, nullable (not computed)
  int tpb;
  int64_t chunk;
  int64_t nchunks;
  int64_t total_groups;
  int32_t* carry_row;      // [nchunks] row of the chunk's carry or -1
  float* carry_val;        // [nchunks x N]
  int accumulate;          // dB += instead of dB =
  int dvals_at_order;      // dvals[q] (At's order) instead of dvals[perm[q]]
};

// Row-balanced sequential reduction (rbsr): lanes [rows][column vectors].
struct RbsrParams {
  int M;
  int N;
  const int32_t* rowptr;
  const int32_t* colind;
  const float* vals;
  const int32_t* vmap;  // nullable: entry e's value is vals[vmap[e]]
  const void* B;
  int64_t ldb;
  float* C;
  int64_t ldc;
  int tpr;            // lanes per row (column vectors per slice)
  int rows_per_block;
  int slice;
  int64_t rows_total; // grid-stride bound (M)
  int accumulate;
};

struct Launch {
  dim3 grid;
  dim3 block;
  size_t shmem;
  hipStream_t stream;
};

// Each returns hipSuccess or the launch error; `ok` false means the (VT, CF, G)
// combination is not compiled.
hipError_t launch_spmm_pr(VType vt, int cf, int g, const PrParams& p, const Launch& l, bool eb, bool& ok);
hipError_t launch_spmm_sr(VType vt, int cf, int g, const SrParams& p, const Launch& l, bool& ok);
hipError_t launch_spmm_rbsr(VType vt, int cf, const RbsrParams& p, const Launch& l, bool& ok);
hipError_t launch_seg_pr(VType vt, int cf, int g, const PrParams& p, const Launch& l, bool& ok);
hipError_t launch_seg_sr(VType vt, int cf, int g, const SrParams& p, const Launch& l, bool& ok);
hipError_t launch_sddmm_sr(VType vt, int cf, int g, const SddmmParams& p, const Launch& l, bool& ok);
// Fused backward (bwd_kernels.hip): l.shmem = 12 * p.tpb, p.tpb <= 256. xt: dC storage,
// yt: B storage (equal, or xt F32 with a 16-bit yt). row_batch: the form that loads B
// rows per batch of entries, compiled for yt BF16, xt F32 or BF16, cf 4, g 8 / 16 / 32.
hipError_t launch_spmm_bwd(VType xt, VType yt, int cf, int g, bool row_batch, const BwdParams& p, const Launch& l,
                           bool& ok);

// Helpers (misc_kernels.hip).
hipError_t launch_fill_zero(float* C, int64_t rows, int64_t N, int64_t ldc, hipStream_t s);
hipError_t launch_partition_rows(const int32_t* rowptr, int M, int64_t nnz, int64_t chunk, int64_t nchunks,
                                 int32_t* chunk_row, int32_t* chunk_prev, hipStream_t s);
// Fix-up pass of the owner write policy: adds each chunk's carry to its row and zeroes
// every row of the entry-free blocks of kZeroRows rows: those flagged in
// zflag[0, zero_blocks(M)) when zflag is given, else those with rowptr[16j] ==
// rowptr[min(16j + 16, M)] when rowptr is given (neither: no zeroing, as when accumulating).
hipError_t launch_carry_fixup(const int32_t* carry_row, const float* carry_val, int64_t nchunks, int N,
                              float* C, int64_t ldc, const uint8_t* zflag, const int32_t* rowptr, int64_t M,
                              hipStream_t s);
// CSR -> CSC transpose as a stable LSD radix sort of the column indices, carrying
// (row << 32 | position) per entry. Tiles of kTransposeTile entries; `passes`
// digits of `dbits` bits cover the `bits` bits of the largest column index.
constexpr int64_t kTransposeTile = 4096;  // 256 lanes x 16 entries

struct TransposePlan {
  int bits = 1, passes = 1, dbits = 1, buckets = 2;
  int64_t ntiles = 0, hist_n = 0, scan_blocks = 0;
  size_t off_keys_a = 0, off_keys_b = 0, off_pay_a = 0, off_pay_b = 0, off_val_a = 0, off_val_b = 0, off_hist = 0,
         off_offs = 0, off_bsums = 0, bytes = 0;
};

inline TransposePlan plan_transpose(int64_t K, int64_t nnz) {
  auto align256 = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
  TransposePlan t;
  int64_t top = K > 1 ? K - 1 : 1;
  t.bits = 0;
  while (top > 0) {
    ++t.bits;
    top >>= 1;
  }
  t.passes = (t.bits + 8) / 9;                      // digits of at most 9 bits
  t.dbits = (t.bits + t.passes - 1) / t.passes;
  t.buckets = 1 << t.dbits;
  t.ntiles = (nnz + kTransposeTile - 1) / kTransposeTile;
  t.hist_n = static_cast<int64_t>(t.buckets) * t.ntiles;
  t.scan_blocks = (t.hist_n + 1023) / 1024;
  const size_t n = static_cast<size_t>(nnz > 0 ? nnz : 1);
  size_t off = 0;
  t.off_keys_a = off; off += align256(n * 4);
  t.off_keys_b = off; off += align256(n * 4);
  t.off_pay_a = off;  off += align256(n * 8);
  t.off_pay_b = off;  off += align256(n * 8);
  t.off_val_a = off;  off += align256(n * 4);
  t.off_val_b = off;  off += align256(n * 4);
  t.off_hist = off;   off += align256(static_cast<size_t>(t.hist_n + 1) * 4);
  t.off_offs = off;   off += align256(static_cast<size_t>(t.hist_n + 1) * 4);
  t.off_bsums = off;  off += align256(static_cast<size_t>(t.scan_blocks + 1) * 4);
  t.bytes = off;
  return t;
}

hipError_t launch_transpose(const int32_t* rowptr, const int32_t* colind, const float* vals, int M, int K,
                            int64_t nnz, int32_t* colptr, int32_t* rowind, int32_t* perm, float* vals_t,
                            void* workspace, const TransposePlan& tp, hipStream_t s);
hipError_t launch_sddmm(VType vt, const int32_t* rowptr, const int32_t* colind, int M, int64_t nnz, int N,
                        const void* X, int64_t ld_x, const void* Y, int64_t ld_y, float* out, hipStream_t s);
hipError_t launch_segment_gather(const float* ddst, int64_t ld_ddst, const int32_t* index, int64_t E, int N,
                                 const int32_t* counts, float* dsrc, int64_t ld_dsrc, hipStream_t s);
hipError_t launch_segment_counts(const int32_t* index, int64_t E, int64_t dst_len, int32_t* counts,
                                 hipStream_t s);
hipError_t launch_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, hipStream_t s);
hipError_t launch_gather_probe(const void* table, int64_t row_bytes, const int32_t* idx, int64_t n, int64_t chunk,
                               float* out, int64_t out_len, hipStream_t s);
hipError_t launch_check_csr(const int32_t* rowptr, const int32_t* colind, int64_t M, int64_t K, int64_t nnz,
                            int32_t* flags, hipStream_t s);
hipError_t launch_check_index(const int32_t* index, int64_t E, int64_t dst_len, int32_t* flags, hipStream_t s);

}  // namespace segscan
