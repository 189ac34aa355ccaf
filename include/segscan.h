/*
 * segscan.h — C ABI of the MI355X (gfx950) SpMM / segment_coo library.
 *
 * Plain C: pointers, sizes and enums only. Every device pointer is a HIP device
 * address (hipMalloc / torch CUDA tensor); `stream` is a hipStream_t passed as
 * void* (NULL = the default stream). All calls are stream-ordered and
 * asynchronous; they never synchronise the device and never allocate memory:
 * scratch space comes in through an explicit `workspace` argument whose size is
 * reported by the matching *_workspace_size call.
 *
 * Arguments are validated on the host before any HIP call; a bad argument
 * returns SEGSCAN_EINVAL / SEGSCAN_EUNSUPPORTED with a message available from
 * segscan_last_error() (thread-local), and nothing is launched.
 *
 * Reference interfaces replaced (zhang677/segScan):
 *   segscan_spmm_csr     <- the eight templated SpMM launchers
 *                           test-cg/include/spmm/spmm.cuh:167-202 (ebpr), :310-343 (rbpr),
 *                           :522-591 (ebsr), :621-637 (rbsr) and
 *                           test-cg/include/spmm/spmm_cg.cuh:176-218 (ebpr_cg),
 *                           :333-374 (rbpr_cg), :553-591 (ebsr_cg), :664-692 (rbsr_cg)
 *   segscan_segment_coo  <- the six segment_coo launchers
 *                           new-cg/include/segscan/segscan.cuh:150-184 (segment_coo),
 *                           :327-355 (segment_coo_new), :817-866 (segment_coo_sr),
 *                           :868-916 (segment_coo_noshmem_sr),
 *                           :918-972 (segment_coo_noshmem_lessatom_sr),
 *                           :974-1008 (segment_coo_sorted_sr)
 *   segscan_csr_transpose <- the host transpose test-cg/include/dataloader/dataloader.hpp:123-144
 *                           (+ compressedRow :109-121), done on the device
 *   segscan_sddmm_csr, segscan_spmm_backward_csr, segscan_segment_gather,
 *   segscan_segment_counts: backward pieces with no reference counterpart (the reference has
 *                           no backward); they serve the gradient of the two ops above.
 */
#ifndef SEGSCAN_H_
#define SEGSCAN_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Every enum ends in a *_MAX_ENUM = 0x7fffffff entry: its values then span all of int32
 * (the type stays int-sized), so a caller's out-of-range value is a defined value that the
 * library rejects with SEGSCAN_EINVAL, not undefined behaviour in the C++ implementation. */
typedef enum segscan_status {
  SEGSCAN_OK = 0,
  SEGSCAN_EINVAL = 1,        /* null pointer, negative size, bad stride, misaligned operand */
  SEGSCAN_EUNSUPPORTED = 2,  /* knob combination outside the compiled set */
  SEGSCAN_EHIP = 3,          /* the HIP runtime reported an error (no device, launch failure) */
  SEGSCAN_EWORKSPACE = 4,    /* workspace missing or smaller than *_workspace_size reported */
  SEGSCAN_STATUS_MAX_ENUM = 0x7fffffff
} segscan_status_t;

/* Storage type of the dense operand (B for SpMM, src for segment_coo).
 * Arithmetic is always fp32; outputs are always fp32. */
typedef enum segscan_dtype {
  SEGSCAN_F32 = 0,
  SEGSCAN_BF16 = 1,
  SEGSCAN_F16 = 2,
  SEGSCAN_DTYPE_MAX_ENUM = 0x7fffffff
} segscan_dtype_t;

/* SpMM variants. EB = element (nnz) balanced, RB = row balanced,
 * PR = parallel (cross-lane) reduction, SR = sequential (per-lane) reduction.
 * The *_CG variants take every knob from segscan_knobs_t (spmm_cg.cuh); the
 * PLAIN_* variants ignore the knobs and use the fixed configuration of the
 * plain launchers in spmm.cuh. AUTO picks variant and knobs with a built-in
 * rule for gfx950 that uses segscan_csr_t.max_row_nnz (per-matrix and per-bucket
 * tables live in the Python autotuner). */
typedef enum segscan_spmm_variant {
  SEGSCAN_SPMM_EBPR = 0,        /* csrspmm_parreduce_nnzbalance_cg */
  SEGSCAN_SPMM_RBPR = 1,        /* csrspmm_parreduce_rowbalance_cg */
  SEGSCAN_SPMM_EBSR = 2,        /* csrspmm_rowcaching_nnzbalance_cg */
  SEGSCAN_SPMM_RBSR = 3,        /* csrspmm_seqreduce_rowbalance_cg */
  SEGSCAN_SPMM_PLAIN_EBPR = 4,  /* csrspmm_parreduce_nnzbalance */
  SEGSCAN_SPMM_PLAIN_RBPR = 5,  /* csrspmm_parreduce_rowbalance */
  SEGSCAN_SPMM_PLAIN_EBSR = 6,  /* csrspmm_rowcaching_nnzbalance */
  SEGSCAN_SPMM_PLAIN_RBSR = 7,  /* csrspmm_seqreduce_rowbalance */
  SEGSCAN_SPMM_AUTO = 8,
  SEGSCAN_SPMM_MAX_ENUM = 0x7fffffff
} segscan_spmm_variant_t;

/* segment_coo variants (new-cg/include/segscan/segscan.cuh). AUTO runs the
 * sequential kernel of SEGSCAN_SEG_LESSATOM_SR (which describe() reports) with
 * owner stores and a gfx950 shape: two columns per lane, ~64 KB of src per chunk,
 * 8 (4-byte src) or 16 (2-byte src) waves per CU, or one chunk per group where
 * that grid would need at most two rounds of chunks. Under AUTO, thread_nz and
 * block_numer / block_denom may set its chunk and grid. */
typedef enum segscan_seg_variant {
  SEGSCAN_SEG_PR = 0,               /* segment_coo */
  SEGSCAN_SEG_PR_NEW = 1,           /* segment_coo_new */
  SEGSCAN_SEG_SR = 2,               /* segment_coo_sr */
  SEGSCAN_SEG_NOSHMEM_SR = 3,       /* segment_coo_noshmem_sr */
  SEGSCAN_SEG_LESSATOM_SR = 4,      /* segment_coo_noshmem_lessatom_sr */
  SEGSCAN_SEG_SORTED_SR = 5,        /* segment_coo_sorted_sr */
  SEGSCAN_SEG_AUTO = 6,
  SEGSCAN_SEG_MAX_ENUM = 0x7fffffff
} segscan_seg_variant_t;

/* Write policy of the sequential-reduction kernels. */
typedef enum segscan_write_policy {
  SEGSCAN_WRITE_DEFAULT = 0,  /* the variant's own policy */
  SEGSCAN_WRITE_ATOMIC = 1,   /* every flushed row is an atomic add (output zero-filled first) */
  SEGSCAN_WRITE_OWNER = 2,    /* rows are stored once by the chunk that owns them (empty rows
                                 zero-filled by their owner); a row that crosses a chunk
                                 boundary gets a carry added by a fix-up pass; needs workspace */
  SEGSCAN_WRITE_LESSATOM = 3, /* rows that start and end inside a chunk are plain stores, the
                                 chunk's first and last row are atomic adds (output zero-filled
                                 first); the reference's segment_coo_noshmem_lessatom_sr policy */
  SEGSCAN_WRITE_MAX_ENUM = 0x7fffffff
} segscan_write_policy_t;

/* Tuning knobs, named as in the reference templates. A field left 0 takes the
 * variant's default (the reference's best static configuration); since 0 means
 * "default", a factor of 0 (size 1) is written as -1.
 *   group_factor     group_size = 1 << group_factor lanes (1..64) that reduce together
 *                    (PR) or that share one dense row panel (SR)
 *   tile_factor      tile_size = 1 << tile_factor (1..64): nnz per tile (PR) and the
 *                    column-panel width per grid slice (PR, RBSR); README "tail_factor"
 *   thread_per_block lanes per workgroup (rounded up to a multiple of 64, <= 1024)
 *   block_numer/block_denom  block_factor: workers = rows (or elements) x numer / denom
 *   coarsen_factor   columns per lane (1, 2, 4, 8); 0 = derived from N as the reference does
 *   thread_nz        nnz tiles per group per chunk (SR variants)
 *   dtile_size, dtile_num, block_dim_y   segment_coo_sorted_sr / segment_coo_new shape knobs
 *   write_policy     segscan_write_policy_t */
typedef struct segscan_knobs {
  int32_t group_factor;
  int32_t tile_factor;
  int32_t thread_per_block;
  int32_t block_numer;
  int32_t block_denom;
  int32_t coarsen_factor;
  int32_t thread_nz;
  int32_t dtile_size;
  int32_t dtile_num;
  int32_t block_dim_y;
  int32_t write_policy;
} segscan_knobs_t;

/* CSR matrix A (M x K), zero-based, int32 indices, rows sorted by construction.
 * vals may be NULL: every value is then 1 (reference __guard_load_default_one,
 * test-cg/include/util/check.cuh:104-111). Mirrors SpMatCsrDescr_t
 * (test-cg/include/dataloader/dataloader.hpp:146-168). */
typedef struct segscan_csr {
  int64_t nrow;
  int64_t ncol;
  int64_t nnz;
  const int32_t* rowptr; /* [nrow + 1], device */
  const int32_t* colind; /* [nnz], device */
  const float* vals;     /* [nnz], device, or NULL */
  int64_t max_row_nnz;   /* longest row, or 0 = unknown; guides SEGSCAN_SPMM_AUTO only
                            (0 <= max_row_nnz <= nnz). Zero-initialise the struct. */
  const int32_t* vals_map; /* NULL, or [nnz] device: entry p's value is vals[vals_map[p]], for
                              values kept in another order (e.g. the transpose's, whose
                              backward then reads and writes them coalesced; vals_map is
                              the inverse of segscan_csr_transpose's perm). SpMM only. */
} segscan_csr_t;

/* Flags. Without SEGSCAN_ACCUMULATE the call computes C = A.B (resp. dst =
 * segment sums): the library writes every output element, so callers need not
 * zero C (the reference requires a pre-zeroed C, check.cuh:71). With it the
 * call computes C += A.B. */
#define SEGSCAN_ACCUMULATE 1u
/* segscan_spmm_backward_csr(_rowmap): dvals is written in At's order (dvals[q] for At entry
 * q) instead of A's (dvals[perm[q]]): coalesced stores, for values kept in At's order. With
 * At.vals set (or all-ones values) the call then reads no perm at all. */
#define SEGSCAN_DVALS_AT_ORDER 2u

/* Launch geometry a call would use; filled by the *_describe functions (host
 * only, no HIP call), for tests and for the autotuner's bookkeeping. */
typedef struct segscan_launch_info {
  int32_t variant;          /* resolved variant (never AUTO) */
  int32_t group_size;
  int32_t tile_size;
  int32_t thread_per_block; /* lanes per workgroup actually launched */
  int32_t coarsen_factor;
  int32_t thread_nz;
  int32_t write_policy;     /* resolved: 1 atomic, 2 owner */
  int32_t chunk;            /* nnz (or elements) per partition chunk, 0 if none */
  int64_t grid_x;
  int64_t grid_y;
  int64_t nchunks;
  size_t workspace_bytes;
} segscan_launch_info_t;

/* ---- SpMM: C[M x N] (=|+=) A[M x K] . B[K x N] --------------------------------------
 * B row stride ldb (elements) in [N, 2^31 - 1]; C row stride ldc >= N. */
segscan_status_t segscan_spmm_describe(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                       segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                       segscan_launch_info_t* info);
segscan_status_t segscan_spmm_workspace_size(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                             segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                             size_t* bytes);
segscan_status_t segscan_spmm_csr(const segscan_csr_t* A, int64_t N,
                                  const void* B, int64_t ldb, segscan_dtype_t b_dtype,
                                  float* C, int64_t ldc,
                                  segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                  uint32_t flags, void* workspace, size_t workspace_bytes, void* stream);

/* Allocation-free SpMM: segscan_spmm_csr without a workspace argument. No counterpart in
 * the reference; it is for callers that cannot allocate (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc). The scratch a plan needs (ebsr
 * carries, ebpr tile rows) comes from a static device array of
 * SEGSCAN_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the call allocates
 * nothing. A plan that needs more (segscan_spmm_workspace_size) returns
 * SEGSCAN_EWORKSPACE before any HIP call. Every call on one GPU shares that GPU's array,
 * so calls that overlap on one GPU (concurrent streams or threads) race; calls ordered
 * on one stream do not. `stream` must belong to the current device. Results equal
 * segscan_spmm_csr's for the same arguments. */
#define SEGSCAN_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
segscan_status_t segscan_spmm_csr_noalloc(const segscan_csr_t* A, int64_t N,
                                          const void* B, int64_t ldb, segscan_dtype_t b_dtype,
                                          float* C, int64_t ldc,
                                          segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                          uint32_t flags, void* stream);

/* ---- segment_coo: dst[s, :] (=|+=) sum_{e : index[e] == s} src[e, :] ----------------
 * index: [E] int32, sorted non-decreasing, values in [0, dst_len). src: E x N, row
 * stride ld_src in [N, 2^31 - 1]. dst: dst_len x N fp32, row stride ld_dst. */
segscan_status_t segscan_segment_coo_describe(int64_t E, int64_t N, int64_t dst_len, segscan_dtype_t src_dtype,
                                              segscan_seg_variant_t variant, const segscan_knobs_t* knobs,
                                              segscan_launch_info_t* info);
segscan_status_t segscan_segment_coo_workspace_size(int64_t E, int64_t N, int64_t dst_len, segscan_dtype_t src_dtype,
                                                    segscan_seg_variant_t variant, const segscan_knobs_t* knobs,
                                                    size_t* bytes);
segscan_status_t segscan_segment_coo(const void* src, int64_t ld_src, segscan_dtype_t src_dtype,
                                     const int32_t* index, int64_t E, int64_t N,
                                     float* dst, int64_t ld_dst, int64_t dst_len,
                                     segscan_seg_variant_t variant, const segscan_knobs_t* knobs,
                                     uint32_t flags, void* workspace, size_t workspace_bytes, void* stream);

/* ---- backward pieces ------------------------------------------------------------------
 * CSR -> CSC of the M x K matrix A: colptr[K+1], rowind[nnz], perm[nnz] with
 * perm[q] = CSR position of CSC entry q; within a column, entries keep CSR
 * (row) order, i.e. the result equals a stable sort of the nnz by column. vals_t
 * may be NULL; otherwise vals_t[q] = vals[perm[q]] (1 when vals is NULL). */
segscan_status_t segscan_csr_transpose_workspace_size(const segscan_csr_t* A, size_t* bytes);
segscan_status_t segscan_csr_transpose(const segscan_csr_t* A, int32_t* colptr, int32_t* rowind,
                                       int32_t* perm, float* vals_t,
                                       void* workspace, size_t workspace_bytes, void* stream);
/* SDDMM on A's pattern: out[p] = sum_j X[row(p), j] * Y[colind[p], j]  (p < nnz).
 * X: M x N (ld_x), Y: K x N (ld_y), both of dtype xy_dtype; out fp32. */
segscan_status_t segscan_sddmm_csr(const segscan_csr_t* A, int64_t N,
                                   const void* X, int64_t ld_x, const void* Y, int64_t ld_y,
                                   segscan_dtype_t xy_dtype, float* out, void* stream);
/* Fused backward of C = A.B (pattern fixed; values and B learnable) in one walk over
 * At = A^T, which gathers each dC row once for both products:
 *   dB[j, :] (=|+=) sum over p with colind[p] == j of vals[p] * dC[row(p), :]   (dB = A^T.dC)
 *   dvals[p]  = sum_n dC[row(p), n] * B[colind[p], n]                          (SDDMM on A's pattern)
 * At and perm are the outputs of segscan_csr_transpose(A): At.rowptr = colptr,
 * At.colind = rowind, At.nrow = K, At.ncol = M. Values: At.vals when non-NULL (At
 * order, e.g. the transpose's vals_t: read coalesced; `vals` is then ignored), else
 * vals in A's (CSR) order read through perm (a random 4-byte read per entry), else
 * all ones. dC: M x N (ld_dc) of dc_dtype, B: K x N (ld_b) of b_dtype: the same
 * type, or an fp32 dC with a 16-bit B (the operand kept in bf16/fp16 storage);
 * dB (K x N, ld_db) and dvals (nnz) are fp32, and either may be NULL (not
 * computed). One group of lanes covers all N columns, so N <= 512 (fp32: N <= 256
 * unless N is a multiple of 8); wider N returns SEGSCAN_EUNSUPPORTED (use
 * segscan_spmm_csr on At plus segscan_sddmm_csr). Knobs: group_factor,
 * coarsen_factor, thread_nz, thread_per_block (<= 256), block_numer / block_denom
 * (workgroups = K * numer / denom), as for SEGSCAN_SPMM_EBSR; NULL = built-in shape
 * (chunks of 16 tiles, 8 where At rows average under 8 entries; with such short rows
 * and a bf16 B, 4 columns per lane and 8-32 lanes, each row's B row is loaded
 * together with the batch of dC rows that reaches it).
 * flags: SEGSCAN_ACCUMULATE adds into dB; dvals is always overwritten. Workspace
 * (only when dB != NULL): segscan_spmm_backward_workspace_size. No counterpart in
 * the reference, whose kernels are forward only. */
segscan_status_t segscan_spmm_backward_workspace_size(const segscan_csr_t* At, int64_t N, segscan_dtype_t dc_dtype,
                                                      const segscan_knobs_t* knobs, size_t* bytes);
segscan_status_t segscan_spmm_backward_csr(const segscan_csr_t* At, const int32_t* perm, const float* vals,
                                           int64_t N, const void* dC, int64_t ld_dc, segscan_dtype_t dc_dtype,
                                           const void* B, int64_t ld_b, segscan_dtype_t b_dtype,
                                           float* dB, int64_t ld_db, float* dvals,
                                           const segscan_knobs_t* knobs, uint32_t flags,
                                           void* workspace, size_t workspace_bytes, void* stream);
/* segscan_spmm_backward_csr with dC given through a row map: A's row r reads
 * dC[dc_map[r], :] (dc_map has At.ncol = M entries, each a valid row of dC, which
 * may have any number of rows; the map is not checked on the device). The
 * backward of C = A.B followed by a per-segment sum pooling P = segment_sum(C,
 * index) takes dC = dP[index] this way: dP (one row per segment) is read in place
 * of an E x N gathered copy, so the gather kernel and its HBM traffic go away.
 * Everything else is as segscan_spmm_backward_csr (same workspace size). No
 * counterpart in the reference. */
segscan_status_t segscan_spmm_backward_csr_rowmap(const segscan_csr_t* At, const int32_t* perm, const float* vals,
                                                  int64_t N, const void* dC, int64_t ld_dc, segscan_dtype_t dc_dtype,
                                                  const int32_t* dc_map, const void* B, int64_t ld_b,
                                                  segscan_dtype_t b_dtype, float* dB, int64_t ld_db, float* dvals,
                                                  const segscan_knobs_t* knobs, uint32_t flags, void* workspace,
                                                  size_t workspace_bytes, void* stream);
/* Gather (backward of segment_coo): dsrc[e, :] = ddst[index[e], :] * scale(index[e]),
 * scale = 1 / max(counts[s], 1) when counts != NULL (mean pooling), else 1. */
segscan_status_t segscan_segment_gather(const float* ddst, int64_t ld_ddst, const int32_t* index, int64_t E,
                                        int64_t N, const int32_t* counts, float* dsrc, int64_t ld_dsrc,
                                        void* stream);
/* inv[perm[i]] = i for i < n (perm a permutation of [0, n)): the vals_map that reads values
 * kept in the transpose's order from A's entries. */
segscan_status_t segscan_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, void* stream);
/* A fixed permutation of per-entry values, out[i] = in[src[i]] (src a permutation of [0, n)),
 * in two streaming passes instead of n random 4-byte reads: e.g. learnable values kept in the
 * transpose's order, made into A's order once per step (src = segscan_invert_permutation(perm))
 * so the forward SpMM reads them coalesced. Creation reads src back, checks it is a
 * permutation (SEGSCAN_EINVAL otherwise), builds the plan's tables on the host and uploads
 * them (8 bytes per value + two 4-byte-per-value buffers, staging and an A-order copy for
 * segscan_spmm_csr_permuted, allocated here); it
 * synchronises `stream`. Runs of one plan must not overlap (they share the staging buffer);
 * in and out must not alias; the current device must be the one the plan was created on. No counterpart in the reference (it has no backward). */
typedef struct segscan_permute_plan* segscan_permute_plan_t;
segscan_status_t segscan_permute_plan_create(const int32_t* src, int64_t n, void* stream,
                                             segscan_permute_plan_t* plan);
segscan_status_t segscan_permute_run(segscan_permute_plan_t plan, const float* in, float* out, void* stream);
segscan_status_t segscan_permute_plan_destroy(segscan_permute_plan_t plan);
/* SpMM whose values arrive through a permutation plan: C (=|+=) A'.B where A' is A's
 * pattern with value p = vals_src[src[p]] (src the plan's permutation, n = A.nnz; A.vals and
 * A.vals_map must be NULL). Equals segscan_permute_run into a temporary followed by
 * segscan_spmm_csr on it, bit for bit, with the same variant, knobs, flags and workspace.
 * When the resolved launch is ebsr and each workgroup round covers a part of one of the
 * plan's 8192-value buckets (thread_nz x group size x workgroups' groups dividing 8192, at
 * least 2048; groups of 8 to 64 lanes), the plan's second pass is fused into the SpMM: each workgroup loads its
 * bucket of staged values into LDS and reads them through 16-bit slots, so the A-order copy
 * is never written or read. Other launches run both passes into a buffer the plan holds.
 * Calls on one plan must not overlap. No counterpart in the reference (its values are
 * never learned). */
segscan_status_t segscan_spmm_csr_permuted(const segscan_csr_t* A, segscan_permute_plan_t plan,
                                           const float* vals_src, int64_t N, const void* B, int64_t ldb,
                                           segscan_dtype_t b_dtype, float* C, int64_t ldc,
                                           segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                           uint32_t flags, void* workspace, size_t workspace_bytes,
                                           void* stream);
/* counts[s] = #{e : index[e] == s}, s < dst_len (sorted index). */
segscan_status_t segscan_segment_counts(const int32_t* index, int64_t E, int64_t dst_len, int32_t* counts,
                                        void* stream);

/* ---- batched graphs: one persistent launch per step ---------------------------------------
 * A batch of graphs as one block-diagonal CSR A (square, M = K nodes; graph g owns rows and
 * columns [graph_offsets[g], graph_offsets[g+1]), every entry's column in its row's graph).
 * One run computes, for node features X (M x N of x_dtype) and the gradient dP (n_graphs x
 * N, fp32) of the per-graph sum pooling:
 *   H = A.X (M x N fp32)   P[g] = sum of the H rows of graph g (n_graphs x N fp32)
 *   dX = A^T.dH and dvals = SDDMM(dH, X) on A's pattern, for dH = dP[graph]
 * in one kernel launch (no fix-up pass, no memset; capturable in a hipGraph). With the
 * block-diagonal structure, dX[j] = colsum(A)[j] * dP[g(j)]; the column sums are taken from
 * A's values at plan creation (segscan_batch_plan_set_values after the values change).
 * H, dX and dvals may be NULL (not computed); dP NULL runs the forward only. N <= 512
 * (fp32: <= 256). Rows of X, H, P, dX and dP must be aligned to the lane's columns (as
 * torch tensors are). The plan borrows A's device arrays (keep them alive) and owns device
 * memory of its own (segscan_batch_plan_describe: plan_bytes), allocated at creation;
 * creation reads rowptr back and synchronises `stream`. Runs of one plan must not overlap
 * (they share its pool accumulator): order them on one stream. No counterpart in the
 * reference (it has no batching); it replaces a per-graph launch loop of its SpMM and
 * segment_coo (spmm_cg.cuh:553-591, segscan.cuh:974-1008). Knobs (NULL = built-in):
 * coarsen_factor, group_factor (the lane shape), thread_nz (entries per group per work
 * item / group size), block_numer / block_denom (scale of the resident grid). */
typedef struct segscan_batch_plan* segscan_batch_plan_t;
typedef struct segscan_batch_info {
  int64_t nitems;          /* work items (row ranges) of the forward */
  int64_t grid;            /* workgroups launched */
  int32_t group_size;
  int32_t coarsen_factor;
  int32_t thread_per_block;
  int64_t item_target;     /* entries + rows per item */
  size_t plan_bytes;       /* device memory held by the plan */
} segscan_batch_info_t;

segscan_status_t segscan_batch_plan_create(const segscan_csr_t* A, const int64_t* graph_offsets, int64_t n_graphs,
                                           int64_t N, segscan_dtype_t x_dtype, const segscan_knobs_t* knobs,
                                           void* stream, segscan_batch_plan_t* plan);
segscan_status_t segscan_batch_plan_set_values(segscan_batch_plan_t plan, const float* vals, void* stream);
segscan_status_t segscan_batch_plan_describe(segscan_batch_plan_t plan, segscan_batch_info_t* info);
segscan_status_t segscan_batch_plan_run(segscan_batch_plan_t plan, const void* X, int64_t ld_x,
                                        const float* dP, int64_t ld_dp, float* H, int64_t ld_h,
                                        float* P, int64_t ld_p, float* dX, int64_t ld_dx, float* dvals,
                                        void* stream);
segscan_status_t segscan_batch_plan_destroy(segscan_batch_plan_t plan);

/* ---- tuning tables and the autotuner ---------------------------------------------------
 * Replace the reference's parameter sweeps (test-cg/src/params.md:1-438, run under Nsight by
 * test-cg/src/test.py:42-59) and its static-table analysis (test-cg/src/static.py:15-35).
 * A table maps keys to a variant and knobs and is stored as JSON (the format
 * segscan_amd/autotune.py reads and writes). Keys start with segscan_tune_prefix(), the
 * GPU architecture and the ROCm release ("gfx950|rocm7.2"); load drops entries of another
 * architecture or release (*dropped counts them), so a table is re-tuned, not misapplied,
 * after a ROCm upgrade. Lookup tries the matrix's own entry (rows, columns, nnz, longest
 * row), then its bucket (N rounded up to a power of two, row-skew class, size class); with
 * neither it returns SEGSCAN_SPMM_AUTO and *found = 0 (1 = bucket, 2 = matrix). Lookup
 * needs A->max_row_nnz (or nnz = 0) and makes no HIP call.
 * segscan_autotune_spmm times the built-in AUTO choice and a fixed candidate list of
 * variants and knobs on the caller's data (C is overwritten; reps timed calls each, median,
 * HIP events on `stream` with a short spin kernel queued before each start event so the
 * host's issue cost overlaps it), returns the fastest and, with a table, records it under
 * the matrix's key. It synchronises `stream`; candidates whose workspace exceeds
 * workspace_bytes are skipped (segscan_autotune_workspace_size: the largest any
 * candidate needs). A->max_row_nnz = 0 is read back from rowptr. */
typedef struct segscan_tune_table* segscan_tune_table_t;
const char* segscan_tune_prefix(void);
segscan_status_t segscan_tune_table_create(segscan_tune_table_t* table);
segscan_status_t segscan_tune_table_destroy(segscan_tune_table_t table);
segscan_status_t segscan_tune_table_load(segscan_tune_table_t table, const char* path, int64_t* loaded,
                                         int64_t* dropped);
segscan_status_t segscan_tune_table_save(segscan_tune_table_t table, const char* path);
segscan_status_t segscan_tune_table_lookup(segscan_tune_table_t table, const segscan_csr_t* A, int64_t N,
                                           segscan_dtype_t b_dtype, segscan_spmm_variant_t* variant,
                                           segscan_knobs_t* knobs, int32_t* found);
segscan_status_t segscan_autotune_workspace_size(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                                 size_t* bytes);
segscan_status_t segscan_autotune_spmm(segscan_tune_table_t table, const segscan_csr_t* A, int64_t N,
                                       const void* B, int64_t ldb, segscan_dtype_t b_dtype, float* C, int64_t ldc,
                                       int32_t reps, void* workspace, size_t workspace_bytes, void* stream,
                                       segscan_spmm_variant_t* variant, segscan_knobs_t* knobs, float* ms);

/* ---- device-side input checks (optional; the compute calls trust their inputs) ---------
 * Writes into *flags_dev (device int32) a bit set: 1 rowptr not monotone or not
 * spanning [0, nnz], 2 column index out of [0, ncol), 4 columns not strictly increasing in a row. */
segscan_status_t segscan_check_csr(const segscan_csr_t* A, int32_t* flags_dev, void* stream);
/* Bit 1: index not sorted, bit 2: index value out of [0, dst_len). */
segscan_status_t segscan_check_index(const int32_t* index, int64_t E, int64_t dst_len, int32_t* flags_dev,
                                     void* stream);

/* ---- gather probe (measurement) -----------------------------------------------------------
 * Reads rows idx[i] (i < n) of a dense table (table_rows x row_bytes, row_bytes 128 or 256)
 * as the SpMM's dense-row gather does: groups of row_bytes / 16 lanes, 16 bytes per lane,
 * 16 rows in flight per group, `chunk` indices per group; writes one checksum per lane to
 * out (segscan_probe_gather_out_len floats). The indices are not checked: each must be
 * below table_rows. It gives the random-row read ceiling the SpMM kernels are compared
 * with; no counterpart in the reference. The timing hook below times it too. */
segscan_status_t segscan_probe_gather_out_len(int64_t row_bytes, int64_t n, int64_t chunk, int64_t* out_len);
segscan_status_t segscan_probe_gather(const void* table, int64_t table_rows, int64_t row_bytes, const int32_t* idx,
                                      int64_t n, int64_t chunk, float* out, int64_t out_len, void* stream);

/* ---- measurement hook ---------------------------------------------------------------------
 * With non-NULL hipEvent_t handles, every later segscan_spmm_csr / segscan_segment_coo /
 * segscan_spmm_backward_csr call on this host thread records `start` on its stream right before its main kernel
 * and `stop` right after it, so the pair times that kernel alone (no zero fill,
 * partition, transpose or fix-up pass). NULL, NULL turns the hook off. */
void segscan_set_timing_events(void* start, void* stop);

/* ---- Matrix Market input (host only) -------------------------------------------------------
 * Replaces read_mtx_file, test-cg/include/dataloader/dataloader.hpp:22-107: coordinate
 * files, 1-based entries made 0-based, symmetric (also skew-symmetric and hermitian)
 * off-diagonal entries mirrored then sorted and de-duplicated; general files sorted
 * by (row, col), duplicates kept. Without SEGSCAN_MTX_KEEP_VALUES every value is 1,
 * as in the reference (dataloader.hpp:154-155); with it the file's values are kept
 * (pattern files: 1, skew-symmetric mirrors: negated, complex: real part).
 * SEGSCAN_MTX_REFERENCE_QUIRKS reproduces read_mtx_file exactly: only "symmetric"
 * files are mirrored, and the CSR keeps only the first <declared nnz> entries of the
 * sorted list (its loop bound, dataloader.hpp:95), which empties the last rows of a
 * mirrored matrix. Host arrays are malloc'd; release them with segscan_host_csr_free.
 * Errors return SEGSCAN_EINVAL with text in segscan_mtx_last_error(). */
#define SEGSCAN_MTX_KEEP_VALUES 1u
#define SEGSCAN_MTX_REFERENCE_QUIRKS 2u
typedef struct segscan_host_csr {
  int64_t nrow;
  int64_t ncol;
  int64_t nnz;
  int32_t* rowptr; /* [nrow + 1], host */
  int32_t* colind; /* [nnz], host */
  float* vals;     /* [nnz], host */
} segscan_host_csr_t;

segscan_status_t segscan_mtx_read(const char* path, uint32_t flags, segscan_host_csr_t* out);
void segscan_host_csr_free(segscan_host_csr_t* csr);
const char* segscan_mtx_last_error(void);

/* ---- misc -------------------------------------------------------------------------------- */
const char* segscan_status_string(segscan_status_t status);
const char* segscan_last_error(void);
int32_t segscan_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SEGSCAN_H_ */
