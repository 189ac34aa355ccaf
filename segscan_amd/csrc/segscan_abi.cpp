// Host side of the C ABI: argument validation, knob resolution (reference
// defaults, plain-launcher configurations), launch geometry, workspace layout
// and kernel dispatch. No HIP call is made before every argument has been
// validated, so every error path here runs without a GPU.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/segscan.h"
#include "abi_internal.hpp"
#include "kernel_params.hpp"

namespace {

using namespace segscan;

thread_local std::string g_last_error;
thread_local hipEvent_t g_ev_start = nullptr;
thread_local hipEvent_t g_ev_stop = nullptr;

void mark_start(hipStream_t s) {
  if (g_ev_start) (void)hipEventRecord(g_ev_start, s);
}
void mark_stop(hipStream_t s) {
  if (g_ev_stop) (void)hipEventRecord(g_ev_stop, s);
}

segscan_status_t fail(segscan_status_t st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

segscan_status_t ok() {
  g_last_error.clear();
  return SEGSCAN_OK;
}

segscan_status_t hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return ok();
  return fail(SEGSCAN_EHIP, "%s: %s", what, hipGetErrorString(e));
}

constexpr int64_t kMaxI32 = 2147483647LL;
constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }
int round_up64(int x) { return (x + 63) / 64 * 64; }

size_t dtype_size(segscan_dtype_t t) { return t == SEGSCAN_F32 ? 4 : 2; }
bool dtype_valid(int t) { return t == SEGSCAN_F32 || t == SEGSCAN_BF16 || t == SEGSCAN_F16; }

// Elements per 128-bit load.
int vec_elems(segscan_dtype_t t) { return static_cast<int>(16 / dtype_size(t)); }

// Largest power-of-two column count <= cap that keeps a lane's columns aligned
// for vector access: divides the leading dimensions and the base addresses.
int aligned_cf(int cap, const void* B, int64_t ldb, size_t es, const float* C, int64_t ldc) {
  int cf = 1;
  for (int c = 2; c <= cap; c <<= 1) {
    const size_t bvec = std::min<size_t>(es * c, 16);
    const size_t cvec = std::min<size_t>(4 * c, 16);
    if (ldb % c) break;
    if (ldc % c) break;
    if (B && reinterpret_cast<uintptr_t>(B) % bvec) break;
    if (C && reinterpret_cast<uintptr_t>(C) % cvec) break;
    cf = c;
  }
  return cf;
}

// Largest power of two <= cap dividing n (the reference's N % 4 / N % 2 rule).
int divisor_cf(int64_t n, int cap) {
  int cf = 1;
  for (int c = 2; c <= cap; c <<= 1)
    if (n % c == 0) cf = c;
  return cf;
}

struct Knobs {
  int group_factor = 0, tile_factor = 0, tpb = 0, bn = 0, bd = 0, cf = 0, thread_nz = 0;
  int dtile_size = 0, dtile_num = 0, block_dim_y = 0, policy = 0;
};

Knobs from_c(const segscan_knobs_t* k) {
  Knobs r;
  if (!k) return r;
  r.group_factor = k->group_factor;
  r.tile_factor = k->tile_factor;
  r.tpb = k->thread_per_block;
  r.bn = k->block_numer;
  r.bd = k->block_denom;
  r.cf = k->coarsen_factor;
  r.thread_nz = k->thread_nz;
  r.dtile_size = k->dtile_size;
  r.dtile_num = k->dtile_num;
  r.block_dim_y = k->block_dim_y;
  r.policy = k->write_policy;
  return r;
}

// Fills zero fields with `d`.
void defaults(Knobs& k, const Knobs& d) {
  if (!k.group_factor) k.group_factor = d.group_factor;
  if (!k.tile_factor) k.tile_factor = d.tile_factor;
  if (!k.tpb) k.tpb = d.tpb;
  if (!k.bn) k.bn = d.bn;
  if (!k.bd) k.bd = d.bd;
  if (!k.cf) k.cf = d.cf;
  if (!k.thread_nz) k.thread_nz = d.thread_nz;
  if (!k.dtile_size) k.dtile_size = d.dtile_size;
  if (!k.dtile_num) k.dtile_num = d.dtile_num;
  if (!k.block_dim_y) k.block_dim_y = d.block_dim_y;
  if (!k.policy) k.policy = d.policy;
}

Knobs mk(int gf, int tf, int tpb, int bn, int bd, int cf, int tnz, int policy) {
  Knobs k;
  k.group_factor = gf;
  k.tile_factor = tf;
  k.tpb = tpb;
  k.bn = bn;
  k.bd = bd;
  k.cf = cf;
  k.thread_nz = tnz;
  k.policy = policy;
  return k;
}

segscan_status_t check_common_knobs(const Knobs& k) {
  if (k.tpb < 1 || k.tpb > 1024) return fail(SEGSCAN_EUNSUPPORTED, "thread_per_block %d outside [1, 1024]", k.tpb);
  if (k.bn < 1 || k.bd < 1) return fail(SEGSCAN_EINVAL, "block_factor %d/%d must be positive", k.bn, k.bd);
  if (k.thread_nz < 1 || k.thread_nz > (1 << 20))
    return fail(SEGSCAN_EUNSUPPORTED, "thread_nz %d outside [1, 2^20]", k.thread_nz);
  return SEGSCAN_OK;
}

// A knob factor f encodes size 1 << f; the value -1 is accepted as an explicit
// "0" (size 1), because 0 in the struct means "default".
int factor_value(int f) { return f < 0 ? 0 : f; }

int64_t workers(int64_t n, const Knobs& k) {
  // The reference computes (int)((float)n * numer / denom); integer arithmetic here.
  int64_t w = n * k.bn / k.bd;
  return std::max<int64_t>(1, w);
}

// ---------------------------------------------------------------------------------------
// SpMM planning

enum Family { kEbpr = 0, kRbpr = 1, kEbsr = 2, kRbsr = 3 };

struct SpmmPlan {
  int variant = 0;
  Family fam = kEbpr;
  int G = 1, T = 1, tpb = 64, CF = 1, tnz = 1, policy = 0;
  int X = 1, P = 1, Y = 1, slice = 1;
  int64_t grid_x = 1, grid_y = 1;
  int block = 64;
  int64_t ntiles = 0, chunk = 0, nchunks = 0, total_groups = 0;
  int tpr = 1, rows_per_block = 1;
  bool zero_fill = false;   // output zero-filled before the kernel
  size_t ws_bytes = 0;
  size_t off_rows = 0, off_prev = 0, off_crow = 0, off_cval = 0;
};

// Reference defaults: README.md:168-210 best static tuples for the cg variants;
// spmm.cuh plain launchers for the plain ones.
segscan_status_t plan_spmm(int64_t M, int64_t nnz, int64_t max_row, int64_t N, segscan_dtype_t bt, int variant,
                           const Knobs& user, int cf_cap, SpmmPlan& pl) {
  if (variant < 0 || variant > SEGSCAN_SPMM_AUTO) return fail(SEGSCAN_EINVAL, "unknown SpMM variant %d", variant);
  const int vec = vec_elems(bt);
  Knobs k = user;
  int v = variant;
  if (v == SEGSCAN_SPMM_AUTO) {
    // Built-in choice for gfx950, fitted to a sweep over uniform, power-law,
    // banded, rectangular and dense-row matrices at N = 4..128 (fp32, bf16):
    //  * rows known to be even (longest row <= 4 x mean + 32): one group of
    //    lanes per row — row-balanced parallel reduction with a group of about
    //    half a row for N <= 8 (and N <= 16 with rows of >= 24 nonzeros), else
    //    a 64-column slice of the dense row per group of lanes with up to 4
    //    columns per lane, whatever the storage type (row-balanced sequential
    //    reduction);
    //  * otherwise (skewed, or the longest row unknown): element-balanced
    //    sequential reduction with 4-byte row slices per lane (one fp32 column,
    //    two 16-bit columns when N >= 64), wider only when N needs more than 64
    //    lanes, so one group of up to 64 lanes covers a dense row; owner
    //    stores, chunks of >= 64 nonzeros with thread_nz the largest power of
    //    two <= 1.125 nnz / 2^18 (within [64 / group, 32]), and up to 2048
    //    workgroups of 256 lanes (32 waves per CU).
    const double mean = M > 0 ? static_cast<double>(nnz) / static_cast<double>(M) : 0.0;
    const bool even = max_row > 0 && static_cast<double>(max_row) <= 4.0 * mean + 32.0;
    int cf = std::min(vec, divisor_cf(N, vec));
    if (even && (N <= 8 || (N <= 16 && mean >= 24.0))) {
      v = SEGSCAN_SPMM_RBPR;
      int gf = 2;
      while (gf < 4 && static_cast<double>(int64_t{1} << (gf + 1)) <= mean / 2.0) ++gf;
      k = mk(gf, gf, 256, 1, 2, 0, 1, 0);
    } else if (even) {
      // 8-byte lane slices for 16-bit B too (16 lanes per 64-column slice, the cooperative
      // walk): banded 120000^2 bf16 N = 64 / 128 45.8 / 83.5 -> 34.5 / 59.5 us against
      // 16-byte slices
      // Narrower slices until a row has >= 8 lanes, the cooperative walk's minimum (N = 16:
      // 2 columns per lane, 8 lanes; geomean 0.935 x the 4-column shape's time over the
      // sweep's seven even matrices).
      v = SEGSCAN_SPMM_RBSR;
      int rcf = std::min(cf, divisor_cf(N, 4));
      while (rcf > 1 && ceil_div(std::min<int64_t>(N, 64), rcf) < 8) rcf >>= 1;
      k = mk(-1, 6, 256, 1, 1, rcf, 1, 0);
    } else {
      v = SEGSCAN_SPMM_EBSR;
      cf = (bt != SEGSCAN_F32 && N >= 64) ? 2 : 1;
      while (cf < 8 && ceil_div(N, cf) > 64) cf <<= 1;
      int lanes = static_cast<int>(std::min<int64_t>(64, ceil_div(N, cf)));
      int g = 1;
      while (g < lanes) g <<= 1;
      int gf = 0;
      while ((1 << gf) < g) ++gf;
      // the largest power of two <= 1.125 nnz / 2^18 (sweep optima: 4.19M nnz -> 16,
      // 2.1M -> 8, 1.05M -> 4), within [64 / g, 32]
      int tnz = std::max(1, 64 / g);
      while (tnz < 32 && (static_cast<int64_t>(tnz) << 19) <= nnz + nnz / 8) tnz <<= 1;
      const int64_t chunk = static_cast<int64_t>(tnz) * g;
      const int64_t gpb = 256 / g;
      const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(ceil_div(nnz, chunk), gpb),
                                                                     256 * 32 * 64 / 256));
      k = mk(gf == 0 ? -1 : gf, 5, 256, 1, 1, cf, tnz, chunk >= 64 ? SEGSCAN_WRITE_OWNER : SEGSCAN_WRITE_LESSATOM);
      // workers(M) = M * bn / bd = blocks
      k.bn = static_cast<int>(blocks);
      k.bd = static_cast<int>(std::max<int64_t>(1, M));
    }
  }
  pl.variant = v;
  switch (v) {
    case SEGSCAN_SPMM_EBPR:
      defaults(k, mk(5, 5, 256, 1, 16, 0, 1, 0));
      pl.fam = kEbpr;
      break;
    case SEGSCAN_SPMM_RBPR:
      defaults(k, mk(3, 3, 256, 1, 2, 0, 1, 0));
      pl.fam = kRbpr;
      break;
    case SEGSCAN_SPMM_EBSR:
      defaults(k, mk(4, 5, 256, 1, 16, 4, 1, 0));
      pl.fam = kEbsr;
      break;
    case SEGSCAN_SPMM_RBSR:
      defaults(k, mk(0, 5, 256, 1, 1, 4, 1, 0));
      pl.fam = kRbsr;
      break;
    case SEGSCAN_SPMM_PLAIN_EBPR:  // spmm.cuh:167-202: warp 32, 256 threads, M tiles
      k = mk(5, 5, 256, 1, 1, 0, 1, 0);
      pl.fam = kEbpr;
      break;
    case SEGSCAN_SPMM_PLAIN_RBPR:  // spmm.cuh:310-343
      k = mk(5, 5, 256, 1, 1, 0, 1, 0);
      pl.fam = kRbpr;
      break;
    case SEGSCAN_SPMM_PLAIN_EBSR: {  // spmm.cuh:522-591: CF by N, ThreadNz 1, CEIL(M, 8) blocks
      int cf = (N >= 512) ? 4 : (N >= 128) ? 2 : 1;
      k = mk(5, 5, 256, 1, 8, cf, 1, SEGSCAN_WRITE_ATOMIC);
      pl.fam = kEbsr;
      break;
    }
    case SEGSCAN_SPMM_PLAIN_RBSR:  // spmm.cuh:621-637: one column per thread, 256-column slices
      k = mk(0, 8, 256, 1, 1, 1, 1, 0);
      pl.fam = kRbsr;
      break;
    default:
      return fail(SEGSCAN_EINVAL, "unknown SpMM variant %d", variant);
  }
  segscan_status_t st = check_common_knobs(k);
  if (st) return st;
  const int gf = factor_value(k.group_factor), tf = factor_value(k.tile_factor);
  if (gf > 6) return fail(SEGSCAN_EUNSUPPORTED, "group_factor %d: groups are at most 64 lanes (wave64)", gf);
  if (tf > 8) return fail(SEGSCAN_EUNSUPPORTED, "tile_factor %d too large", tf);
  pl.G = 1 << gf;
  pl.T = 1 << tf;
  if (k.cf && !(k.cf == 1 || k.cf == 2 || k.cf == 4 || k.cf == 8))
    return fail(SEGSCAN_EUNSUPPORTED, "coarsen_factor %d not in {1,2,4,8}", k.cf);
  if (k.policy < 0 || k.policy > SEGSCAN_WRITE_LESSATOM)
    return fail(SEGSCAN_EINVAL, "unknown write policy %d", k.policy);
  pl.tnz = k.thread_nz;

  switch (pl.fam) {
    case kEbpr:
    case kRbpr: {
      if (pl.T > 64) return fail(SEGSCAN_EUNSUPPORTED, "tile_size %d: tiles are at most one wave (64)", pl.T);
      if (pl.G > pl.T) return fail(SEGSCAN_EUNSUPPORTED, "group_size %d exceeds tile_size %d", pl.G, pl.T);
      int cf = k.cf ? k.cf : divisor_cf(N, vec);
      if (cf == 8 && bt == SEGSCAN_F32) return fail(SEGSCAN_EUNSUPPORTED, "coarsen_factor 8 needs a 16-bit dense operand");
      cf = std::min(cf, cf_cap);
      pl.CF = cf;
      pl.X = pl.T;
      pl.slice = std::max(pl.T, cf);  // columns per blockIdx.y (reference: tile_size)
      pl.P = static_cast<int>(ceil_div(std::min<int64_t>(N, pl.slice), cf));
      if (pl.P * pl.X > 1024) {  // narrower slices instead of an oversized workgroup
        pl.P = 1024 / pl.X;
        pl.slice = pl.P * cf;
      }
      pl.Y = static_cast<int>(ceil_div(std::max(1, k.tpb / pl.T), pl.P));
      const int lanes = pl.P * pl.X * pl.Y;
      if (lanes > 1024)
        return fail(SEGSCAN_EUNSUPPORTED, "workgroup of %d lanes (panels %d x tile %d x tiles %d) exceeds 1024",
                    lanes, pl.P, pl.X, pl.Y);
      pl.block = round_up64(lanes);
      pl.tpb = pl.block;
      pl.grid_y = ceil_div(N, pl.slice);
      if (pl.fam == kEbpr) {
        pl.ntiles = ceil_div(nnz, pl.T);
        pl.grid_x = std::min(ceil_div(workers(M, k), pl.Y), std::max<int64_t>(1, ceil_div(pl.ntiles, pl.Y)));
        pl.zero_fill = true;
        pl.policy = SEGSCAN_WRITE_ATOMIC;
        pl.off_rows = 0;
        pl.ws_bytes = align_up(static_cast<size_t>(std::max<int64_t>(pl.ntiles, 1)) * 4);
      } else {
        pl.grid_x = std::min(ceil_div(workers(M, k), pl.Y), std::max<int64_t>(1, ceil_div(M, pl.Y)));
        const bool store = pl.G == pl.T;
        pl.zero_fill = !store;
        pl.policy = store ? SEGSCAN_WRITE_OWNER : SEGSCAN_WRITE_ATOMIC;
        pl.ws_bytes = 0;
      }
      break;
    }
    case kEbsr: {
      if (bt == SEGSCAN_F32 && k.cf == 8 && vec < 8) { /* fp32 CF 8 = two 128-bit loads: allowed */ }
      // reference: coarsen = min(CEIL(N, group_size), CoarsenFactor) (spmm_cg.cuh:557)
      int cf = static_cast<int>(std::min<int64_t>(ceil_div(N, pl.G), k.cf ? k.cf : 4));
      int c2 = 1;
      while (c2 * 2 <= cf) c2 <<= 1;  // power of two not above it
      cf = std::min(c2, cf_cap);
      pl.CF = cf;
      pl.tpb = round_up64(std::max(k.tpb, pl.G));
      pl.block = pl.tpb;
      pl.chunk = static_cast<int64_t>(pl.tnz) * pl.G;
      pl.nchunks = ceil_div(nnz, pl.chunk);
      const int gpb = pl.tpb / pl.G;
      pl.grid_x = std::min(workers(M, k), std::max<int64_t>(1, ceil_div(pl.nchunks, gpb)));
      pl.total_groups = pl.grid_x * gpb;
      pl.grid_y = ceil_div(N, static_cast<int64_t>(pl.G) * cf);
      int policy = k.policy;
      if (policy == SEGSCAN_WRITE_DEFAULT) policy = (pl.chunk >= 128) ? SEGSCAN_WRITE_OWNER : SEGSCAN_WRITE_LESSATOM;
      pl.policy = policy;
      pl.zero_fill = policy != SEGSCAN_WRITE_OWNER;
      size_t off = 0;  // the kernel finds chunk start rows itself; only the owner carries need space
      if (policy == SEGSCAN_WRITE_OWNER) {
        pl.off_crow = off;
        off += align_up(static_cast<size_t>(std::max<int64_t>(pl.nchunks, 1)) * 4);
        pl.off_cval = off;
        off += align_up(static_cast<size_t>(std::max<int64_t>(pl.nchunks, 1)) * static_cast<size_t>(N) * 4);
      }
      pl.ws_bytes = off;
      break;
    }
    case kRbsr: {
      int cf = k.cf ? k.cf : 1;
      cf = std::min(cf, cf_cap);
      pl.CF = cf;
      pl.slice = std::max(pl.T, cf);
      pl.tpr = static_cast<int>(ceil_div(std::min<int64_t>(N, pl.slice), cf));
      pl.rows_per_block = static_cast<int>(std::max<int64_t>(1, ceil_div(k.tpb, pl.tpr)));
      const int lanes = pl.tpr * pl.rows_per_block;
      if (lanes > 1024) return fail(SEGSCAN_EUNSUPPORTED, "workgroup of %d lanes exceeds 1024", lanes);
      pl.block = round_up64(lanes);
      pl.tpb = pl.block;
      pl.grid_x = std::min(ceil_div(workers(M, k), pl.rows_per_block), std::max<int64_t>(1, ceil_div(M, pl.rows_per_block)));
      pl.grid_y = ceil_div(N, pl.slice);
      pl.zero_fill = false;
      pl.policy = SEGSCAN_WRITE_OWNER;
      pl.ws_bytes = 0;
      break;
    }
  }
  if (pl.grid_y > 65535) return fail(SEGSCAN_EUNSUPPORTED, "N = %lld needs %lld column slices (> 65535)", (long long)N,
                                     (long long)pl.grid_y);
  if (pl.grid_x > kMaxI32) return fail(SEGSCAN_EUNSUPPORTED, "grid too large");
  return SEGSCAN_OK;
}

segscan_status_t validate_csr(const segscan_csr_t* A, bool need_pointers) {
  if (!A) return fail(SEGSCAN_EINVAL, "A is NULL");
  if (A->nrow < 0 || A->ncol < 0 || A->nnz < 0) return fail(SEGSCAN_EINVAL, "negative matrix dimension");
  if (A->nrow >= kMaxI32) return fail(SEGSCAN_EINVAL, "nrow %lld exceeds int32 indexing", (long long)A->nrow);
  if (A->ncol > kMaxI32) return fail(SEGSCAN_EINVAL, "ncol %lld exceeds int32 indexing", (long long)A->ncol);
  if (A->nnz > kMaxI32) return fail(SEGSCAN_EINVAL, "nnz %lld exceeds int32 indexing", (long long)A->nnz);
  if (A->nrow == 0 && A->nnz > 0) return fail(SEGSCAN_EINVAL, "nnz > 0 with no rows");
  if (A->ncol == 0 && A->nnz > 0) return fail(SEGSCAN_EINVAL, "nnz > 0 with no columns");
  if (A->max_row_nnz < 0 || A->max_row_nnz > A->nnz)
    return fail(SEGSCAN_EINVAL, "max_row_nnz %lld outside [0, nnz]", (long long)A->max_row_nnz);
  if (A->vals_map && !A->vals) return fail(SEGSCAN_EINVAL, "vals_map given without vals");
  if (need_pointers) {
    if (A->nrow > 0 && !A->rowptr) return fail(SEGSCAN_EINVAL, "rowptr is NULL");
    if (A->nnz > 0 && !A->colind) return fail(SEGSCAN_EINVAL, "colind is NULL");
  }
  return SEGSCAN_OK;
}

void fill_info(const SpmmPlan& pl, segscan_launch_info_t* info) {
  if (!info) return;
  std::memset(info, 0, sizeof(*info));
  info->variant = pl.variant;
  info->group_size = pl.G;
  info->tile_size = pl.T;
  info->thread_per_block = pl.block;
  info->coarsen_factor = pl.CF;
  info->thread_nz = pl.tnz;
  info->write_policy = pl.policy;
  info->chunk = static_cast<int32_t>(pl.fam == kEbpr ? pl.T : pl.chunk);
  info->grid_x = pl.grid_x;
  info->grid_y = pl.grid_y;
  info->nchunks = pl.fam == kEbpr ? pl.ntiles : pl.nchunks;
  info->workspace_bytes = pl.ws_bytes;
}

VType vtype(segscan_dtype_t t) { return static_cast<VType>(static_cast<int>(t)); }

// ---------------------------------------------------------------------------------------
// segment_coo planning

struct SegPlan {
  int variant = 0;
  bool pr = true;
  int G = 1, T = 1, CF = 1, tnz = 1, policy = SEGSCAN_WRITE_ATOMIC;
  int X = 1, P = 1, Y = 1, slice = 1, block = 64, tpb = 64;
  int64_t grid_x = 1, grid_y = 1, ntiles = 0, chunk = 0, nchunks = 0, total_groups = 0;
  bool zero_fill = true;
  size_t ws_bytes = 0, off_crow = 0, off_cval = 0, off_zflag = 0;
};

segscan_status_t plan_seg(int64_t E, int64_t N, int64_t dst_len, segscan_dtype_t st, int variant, const Knobs& user, int cf_cap,
                          SegPlan& pl) {
  if (variant < 0 || variant > SEGSCAN_SEG_AUTO) return fail(SEGSCAN_EINVAL, "unknown segment_coo variant %d", variant);
  const int vec = vec_elems(st);
  Knobs k = user;
  int v = variant;
  if (v == SEGSCAN_SEG_AUTO) {
    // Streaming src through the sequential kernel's full-tile loop: a group of
    // lanes per src row with two columns per lane (wider only when N needs more
    // than 64 lanes), owner chunks of ~64 KB of src (256 fp32 / 512 bf16 rows of
    // 64 columns), 8 waves per CU for 4-byte src and 16 for 2-byte src. On the
    // E = 4.19M, N = 64 power-law index: fp32 235 us, bf16 136 us, against 253 /
    // 150 us for the previous dedicated kernel.
    v = SEGSCAN_SEG_LESSATOM_SR;
    int cf = std::min(2, divisor_cf(N, vec));
    while (cf < 8 && ceil_div(N, cf) > 64) cf <<= 1;
    int lanes = static_cast<int>(std::min<int64_t>(64, ceil_div(N, cf)));
    int g = 1;
    while (g < lanes) g <<= 1;
    int gf = 0;
    while ((1 << gf) < g) ++gf;
    const int64_t row_bytes = std::max<int64_t>(1, N * static_cast<int64_t>(dtype_size(st)));
    const int64_t gpb = 256 / g;
    const int64_t waves_per_cu = dtype_size(st) == 4 ? 8 : 16;
    // ~64 KB of src per chunk, but short enough that every group of the grid gets one
    const int64_t groups = 256 * waves_per_cu * 64 / g;
    int64_t rows = std::min<int64_t>(1024, std::max<int64_t>(64, 65536 / row_bytes));
    rows = std::max<int64_t>(g, std::min<int64_t>(rows, std::max<int64_t>(64, ceil_div(E, groups))));
    const int tnz = static_cast<int>(std::max<int64_t>(1, rows / g));
    const int64_t chunk = static_cast<int64_t>(tnz) * g;
    // At most waves_per_cu waves per CU, each group walking several chunks in turn;
    // but chunks that would need at most two such rounds all start at once (one
    // per group, up to twice the waves), since a partly filled last round is the
    // kernel's tail. N = 64, Zipf(2.1) runs: E = 0.7M fp32 63.3 -> 48.8 us, E = 1.5M
    // bf16 80.8 -> 64.2 us, E = 3M bf16 181 -> 141 us; the 4.19M index is unchanged.
    const int64_t need = ceil_div(ceil_div(E, chunk), gpb);
    const int64_t cap = 256 * waves_per_cu * 64 / 256;
    const int64_t blocks = std::max<int64_t>(1, need <= 2 * cap ? need : cap);
    k = mk(gf == 0 ? -1 : gf, 5, 256, 1, 1, cf, tnz, SEGSCAN_WRITE_OWNER);
    k.bn = 1;
    k.bd = static_cast<int>(std::max<int64_t>(1, E / blocks));  // workers(E) = E / bd = blocks
    // the chunk length and grid may be given explicitly (thread_nz, block_numer / block_denom)
    if (user.thread_nz > 0) k.thread_nz = user.thread_nz;
    if (user.bd > 0) {
      k.bd = user.bd;
      k.bn = user.bn > 0 ? user.bn : 1;
    }
  }
  pl.variant = v;
  switch (v) {
    case SEGSCAN_SEG_PR:  // main.cu:49 segment_coo<float,int,5,256,5,1,16>
      defaults(k, mk(5, 5, 256, 1, 16, 0, 1, 0));
      break;
    case SEGSCAN_SEG_PR_NEW:  // main.cu:50 segment_coo_new<float,int,5,256,5,2>
      defaults(k, mk(5, 5, 256, 1, 1, 0, 1, 0));
      if (!k.block_dim_y) k.block_dim_y = 2;
      break;
    case SEGSCAN_SEG_SR:  // main.cu:51 segment_coo_sr<float,int,4,256,4,1,1,16>
      defaults(k, mk(4, 5, 256, 1, 16, 4, 1, SEGSCAN_WRITE_ATOMIC));
      break;
    case SEGSCAN_SEG_NOSHMEM_SR:  // main.cu:52 segment_coo_noshmem_sr<float,int,4,256,4,2,1,16>
      defaults(k, mk(4, 5, 256, 1, 16, 4, 2, SEGSCAN_WRITE_ATOMIC));
      break;
    case SEGSCAN_SEG_LESSATOM_SR:  // main.cu:53 segment_coo_noshmem_lessatom_sr<float,int,4,256,4,1,1,16>
      defaults(k, mk(4, 5, 256, 1, 16, 4, 1, SEGSCAN_WRITE_LESSATOM));
      break;
    case SEGSCAN_SEG_SORTED_SR:  // main.cu:54 segment_coo_sorted_sr<float,int,2,4,8,2,16>
      defaults(k, mk(0, 5, 256, 1, 1, 2, 4, SEGSCAN_WRITE_LESSATOM));
      if (!k.dtile_size) k.dtile_size = 8;
      if (!k.dtile_num) k.dtile_num = 2;
      if (!k.block_dim_y) k.block_dim_y = 16;
      break;
    default:
      return fail(SEGSCAN_EINVAL, "unknown segment_coo variant %d", variant);
  }
  segscan_status_t s = check_common_knobs(k);
  if (s) return s;
  const int gf = factor_value(k.group_factor), tf = factor_value(k.tile_factor);
  if (gf > 6) return fail(SEGSCAN_EUNSUPPORTED, "group_factor %d: groups are at most 64 lanes (wave64)", gf);
  if (tf > 10) return fail(SEGSCAN_EUNSUPPORTED, "tile_factor %d too large", tf);
  if (k.cf && !(k.cf == 1 || k.cf == 2 || k.cf == 4 || k.cf == 8))
    return fail(SEGSCAN_EUNSUPPORTED, "coarsen_factor %d not in {1,2,4,8}", k.cf);
  if (k.policy < 0 || k.policy > SEGSCAN_WRITE_LESSATOM) return fail(SEGSCAN_EINVAL, "unknown write policy %d", k.policy);
  pl.G = 1 << gf;
  pl.T = 1 << tf;
  pl.tnz = k.thread_nz;

  if (v == SEGSCAN_SEG_PR || v == SEGSCAN_SEG_PR_NEW) {
    pl.pr = true;
    int cf = k.cf ? k.cf : divisor_cf(N, vec);
    if (cf == 8 && st == SEGSCAN_F32) return fail(SEGSCAN_EUNSUPPORTED, "coarsen_factor 8 needs a 16-bit src");
    cf = std::min(cf, cf_cap);
    pl.CF = cf;
    if (v == SEGSCAN_SEG_PR) {
      if (pl.T > 64) return fail(SEGSCAN_EUNSUPPORTED, "tile_size %d: tiles are at most one wave (64)", pl.T);
      if (pl.G > pl.T) return fail(SEGSCAN_EUNSUPPORTED, "group_size %d exceeds tile_size %d", pl.G, pl.T);
      pl.X = pl.T;
      pl.slice = std::max(pl.T, cf);
      pl.P = static_cast<int>(ceil_div(std::min<int64_t>(N, pl.slice), cf));
      if (pl.P * pl.X > 1024) {
        pl.P = 1024 / pl.X;
        pl.slice = pl.P * cf;
      }
      pl.Y = static_cast<int>(ceil_div(std::max(1, k.tpb / pl.T), pl.P));
      pl.ntiles = ceil_div(E, pl.X);
      pl.grid_x = std::min(ceil_div(workers(E, k), pl.Y), std::max<int64_t>(1, ceil_div(pl.ntiles, pl.Y)));
    } else {
      // segscan.cuh:327-355: x = elements, y = column panels of CF
      const int max_bdy = std::max(1, k.tpb / pl.T);
      const int64_t ntile_y = ceil_div(N, cf);
      const int bdy = static_cast<int>(std::min<int64_t>(ntile_y, std::min(k.block_dim_y, max_bdy)));
      int bdx = static_cast<int>(ceil_div(k.tpb, bdy));
      if (bdx % pl.G) bdx = static_cast<int>(ceil_div(bdx, pl.G) * pl.G);  // whole groups
      pl.X = bdx;
      pl.P = bdy;
      pl.Y = 1;
      pl.slice = bdy * cf;
      pl.ntiles = ceil_div(E, pl.X);
      pl.grid_x = std::max<int64_t>(1, pl.ntiles);
    }
    const int lanes = pl.P * pl.X * pl.Y;
    if (lanes > 1024) return fail(SEGSCAN_EUNSUPPORTED, "workgroup of %d lanes exceeds 1024", lanes);
    pl.block = round_up64(lanes);
    pl.tpb = pl.block;
    pl.grid_y = ceil_div(N, pl.slice);
    pl.policy = SEGSCAN_WRITE_ATOMIC;
    pl.zero_fill = true;
    pl.ws_bytes = 0;
  } else {
    pl.pr = false;
    int cf, tpb;
    if (v == SEGSCAN_SEG_SORTED_SR) {
      // segscan.cuh:974-1008: x lanes = min(DtileNum, CEIL(N, Dtile*CF)) * Dtile over
      // columns, y = BlockDimY element chunks of ThreadNz.
      if (!is_pow2(k.dtile_size) || k.dtile_num < 1)
        return fail(SEGSCAN_EUNSUPPORTED, "dtile_size %d must be a power of two, dtile_num >= 1", k.dtile_size);
      cf = std::min(k.cf ? k.cf : 2, 4);
      int c2 = 1;
      while (c2 * 2 <= cf) c2 <<= 1;
      cf = c2;
      const int bdx = static_cast<int>(std::min<int64_t>(k.dtile_num, ceil_div(N, static_cast<int64_t>(k.dtile_size) * cf))) *
                      k.dtile_size;
      int g = 1;
      while (g < bdx) g <<= 1;
      if (g > 64) return fail(SEGSCAN_EUNSUPPORTED, "dtile_size * dtile_num = %d lanes exceeds a wave", bdx);
      pl.G = g;
      const int bdy = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(k.block_dim_y, ceil_div(E, pl.tnz))));
      tpb = g * bdy;
      if (tpb > 1024) return fail(SEGSCAN_EUNSUPPORTED, "workgroup of %d lanes exceeds 1024", tpb);
      pl.chunk = pl.tnz;
    } else {
      // segscan.cuh:817-866: coarsen = min(CEIL(N, group), CoarsenFactor)
      cf = static_cast<int>(std::min<int64_t>(ceil_div(N, pl.G), k.cf ? k.cf : 4));
      int c2 = 1;
      while (c2 * 2 <= cf) c2 <<= 1;
      cf = c2;
      tpb = std::max(k.tpb, pl.G);
      pl.chunk = static_cast<int64_t>(pl.tnz) * pl.G;
    }
    pl.CF = std::min(cf, cf_cap);
    pl.tpb = round_up64(tpb);
    pl.block = pl.tpb;
    pl.nchunks = ceil_div(E, pl.chunk);
    const int gpb = pl.tpb / pl.G;
    const int64_t want = (v == SEGSCAN_SEG_SORTED_SR) ? ceil_div(pl.nchunks, gpb) : workers(E, k);
    pl.grid_x = std::min(want, std::max<int64_t>(1, ceil_div(pl.nchunks, gpb)));
    pl.total_groups = pl.grid_x * gpb;
    pl.grid_y = ceil_div(N, static_cast<int64_t>(pl.G) * pl.CF);
    pl.policy = k.policy ? k.policy : SEGSCAN_WRITE_ATOMIC;
    pl.zero_fill = pl.policy != SEGSCAN_WRITE_OWNER;
    size_t off = 0;
    if (pl.policy == SEGSCAN_WRITE_OWNER) {
      pl.off_crow = off;
      off += align_up(static_cast<size_t>(std::max<int64_t>(pl.nchunks, 1)) * 4);
      pl.off_cval = off;
      off += align_up(static_cast<size_t>(std::max<int64_t>(pl.nchunks, 1)) * static_cast<size_t>(N) * 4);
      pl.off_zflag = off;  // one flag per block of kZeroRows segments
      off += align_up(static_cast<size_t>(zero_blocks(dst_len)));
    }
    pl.ws_bytes = off;
  }
  if (pl.grid_y > 65535) return fail(SEGSCAN_EUNSUPPORTED, "N = %lld needs too many column slices", (long long)N);
  if (pl.grid_x > kMaxI32) return fail(SEGSCAN_EUNSUPPORTED, "grid too large");
  return SEGSCAN_OK;
}

void fill_info(const SegPlan& pl, segscan_launch_info_t* info) {
  if (!info) return;
  std::memset(info, 0, sizeof(*info));
  info->variant = pl.variant;
  info->group_size = pl.G;
  info->tile_size = pl.pr ? pl.X : pl.T;
  info->thread_per_block = pl.block;
  info->coarsen_factor = pl.CF;
  info->thread_nz = pl.tnz;
  info->write_policy = pl.policy;
  info->chunk = static_cast<int32_t>(pl.pr ? pl.X : pl.chunk);
  info->grid_x = pl.grid_x;
  info->grid_y = pl.grid_y;
  info->nchunks = pl.pr ? pl.ntiles : pl.nchunks;
  info->workspace_bytes = pl.ws_bytes;
}

}  // namespace

// =========================================================================================
// C ABI

extern "C" {

int32_t segscan_abi_version(void) { return 3; }

void segscan_set_timing_events(void* start, void* stop) {
  g_ev_start = static_cast<hipEvent_t>(start);
  g_ev_stop = static_cast<hipEvent_t>(stop);
}

const char* segscan_status_string(segscan_status_t s) {
  switch (s) {
    case SEGSCAN_OK: return "SEGSCAN_OK";
    case SEGSCAN_EINVAL: return "SEGSCAN_EINVAL: invalid argument";
    case SEGSCAN_EUNSUPPORTED: return "SEGSCAN_EUNSUPPORTED: unsupported knob combination";
    case SEGSCAN_EHIP: return "SEGSCAN_EHIP: HIP runtime error";
    case SEGSCAN_EWORKSPACE: return "SEGSCAN_EWORKSPACE: workspace missing or too small";
    default: break;
  }
  return "unknown segscan status";
}

const char* segscan_last_error(void) { return g_last_error.c_str(); }

segscan_status_t segscan_spmm_describe(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                       segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                       segscan_launch_info_t* info) {
  segscan_status_t st = validate_csr(A, false);
  if (st) return st;
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (!dtype_valid(b_dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)b_dtype);
  SpmmPlan pl;
  st = plan_spmm(A->nrow, A->nnz, A->max_row_nnz, std::max<int64_t>(N, 1), b_dtype, variant, from_c(knobs), 8, pl);
  if (st) return st;
  fill_info(pl, info);
  return ok();
}

segscan_status_t segscan_spmm_workspace_size(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                             segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                             size_t* bytes) {
  if (!bytes) return fail(SEGSCAN_EINVAL, "bytes is NULL");
  segscan_launch_info_t info;
  segscan_status_t st = segscan_spmm_describe(A, N, b_dtype, variant, knobs, &info);
  if (st) return st;
  *bytes = info.workspace_bytes;
  return ok();
}

segscan_status_t segscan_spmm_csr(const segscan_csr_t* A, int64_t N, const void* B, int64_t ldb,
                                  segscan_dtype_t b_dtype, float* C, int64_t ldc, segscan_spmm_variant_t variant,
                                  const segscan_knobs_t* knobs, uint32_t flags, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  return segscan::spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, workspace, workspace_bytes,
                                stream, nullptr);
}

segscan_status_t segscan_spmm_csr_permuted(const segscan_csr_t* A, segscan_permute_plan_t plan,
                                           const float* vals_src, int64_t N, const void* B, int64_t ldb,
                                           segscan_dtype_t b_dtype, float* C, int64_t ldc,
                                           segscan_spmm_variant_t variant, const segscan_knobs_t* knobs,
                                           uint32_t flags, void* workspace, size_t workspace_bytes,
                                           void* stream) {
  using segscan::spmm_csr_impl;
  if (!A) return fail(SEGSCAN_EINVAL, "A is NULL");
  if (!plan) return fail(SEGSCAN_EINVAL, "plan is NULL");
  if (A->vals || A->vals_map)
    return fail(SEGSCAN_EINVAL, "A.vals and A.vals_map must be NULL: the values come from vals_src through the plan");
  const segscan::PermuteBuffers pb = segscan::permute_buffers(plan);
  if (pb.n != A->nnz)
    return fail(SEGSCAN_EINVAL, "plan permutes %lld values, A has %lld entries", (long long)pb.n, (long long)A->nnz);
  if (!vals_src) return fail(SEGSCAN_EINVAL, "vals_src is NULL");
  if (reinterpret_cast<uintptr_t>(vals_src) % 4) return fail(SEGSCAN_EINVAL, "vals_src is not 4-byte aligned");
  int device = 0;
  const hipError_t de = hipGetDevice(&device);
  if (de != hipSuccess) return hip_status(de, "current device");
  if (device != pb.device) return fail(SEGSCAN_EINVAL, "the current device (%d) is not the plan's (%d)", device, pb.device);
  // validation and planning first: nothing is launched for a call that would fail
  segscan::StagedValues sv;
  sv.mid = pb.mid;
  sv.slot = pb.slot2;
  size_t need = 0;
  segscan_status_t st = spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, workspace,
                                      workspace_bytes, stream, &need, &sv);
  if (st) return st;
  if (A->nrow == 0 || N == 0)
    return spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, workspace, workspace_bytes, stream,
                         nullptr);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(SEGSCAN_EWORKSPACE, "workspace of %zu bytes required", need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = segscan::permute_pass1(plan, vals_src, s);
  if (e != hipSuccess) return hip_status(e, "permute pass 1");
  if (sv.fused)
    return spmm_csr_impl(A, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, workspace, workspace_bytes, stream,
                         nullptr, &sv);
  e = segscan::permute_pass2(plan, pb.aux, s);
  if (e != hipSuccess) return hip_status(e, "permute pass 2");
  segscan_csr_t a = *A;
  a.vals = pb.aux;
  return spmm_csr_impl(&a, N, B, ldb, b_dtype, C, ldc, variant, knobs, flags, workspace, workspace_bytes, stream,
                       nullptr);
}

}  // extern "C"

namespace segscan {

// segscan_spmm_csr's body. With plan_bytes non-NULL the call stops after validation and
// planning (no HIP call) and reports the workspace the launch would need (0 when the
// call launches nothing or needs no workspace).
segscan_status_t spmm_csr_impl(const segscan_csr_t* A, int64_t N, const void* B, int64_t ldb,
                               segscan_dtype_t b_dtype, float* C, int64_t ldc, segscan_spmm_variant_t variant,
                               const segscan_knobs_t* knobs, uint32_t flags, void* workspace, size_t workspace_bytes,
                               void* stream, size_t* plan_bytes, StagedValues* staged) {
  segscan_status_t st = validate_csr(A, true);
  if (st) return st;
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (!dtype_valid(b_dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)b_dtype);
  if (flags & ~SEGSCAN_ACCUMULATE) return fail(SEGSCAN_EINVAL, "unknown flags 0x%x", flags);
  const int64_t M = A->nrow, nnz = A->nnz;
  if (M > 0 && N > 0) {
    if (!C) return fail(SEGSCAN_EINVAL, "C is NULL");
    if (ldc < N) return fail(SEGSCAN_EINVAL, "ldc %lld < N %lld", (long long)ldc, (long long)N);
    if (reinterpret_cast<uintptr_t>(C) % 4) return fail(SEGSCAN_EINVAL, "C is not 4-byte aligned");
  }
  if (nnz > 0 && N > 0) {
    if (!B) return fail(SEGSCAN_EINVAL, "B is NULL");
    if (ldb < N) return fail(SEGSCAN_EINVAL, "ldb %lld < N %lld", (long long)ldb, (long long)N);
    if (ldb > kMaxI32) return fail(SEGSCAN_EUNSUPPORTED, "ldb %lld beyond 2^31 - 1", (long long)ldb);
    if (reinterpret_cast<uintptr_t>(B) % dtype_size(b_dtype)) return fail(SEGSCAN_EINVAL, "B is misaligned");
  }
  const bool accumulate = flags & SEGSCAN_ACCUMULATE;
  const int cap = aligned_cf(8, B, std::max<int64_t>(ldb, 1), dtype_size(b_dtype), C, std::max<int64_t>(ldc, 1));
  SpmmPlan pl;
  st = plan_spmm(M, nnz, A->max_row_nnz, std::max<int64_t>(N, 1), b_dtype, variant, from_c(knobs), cap, pl);
  if (st) return st;
  if (staged) {
    // a workgroup round spans chunk x groups entries from a multiple of that span; the
    // staged-value kernels are built for groups of 8 to 64 lanes
    const int64_t span = pl.chunk * (pl.tpb / pl.G);
    staged->fused = pl.fam == kEbsr && pl.G >= 8 && span >= 2048 && kPermBucket % span == 0;
    if (!plan_bytes && !staged->fused)
      return fail(SEGSCAN_EUNSUPPORTED, "staged values need an ebsr launch whose rounds cover part of one bucket");
  }
  if (plan_bytes) {
    *plan_bytes = (M == 0 || N == 0 || nnz == 0) ? 0 : pl.ws_bytes;
    return ok();
  }
  if (M == 0 || N == 0) return ok();
  if (pl.ws_bytes > 0 && nnz > 0) {
    if (!workspace) return fail(SEGSCAN_EWORKSPACE, "workspace of %zu bytes required", pl.ws_bytes);
    if (workspace_bytes < pl.ws_bytes)
      return fail(SEGSCAN_EWORKSPACE, "workspace %zu bytes < %zu required", workspace_bytes, pl.ws_bytes);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (nnz == 0) {
    if (!accumulate) return hip_status(launch_fill_zero(C, M, N, ldc, s), "zero fill");
    return ok();
  }
  if (pl.zero_fill && !accumulate) {
    hipError_t e = launch_fill_zero(C, M, N, ldc, s);
    if (e != hipSuccess) return hip_status(e, "zero fill");
  }
  char* ws = static_cast<char*>(workspace);
  const VType vt = vtype(b_dtype);
  bool compiled = true;
  hipError_t e = hipSuccess;
  Launch l;
  l.grid = dim3(static_cast<unsigned>(pl.grid_x), static_cast<unsigned>(pl.grid_y));
  l.block = dim3(pl.block);
  l.shmem = 0;
  l.stream = s;
  switch (pl.fam) {
    case kEbpr:
    case kRbpr: {
      PrParams p{};
      p.M = static_cast<int>(M);
      p.N = static_cast<int>(N);
      p.nnz = nnz;
      p.rowptr = A->rowptr;
      p.colind = A->colind;
      p.vals = A->vals;
      p.vmap = A->vals_map;
      p.B = B;
      p.ldb = ldb;
      p.C = C;
      p.ldc = ldc;
      p.X = pl.X;
      p.P = pl.P;
      p.Y = pl.Y;
      p.G = pl.G;
      p.slice = pl.slice;
      p.ntiles = pl.ntiles;
      p.row_store = pl.policy == SEGSCAN_WRITE_OWNER;
      p.accumulate = accumulate;
      if (pl.fam == kEbpr) {
        p.tile_row = reinterpret_cast<int32_t*>(ws + pl.off_rows);
        e = launch_partition_rows(A->rowptr, p.M, nnz, pl.T, pl.ntiles, const_cast<int32_t*>(p.tile_row), nullptr, s);
        if (e != hipSuccess) return hip_status(e, "partition");
      }
      mark_start(s);
      e = launch_spmm_pr(vt, pl.CF, pl.G, p, l, pl.fam == kEbpr, compiled);
      mark_stop(s);
      break;
    }
    case kEbsr: {
      SrParams p{};
      p.M = static_cast<int>(M);
      p.N = static_cast<int>(N);
      p.nnz = nnz;
      p.rowptr = A->rowptr;
      p.colind = A->colind;
      p.vals = A->vals;
      p.vmap = A->vals_map;
      p.B = B;
      p.ldb = ldb;
      p.C = C;
      p.ldc = ldc;
      p.tpb = pl.tpb;
      p.chunk = pl.chunk;
      p.nchunks = pl.nchunks;
      p.total_groups = pl.total_groups;
      p.policy = pl.policy;
      p.accumulate = accumulate;
      if (pl.policy == SEGSCAN_WRITE_OWNER) {
        p.carry_row = reinterpret_cast<int32_t*>(ws + pl.off_crow);
        p.carry_val = reinterpret_cast<float*>(ws + pl.off_cval);
      }
      // chunk start rows are found by the kernel itself (no partition pass)
      l.shmem = static_cast<size_t>(pl.tpb) * 12;
      if (staged) {
        p.vmid = staged->mid;
        p.vslot = staged->slot;
        l.shmem += static_cast<size_t>(kPermBucket) * 4;
      }
      mark_start(s);
      e = launch_spmm_sr(vt, pl.CF, pl.G, p, l, compiled);
      mark_stop(s);
      if (e == hipSuccess && compiled && pl.policy == SEGSCAN_WRITE_OWNER)
        e = launch_carry_fixup(p.carry_row, p.carry_val, pl.nchunks, p.N, C, ldc, nullptr,
                               accumulate ? nullptr : A->rowptr, M, s);
      break;
    }
    case kRbsr: {
      RbsrParams p{};
      p.M = static_cast<int>(M);
      p.N = static_cast<int>(N);
      p.rowptr = A->rowptr;
      p.colind = A->colind;
      p.vals = A->vals;
      p.vmap = A->vals_map;
      p.B = B;
      p.ldb = ldb;
      p.C = C;
      p.ldc = ldc;
      p.tpr = pl.tpr;
      p.rows_per_block = pl.rows_per_block;
      p.slice = pl.slice;
      p.rows_total = M;
      p.accumulate = accumulate;
      mark_start(s);
      e = launch_spmm_rbsr(vt, pl.CF, p, l, compiled);
      mark_stop(s);
      break;
    }
  }
  if (!compiled)
    return fail(SEGSCAN_EUNSUPPORTED, "no compiled kernel for dtype %d, coarsen %d, group %d", (int)b_dtype, pl.CF, pl.G);
  return hip_status(e, "SpMM launch");
}

segscan_status_t set_error(segscan_status_t st, const char* msg) { return fail(st, "%s", msg); }

}  // namespace segscan

extern "C" {

segscan_status_t segscan_segment_coo_describe(int64_t E, int64_t N, int64_t dst_len, segscan_dtype_t src_dtype,
                                              segscan_seg_variant_t variant, const segscan_knobs_t* knobs,
                                              segscan_launch_info_t* info) {
  if (E < 0 || E > kMaxI32) return fail(SEGSCAN_EINVAL, "E = %lld out of range", (long long)E);
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (dst_len < 0 || dst_len >= kMaxI32) return fail(SEGSCAN_EINVAL, "dst_len = %lld out of range", (long long)dst_len);
  if (!dtype_valid(src_dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)src_dtype);
  SegPlan pl;
  segscan_status_t st = plan_seg(E, std::max<int64_t>(N, 1), dst_len, src_dtype, variant, from_c(knobs), 8, pl);
  if (st) return st;
  fill_info(pl, info);
  return ok();
}

segscan_status_t segscan_segment_coo_workspace_size(int64_t E, int64_t N, int64_t dst_len, segscan_dtype_t src_dtype,
                                                    segscan_seg_variant_t variant, const segscan_knobs_t* knobs,
                                                    size_t* bytes) {
  if (!bytes) return fail(SEGSCAN_EINVAL, "bytes is NULL");
  segscan_launch_info_t info;
  segscan_status_t st = segscan_segment_coo_describe(E, N, dst_len, src_dtype, variant, knobs, &info);
  if (st) return st;
  *bytes = info.workspace_bytes;
  return ok();
}

segscan_status_t segscan_segment_coo(const void* src, int64_t ld_src, segscan_dtype_t src_dtype, const int32_t* index,
                                     int64_t E, int64_t N, float* dst, int64_t ld_dst, int64_t dst_len,
                                     segscan_seg_variant_t variant, const segscan_knobs_t* knobs, uint32_t flags,
                                     void* workspace, size_t workspace_bytes, void* stream) {
  if (E < 0 || E > kMaxI32) return fail(SEGSCAN_EINVAL, "E = %lld out of range", (long long)E);
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (dst_len < 0 || dst_len >= kMaxI32) return fail(SEGSCAN_EINVAL, "dst_len = %lld out of range", (long long)dst_len);
  if (!dtype_valid(src_dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)src_dtype);
  if (flags & ~SEGSCAN_ACCUMULATE) return fail(SEGSCAN_EINVAL, "unknown flags 0x%x", flags);
  if (E > 0 && dst_len == 0) return fail(SEGSCAN_EINVAL, "elements present but dst_len is 0");
  if (E > 0 && N > 0) {
    if (!src) return fail(SEGSCAN_EINVAL, "src is NULL");
    if (!index) return fail(SEGSCAN_EINVAL, "index is NULL");
    if (ld_src < N) return fail(SEGSCAN_EINVAL, "ld_src %lld < N %lld", (long long)ld_src, (long long)N);
    if (ld_src > kMaxI32) return fail(SEGSCAN_EUNSUPPORTED, "ld_src %lld beyond 2^31 - 1", (long long)ld_src);
    if (reinterpret_cast<uintptr_t>(src) % dtype_size(src_dtype)) return fail(SEGSCAN_EINVAL, "src is misaligned");
  }
  if (dst_len > 0 && N > 0) {
    if (!dst) return fail(SEGSCAN_EINVAL, "dst is NULL");
    if (ld_dst < N) return fail(SEGSCAN_EINVAL, "ld_dst %lld < N %lld", (long long)ld_dst, (long long)N);
    if (reinterpret_cast<uintptr_t>(dst) % 4) return fail(SEGSCAN_EINVAL, "dst is not 4-byte aligned");
  }
  const bool accumulate = flags & SEGSCAN_ACCUMULATE;
  const int cap = aligned_cf(8, src, std::max<int64_t>(ld_src, 1), dtype_size(src_dtype), dst, std::max<int64_t>(ld_dst, 1));
  SegPlan pl;
  segscan_status_t st = plan_seg(E, std::max<int64_t>(N, 1), dst_len, src_dtype, variant, from_c(knobs), cap, pl);
  if (st) return st;
  if (dst_len == 0 || N == 0) return ok();
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (E == 0) {
    if (!accumulate) return hip_status(launch_fill_zero(dst, dst_len, N, ld_dst, s), "zero fill");
    return ok();
  }
  if (pl.ws_bytes > 0) {
    if (!workspace) return fail(SEGSCAN_EWORKSPACE, "workspace of %zu bytes required", pl.ws_bytes);
    if (workspace_bytes < pl.ws_bytes)
      return fail(SEGSCAN_EWORKSPACE, "workspace %zu bytes < %zu required", workspace_bytes, pl.ws_bytes);
  }
  if (pl.zero_fill && !accumulate) {
    hipError_t e = launch_fill_zero(dst, dst_len, N, ld_dst, s);
    if (e != hipSuccess) return hip_status(e, "zero fill");
  }
  const VType vt = vtype(src_dtype);
  bool compiled = true;
  hipError_t e;
  Launch l;
  l.grid = dim3(static_cast<unsigned>(pl.grid_x), static_cast<unsigned>(pl.grid_y));
  l.block = dim3(pl.block);
  l.shmem = 0;
  l.stream = s;
  char* ws = static_cast<char*>(workspace);
  if (pl.pr) {
    PrParams p{};
    p.M = static_cast<int>(dst_len);
    p.N = static_cast<int>(N);
    p.nnz = E;
    p.index = index;
    p.B = src;
    p.ldb = ld_src;
    p.C = dst;
    p.ldc = ld_dst;
    p.X = pl.X;
    p.P = pl.P;
    p.Y = pl.Y;
    p.G = pl.G;
    p.slice = pl.slice;
    p.ntiles = pl.ntiles;
    p.accumulate = accumulate;
    mark_start(s);
    e = launch_seg_pr(vt, pl.CF, pl.G, p, l, compiled);
    mark_stop(s);
  } else {
    SrParams p{};
    p.M = static_cast<int>(dst_len);
    p.N = static_cast<int>(N);
    p.nnz = E;
    p.index = index;
    p.B = src;
    p.ldb = ld_src;
    p.C = dst;
    p.ldc = ld_dst;
    p.tpb = pl.tpb;
    p.chunk = pl.chunk;
    p.nchunks = pl.nchunks;
    p.total_groups = pl.total_groups;
    p.policy = pl.policy;
    p.accumulate = accumulate;
    if (pl.policy == SEGSCAN_WRITE_OWNER) {
      p.carry_row = reinterpret_cast<int32_t*>(ws + pl.off_crow);
      p.carry_val = reinterpret_cast<float*>(ws + pl.off_cval);
      p.zflag = accumulate ? nullptr : reinterpret_cast<uint8_t*>(ws + pl.off_zflag);
    }
    mark_start(s);
    l.shmem = static_cast<size_t>(pl.tpb) * 12;
    e = launch_seg_sr(vt, pl.CF, pl.G, p, l, compiled);
    mark_stop(s);
    if (e == hipSuccess && compiled && pl.policy == SEGSCAN_WRITE_OWNER)
      e = launch_carry_fixup(p.carry_row, p.carry_val, pl.nchunks, p.N, dst, ld_dst, p.zflag, nullptr, dst_len, s);
  }
  if (!compiled)
    return fail(SEGSCAN_EUNSUPPORTED, "no compiled kernel for dtype %d, coarsen %d, group %d", (int)src_dtype, pl.CF, pl.G);
  return hip_status(e, "segment_coo launch");
}

// ---- backward pieces -------------------------------------------------------------------

segscan_status_t segscan_csr_transpose_workspace_size(const segscan_csr_t* A, size_t* bytes) {
  if (!bytes) return fail(SEGSCAN_EINVAL, "bytes is NULL");
  segscan_status_t st = validate_csr(A, false);
  if (st) return st;
  *bytes = plan_transpose(A->ncol, A->nnz).bytes;
  return ok();
}

segscan_status_t segscan_csr_transpose(const segscan_csr_t* A, int32_t* colptr, int32_t* rowind, int32_t* perm,
                                       float* vals_t, void* workspace, size_t workspace_bytes, void* stream) {
  segscan_status_t st = validate_csr(A, true);
  if (st) return st;
  if (!colptr) return fail(SEGSCAN_EINVAL, "colptr is NULL");
  if (A->nnz > 0 && (!rowind || !perm)) return fail(SEGSCAN_EINVAL, "rowind / perm is NULL");
  if (A->vals_map) return fail(SEGSCAN_EUNSUPPORTED, "transpose reads values in A's order (vals_map must be NULL)");
  size_t need = 0;
  segscan_csr_transpose_workspace_size(A, &need);
  if (!workspace) return fail(SEGSCAN_EWORKSPACE, "workspace of %zu bytes required", need);
  if (workspace_bytes < need) return fail(SEGSCAN_EWORKSPACE, "workspace %zu bytes < %zu required", workspace_bytes, need);
  const int64_t K = A->ncol;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = launch_transpose(A->rowptr, A->colind, A->vals, static_cast<int>(A->nrow), static_cast<int>(K), A->nnz,
                                  colptr, rowind, perm, vals_t, workspace, plan_transpose(K, A->nnz), s);
  return hip_status(e, "transpose");
}

segscan_status_t segscan_sddmm_csr(const segscan_csr_t* A, int64_t N, const void* X, int64_t ld_x, const void* Y,
                                   int64_t ld_y, segscan_dtype_t xy_dtype, float* out, void* stream) {
  segscan_status_t st = validate_csr(A, true);
  if (st) return st;
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N out of range");
  if (!dtype_valid(xy_dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)xy_dtype);
  if (A->nnz == 0) return ok();
  if (!out) return fail(SEGSCAN_EINVAL, "out is NULL");
  if (N > 0) {
    if (!X || !Y) return fail(SEGSCAN_EINVAL, "X / Y is NULL");
    if (ld_x < N || ld_y < N) return fail(SEGSCAN_EINVAL, "leading dimension < N");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (N == 0) {
    return hip_status(launch_fill_zero(out, 1, A->nnz, A->nnz, s), "sddmm zero");
  }
  hipError_t e = launch_sddmm(vtype(xy_dtype), A->rowptr, A->colind, static_cast<int>(A->nrow), A->nnz,
                              static_cast<int>(N), X, ld_x, Y, ld_y, out, s);
  return hip_status(e, "sddmm");
}

// Fused backward plan: one group of G lanes x CF columns covers all N columns
// (G * CF >= N); CF <= 4 by default (as the forward's element-balanced rule),
// widened to 8 only when N needs more than 64 lanes. Chunks of 16 tiles (16 G
// entries) while that leaves >= 2048 chunks, else down to 64 entries: a chunk
// boundary inside an At row costs a carry of N floats and N fix-up atomics, and
// on a skewed transpose (131072-node power-law graph, N = 64 / 128) chunks of
// 16 tiles ran 11 % / 10 % faster than chunks of 8 / 4 tiles; on an even one
// (16 entries per row) 8 and 16 tiles tie. <= 1024 workgroups of 256 lanes,
// the rest grid-strided.
struct BwdPlan {
  int G = 1, CF = 1, tpb = 256;
  int64_t chunk = 0, nchunks = 0, grid = 1, total_groups = 0;
  size_t off_crow = 0, off_cval = 0, ws_bytes = 0;
};

segscan_status_t plan_bwd(int64_t K, int64_t nnz, int64_t N, segscan_dtype_t dt, const Knobs& k, int cf_cap,
                          BwdPlan& pl) {
  const int vec = vec_elems(dt);
  int cf = k.cf;
  if (cf) {
    if (!(cf == 1 || cf == 2 || cf == 4 || cf == 8)) return fail(SEGSCAN_EUNSUPPORTED, "coarsen_factor %d not in {1,2,4,8}", cf);
  } else {
    cf = std::min(std::min(vec, 4), divisor_cf(N, vec));
    while (cf < 8 && ceil_div(N, cf) > 64 && N % (cf * 2) == 0) cf <<= 1;
  }
  cf = std::min(cf, cf_cap);
  int G;
  if (k.group_factor) {
    const int gf = factor_value(k.group_factor);
    if (gf > 6) return fail(SEGSCAN_EUNSUPPORTED, "group_factor %d: groups are at most 64 lanes (wave64)", gf);
    G = 1 << gf;
  } else {
    G = 1;
    while (G < 64 && static_cast<int64_t>(G) * cf < N) G <<= 1;
  }
  if (static_cast<int64_t>(G) * cf < N)
    return fail(SEGSCAN_EUNSUPPORTED,
                "fused backward needs group_size x coarsen_factor >= N (%d x %d < %lld); use the transposed SpMM "
                "and SDDMM separately",
                G, cf, (long long)N);
  int tnz = k.thread_nz;
  if (!tnz) {
    // 16 tiles per chunk: a chunk boundary inside a long At row costs a carry of N
    // floats and N fix-up atomics. Where At rows average under 8 entries carries are
    // rare, and 8 tiles narrow the span of dC that the grid gathers from at once
    // (block-diagonal batch, 2.39M rows, 9.58M entries, N = 64: 917 -> 813 us).
    const int max_tnz = nnz < 8 * K ? 8 : 16;
    tnz = std::max(1, 64 / G);
    while (tnz < max_tnz && ceil_div(nnz, static_cast<int64_t>(2 * tnz) * G) >= 2048) tnz <<= 1;
  }
  if (tnz < 1 || tnz > (1 << 20)) return fail(SEGSCAN_EUNSUPPORTED, "thread_nz %d outside [1, 2^20]", tnz);
  int tpb = k.tpb ? k.tpb : 256;
  tpb = round_up64(std::max(tpb, G));
  if (tpb > 256) return fail(SEGSCAN_EUNSUPPORTED, "thread_per_block %d: the fused backward takes at most 256", tpb);
  pl.G = G;
  pl.CF = cf;
  pl.tpb = tpb;
  pl.chunk = static_cast<int64_t>(tnz) * G;
  pl.nchunks = ceil_div(std::max<int64_t>(nnz, 1), pl.chunk);
  const int64_t gpb = tpb / G;
  const int64_t need = std::max<int64_t>(1, ceil_div(pl.nchunks, gpb));
  int64_t grid = std::min<int64_t>(need, 1024);
  if (k.bn || k.bd) {
    if (k.bn < 0 || k.bd < 0) return fail(SEGSCAN_EINVAL, "block_factor %d/%d must be positive", k.bn, k.bd);
    Knobs w = k;
    if (!w.bn) w.bn = 1;
    if (!w.bd) w.bd = 1;
    grid = std::min(need, workers(K, w));
  }
  if (grid > kMaxI32) return fail(SEGSCAN_EUNSUPPORTED, "grid too large");
  pl.grid = grid;
  pl.total_groups = grid * gpb;
  pl.off_crow = 0;
  pl.off_cval = align_up(static_cast<size_t>(pl.nchunks) * 4);
  pl.ws_bytes = pl.off_cval + align_up(static_cast<size_t>(pl.nchunks) * static_cast<size_t>(std::max<int64_t>(N, 1)) * 4);
  return SEGSCAN_OK;
}

segscan_status_t segscan_spmm_backward_workspace_size(const segscan_csr_t* At, int64_t N, segscan_dtype_t dc_dtype,
                                                      const segscan_knobs_t* knobs, size_t* bytes) {
  const segscan_dtype_t dtype = dc_dtype;
  if (!bytes) return fail(SEGSCAN_EINVAL, "bytes is NULL");
  segscan_status_t st = validate_csr(At, false);
  if (st) return st;
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (!dtype_valid(dtype)) return fail(SEGSCAN_EINVAL, "unknown dtype %d", (int)dtype);
  BwdPlan pl;
  st = plan_bwd(At->nrow, At->nnz, std::max<int64_t>(N, 1), dtype, from_c(knobs), 8, pl);
  if (st) return st;
  *bytes = pl.ws_bytes;
  return ok();
}

// Both fused-backward entry points: dc_map NULL (dC row = A's row) or a row map.
static segscan_status_t spmm_backward_impl(const segscan_csr_t* At, const int32_t* perm, const float* vals, int64_t N,
                                    const void* dC, int64_t ld_dc, segscan_dtype_t dc_dtype, const int32_t* dc_map,
                                    const void* B, int64_t ld_b, segscan_dtype_t b_dtype, float* dB, int64_t ld_db,
                                    float* dvals, const segscan_knobs_t* knobs, uint32_t flags, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  segscan_status_t st = validate_csr(At, true);
  if (st) return st;
  if (N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "N = %lld out of range", (long long)N);
  if (!dtype_valid(dc_dtype) || !dtype_valid(b_dtype))
    return fail(SEGSCAN_EINVAL, "unknown dtype %d / %d", (int)dc_dtype, (int)b_dtype);
  if (dc_dtype != b_dtype && dc_dtype != SEGSCAN_F32)
    return fail(SEGSCAN_EUNSUPPORTED, "dC and B must share a dtype, or dC be fp32");
  if (flags & ~(SEGSCAN_ACCUMULATE | SEGSCAN_DVALS_AT_ORDER)) return fail(SEGSCAN_EINVAL, "unknown flags 0x%x", flags);
  if (At->vals_map) return fail(SEGSCAN_EINVAL, "At.vals_map must be NULL (At.vals are in At's order)");
  const bool dvals_at = flags & SEGSCAN_DVALS_AT_ORDER;
  const int64_t K = At->nrow, nnz = At->nnz;
  const size_t es = dtype_size(dc_dtype), es_b = dtype_size(b_dtype);
  if (dB && K > 0 && N > 0) {
    if (ld_db < N) return fail(SEGSCAN_EINVAL, "ld_db %lld < N %lld", (long long)ld_db, (long long)N);
    if (reinterpret_cast<uintptr_t>(dB) % 4) return fail(SEGSCAN_EINVAL, "dB is not 4-byte aligned");
  }
  if (nnz > 0) {
    // perm is read for values in A's order and for dvals in A's order
    if (!perm && ((!At->vals && vals) || (dvals && !dvals_at))) return fail(SEGSCAN_EINVAL, "perm is NULL");
    if (N > 0) {
      if (!dC || !B) return fail(SEGSCAN_EINVAL, "dC / B is NULL");
      if (ld_dc < N || ld_b < N) return fail(SEGSCAN_EINVAL, "leading dimension < N");
      if (reinterpret_cast<uintptr_t>(dC) % es || reinterpret_cast<uintptr_t>(B) % es_b)
        return fail(SEGSCAN_EINVAL, "dC / B is misaligned");
    }
  }
  const bool accumulate = flags & SEGSCAN_ACCUMULATE;
  const int64_t ldd = dB ? std::max<int64_t>(ld_db, 1) : std::max<int64_t>(ld_dc, 1);
  const int cap = std::min(aligned_cf(8, dC, std::max<int64_t>(ld_dc, 1), es, dB, ldd),
                           aligned_cf(8, B, std::max<int64_t>(ld_b, 1), es_b, nullptr, std::max<int64_t>(ld_b, 1)));
  BwdPlan pl;
  st = plan_bwd(K, nnz, std::max<int64_t>(N, 1), dc_dtype, from_c(knobs), cap, pl);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (N == 0) {
    if (dvals && nnz > 0) return hip_status(launch_fill_zero(dvals, 1, nnz, nnz, s), "dvals zero");
    return ok();
  }
  if (nnz == 0) {
    if (dB && !accumulate && K > 0) return hip_status(launch_fill_zero(dB, K, N, ld_db, s), "dB zero fill");
    return ok();
  }
  if (!dB && !dvals) return ok();
  if (dB) {
    if (!workspace) return fail(SEGSCAN_EWORKSPACE, "workspace of %zu bytes required", pl.ws_bytes);
    if (workspace_bytes < pl.ws_bytes)
      return fail(SEGSCAN_EWORKSPACE, "workspace %zu bytes < %zu required", workspace_bytes, pl.ws_bytes);
  }
  char* ws = static_cast<char*>(workspace);
  BwdParams p{};
  p.K = static_cast<int>(K);
  p.N = static_cast<int>(N);
  p.nnz = nnz;
  p.rowptr = At->rowptr;
  p.colind = At->colind;
  p.perm = perm;
  p.dc_map = dc_map;
  p.vals_t = At->vals;
  p.vals = At->vals ? nullptr : vals;
  p.dC = dC;
  p.ld_dc = ld_dc;
  p.B = B;
  p.ld_b = ld_b;
  p.dB = dB;
  p.ld_db = ld_db;
  p.dvals = dvals;
  p.tpb = pl.tpb;
  p.chunk = pl.chunk;
  p.nchunks = pl.nchunks;
  p.total_groups = pl.total_groups;
  p.carry_row = dB ? reinterpret_cast<int32_t*>(ws + pl.off_crow) : nullptr;
  p.carry_val = dB ? reinterpret_cast<float*>(ws + pl.off_cval) : nullptr;
  p.accumulate = accumulate;
  p.dvals_at_order = dvals_at;
  Launch l;
  l.grid = dim3(static_cast<unsigned>(pl.grid));
  l.block = dim3(pl.tpb);
  l.shmem = static_cast<size_t>(pl.tpb) * 12;
  l.stream = s;
  // At rows averaging under 8 entries: with a bf16 B (8 rows in registers per batch),
  // load the B row of each new row with the batch's dC rows (block-diagonal batch,
  // 9.58M entries over 2.39M rows, N = 64: 815 -> 761 us with an fp32 dC, 759 -> 636 us
  // with a bf16 dC); with longer rows the one-row-ahead form is faster
  const bool row_batch = At->nnz < 8 * At->nrow && b_dtype == SEGSCAN_BF16 &&
                         (dc_dtype == SEGSCAN_F32 || dc_dtype == SEGSCAN_BF16) && pl.CF == 4 &&
                         (pl.G == 8 || pl.G == 16 || pl.G == 32);
  bool compiled = true;
  mark_start(s);
  hipError_t e = launch_spmm_bwd(vtype(dc_dtype), vtype(b_dtype), pl.CF, pl.G, row_batch, p, l, compiled);
  mark_stop(s);
  if (!compiled)
    return fail(SEGSCAN_EUNSUPPORTED, "no compiled backward kernel for coarsen %d, group %d", pl.CF, pl.G);
  if (e == hipSuccess && dB)
    e = launch_carry_fixup(p.carry_row, p.carry_val, pl.nchunks, p.N, dB, ld_db, nullptr, accumulate ? nullptr : At->rowptr,
                           K, s);
  return hip_status(e, "fused backward");
}

segscan_status_t segscan_spmm_backward_csr(const segscan_csr_t* At, const int32_t* perm, const float* vals, int64_t N,
                                           const void* dC, int64_t ld_dc, segscan_dtype_t dc_dtype, const void* B,
                                           int64_t ld_b, segscan_dtype_t b_dtype, float* dB, int64_t ld_db,
                                           float* dvals, const segscan_knobs_t* knobs, uint32_t flags,
                                           void* workspace, size_t workspace_bytes, void* stream) {
  return spmm_backward_impl(At, perm, vals, N, dC, ld_dc, dc_dtype, nullptr, B, ld_b, b_dtype, dB, ld_db, dvals, knobs,
                            flags, workspace, workspace_bytes, stream);
}

segscan_status_t segscan_spmm_backward_csr_rowmap(const segscan_csr_t* At, const int32_t* perm, const float* vals,
                                                  int64_t N, const void* dC, int64_t ld_dc, segscan_dtype_t dc_dtype,
                                                  const int32_t* dc_map, const void* B, int64_t ld_b,
                                                  segscan_dtype_t b_dtype, float* dB, int64_t ld_db, float* dvals,
                                                  const segscan_knobs_t* knobs, uint32_t flags, void* workspace,
                                                  size_t workspace_bytes, void* stream) {
  if (!dc_map && At && At->nnz > 0 && N > 0) return fail(SEGSCAN_EINVAL, "dc_map is NULL");
  return spmm_backward_impl(At, perm, vals, N, dC, ld_dc, dc_dtype, dc_map, B, ld_b, b_dtype, dB, ld_db, dvals, knobs,
                            flags, workspace, workspace_bytes, stream);
}

segscan_status_t segscan_segment_gather(const float* ddst, int64_t ld_ddst, const int32_t* index, int64_t E, int64_t N,
                                        const int32_t* counts, float* dsrc, int64_t ld_dsrc, void* stream) {
  if (E < 0 || E > kMaxI32 || N < 0 || N > kMaxI32) return fail(SEGSCAN_EINVAL, "size out of range");
  if (E == 0 || N == 0) return ok();
  if (!ddst || !index || !dsrc) return fail(SEGSCAN_EINVAL, "NULL pointer");
  if (ld_ddst < N || ld_dsrc < N) return fail(SEGSCAN_EINVAL, "leading dimension < N");
  return hip_status(launch_segment_gather(ddst, ld_ddst, index, E, static_cast<int>(N), counts, dsrc, ld_dsrc,
                                          static_cast<hipStream_t>(stream)),
                    "segment gather");
}

segscan_status_t segscan_probe_gather_out_len(int64_t row_bytes, int64_t n, int64_t chunk, int64_t* out_len) {
  if (!out_len || n < 0 || chunk < 1 || (row_bytes != 128 && row_bytes != 256))
    return fail(SEGSCAN_EINVAL, "row_bytes must be 128 or 256, chunk >= 1, n >= 0");
  const int64_t groups = (n + chunk - 1) / chunk;
  *out_len = std::max<int64_t>(1, (groups * (row_bytes / 16) + 255) / 256) * 256;
  return ok();
}

segscan_status_t segscan_probe_gather(const void* table, int64_t table_rows, int64_t row_bytes, const int32_t* idx,
                                      int64_t n, int64_t chunk, float* out, int64_t out_len, void* stream) {
  int64_t need = 0;
  segscan_status_t st = segscan_probe_gather_out_len(row_bytes, n, chunk, &need);
  if (st) return st;
  if (n == 0) return ok();
  if (!table || !idx || !out || table_rows < 1) return fail(SEGSCAN_EINVAL, "NULL pointer or empty table");
  if (reinterpret_cast<uintptr_t>(table) % 16) return fail(SEGSCAN_EINVAL, "table is not 16-byte aligned");
  if (out_len < need) return fail(SEGSCAN_EINVAL, "out holds %lld floats, %lld needed", (long long)out_len,
                                  (long long)need);
  mark_start(static_cast<hipStream_t>(stream));
  hipError_t e = launch_gather_probe(table, row_bytes, idx, n, chunk, out, out_len, static_cast<hipStream_t>(stream));
  mark_stop(static_cast<hipStream_t>(stream));
  return hip_status(e, "gather probe");
}

segscan_status_t segscan_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, void* stream) {
  if (n < 0 || n > kMaxI32) return fail(SEGSCAN_EINVAL, "n out of range");
  if (n == 0) return ok();
  if (!perm || !inv) return fail(SEGSCAN_EINVAL, "NULL pointer");
  return hip_status(launch_invert_permutation(perm, n, inv, static_cast<hipStream_t>(stream)), "invert permutation");
}

segscan_status_t segscan_segment_counts(const int32_t* index, int64_t E, int64_t dst_len, int32_t* counts,
                                        void* stream) {
  if (E < 0 || E > kMaxI32 || dst_len < 0 || dst_len >= kMaxI32) return fail(SEGSCAN_EINVAL, "size out of range");
  if (dst_len == 0) return ok();
  if (!counts || (E > 0 && !index)) return fail(SEGSCAN_EINVAL, "NULL pointer");
  return hip_status(launch_segment_counts(index, E, dst_len, counts, static_cast<hipStream_t>(stream)), "counts");
}

segscan_status_t segscan_check_csr(const segscan_csr_t* A, int32_t* flags_dev, void* stream) {
  segscan_status_t st = validate_csr(A, true);
  if (st) return st;
  if (!flags_dev) return fail(SEGSCAN_EINVAL, "flags_dev is NULL");
  return hip_status(launch_check_csr(A->rowptr, A->colind, A->nrow, A->ncol, A->nnz, flags_dev,
                                     static_cast<hipStream_t>(stream)),
                    "check_csr");
}

segscan_status_t segscan_check_index(const int32_t* index, int64_t E, int64_t dst_len, int32_t* flags_dev,
                                     void* stream) {
  if (E < 0 || E > kMaxI32 || dst_len < 0) return fail(SEGSCAN_EINVAL, "size out of range");
  if (!flags_dev || (E > 0 && !index)) return fail(SEGSCAN_EINVAL, "NULL pointer");
  return hip_status(launch_check_index(index, E, dst_len, flags_dev, static_cast<hipStream_t>(stream)), "check_index");
}

}  // extern "C"
