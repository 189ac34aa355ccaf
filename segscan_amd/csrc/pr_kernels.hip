// Parallel-reduction kernels: element-balanced SpMM (ebpr), row-balanced SpMM
// (rbpr) and the element-parallel segment_coo / segment_coo_new.
//
// Work decomposition follows the reference templates
// (test-cg/include/spmm/spmm_cg.cuh:14-218, :220-374;
//  new-cg/include/segscan/segscan.cuh:7-184, :186-355): a tile of X lanes holds X
// consecutive nnz (or one row, for rbpr), each lane loads CF contiguous columns
// of one dense row, and aligned groups of G lanes reduce together. What is
// MI355X-specific: wave64 groups of up to 64 lanes reduced with DPP row shifts
// (G <= 16) or ds_bpermute (G = 32, 64) using one ballot of segment ends; the CSR
// row of each nnz comes from a per-tile partition entry plus a coalesced
// 64-entry rowptr window (no per-lane 18-step global binary search); 64-bit
// offsets; padding lanes neither load B nor write C.
#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

// Lane layout helper: flat lane -> (y tile, panel, lane in tile).
struct PrLane {
  int y, px, t;
  bool exists;
};

__device__ __forceinline__ PrLane pr_lane(const PrParams& p) {
  PrLane r;
  const int L = threadIdx.x;
  const int per_y = p.P * p.X;
  r.y = L / per_y;
  const int rem = L - r.y * per_y;
  r.px = rem / p.X;
  r.t = rem - r.px * p.X;
  r.exists = r.y < p.Y;
  return r;
}

// Element-balanced parallel reduction. SEG = segment_coo (row = index[e],
// dense row = e, value 1) instead of CSR SpMM.
template <typename VT, int CF, int G, bool SEG>
__global__ void __launch_bounds__(1024) pr_eb_kernel(PrParams p) {
  const PrLane ln = pr_lane(p);
  const int col0 = blockIdx.y * p.slice + ln.px * CF;
  const int nv = ln.exists ? min(CF, p.N - col0) : 0;
  // A wave whose panels all lie past N has nothing to do (wave-uniform exit).
  if (__ballot(nv > 0) == 0) return;
  const VT* __restrict__ B = static_cast<const VT*>(p.B);
  const int64_t tiles_per_iter = static_cast<int64_t>(gridDim.x) * p.Y;

  for (int64_t it_base = static_cast<int64_t>(blockIdx.x) * p.Y;; it_base += tiles_per_iter) {
    const int64_t tile = it_base + ln.y;
    const bool tile_valid = ln.exists && tile < p.ntiles;
    // Tiles grow with the lane index, so the wave's first lane holds its lowest tile.
    if (__ballot(tile_valid) == 0) break;
    const int64_t e = tile * p.X + ln.t;
    const bool valid = tile_valid && e < p.nnz && nv > 0;

    int row;
    if constexpr (SEG) {
      row = valid ? p.index[e] : INT32_MAX;
    } else {
      const int64_t wave_tile = __builtin_amdgcn_readfirstlane(static_cast<int>(tile));
      const int r0 = p.tile_row[wave_tile];
      const int64_t e_lo = wave_tile * p.X;
      const int e_c = static_cast<int>((tile_valid && e < p.nnz) ? e : e_lo);
      row = wave_row_lookup(p.rowptr, p.M, r0, e_c);
      if (!valid) row = INT32_MAX;
    }

    float c[CF];
    if (valid) {
      int64_t k;
      float v;
      if constexpr (SEG) {
        k = e;
        v = 1.0f;
      } else {
        k = p.colind[e];
        v = p.vals ? p.vals[p.vmap ? p.vmap[e] : e] : 1.0f;
      }
      load_cols_any<VT, CF>(B + k * p.ldb + col0, nv, c);
#pragma unroll
      for (int i = 0; i < CF; ++i) c[i] *= v;
    } else {
#pragma unroll
      for (int i = 0; i < CF; ++i) c[i] = 0.0f;
    }

    const bool head = group_segmented_reduce<G, CF>(c, row);
    if (head && row != INT32_MAX) {
      atomic_add_cols<CF>(p.C + static_cast<int64_t>(row) * p.ldc + col0, nv, c);
    }
  }
}

// Row-balanced parallel reduction: a tile of X lanes strides over one row's nnz,
// groups of G lanes reduce; with G == X the row total is stored, with G < X the
// X/G group partials are added atomically (the reference's rbpr_cg, :278).
template <typename VT, int CF, int G>
__global__ void __launch_bounds__(1024) pr_rb_kernel(PrParams p) {
  const PrLane ln = pr_lane(p);
  const int col0 = blockIdx.y * p.slice + ln.px * CF;
  const int nv = ln.exists ? min(CF, p.N - col0) : 0;
  // Groups share (y, px), so a lane without columns belongs to a group without
  // columns; the whole group leaves together.
  if (nv <= 0) return;
  const VT* __restrict__ B = static_cast<const VT*>(p.B);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * p.Y;
  const int g_lane = __lane_id() & (G - 1);

  for (int64_t r = static_cast<int64_t>(blockIdx.x) * p.Y + ln.y; r < p.M; r += stride) {
    const int start = p.rowptr[r];
    const int end = p.rowptr[r + 1];
    float c[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) c[i] = 0.0f;
    for (int jj = start + ln.t; jj < end; jj += p.X) {
      const int64_t k = p.colind[jj];
      const float v = p.vals ? p.vals[p.vmap ? p.vmap[jj] : jj] : 1.0f;
      float b[CF];
      load_cols_any<VT, CF>(B + k * p.ldb + col0, nv, b);
#pragma unroll
      for (int i = 0; i < CF; ++i) c[i] += v * b[i];
    }
    group_reduce_to_first<G, CF>(c);
    if (g_lane == 0) {
      float* dst = p.C + r * p.ldc + col0;
      if (p.row_store) store_cols<CF>(dst, nv, c, p.accumulate != 0);
      else atomic_add_cols<CF>(dst, nv, c);
    }
  }
}

// ---- instantiation and dispatch ----------------------------------------------------------

template <typename VT, int CF, int G>
static hipError_t launch_pr_t(const PrParams& p, const Launch& l, int kind) {
  if (kind == 0) pr_eb_kernel<VT, CF, G, false><<<l.grid, l.block, 0, l.stream>>>(p);
  else if (kind == 1) pr_rb_kernel<VT, CF, G><<<l.grid, l.block, 0, l.stream>>>(p);
  else pr_eb_kernel<VT, CF, G, true><<<l.grid, l.block, 0, l.stream>>>(p);
  return hipGetLastError();
}

template <typename VT, int CF>
static hipError_t launch_pr_g(int g, const PrParams& p, const Launch& l, int kind, bool& ok) {
  ok = true;
  switch (g) {
    case 1: return launch_pr_t<VT, CF, 1>(p, l, kind);
    case 2: return launch_pr_t<VT, CF, 2>(p, l, kind);
    case 4: return launch_pr_t<VT, CF, 4>(p, l, kind);
    case 8: return launch_pr_t<VT, CF, 8>(p, l, kind);
    case 16: return launch_pr_t<VT, CF, 16>(p, l, kind);
    case 32: return launch_pr_t<VT, CF, 32>(p, l, kind);
    case 64: return launch_pr_t<VT, CF, 64>(p, l, kind);
    default: ok = false; return hipSuccess;
  }
}

template <typename VT>
static hipError_t launch_pr_cf(int cf, int g, const PrParams& p, const Launch& l, int kind, bool& ok) {
  switch (cf) {
    case 1: return launch_pr_g<VT, 1>(g, p, l, kind, ok);
    case 2: return launch_pr_g<VT, 2>(g, p, l, kind, ok);
    case 4: return launch_pr_g<VT, 4>(g, p, l, kind, ok);
    case 8:
      if constexpr (sizeof(VT) == 2) return launch_pr_g<VT, 8>(g, p, l, kind, ok);
      [[fallthrough]];
    default: ok = false; return hipSuccess;
  }
}

static hipError_t launch_pr_vt(VType vt, int cf, int g, const PrParams& p, const Launch& l, int kind, bool& ok) {
  switch (vt) {
    case VType::F32: return launch_pr_cf<float>(cf, g, p, l, kind, ok);
    case VType::BF16: return launch_pr_cf<bf16_t>(cf, g, p, l, kind, ok);
    case VType::F16: return launch_pr_cf<f16_t>(cf, g, p, l, kind, ok);
  }
  ok = false;
  return hipSuccess;
}

hipError_t launch_spmm_pr(VType vt, int cf, int g, const PrParams& p, const Launch& l, bool eb, bool& ok) {
  return launch_pr_vt(vt, cf, g, p, l, eb ? 0 : 1, ok);
}

hipError_t launch_seg_pr(VType vt, int cf, int g, const PrParams& p, const Launch& l, bool& ok) {
  return launch_pr_vt(vt, cf, g, p, l, 2, ok);
}

}  // namespace segscan
