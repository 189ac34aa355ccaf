// Row-balanced sequential reduction (rbsr): one lane computes one row x CF
// columns, walking the row's nnz in order — the reference's
// csrspmm_seqreduce_rowbalance (test-cg/include/spmm/spmm.cuh:593-637) and
// csrspmm_seqreduce_rowbalance_cg (spmm_cg.cuh:593-692).
//
// The knobs keep their meaning (tile = columns per grid slice, CF columns per
// lane, thread_per_block lanes, block_factor rows of work), but lanes are laid
// out column-fastest ([row][column vector]) so that the lanes of one row read
// one dense row with consecutive 128-bit loads; the reference's cg variant puts
// rows fastest, which scatters every load instruction over as many rows as
// lanes. Values are fp32 (the reference's cg kernel reads them into an int,
// spmm_cg.cuh:608, which truncates non-integer A).
#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

template <typename VT, int CF>
__global__ void __launch_bounds__(1024) rbsr_kernel(RbsrParams p) {
  // Dense-row loads per batch. 4: banded 120000^2 N = 64 fp32 57 -> 43 us against
  // the two-at-a-time loop; 8 and 16 spill under the 128-VGPR cap of 1024-lane
  // workgroups and lose.
  constexpr int RB = 4;
  const int L = threadIdx.x;
  const int yr = L / p.tpr;
  const int xv = L - yr * p.tpr;
  if (yr >= p.rows_per_block) return;
  const int col0 = blockIdx.y * p.slice + xv * CF;
  const int nv = min(CF, p.N - col0);
  if (nv <= 0) return;
  const bool full = nv == CF;
  const VT* __restrict__ B = static_cast<const VT*>(p.B);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * p.rows_per_block;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * p.rows_per_block + yr; r < p.rows_total; r += stride) {
    const int start = p.rowptr[r];
    const int end = p.rowptr[r + 1];
    float acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) acc[i] = 0.0f;
    int q = start;
    // Batches of RB nnz: RB dense-row loads in flight before any is consumed.
    for (; q + RB <= end; q += RB) {
      int k[RB];
      float v[RB];
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        k[u] = p.colind[q + u];
        v[u] = p.vals ? p.vals[p.vmap ? p.vmap[q + u] : q + u] : 1.0f;
      }
      RawCols<VT, CF> raw[RB];
      if (full) {
#pragma unroll
        for (int u = 0; u < RB; ++u) load_raw<VT, CF>(B + static_cast<int64_t>(k[u]) * p.ldb + col0, raw[u]);
      }
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        float b[CF];
        if (full) widen<VT, CF>(raw[u], b);
        else load_cols_partial<VT, CF>(B + static_cast<int64_t>(k[u]) * p.ldb + col0, nv, b);
#pragma unroll
        for (int i = 0; i < CF; ++i) acc[i] += v[u] * b[i];
      }
    }
    // The rest two at a time.
    for (; q + 1 < end; q += 2) {
      const int64_t k0 = p.colind[q];
      const int64_t k1 = p.colind[q + 1];
      const float v0 = p.vals ? p.vals[p.vmap ? p.vmap[q] : q] : 1.0f;
      const float v1 = p.vals ? p.vals[p.vmap ? p.vmap[q + 1] : q + 1] : 1.0f;
      float b0[CF], b1[CF];
      load_cols_any<VT, CF>(B + k0 * p.ldb + col0, nv, b0);
      load_cols_any<VT, CF>(B + k1 * p.ldb + col0, nv, b1);
#pragma unroll
      for (int i = 0; i < CF; ++i) acc[i] += v0 * b0[i];
#pragma unroll
      for (int i = 0; i < CF; ++i) acc[i] += v1 * b1[i];
    }
    if (q < end) {
      const int64_t k0 = p.colind[q];
      const float v0 = p.vals ? p.vals[p.vmap ? p.vmap[q] : q] : 1.0f;
      float b0[CF];
      load_cols_any<VT, CF>(B + k0 * p.ldb + col0, nv, b0);
#pragma unroll
      for (int i = 0; i < CF; ++i) acc[i] += v0 * b0[i];
    }
    store_cols<CF>(p.C + r * p.ldc + col0, nv, acc, p.accumulate != 0);
  }
}

// The same walk when a row's lanes are a power of two in [8, 64] (so rows never straddle a
// wave): the row's lanes load the column and value of tpr consecutive entries at once (one
// each) and hand them round with shuffles, instead of every lane loading every entry's
// metadata; the next tile's metadata is in flight while the current tile's dense rows are
// gathered, RB at a time. Entries are summed in the same order as rbsr_kernel.
template <typename VT, int CF>
__global__ void __launch_bounds__(1024) rbsr_coop_kernel(RbsrParams p) {
  constexpr int RB = 8;  // dense-row loads per batch
  const int L = threadIdx.x;
  const int tpr = p.tpr;
  const int yr = L / tpr;
  const int xv = L - yr * tpr;
  if (yr >= p.rows_per_block) return;  // whole rows (tpr divides the wave)
  const int rbase = __lane_id() - xv;  // first lane of this row in the wave
  const int col0 = blockIdx.y * p.slice + xv * CF;
  const int nv = min(CF, p.N - col0);  // <= 0: the lane only carries metadata
  const bool full = nv == CF;
  const VT* __restrict__ B = static_cast<const VT*>(p.B);
  const uint32_t ldb32 = static_cast<uint32_t>(p.ldb);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * p.rows_per_block;
  auto meta = [&](int e, int end, int& k, float& v) {
    k = 0;
    v = 0.0f;
    if (e < end) {
      k = p.colind[e];
      v = p.vals ? p.vals[p.vmap ? p.vmap[e] : e] : 1.0f;
    }
  };
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * p.rows_per_block + yr; r < p.rows_total; r += stride) {
    const int start = p.rowptr[r];
    const int end = p.rowptr[r + 1];
    float acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) acc[i] = 0.0f;
    int mk;
    float mv;
    meta(start + xv, end, mk, mv);
    for (int q = start; q < end; q += tpr) {
      const int cnt = min(tpr, end - q);
      int nk;
      float nvv;
      meta(q + tpr + xv, end, nk, nvv);  // next tile, in flight during this one
      for (int b = 0; b < cnt; b += RB) {
        RawCols<VT, CF> raw[RB];
        int kk[RB];
        float vv[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
          kk[u] = __shfl(mk, rbase + b + u);
          vv[u] = __shfl(mv, rbase + b + u);
          // slots past the tile read dense row 0 (always valid) and are never summed:
          // one branch-free batch keeps all RB loads in flight together
          if (b + u >= cnt) kk[u] = 0;
        }
        if (full) {
#pragma unroll
          for (int u = 0; u < RB; ++u)
            load_raw<VT, CF>(B + static_cast<uint64_t>(static_cast<uint32_t>(kk[u])) * ldb32 + col0, raw[u]);
        }
#pragma unroll
        for (int u = 0; u < RB; ++u) {
          if (b + u < cnt && nv > 0) {
            float bb[CF];
            if (full) widen<VT, CF>(raw[u], bb);
            else load_cols_partial<VT, CF>(B + static_cast<int64_t>(kk[u]) * p.ldb + col0, nv, bb);
#pragma unroll
            for (int i = 0; i < CF; ++i) acc[i] += vv[u] * bb[i];
          }
        }
      }
      mk = nk;
      mv = nvv;
    }
    if (nv > 0) store_cols<CF>(p.C + r * p.ldc + col0, nv, acc, p.accumulate != 0);
  }
}

template <typename VT, int CF>
static hipError_t launch_rbsr_t(const RbsrParams& p, const Launch& l) {
  // With 8 columns per lane the widened rows of a batch take the cooperative kernel to
  // 113-120 VGPRs (4 waves per SIMD) and it is no faster; rbsr_kernel keeps that shape.
  if constexpr (CF <= 4) {
    if (p.tpr >= 8 && p.tpr <= 64 && (p.tpr & (p.tpr - 1)) == 0) {
      rbsr_coop_kernel<VT, CF><<<l.grid, l.block, 0, l.stream>>>(p);
      return hipGetLastError();
    }
  }
  rbsr_kernel<VT, CF><<<l.grid, l.block, 0, l.stream>>>(p);
  return hipGetLastError();
}

template <typename VT>
static hipError_t launch_rbsr_cf(int cf, const RbsrParams& p, const Launch& l, bool& ok) {
  ok = true;
  switch (cf) {
    case 1: return launch_rbsr_t<VT, 1>(p, l);
    case 2: return launch_rbsr_t<VT, 2>(p, l);
    case 4: return launch_rbsr_t<VT, 4>(p, l);
    case 8: return launch_rbsr_t<VT, 8>(p, l);
    default: ok = false; return hipSuccess;
  }
}

hipError_t launch_spmm_rbsr(VType vt, int cf, const RbsrParams& p, const Launch& l, bool& ok) {
  switch (vt) {
    case VType::F32: return launch_rbsr_cf<float>(cf, p, l, ok);
    case VType::BF16: return launch_rbsr_cf<bf16_t>(cf, p, l, ok);
    case VType::F16: return launch_rbsr_cf<f16_t>(cf, p, l, ok);
  }
  ok = false;
  return hipSuccess;
}

}  // namespace segscan
