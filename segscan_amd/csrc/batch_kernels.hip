// Batched step over many graphs in one persistent launch.
//
// A is the block-diagonal CSR of the batch (graph g owns rows and columns
// [offsets[g], offsets[g+1])). One run computes, for node features X (K x N) and
// the gradient dP of a per-graph sum pooling:
//   H     = A . X                          (forward SpMM, fp32)
//   P[g]  = sum of H rows of graph g       (pooling)
//   dX    = A^T . dH with dH = dP[graph]   (backward)
//   dvals = SDDMM(dH, X) on A's pattern    (backward, one value per entry)
// Because A is block diagonal, every entry of column j lies in a row of graph
// g(j), so A^T . dP[graph] = colsum(A)[j] * dP[g(j)] (a streaming pass over the
// nodes, colsum fixed with the values at plan time), and dvals[p] =
// <dP[g(row p)], X[colind[p]]> uses the X row the forward already gathers.
//
// Work: the plan cuts each graph's rows into "items" of about the same nnz + rows
// (an item never spans two graphs), and hands whole graphs to the XCDs: a slot table
// maps workgroup w's k-th item to an item of a graph placed on w's XCD, so each graph's
// dense rows are gathered through a single L2. A workgroup takes an item and splits its entries
// evenly over its groups of G lanes (G * CF >= N columns); each group walks its
// entries in order (16 dense-row loads in flight; full tiles predicate-free), stores
// the rows it owns, and leaves the partials of rows crossing its boundaries in LDS,
// where the owner adds them after one workgroup barrier: items end at row boundaries,
// so no partial leaves the workgroup and there is no fix-up pass. Per-entry dot
// products for dvals go through a streaming butterfly across the group (about one
// cross-lane exchange per entry, log2(16) values pending). The workgroup also writes
// dX for the item's nodes and zeroes its empty rows of H. Pool partials are combined
// per item and added atomically into an accumulator that stays zero between runs:
// the last workgroup to finish (a system-scope counter, incremented only after each
// workgroup's returning atomics have completed) moves it into P with atomic
// exchanges, zeroing it for the next run. No workgroup ever waits on another.
#include <algorithm>
#include <cstdio>
#include <type_traits>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/segscan.h"
#include "abi_internal.hpp"
#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

constexpr int kBatchTpb = 256;

struct BatchParams {
  int M;                   // nodes (rows = columns of A)
  int N;
  int n_graphs;
  int nitems;
  int nslots;              // work slots: slot s runs on the XCD of workgroup s % grid
  const int32_t* item_of_slot;  // [nslots] item of each slot, -1 for none
  const int32_t* rowptr;
  const int32_t* colind;
  const float* vals;       // nullable: ones
  const int32_t* offsets;  // [n_graphs + 1]
  const int32_t* graph_of; // [M]
  const int32_t* item_row; // [nitems + 1] first row of each item
  const int4* sub;         // [nitems * NG + 1] per group of each item: {first entry, first row |
                           // 1 << 31 when that row started before, the item's graph, 0}
  const float* colsum;     // [M]
  const void* X;
  int64_t ld_x;
  const float* dP;         // nullable: forward only
  int64_t ld_dp;
  float* H;                // nullable
  int64_t ld_h;
  float* P;
  int64_t ld_p;
  float* dX;               // nullable
  int64_t ld_dx;
  float* dvals;            // nullable
  float* P_acc;            // [n_graphs x N], zero between runs
  int32_t* done;           // zero between runs
};

// Adds v at dst with a returning system-scope atomic: the returned value arrives only
// once the add has been performed where every XCD sees it.
__device__ __forceinline__ void coherent_add(float* dst, float v) {
  const float old = __hip_atomic_fetch_add(dst, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  asm volatile("" ::"v"(old));
}

// Stores a lane's CF fp32 columns as non-temporal writes: the outputs are written once
// and never read here, so they should not push the gathered X rows out of L2.
template <int CF>
__device__ __forceinline__ void store_cols_nt(float* __restrict__ p, int nv, const float (&v)[CF]) {
  if (nv >= CF) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    if constexpr (CF == 1) {
      __builtin_nontemporal_store(v[0], p);
    } else if constexpr (CF == 2) {
      __builtin_nontemporal_store(f2{v[0], v[1]}, reinterpret_cast<f2*>(p));
    } else {
#pragma unroll
      for (int i = 0; i < CF; i += 4)
        __builtin_nontemporal_store(f4{v[i], v[i + 1], v[i + 2], v[i + 3]}, reinterpret_cast<f4*>(p + i));
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < CF; ++i) {
    if (i < nv) __builtin_nontemporal_store(v[i], p + i);
  }
}

template <int CF>
__device__ __forceinline__ void load_f32_cols(const float* __restrict__ p, int nv, float (&out)[CF]) {
  if (nv >= CF) {
    load_cols<float, CF>(p, out);
  } else {
#pragma unroll
    for (int i = 0; i < CF; ++i) out[i] = (i < nv) ? p[i] : 0.0f;
  }
}


// Waves per SIMD the register allocator is held to: with up to 8 bytes of a dense row
// per lane the kernel fits 128 VGPRs with a few spills, and 4 waves beat 3 unspilled
// ones (256 graphs, bf16 N = 64: 0.60 against 0.64 ms); wider slices, and fp16 at 8
// bytes (its widening takes more registers), spill too much and stay uncapped.
template <typename VT, int CF>
constexpr int kBatchWaves = (sizeof(VT) * CF <= 8 && !(std::is_same_v<VT, f16_t> && CF == 4)) ? 4 : 1;

// FULL: dP, dvals and dX all given (the whole step), fixed at compile time: one replay of the
// 256-graph bf16 step 0.5135 -> 0.5080 ms over three pairs.
template <typename VT, int CF, int G, bool FULL = false>
__global__ void __launch_bounds__(kBatchTpb) __attribute__((amdgpu_waves_per_eu(kBatchWaves<VT, CF>)))
batch_step_kernel(BatchParams p) {
  constexpr int NG = kBatchTpb / G;   // groups per workgroup
  constexpr int U = (G < 16) ? G : 16;  // entries per batch of dense-row loads
  constexpr int W = G * CF;           // columns one group covers (>= N)
  constexpr int kLevels = __builtin_ctz(U);  // butterfly levels per batch
  __shared__ int32_t s_k[kBatchTpb];
  __shared__ float s_v[kBatchTpb];
  __shared__ int32_t s_r[kBatchTpb];
  __shared__ float s_head[NG][W];
  __shared__ float s_tail[NG][W];
  __shared__ float s_pool[NG][W];
  __shared__ int32_t s_head_row[NG];
  __shared__ int32_t s_tail_row[NG];
  __shared__ int32_t s_pool_g[NG];
  __shared__ int32_t s_last;

  const int L = threadIdx.x;
  const int g = L / G;
  const int gl = L - g * G;
  const int wave_gbase = __lane_id() - gl;
  const int col0 = gl * CF;
  const int nv = min(CF, p.N - col0);
  const bool full = nv == CF;
  const VT* __restrict__ X = static_cast<const VT*>(p.X);
  const bool want_bwd = FULL || p.dP != nullptr;
  const bool want_dvals = FULL || (want_bwd && p.dvals != nullptr);
  const bool want_dx = FULL || (want_bwd && p.dX != nullptr);
  const uint32_t ldx32 = static_cast<uint32_t>(p.ld_x);
  auto x_row = [&](int k) { return X + static_cast<uint64_t>(static_cast<uint32_t>(k)) * ldx32 + col0; };

  for (int slot = blockIdx.x; slot < p.nslots; slot += gridDim.x) {
    const int it = p.item_of_slot[slot];
    if (it < 0) continue;  // uniform over the workgroup
    // this group's share of the item, laid out at plan time
    const int4 sub = p.sub[static_cast<int64_t>(it) * NG + g];
    const int64_t s0 = sub.x;
    const int64_t s1 = p.sub[static_cast<int64_t>(it) * NG + g + 1].x;
    const int64_t clen = s1 - s0;
    const int first_row = sub.y & 0x7fffffff;
    const bool starts_before = sub.y < 0;

    int head_row = -1, tail_row = -1;
    float pool_acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) pool_acc[i] = 0.0f;
    int cur = first_row;
    const int pg = sub.z;
    float dp[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) dp[i] = 0.0f;
    if (want_bwd && nv > 0) load_f32_cols<CF>(p.dP + static_cast<int64_t>(pg) * p.ld_dp + col0, nv, dp);
    float acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) acc[i] = 0.0f;

    // A finished row's partial: owned rows are stored; the row that started before s0
    // (head) and the row that continues past s1 (tail) go to LDS.
    auto flush = [&](int row, bool last) {
      const bool is_head = row == first_row && starts_before;
      const bool is_tail = !is_head && last && p.rowptr[row + 1] > s1;
      if (is_head) {
        head_row = row;
        if (nv > 0) {
#pragma unroll
          for (int i = 0; i < CF; ++i) s_head[g][col0 + i] = acc[i];
        }
      } else if (is_tail) {
        tail_row = row;
        if (nv > 0) {
#pragma unroll
          for (int i = 0; i < CF; ++i) s_tail[g][col0 + i] = acc[i];
        }
      } else if (p.H && nv > 0) {
        store_cols_nt<CF>(p.H + static_cast<int64_t>(row) * p.ld_h + col0, nv, acc);
      }
#pragma unroll
      for (int i = 0; i < CF; ++i) pool_acc[i] += acc[i];
    };

    // Next row r (items never span graphs, so the row stays in graph pg): no load here,
    // which keeps the batch's dense-row loads in flight across row changes.
    auto next_row = [&](int r) {
      flush(cur, false);
      cur = r;
#pragma unroll
      for (int i = 0; i < CF; ++i) acc[i] = 0.0f;
    };

    // metadata of the next tile: column and value of the lane's entry, and the rowptr
    // window after r_hint
    int r_hint = first_row;
    int pk = 0, prp = INT32_MAX;
    float pv = 0.0f;
    auto prefetch = [&](int64_t off_n) {
      const bool vn = off_n + gl < clen;
      const int64_t en = s0 + off_n + gl;
      // read once: non-temporal, leaving L2 to the X rows the gather reuses (one replay
      // 0.5238 -> 0.5194 ms over four pairs)
      pk = vn ? __builtin_nontemporal_load(p.colind + en) : 0;
      pv = vn ? (p.vals ? __builtin_nontemporal_load(p.vals + en) : 1.0f) : 0.0f;
      if (off_n < clen) {
        const int jj = r_hint + 1 + gl;
        prp = (jj <= p.M) ? p.rowptr[jj] : INT32_MAX;
      }
    };
    prefetch(0);

    // Stages the tile at `off` (rows found; column, value and row of each entry in the
    // group's LDS slots) and returns the group's entry count in it.
    const int sbase = g * G;
    auto stage = [&](int64_t off) -> int {
      const bool valid = off + gl < clen;
      const int cnt = static_cast<int>(max<int64_t>(0, min<int64_t>(G, clen - off)));
      const int row = group_row_find<G>(p.rowptr, p.M, r_hint, prp, static_cast<int>(s0 + off + gl), valid);
      const int last_row = __shfl(row, wave_gbase + max(cnt, 1) - 1);
      if (cnt > 0) r_hint = last_row;
      s_k[L] = pk;
      s_v[L] = pv;
      s_r[L] = valid ? row : INT32_MAX;
      // read back only by lanes of the same wave
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      return cnt;
    };

    // Entry (row r, value v, dense row b): the row sum, and its dvals partial d.
    auto consume = [&](int r, float v, const float (&b)[CF], float& d) {
      if (r != cur) next_row(r);
#pragma unroll
      for (int i = 0; i < CF; ++i) {
        acc[i] += v * b[i];
        d += dp[i] * b[i];
      }
    };
    // Entry u of a batch enters the butterfly at level 0; an odd index at a level
    // combines with the value pending there.
    float lvl[kLevels + 1];
    auto feed = [&](int u, float d) {
#pragma unroll
      for (int h = 0; h <= kLevels; ++h) {
        if (h == kLevels || ((u >> h) & 1) == 0) {
          lvl[h] = d;
          break;
        }
        d = butterfly<G>(lvl[h], d, G >> (h + 1));
      }
    };
    auto emit = [&](int64_t off, int pp, int cnt) {
      float tot = lvl[kLevels];
#pragma unroll
      for (int m = (G >> (kLevels + 1)); m >= 1; m >>= 1) tot += __shfl_xor(tot, m, G);
      const int eu = butterfly_entry<G, U>(gl);
      // written once, read by no later pass of the step: non-temporal (one replay
      // 0.5177 -> 0.5126 ms over four pairs)
      if ((gl & ((G >> kLevels) - 1)) == 0 && pp + eu < cnt) __builtin_nontemporal_store(tot, p.dvals + s0 + off + pp + eu);
    };

    // A tile full for every group of the wave, in a wave whose lanes all hold CF
    // columns: no per-entry predicates; staged metadata read four entries at a time.
    auto fast_tile = [&](int64_t off) {
      for (int pp = 0; pp < G; pp += U) {
        RawCols<VT, CF> raw[U];
#pragma unroll
        for (int u = 0; u < U; u += 4) {
          const int4 kk = *reinterpret_cast<const int4*>(s_k + sbase + pp + u);
          load_raw<VT, CF>(x_row(kk.x), raw[u]);
          load_raw<VT, CF>(x_row(kk.y), raw[u + 1]);
          load_raw<VT, CF>(x_row(kk.z), raw[u + 2]);
          load_raw<VT, CF>(x_row(kk.w), raw[u + 3]);
        }
        if (pp == 0) prefetch(off + G);
#pragma unroll
        for (int u = 0; u < U; u += 4) {
          const int4 rr = *reinterpret_cast<const int4*>(s_r + sbase + pp + u);
          const float4 vv = *reinterpret_cast<const float4*>(s_v + sbase + pp + u);
          const int ra[4] = {rr.x, rr.y, rr.z, rr.w};
          const float va[4] = {vv.x, vv.y, vv.z, vv.w};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            float b[CF];
            widen<VT, CF>(raw[u + q], b);
            float d = 0.0f;
            consume(ra[q], va[q], b, d);
            if (want_dvals) feed(u + q, d);
          }
        }
        if (want_dvals) emit(off, pp, G);
      }
    };

    // Any tile: per-entry predicates for partial tiles and partial column vectors.
    auto slow_tile = [&](int64_t off, int cnt) {
      for (int pp = 0; pp < cnt; pp += U) {
        RawCols<VT, CF> raw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (full && pp + u < cnt) load_raw<VT, CF>(x_row(s_k[sbase + pp + u]), raw[u]);
        }
        if (pp == 0) prefetch(off + G);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float d = 0.0f;
          if (pp + u < cnt) {
            const int r = s_r[sbase + pp + u];
            const float v = s_v[sbase + pp + u];
            float b[CF];
            if (full) {
              widen<VT, CF>(raw[u], b);
            } else if (nv > 0) {
              load_cols_partial<VT, CF>(x_row(s_k[sbase + pp + u]), nv, b);
            } else {
#pragma unroll
              for (int i = 0; i < CF; ++i) b[i] = 0.0f;
            }
            consume(r, v, b, d);
          }
          if (want_dvals) feed(u, d);
        }
        if (want_dvals) emit(off, pp, cnt);
      }
    };

    // Full tiles run in the fast loop; the first tile not full for every group of the
    // wave hands over, already staged, to the predicated loop (separate regions for the
    // register allocator; tile counts and exits are wave-uniform).
    const bool lanes_full = __ballot(!full) == 0;
    int64_t off = 0;
    int cnt = 0;
    bool staged = false;
    if (lanes_full) {
      for (;; off += G) {
        if (__ballot(off < clen) == 0) break;
        cnt = stage(off);
        if (__ballot(cnt != G) != 0) {
          staged = true;
          break;
        }
        fast_tile(off);
        __builtin_amdgcn_wave_barrier();  // LDS slots are rewritten by the next tile
      }
    }
    for (;; off += G) {
      if (!staged) {
        if (__ballot(off < clen) == 0) break;
        cnt = stage(off);
      }
      staged = false;
      slow_tile(off, cnt);
      __builtin_amdgcn_wave_barrier();
    }
    if (clen > 0) flush(cur, true);

    // The item's rows are nodes of graph pg: dX[j] = colsum[j] * dP[pg] (dP[pg] is in
    // dp), and empty rows of H are zero. Group g takes rows r0 + g + NG q; its lanes load
    // the column sums and empty-row flags of G such rows at once, then the group stores
    // those rows back to back with the values shuffled out (one load round trip per G rows).
    if (p.H || want_dx) {
      const int r0 = p.item_row[it], r1 = p.item_row[it + 1];
      const int nrows_g = r1 - r0 > g ? (r1 - r0 - g + NG - 1) / NG : 0;  // this group's rows
      for (int q0 = 0; q0 < nrows_g; q0 += G) {
        const int jl = r0 + g + (q0 + gl) * NG;
        float csl = 0.0f;
        int eml = 0;
        if (q0 + gl < nrows_g) {
          if (want_dx) csl = p.colsum[jl];
          if (p.H) eml = p.rowptr[jl] == p.rowptr[jl + 1];
        }
        const int cnt = min(G, nrows_g - q0);
        for (int q = 0; q < cnt; ++q) {
          const int jq = r0 + g + (q0 + q) * NG;
          const float cs = __shfl(csl, wave_gbase + q);
          const int em = __shfl(eml, wave_gbase + q);
          if (nv > 0) {
            if (em) zero_cols<CF>(p.H + static_cast<int64_t>(jq) * p.ld_h + col0, nv);
            if (want_dx) {
              float d[CF];
#pragma unroll
              for (int i = 0; i < CF; ++i) d[i] = cs * dp[i];
              store_cols_nt<CF>(p.dX + static_cast<int64_t>(jq) * p.ld_dx + col0, nv, d);
            }
          }
        }
      }
    }
    if (gl == 0) {
      s_head_row[g] = clen > 0 ? head_row : -2;  // -2: a group without entries (items shorter than NG)
      s_tail_row[g] = tail_row;
      s_pool_g[g] = clen > 0 ? pg : -1;
    }
    if (nv > 0) {
#pragma unroll
      for (int i = 0; i < CF; ++i) s_pool[g][col0 + i] = pool_acc[i];
    }
    __syncthreads();

    // rows that cross group boundaries: the owner (tail) adds the heads that follow
    const int t = s_tail_row[g];
    if (t >= 0 && nv > 0) {
      float tot[CF];
#pragma unroll
      for (int i = 0; i < CF; ++i) tot[i] = s_tail[g][col0 + i];
      for (int k2 = g + 1; k2 < NG; ++k2) {
        const int hr = s_head_row[k2];
        if (hr == -2) continue;
        if (hr != t) break;
#pragma unroll
        for (int i = 0; i < CF; ++i) tot[i] += s_head[k2][col0 + i];
      }
      if (p.H) store_cols_nt<CF>(p.H + static_cast<int64_t>(t) * p.ld_h + col0, nv, tot);
    }
    // pool partials of the item's open graphs: the first group of each run of one graph
    // adds the run's sum
    const int mg = s_pool_g[g];
    if (mg >= 0 && nv > 0) {
      int prev = g - 1;
      while (prev >= 0 && s_pool_g[prev] < 0) --prev;
      if (prev < 0 || s_pool_g[prev] != mg) {
        float tot[CF];
#pragma unroll
        for (int i = 0; i < CF; ++i) tot[i] = s_pool[g][col0 + i];
        for (int k2 = g + 1; k2 < NG; ++k2) {
          const int gk = s_pool_g[k2];
          if (gk < 0) continue;
          if (gk != mg) break;
#pragma unroll
          for (int i = 0; i < CF; ++i) tot[i] += s_pool[k2][col0 + i];
        }
#pragma unroll
        for (int i = 0; i < CF; ++i) {
          if (i < nv) coherent_add(p.P_acc + static_cast<int64_t>(mg) * p.N + col0 + i, tot[i]);
        }
      }
    }
    __syncthreads();  // LDS is reused by the next item
  }

  // ---- the last workgroup moves the pool accumulator into P ----
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this lane's pool atomics have completed
  __syncthreads();
  if (L == 0) {
    const int old = __hip_atomic_fetch_add(p.done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    s_last = old == static_cast<int>(gridDim.x) - 1;
  }
  __syncthreads();
  if (s_last) {
    const int64_t total = static_cast<int64_t>(p.n_graphs) * p.N;
    for (int64_t i = L; i < total; i += kBatchTpb) {
      const int64_t gi = i / p.N;
      const float v = __hip_atomic_exchange(p.P_acc + i, 0.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      p.P[gi * p.ld_p + (i - gi * p.N)] = v;
    }
    if (L == 0) __hip_atomic_store(p.done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---- set-up kernels ----

__global__ void batch_graph_of_kernel(const int32_t* __restrict__ offsets, int n_graphs, int32_t* __restrict__ graph_of) {
  const int g = blockIdx.x;
  if (g >= n_graphs) return;
  for (int j = offsets[g] + threadIdx.x; j < offsets[g + 1]; j += blockDim.x) graph_of[j] = g;
}

// Bit 1: an entry's column lies in another graph than its row (or outside [0, M)).
__global__ void batch_check_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind, int M,
                                   const int32_t* __restrict__ graph_of, int32_t* __restrict__ flags) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < M; r += stride) {
    const int gr = graph_of[r];
    for (int q = rowptr[r]; q < rowptr[r + 1]; ++q) {
      const int c = colind[q];
      if (c < 0 || c >= M || graph_of[c] != gr) {
        atomicOr(flags, 1);
        break;
      }
    }
  }
}

// colsum[j] = sum of column j's values, in the transpose's (row) order.
__global__ void batch_colsum_kernel(const int32_t* __restrict__ colptr, const float* __restrict__ vals_t, int K,
                                    float* __restrict__ colsum) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < K; j += stride) {
    float s = 0.0f;
    for (int q = colptr[j]; q < colptr[j + 1]; ++q) s += vals_t ? vals_t[q] : 1.0f;
    colsum[j] = s;
  }
}

namespace {

using KernelFn = void (*)(BatchParams);

template <typename VT, int CF, bool FULL>
KernelFn pick_g(int g) {
  switch (g) {
    case 8: return batch_step_kernel<VT, CF, 8, FULL>;
    case 16: return batch_step_kernel<VT, CF, 16, FULL>;
    case 32: return batch_step_kernel<VT, CF, 32, FULL>;
    case 64: return batch_step_kernel<VT, CF, 64, FULL>;
    default: return nullptr;
  }
}

template <typename VT, bool FULL>
KernelFn pick_cf(int cf, int g) {
  switch (cf) {
    case 1: return pick_g<VT, 1, FULL>(g);
    case 2: return pick_g<VT, 2, FULL>(g);
    case 4: return pick_g<VT, 4, FULL>(g);
    case 8:
      if constexpr (sizeof(VT) == 2) return pick_g<VT, 8, FULL>(g);
      return nullptr;
    default: return nullptr;
  }
}

// full: the kernel form for a run that passes dP, dvals and dX
template <bool FULL>
KernelFn pick_kernel(segscan_dtype_t t, int cf, int g) {
  switch (t) {
    case SEGSCAN_F32: return pick_cf<float, FULL>(cf, g);
    case SEGSCAN_BF16: return pick_cf<bf16_t, FULL>(cf, g);
    case SEGSCAN_F16: return pick_cf<f16_t, FULL>(cf, g);
    default: return nullptr;
  }
}

}  // namespace
}  // namespace segscan

// =========================================================================================
// C ABI: the batch plan

struct segscan_batch_plan {
  int device = 0;
  int64_t M = 0, nnz = 0, n_graphs = 0, N = 0;
  segscan_dtype_t xt = SEGSCAN_F32;
  const int32_t* rowptr = nullptr;
  const int32_t* colind = nullptr;
  const float* vals = nullptr;
  int CF = 1, G = 64;
  int nitems = 0, grid = 1, nslots = 0;
  int64_t target = 0;   // entries + rows per item
  segscan::KernelFn kernel = nullptr;
  segscan::KernelFn kernel_full = nullptr;  // the same with dP, dvals and dX fixed as given
  // plan-owned device memory (one allocation)
  void* mem = nullptr;
  size_t mem_bytes = 0;
  int32_t* offsets = nullptr;
  int32_t* graph_of = nullptr;
  int4* sub = nullptr;
  int32_t* item_row = nullptr;
  int32_t* item_of_slot = nullptr;
  float* colsum = nullptr;
  float* P_acc = nullptr;
  int32_t* done = nullptr;
};

namespace {

using segscan::set_error;

constexpr int64_t kMaxI32 = 2147483647LL;

size_t a256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

segscan_status_t hip_fail(hipError_t e, const char* what) {
  char msg[256];
  std::snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
  return set_error(SEGSCAN_EHIP, msg);
}

segscan_status_t einval(const char* msg) { return set_error(SEGSCAN_EINVAL, msg); }

// Column sums of A's values, in the transpose's order (deterministic), into plan->colsum.
segscan_status_t compute_colsum(segscan_batch_plan* pl, hipStream_t s) {
  segscan_csr_t A{};
  A.nrow = pl->M;
  A.ncol = pl->M;
  A.nnz = pl->nnz;
  A.rowptr = pl->rowptr;
  A.colind = pl->colind;
  A.vals = pl->vals;
  size_t ws_bytes = 0;
  segscan_status_t st = segscan_csr_transpose_workspace_size(&A, &ws_bytes);
  if (st != SEGSCAN_OK) return st;
  const size_t n = static_cast<size_t>(std::max<int64_t>(pl->nnz, 1));
  const size_t bytes = a256((pl->M + 1) * 4) + 3 * a256(n * 4) + a256(ws_bytes);
  char* tmp = nullptr;
  hipError_t e = hipMalloc(&tmp, bytes);
  if (e != hipSuccess) return hip_fail(e, "batch plan: transpose scratch");
  int32_t* colptr = reinterpret_cast<int32_t*>(tmp);
  int32_t* rowind = reinterpret_cast<int32_t*>(tmp + a256((pl->M + 1) * 4));
  int32_t* perm = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(rowind) + a256(n * 4));
  float* vals_t = reinterpret_cast<float*>(reinterpret_cast<char*>(perm) + a256(n * 4));
  void* ws = reinterpret_cast<char*>(vals_t) + a256(n * 4);
  st = segscan_csr_transpose(&A, colptr, rowind, perm, pl->vals ? vals_t : nullptr, ws, ws_bytes, s);
  if (st == SEGSCAN_OK) {
    const int blocks = static_cast<int>(std::min<int64_t>((pl->M + 255) / 256, 65536));
    segscan::batch_colsum_kernel<<<std::max(blocks, 1), 256, 0, s>>>(colptr, pl->vals ? vals_t : nullptr,
                                                                      static_cast<int>(pl->M), pl->colsum);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = hip_fail(e, "batch plan: column sums");
  }
  hipError_t ef = hipFree(tmp);
  if (st == SEGSCAN_OK && ef != hipSuccess) st = hip_fail(ef, "batch plan: free");
  return st;
}

}  // namespace

extern "C" {

segscan_status_t segscan_batch_plan_create(const segscan_csr_t* A, const int64_t* graph_offsets, int64_t n_graphs,
                                           int64_t N, segscan_dtype_t x_dtype, const segscan_knobs_t* knobs,
                                           void* stream, segscan_batch_plan_t* plan) {
  if (!plan) return einval("plan is NULL");
  *plan = nullptr;
  if (!A) return einval("A is NULL");
  if (A->nrow < 0 || A->nrow >= kMaxI32 || A->nrow != A->ncol) return einval("A must be square with < 2^31 rows");
  if (A->nnz < 0 || A->nnz > kMaxI32) return einval("nnz out of range");
  if (A->nrow > 0 && !A->rowptr) return einval("rowptr is NULL");
  if (A->nnz > 0 && !A->colind) return einval("colind is NULL");
  if (A->vals_map) return set_error(SEGSCAN_EUNSUPPORTED, "the batch plan reads values in A's order (vals_map NULL)");
  if (!graph_offsets) return einval("graph_offsets is NULL");
  if (n_graphs < 1 || n_graphs >= kMaxI32) return einval("n_graphs out of range");
  if (graph_offsets[0] != 0 || graph_offsets[n_graphs] != A->nrow)
    return einval("graph_offsets must start at 0 and end at the number of rows");
  for (int64_t g = 0; g < n_graphs; ++g)
    if (graph_offsets[g + 1] < graph_offsets[g]) return einval("graph_offsets must be non-decreasing");
  if (N < 1 || N > 512) return set_error(SEGSCAN_EUNSUPPORTED, "N must be in [1, 512]");
  if (x_dtype != SEGSCAN_F32 && x_dtype != SEGSCAN_BF16 && x_dtype != SEGSCAN_F16) return einval("unknown dtype");
  if (A->nrow == 0) return einval("empty batch");

  // lane shape: 4-byte slices per lane (two 16-bit columns from N = 64 on), wider when N
  // needs more than 64 lanes; one group covers all N columns
  const int es = x_dtype == SEGSCAN_F32 ? 4 : 2;
  int cf = 8 / es;  // 8 bytes of a dense row per lane
  while (cf > 1 && (N + cf - 1) / cf < 8) cf /= 2;
  if (knobs && knobs->coarsen_factor > 0) cf = knobs->coarsen_factor;
  while ((N + cf - 1) / cf > 64 && cf < (es == 2 ? 8 : 4)) cf *= 2;
  int lanes = static_cast<int>((N + cf - 1) / cf);
  int g = 8;
  while (g < lanes) g *= 2;
  if (knobs && knobs->group_factor != 0) {
    const int gk = knobs->group_factor < 0 ? 1 : (1 << knobs->group_factor);
    if (gk < lanes) return set_error(SEGSCAN_EUNSUPPORTED, "group_factor gives fewer lanes than N / coarsen_factor");
    g = gk;
  }
  if (g > 64 || g < 8 || static_cast<int64_t>(g) * cf < N)
    return set_error(SEGSCAN_EUNSUPPORTED, "N does not fit one group of at most 64 lanes");
  segscan::KernelFn kern = segscan::pick_kernel<false>(x_dtype, cf, g);
  if (!kern) return set_error(SEGSCAN_EUNSUPPORTED, "no compiled batch kernel for this dtype / coarsen / group");
  // entries per group per item
  int tnz = 8;
  if (knobs && knobs->thread_nz > 0) tnz = knobs->thread_nz;
  if (tnz > 4096) return einval("thread_nz too large");

  hipStream_t s = static_cast<hipStream_t>(stream);
  auto* pl = new segscan_batch_plan();
  pl->M = A->nrow;
  pl->nnz = A->nnz;
  pl->n_graphs = n_graphs;
  pl->N = N;
  pl->xt = x_dtype;
  pl->rowptr = A->rowptr;
  pl->colind = A->colind;
  pl->vals = A->vals;
  pl->CF = cf;
  pl->G = g;
  pl->kernel = kern;
  pl->kernel_full = segscan::pick_kernel<true>(x_dtype, cf, g);
  const int ngroups = segscan::kBatchTpb / g;
  pl->target = static_cast<int64_t>(tnz) * g * ngroups;

  segscan_status_t st = SEGSCAN_OK;
  hipError_t e = hipGetDevice(&pl->device);
  // items: contiguous rows of about `target` entries + rows (a longer row is an item of its own)
  std::vector<int32_t> rp(static_cast<size_t>(pl->M) + 1);
  if (e == hipSuccess) e = hipMemcpyAsync(rp.data(), A->rowptr, rp.size() * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "batch plan: read rowptr");
  }
  if (rp[0] != 0 || rp[pl->M] != pl->nnz) {
    delete pl;
    return einval("rowptr must span [0, nnz]");
  }
  std::vector<int32_t> items;
  std::vector<int32_t> item_graph;
  for (int64_t gi = 0; gi < n_graphs; ++gi) {
    // items never span graphs: the kernel's walk then needs no graph lookup
    int64_t acc = 0;
    for (int64_t r = graph_offsets[gi]; r < graph_offsets[gi + 1]; ++r) {
      const int64_t c = static_cast<int64_t>(rp[r + 1]) - rp[r] + 1;
      if (r == graph_offsets[gi] || acc + c > pl->target) {
        items.push_back(static_cast<int32_t>(r));
        item_graph.push_back(static_cast<int32_t>(gi));
        acc = 0;
      }
      acc += c;
    }
  }
  items.push_back(static_cast<int32_t>(pl->M));
  pl->nitems = static_cast<int>(items.size()) - 1;
  // per group of each item: its first entry, the row holding it (flagged when the row
  // started in an earlier group), and that row's graph with the graph's end row
  std::vector<int4> sub(static_cast<size_t>(pl->nitems) * ngroups + 1);
  for (int it = 0; it < pl->nitems; ++it) {
    const int r0 = items[it], r1 = items[it + 1];
    const int64_t e0 = rp[r0], e1 = rp[r1], len = e1 - e0;
    for (int gi = 0; gi < ngroups; ++gi) {
      const int64_t s0 = e0 + len * gi / ngroups;
      int4 rec = make_int4(static_cast<int>(s0), r0, 0, 0);
      if (s0 < e1) {
        // first row j - 1 with rp[j] > s0, j in [r0 + 1, r1]
        const int j = static_cast<int>(std::upper_bound(rp.begin() + r0 + 1, rp.begin() + r1 + 1, static_cast<int32_t>(s0)) -
                                       rp.begin());
        const int fr = j - 1;
        rec.y = fr | (rp[fr] < s0 ? static_cast<int>(0x80000000u) : 0);
      }
      rec.z = item_graph[it];
      sub[static_cast<size_t>(it) * ngroups + gi] = rec;
    }
  }
  sub.back() = make_int4(static_cast<int>(pl->nnz), 0, 0, 0);

  // persistent grid: as many workgroups as stay resident, at most one per item
  int cus = 0, per_cu = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), segscan::kBatchTpb, 0);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "batch plan: occupancy");
  }
  int64_t grid = static_cast<int64_t>(std::max(cus, 1)) * std::max(per_cu, 1);
  if (knobs && knobs->block_numer > 0 && knobs->block_denom > 0)
    grid = std::max<int64_t>(1, grid * knobs->block_numer / knobs->block_denom);
  grid = std::min<int64_t>(grid, pl->nitems);
  // XCD-aware slots: workgroups are dispatched to the XCDs round robin, so with a grid
  // that is a multiple of the XCD count, slot s (run by workgroup s % grid) always lands
  // on XCD s % kXcds. Whole graphs go to one XCD (largest first, to the XCD with the
  // least work), so a graph's dense rows are gathered through one L2, not eight (256
  // graphs, bf16 N = 64: L2 hit rate 37 -> 46 %, 1.40 -> 1.03 GB read past L2).
  std::vector<int32_t> slot_item;
  if (grid >= segscan::kXcds) {
    grid -= grid % segscan::kXcds;
    std::vector<int64_t> g_items(static_cast<size_t>(n_graphs), 0);
    for (int it = 0; it < pl->nitems; ++it) g_items[item_graph[it]]++;
    std::vector<int32_t> gorder(static_cast<size_t>(n_graphs));
    for (int64_t gi = 0; gi < n_graphs; ++gi) gorder[gi] = static_cast<int32_t>(gi);
    std::stable_sort(gorder.begin(), gorder.end(), [&](int32_t a, int32_t b) { return g_items[a] > g_items[b]; });
    std::vector<std::vector<int32_t>> xlist(segscan::kXcds);
    std::vector<int64_t> load(segscan::kXcds, 0);
    std::vector<int32_t> g_first(static_cast<size_t>(n_graphs), -1);
    for (int it = pl->nitems - 1; it >= 0; --it) g_first[item_graph[it]] = it;
    for (int32_t gi : gorder) {
      if (g_items[gi] == 0) continue;
      const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      for (int64_t k = 0; k < g_items[gi]; ++k) xlist[x].push_back(g_first[gi] + static_cast<int32_t>(k));
      load[x] += g_items[gi];
    }
    const int64_t longest = *std::max_element(load.begin(), load.end());
    slot_item.assign(static_cast<size_t>(longest) * segscan::kXcds, -1);
    for (int x = 0; x < segscan::kXcds; ++x)
      for (size_t j = 0; j < xlist[x].size(); ++j) slot_item[j * segscan::kXcds + x] = xlist[x][j];
  } else {
    slot_item.resize(static_cast<size_t>(pl->nitems));
    for (int it = 0; it < pl->nitems; ++it) slot_item[it] = it;
  }
  pl->nslots = static_cast<int>(slot_item.size());
  pl->grid = static_cast<int>(std::max<int64_t>(grid, 1));


  // plan memory: offsets, graph_of, group layout of the items, colsum, pool accumulator + counter, flags
  const size_t off_offsets = 0;
  const size_t off_graph = off_offsets + a256((n_graphs + 1) * 4);
  const size_t off_items = off_graph + a256(pl->M * 4);
  const size_t off_rows = off_items + a256(sub.size() * sizeof(int4));
  const size_t off_slots = off_rows + a256(items.size() * 4);
  const size_t off_colsum = off_slots + a256(slot_item.size() * 4);
  const size_t off_pacc = off_colsum + a256(pl->M * 4);
  const size_t off_done = off_pacc + a256(static_cast<size_t>(n_graphs) * N * 4);
  const size_t off_flags = off_done + 256;
  pl->mem_bytes = off_flags + 256;
  e = hipMalloc(&pl->mem, pl->mem_bytes);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "batch plan: allocation");
  }
  char* base = static_cast<char*>(pl->mem);
  pl->offsets = reinterpret_cast<int32_t*>(base + off_offsets);
  pl->graph_of = reinterpret_cast<int32_t*>(base + off_graph);
  pl->sub = reinterpret_cast<int4*>(base + off_items);
  pl->item_row = reinterpret_cast<int32_t*>(base + off_rows);
  pl->item_of_slot = reinterpret_cast<int32_t*>(base + off_slots);
  pl->colsum = reinterpret_cast<float*>(base + off_colsum);
  pl->P_acc = reinterpret_cast<float*>(base + off_pacc);
  pl->done = reinterpret_cast<int32_t*>(base + off_done);
  int32_t* flags = reinterpret_cast<int32_t*>(base + off_flags);

  std::vector<int32_t> off32(static_cast<size_t>(n_graphs) + 1);
  for (int64_t i = 0; i <= n_graphs; ++i) off32[i] = static_cast<int32_t>(graph_offsets[i]);
  e = hipMemcpyAsync(pl->offsets, off32.data(), off32.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(pl->sub, sub.data(), sub.size() * sizeof(int4), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(pl->item_row, items.data(), items.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pl->item_of_slot, slot_item.data(), slot_item.size() * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(pl->P_acc, 0, off_flags + 256 - off_pacc, s);
  if (e == hipSuccess) {
    segscan::batch_graph_of_kernel<<<static_cast<unsigned>(n_graphs), 256, 0, s>>>(pl->offsets,
                                                                                 static_cast<int>(n_graphs),
                                                                                 pl->graph_of);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const int blocks = static_cast<int>(std::min<int64_t>((pl->M + 255) / 256, 65536));
    segscan::batch_check_kernel<<<blocks, 256, 0, s>>>(A->rowptr, A->colind, static_cast<int>(pl->M), pl->graph_of,
                                                       flags);
    e = hipGetLastError();
  }
  int32_t hflags = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) st = hip_fail(e, "batch plan: set-up");
  else if (hflags) st = einval("A is not block diagonal over graph_offsets (an entry's column lies in another graph)");
  if (st == SEGSCAN_OK) st = compute_colsum(pl, s);
  if (st != SEGSCAN_OK) {
    (void)hipFree(pl->mem);
    delete pl;
    return st;
  }
  *plan = pl;
  set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_batch_plan_set_values(segscan_batch_plan_t plan, const float* vals, void* stream) {
  if (!plan) return einval("plan is NULL");
  plan->vals = vals;
  hipError_t e = hipSetDevice(plan->device);
  if (e != hipSuccess) return hip_fail(e, "batch plan: device");
  segscan_status_t st = compute_colsum(plan, static_cast<hipStream_t>(stream));
  if (st == SEGSCAN_OK) set_error(SEGSCAN_OK, "");
  return st;
}

segscan_status_t segscan_batch_plan_describe(segscan_batch_plan_t plan, segscan_batch_info_t* info) {
  if (!plan || !info) return einval("plan or info is NULL");
  std::memset(info, 0, sizeof(*info));
  info->nitems = plan->nitems;
  info->grid = plan->grid;
  info->group_size = plan->G;
  info->coarsen_factor = plan->CF;
  info->thread_per_block = segscan::kBatchTpb;
  info->item_target = plan->target;
  info->plan_bytes = plan->mem_bytes;
  return SEGSCAN_OK;
}

segscan_status_t segscan_batch_plan_run(segscan_batch_plan_t plan, const void* X, int64_t ld_x, const float* dP,
                                        int64_t ld_dp, float* H, int64_t ld_h, float* P, int64_t ld_p, float* dX,
                                        int64_t ld_dx, float* dvals, void* stream) {
  if (!plan) return einval("plan is NULL");
  const int64_t N = plan->N;
  if (!X || ld_x < N || ld_x > kMaxI32) return einval("X is NULL or ld_x out of [N, 2^31)");
  if (!P || ld_p < N) return einval("P is NULL or ld_p < N");
  if (H && ld_h < N) return einval("ld_h < N");
  if (dP && ld_dp < N) return einval("ld_dp < N");
  if (!dP && (dX || dvals)) return einval("dX / dvals need dP");
  if (dX && ld_dx < N) return einval("ld_dx < N");
  const int cf = plan->CF;
  const size_t es = plan->xt == SEGSCAN_F32 ? 4 : 2;
  const size_t xvec = std::min<size_t>(es * cf, 16), fvec = std::min<size_t>(4 * cf, 16);
  if (ld_x % cf || reinterpret_cast<uintptr_t>(X) % xvec) return einval("X rows must be aligned to the lane's columns");
  auto fok = [&](const float* q, int64_t ld) { return !q || (ld % cf == 0 && reinterpret_cast<uintptr_t>(q) % fvec == 0); };
  if (!fok(H, ld_h) || !fok(P, ld_p) || !fok(dX, ld_dx) || !fok(dP, ld_dp))
    return einval("H, P, dX and dP rows must be aligned to the lane's columns");
  segscan::BatchParams p{};
  p.M = static_cast<int>(plan->M);
  p.N = static_cast<int>(N);
  p.n_graphs = static_cast<int>(plan->n_graphs);
  p.nitems = plan->nitems;
  p.nslots = plan->nslots;
  p.item_of_slot = plan->item_of_slot;
  p.rowptr = plan->rowptr;
  p.colind = plan->colind;
  p.vals = plan->vals;
  p.offsets = plan->offsets;
  p.graph_of = plan->graph_of;
  p.sub = plan->sub;
  p.item_row = plan->item_row;
  p.colsum = plan->colsum;
  p.X = X;
  p.ld_x = ld_x;
  p.dP = dP;
  p.ld_dp = ld_dp;
  p.H = H;
  p.ld_h = ld_h;
  p.P = P;
  p.ld_p = ld_p;
  p.dX = dX;
  p.ld_dx = ld_dx;
  p.dvals = dvals;
  p.P_acc = plan->P_acc;
  p.done = plan->done;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const segscan::KernelFn kern = (dP && dvals && dX) ? plan->kernel_full : plan->kernel;
  hipLaunchKernelGGL(kern, dim3(plan->grid), dim3(segscan::kBatchTpb), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "batch step launch");
  set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_batch_plan_destroy(segscan_batch_plan_t plan) {
  if (!plan) return SEGSCAN_OK;
  hipError_t e = hipSuccess;
  if (plan->mem) e = hipFree(plan->mem);
  delete plan;
  if (e != hipSuccess) return hip_fail(e, "batch plan: free");
  return SEGSCAN_OK;
}

}  // extern "C"
