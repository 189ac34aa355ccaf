// Fused backward of C = A.B for a fixed pattern, in one walk over At = A^T:
//
//   dB[j, :]       = sum over the entries q of row j of At of  vals[perm[q]] * dC[rowind[q], :]
//   dvals[perm[q]] = <dC[rowind[q], :], B[j, :]>
//
// Both products need the same dC row for every entry, so one gather feeds the
// transposed SpMM (dB = A^T.dC) and the SDDMM on A's pattern (dvals), which the
// separate path reads twice (once per kernel, each a random row gather).
//
// The walk is the element-balanced sequential walk of the forward SpMM: a group
// of G lanes covers all N columns (G * CF >= N, so each dot product ends inside
// the group) and walks a chunk of consecutive At entries; rows come from the
// group's rowptr window (group_row_find). Per tile of G entries every lane
// stages its entry's dC row, At row and A position in LDS and holds its entry's
// value: read in At order when the caller has the values in that order (one
// tile ahead, coalesced), else through perm from A's order (a random 4-byte
// read, which costs a whole cache line of fabric traffic). Per batch of U entries the group issues U dC-row loads
// before consuming any. The B row of the current At row stays in registers and
// the next one is in flight (At rows are consecutive unless empty ones lie in
// between, which are then loaded on demand). dB rows use the owner policy (a chunk
// stores the rows it owns, zeroes its empty ones but whole entry-free blocks of
// them (left to the fix-up pass), and sends a first row
// that began in an earlier chunk to carry_val for the fix-up pass). Each dvals
// entry is written once (perm is a bijection) after a DPP reduction across the
// group. Where At rows are short (a few entries each) the one-row-ahead B load
// arrives too late; the row-batch form (RB) instead issues the B row of every row
// that starts inside a batch together with the batch's dC loads.
#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

// XT: storage type of dC (gathered); YT: storage type of B (one row per At row).
// RB: B rows loaded per batch of entries (short At rows) instead of one row ahead.
template <typename XT, typename YT, int CF, int G, bool RB>
__device__ __forceinline__ void bwd_walk(const BwdParams& p, int32_t* bwd_smem) {
  int32_t* s_k = bwd_smem;               // [tpb] dC row of the entry (a row of A)
  int32_t* s_r = bwd_smem + p.tpb;       // [tpb] At row of the entry (a column of A)
  int32_t* s_p = bwd_smem + 2 * p.tpb;   // [tpb] position of the entry in A

  constexpr int U = (G < 8) ? G : 8;  // entries per batch of loads
  const int L = threadIdx.x;
  const int g = L / G;
  const int g_lane = L - g * G;
  const int wave_gbase = __lane_id() - g_lane;
  const int groups_per_block = p.tpb / G;
  const int64_t ggid = static_cast<int64_t>(blockIdx.x) * groups_per_block + g;
  const int col0 = g_lane * CF;
  const int nv = min(CF, p.N - col0);
  const bool full = nv == CF;
  const XT* __restrict__ X = static_cast<const XT*>(p.dC);  // gathered by At's column index
  const YT* __restrict__ Y = static_cast<const YT*>(p.B);   // row j of At
  const int sbase = g * G;
  const bool want_db = p.dB != nullptr;
  const bool want_dv = p.dvals != nullptr;
  const bool acc_mode = p.accumulate != 0;
  const bool need_perm = (!p.vals_t && p.vals) || (want_dv && !p.dvals_at_order);
  // empty dB rows: partial blocks zeroed here, entry-free blocks by the fix-up pass
  const bool zrun = want_db && !acc_mode;

  for (int64_t c = ggid;; c += p.total_groups) {
    if (__ballot(c < p.nchunks) == 0) break;
    const bool chunk_valid = c < p.nchunks;
    const int64_t c0 = c * p.chunk;
    const int64_t clen = chunk_valid ? min(p.chunk, p.nnz - c0) : 0;

    // Rows of entries c0 and c0 - 1 (the second search only when a row starts at c0).
    int first_row = 0, prev_row = -1;
    if (chunk_valid) {
      if constexpr (G >= 8) first_row = static_cast<int>(group_search_gt<G>(p.rowptr, 1, p.K, c0)) - 1;
      else first_row = row_of_nnz(p.rowptr, p.K, c0);
      if (c0 == 0) {
        prev_row = -1;
      } else if (p.rowptr[first_row] < c0) {
        prev_row = first_row;
      } else {
        if constexpr (G >= 8) prev_row = static_cast<int>(group_search_gt<G>(p.rowptr, 1, p.K, c0 - 1)) - 1;
        else prev_row = row_of_nnz(p.rowptr, p.K, c0 - 1);
      }
    }
    const bool is_carry = chunk_valid && c0 > 0 && prev_row == first_row;
    if (chunk_valid && want_db) {
      if (g_lane == 0) p.carry_row[c] = is_carry ? first_row : -1;
      if (zrun) {
        // the run between the previous chunk's last row (-1 for chunk 0) and this one's first
        owner_empty_run<CF, G, false>(p.dB, p.ld_db, col0, nv, p.K, prev_row, first_row, nullptr, false, g_lane);
      }
    }
    int cur = first_row;
    float acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) acc[i] = 0.0f;

    // B row of the current At row (y) and of the next one (ynx, in flight).
    float y[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) y[i] = 0.0f;
    RawCols<YT, CF> ynx;
    auto load_y = [&](int r) {
      if (nv > 0) load_cols_any<YT, CF>(Y + static_cast<int64_t>(r) * p.ld_b + col0, nv, y);
    };
    auto issue_next = [&](int r) {
      if (full && r + 1 < p.K) load_raw<YT, CF>(Y + static_cast<int64_t>(r + 1) * p.ld_b + col0, ynx);
    };
    if (want_dv && chunk_valid && clen > 0) {
      load_y(first_row);
      if constexpr (!RB) issue_next(first_row);
    }

    auto flush = [&](int row) {
      if (!want_db || nv <= 0) return;
      if (row == first_row && is_carry) store_cols<CF>(p.carry_val + c * p.N + col0, nv, acc, false);
      else store_cols<CF>(p.dB + static_cast<int64_t>(row) * p.ld_db + col0, nv, acc, acc_mode);
    };

    // Metadata of the next tile, one tile ahead: dC row, A position (and the
    // At-order value) of the lane's entry, and the rowptr window after r_hint.
    int r_hint = first_row;
    int pk = 0, ppos = 0, prp = INT32_MAX;
    float pvt = 1.0f;
    auto prefetch = [&](int64_t off_n) {
      const bool vn = off_n + g_lane < clen;
      const int64_t en = c0 + off_n + g_lane;
      pk = vn ? p.colind[en] : 0;
      if (p.dc_map != nullptr && vn) pk = p.dc_map[pk];  // dC row of A's row pk
      // A position: for values read through perm, and for dvals in A's order
      ppos = (vn && need_perm) ? p.perm[en] : 0;
      if (p.vals_t) pvt = vn ? p.vals_t[en] : 1.0f;
      if (off_n < clen) {
        const int jj = r_hint + 1 + g_lane;
        prp = (jj <= p.K) ? p.rowptr[jj] : INT32_MAX;
      }
    };
    prefetch(0);

    const int ntile = static_cast<int>((p.chunk + G - 1) / G);  // wave-uniform
    for (int j = 0; j < ntile; ++j) {
      const int64_t off = static_cast<int64_t>(j) * G;
      if (__ballot(off < clen) == 0) break;
      const bool valid = off + g_lane < clen;
      const int cnt = static_cast<int>(max<int64_t>(0, min<int64_t>(G, clen - off)));
      const int row = group_row_find<G>(p.rowptr, p.K, r_hint, prp, static_cast<int>(c0 + off + g_lane), valid);
      const int last_row = __shfl(row, wave_gbase + max(cnt, 1) - 1);
      if (cnt > 0) r_hint = last_row;
      const int pos = ppos;
      // The value of the lane's entry; read back by shuffles when consumed.
      const float myv = p.vals_t ? pvt : ((valid && p.vals) ? p.vals[pos] : 1.0f);
      s_k[L] = pk;
      s_r[L] = valid ? row : INT32_MAX;
      s_p[L] = p.dvals_at_order ? static_cast<int>(c0 + off + g_lane) : pos;  // where dvals goes
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      for (int pp = 0; pp < cnt; pp += U) {
        RawCols<XT, CF> xr[U];
        RawCols<YT, CF> yr[RB ? U : 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (full && pp + u < cnt) {
            load_raw<XT, CF>(X + static_cast<int64_t>(s_k[sbase + pp + u]) * p.ld_dc + col0, xr[u]);
            if constexpr (RB) {
              // the B row of an entry that starts a row (the consume loop's test below)
              const int ru = s_r[sbase + pp + u];
              const int rp = u == 0 ? cur : s_r[sbase + pp + u - 1];
              if (want_dv && ru != rp) load_raw<YT, CF>(Y + static_cast<int64_t>(ru) * p.ld_b + col0, yr[u]);
            }
          }
        }
        if (pp == 0) prefetch(off + G);
        float part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          part[u] = 0.0f;
          // every lane takes part in the shuffle; only entries < cnt are used
          const float v = __shfl(myv, wave_gbase + min(pp + u, G - 1));
          if (pp + u < cnt) {
            const int r = s_r[sbase + pp + u];
            if (r != cur) {
              flush(cur);
              if (zrun) {
                owner_empty_run<CF, G, false>(p.dB, p.ld_db, col0, nv, p.K, cur, r, nullptr, false, g_lane);
              }
              if (want_dv) {
                if constexpr (RB) {
                  if (full) widen<YT, CF>(yr[u], y);
                  else load_y(r);
                } else {
                  if (full && r == cur + 1) widen<YT, CF>(ynx, y);
                  else load_y(r);
                  issue_next(r);
                }
              }
              cur = r;
#pragma unroll
              for (int i = 0; i < CF; ++i) acc[i] = 0.0f;
            }
            float xa[CF];
            if (full) {
              widen<XT, CF>(xr[u], xa);
            } else if (nv > 0) {
              load_cols_partial<XT, CF>(X + static_cast<int64_t>(s_k[sbase + pp + u]) * p.ld_dc + col0, nv, xa);
            } else {
#pragma unroll
              for (int i = 0; i < CF; ++i) xa[i] = 0.0f;
            }
#pragma unroll
            for (int i = 0; i < CF; ++i) {
              acc[i] += v * xa[i];
              part[u] += xa[i] * y[i];
            }
          }
        }
        if (want_dv) {
          // The batch's U dot products reduced across the group by a butterfly (U - 1 pair
          // exchanges plus log2(G / U) final ones, instead of log2 G shifts per entry);
          // lane l then holds the total of entry butterfly_entry(l) and the lanes that
          // hold one store it.
          constexpr int kLevels = __builtin_ctz(U);
          float lvl[kLevels + 1];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            float d = part[u];
#pragma unroll
            for (int h = 0; h <= kLevels; ++h) {
              if (h == kLevels || ((u >> h) & 1) == 0) {
                lvl[h] = d;
                break;
              }
              d = butterfly<G>(lvl[h], d, G >> (h + 1));
            }
          }
          float tot = lvl[kLevels];
#pragma unroll
          for (int m = (G >> (kLevels + 1)); m >= 1; m >>= 1) tot += __shfl_xor(tot, m, G);
          const int eu = butterfly_entry<G, U>(g_lane);
          // non-temporal: a scattered 4-byte store would otherwise hold an L2 line that the
          // dC gathers reuse (power-law 262144^2, N = 64: 218 -> 196 us with At-order
          // values); non-temporal value loads measured slower and are not used
          if ((g_lane & ((G >> kLevels) - 1)) == 0 && pp + eu < cnt)
            __builtin_nontemporal_store(tot, p.dvals + s_p[sbase + pp + eu]);
        }
      }
      __builtin_amdgcn_wave_barrier();  // LDS slots are rewritten by the next tile
    }
    if (chunk_valid && clen > 0) {
      flush(cur);
      // the last chunk owns the run after the last row with entries
      if (zrun && c == p.nchunks - 1) owner_empty_run<CF, G, false>(p.dB, p.ld_db, col0, nv, p.K, cur, p.K, nullptr, false, g_lane);
    }
  }
}

template <typename XT, typename YT, int CF, int G>
__global__ void __launch_bounds__(256) bwd_sr_kernel(BwdParams p) {
  extern __shared__ int32_t bwd_smem[];
  bwd_walk<XT, YT, CF, G, false>(p, bwd_smem);
}

// Row-batch form, bf16 B only (a batch's 8 B rows take 2 VGPRs each). Held to 4
// waves per SIMD (at most 128 VGPRs): with an fp32 dC it would otherwise take 130
// and drop to 3 waves, which costs more than the early B loads gain.
template <typename XT, typename YT, int CF, int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) bwd_sr_rowbatch_kernel(BwdParams p) {
  extern __shared__ int32_t bwd_smem[];
  bwd_walk<XT, YT, CF, G, true>(p, bwd_smem);
}

template <typename XT, typename YT, int CF, int G>
static hipError_t launch_bwd_t(const BwdParams& p, const Launch& l) {
  bwd_sr_kernel<XT, YT, CF, G><<<l.grid, l.block, l.shmem, l.stream>>>(p);
  return hipGetLastError();
}

template <typename XT, int G>
static hipError_t launch_bwd_rb(const BwdParams& p, const Launch& l) {
  bwd_sr_rowbatch_kernel<XT, bf16_t, 4, G><<<l.grid, l.block, l.shmem, l.stream>>>(p);
  return hipGetLastError();
}

template <typename XT, typename YT, int CF>
static hipError_t launch_bwd_g(int g, const BwdParams& p, const Launch& l, bool& ok) {
  ok = true;
  switch (g) {
    case 1: return launch_bwd_t<XT, YT, CF, 1>(p, l);
    case 2: return launch_bwd_t<XT, YT, CF, 2>(p, l);
    case 4: return launch_bwd_t<XT, YT, CF, 4>(p, l);
    case 8: return launch_bwd_t<XT, YT, CF, 8>(p, l);
    case 16: return launch_bwd_t<XT, YT, CF, 16>(p, l);
    case 32: return launch_bwd_t<XT, YT, CF, 32>(p, l);
    case 64: return launch_bwd_t<XT, YT, CF, 64>(p, l);
    default: ok = false; return hipSuccess;
  }
}

template <typename XT, typename YT>
static hipError_t launch_bwd_cf(int cf, int g, const BwdParams& p, const Launch& l, bool& ok) {
  switch (cf) {
    case 1: return launch_bwd_g<XT, YT, 1>(g, p, l, ok);
    case 2: return launch_bwd_g<XT, YT, 2>(g, p, l, ok);
    case 4: return launch_bwd_g<XT, YT, 4>(g, p, l, ok);
    case 8: return launch_bwd_g<XT, YT, 8>(g, p, l, ok);
    default: ok = false; return hipSuccess;
  }
}

// dC and B share a storage type, or dC is fp32 and B 16-bit (the gradient kept in
// fp32 against a 16-bit stored operand).
hipError_t launch_spmm_bwd(VType xt, VType yt, int cf, int g, bool row_batch, const BwdParams& p, const Launch& l,
                           bool& ok) {
  if (row_batch) {
    ok = yt == VType::BF16 && (xt == VType::F32 || xt == VType::BF16) && cf == 4 && (g == 8 || g == 16 || g == 32);
    if (!ok) return hipSuccess;
    if (xt == VType::F32)
      return g == 8 ? launch_bwd_rb<float, 8>(p, l) : g == 16 ? launch_bwd_rb<float, 16>(p, l) : launch_bwd_rb<float, 32>(p, l);
    return g == 8 ? launch_bwd_rb<bf16_t, 8>(p, l) : g == 16 ? launch_bwd_rb<bf16_t, 16>(p, l) : launch_bwd_rb<bf16_t, 32>(p, l);
  }
  if (xt == yt) {
    switch (xt) {
      case VType::F32: return launch_bwd_cf<float, float>(cf, g, p, l, ok);
      case VType::BF16: return launch_bwd_cf<bf16_t, bf16_t>(cf, g, p, l, ok);
      case VType::F16: return launch_bwd_cf<f16_t, f16_t>(cf, g, p, l, ok);
    }
  } else if (xt == VType::F32) {
    if (yt == VType::BF16) return launch_bwd_cf<float, bf16_t>(cf, g, p, l, ok);
    if (yt == VType::F16) return launch_bwd_cf<float, f16_t>(cf, g, p, l, ok);
  }
  ok = false;
  return hipSuccess;
}

}  // namespace segscan
