// Sequential-reduction kernels: element-balanced SpMM (ebsr, the reference's
// "row caching" kernels test-cg/include/spmm/spmm.cuh:345-591 and
// spmm_cg.cuh:376-591) and the segment_coo_sr family
// (new-cg/include/segscan/segscan.cuh:357-815).
//
// A group of G lanes covers G*CF contiguous columns (one 128-bit load per lane
// at CF=4 fp32 or CF=8 bf16/f16, so a 16-lane group reads a 256-byte dense row
// in one instruction) and walks a chunk of consecutive nnz in order,
// accumulating in registers and flushing when the row changes. Per tile of G
// nnz every lane stages one (column, value, row) triple in LDS; the group then
// reads them back as broadcast LDS loads and issues up to 16 dense-row loads
// before consuming any, while the next tile's column, value and rowptr window
// are already in flight (one-tile software pipeline), so that each group keeps
// many MALL/HBM requests outstanding. Rows come from a per-group rowptr window
// (group_row_find), not from a per-nnz binary search.
//
// Write policies:
//  * atomic: every flushed row is an atomic add (output zero-filled beforehand);
//  * owner:  chunk c owns rows (row(c0-1), row(c1-1)] — including empty rows
//            (zeroed as owner_empty_run describes) — and stores each once; a row
//            that started in an earlier chunk is a carry, written to carry_val[c]
//            and added by a fix-up pass. No atomics on the main path and no
//            zero-fill pass over the output;
//  * lessatom: rows that start and end inside the chunk are plain stores, the
//            chunk's first and last rows atomic adds (output zero-filled first).
#pragma once

#include <type_traits>

#include "device_common.hpp"
#include "kernel_params.hpp"

namespace segscan {

// VB: values from a permutation plan's staged buckets (SrParams::vslot); SpMM only.
// OWN: the owner write policy without accumulation, fixed at compile time (the tuned
// points' case): the flush path then holds one store form (tuned fp32 kernel 167.7 -> 166.4
// us, bf16 0.1156 -> 0.1141 ms per call, fwd + bwd step 0.3494 -> 0.3472 ms, three pairs);
// SpMM only.
template <typename VT, int CF, int G, bool SEG, bool VB = false, bool OWN = false>
__global__ void __launch_bounds__(1024) sr_kernel(SrParams p) {
  static_assert(!(SEG && VB), "staged values are an SpMM mode");
  static_assert(!(SEG && OWN), "the fixed owner form is an SpMM mode");
  extern __shared__ int32_t sr_smem[];
  int32_t* s_k = sr_smem;                                  // [tpb] column (SpMM)
  float* s_v = reinterpret_cast<float*>(sr_smem + p.tpb);  // [tpb] value (SpMM)
  int32_t* s_r = sr_smem + 2 * p.tpb;                      // [tpb] row / segment
  float* s_bkt = reinterpret_cast<float*>(sr_smem + 3 * p.tpb);  // [kPermBucket] staged values (vslot)

  constexpr int U = (G < 16) ? G : 16;  // dense-row loads issued back to back per batch
  // Full tiles take the predicate-free loop when batches split into fours. (With
  // 16-byte row slices the fp32 kernel spills a few registers under the 128-VGPR
  // cap of 1024-lane workgroups and is still ~1 % faster than without the loop.)
  constexpr bool kFastTiles = U % 4 == 0;
  const int L = threadIdx.x;
  const int g = L / G;
  const int g_lane = L - g * G;
  const int wave_gbase = __lane_id() - g_lane;  // first lane of this group in the wave
  const int groups_per_block = p.tpb / G;
  const int64_t ggid = static_cast<int64_t>(blockIdx.x) * groups_per_block + g;
  const int col0 = blockIdx.y * (G * CF) + g_lane * CF;
  const int nv = min(CF, p.N - col0);
  const bool full = nv == CF;
  const VT* __restrict__ B = static_cast<const VT*>(p.B);
  const VT* __restrict__ Bc = B + col0;  // only dereferenced by lanes with nv > 0
  // Dense row k of this lane: row index and leading dimension are both below 2^31,
  // so one 32 x 32 -> 64-bit multiply-add forms the address.
  const uint32_t ldb32 = static_cast<uint32_t>(p.ldb);
  auto row_at = [&](int64_t k) { return Bc + static_cast<uint64_t>(static_cast<uint32_t>(k)) * ldb32; };
  // segment_coo streams its src rows once, in order: non-temporal loads (src 1.07 GB,
  // N = 64 fp32: ~10 % faster); the SpMM gathers reuse B rows and keep plain loads
  auto load_row = [&](const VT* q, RawCols<VT, CF>& r) {
    if constexpr (SEG) load_raw_stream<VT, CF>(q, r);
    else load_raw<VT, CF>(q, r);
  };
  // Every lane of the wave holds CF columns: full tiles can skip per-lane column checks.
  const bool lanes_full = __ballot(!full) == 0;
  const int sbase = g * G;  // this group's LDS slots
  const bool acc_mode = !OWN && p.accumulate != 0;
  const bool owner = OWN || p.policy == 2;
  // empty runs: rows zeroed here, whole blocks flagged for the fix-up pass
  // (segments flag entry-free blocks; for a CSR matrix the fix-up pass reads rowptr)
  const bool zrun = owner && !acc_mode;
  const bool zflags = SEG && zrun && blockIdx.y == 0;
  // Values from a permutation plan's staged buckets: each workgroup round covers a part of
  // one bucket (checked by the host), which the workgroup loads into LDS; rounds are then
  // workgroup-uniform, and waves without a chunk in a round only join its barriers.
  constexpr bool vbucket = VB;

  for (int64_t c = ggid;; c += p.total_groups) {
    if (vbucket) {
      const int64_t cw = c - g;  // the workgroup's first chunk this round
      if (cw >= p.nchunks) break;
      __syncthreads();  // every lane is done with the previous round's bucket
      const int64_t b0 = (cw * p.chunk / kPermBucket) * kPermBucket;
      const int blen = static_cast<int>(min<int64_t>(kPermBucket, p.nnz - b0));
      // read once per workgroup: non-temporal, so the bucket does not take L2 room from the
      // B rows the gather reuses (fwd + bwd step 0.3536 -> 0.3522 ms over five pairs)
      typedef float f32x4 __attribute__((ext_vector_type(4)));
      const f32x4* src4 = reinterpret_cast<const f32x4*>(p.vmid + b0);
      for (int i = L; i < (blen >> 2); i += p.tpb)
        reinterpret_cast<f32x4*>(s_bkt)[i] = __builtin_nontemporal_load(src4 + i);
      for (int i = (blen & ~3) + L; i < blen; i += p.tpb) s_bkt[i] = p.vmid[b0 + i];
      // the barrier that publishes the bucket follows the chunk start search
    } else if (__ballot(c < p.nchunks) == 0) {
      // Chunks grow with the lane index; the wave's first lane has the lowest one.
      break;
    }
    const bool chunk_valid = c < p.nchunks;
    const int64_t c0 = c * p.chunk;
    const int64_t clen = chunk_valid ? min(p.chunk, p.nnz - c0) : 0;

    // Chunk entry state.
    int first_row = 0, prev_row = -1;
    if (chunk_valid) {
      if constexpr (SEG) {
        first_row = p.index[c0];
        prev_row = (c0 > 0) ? p.index[c0 - 1] : -1;
      } else {
        // Rows of nnz c0 and c0 - 1: a G-ary search by the group (a binary search
        // for groups under 8 lanes); the second search only when a row starts
        // exactly at c0.
        if constexpr (G >= 8) first_row = static_cast<int>(group_search_gt<G>(p.rowptr, 1, p.M, c0)) - 1;
        else first_row = row_of_nnz(p.rowptr, p.M, c0);
        if (c0 == 0) {
          prev_row = -1;
        } else if (p.rowptr[first_row] < c0) {
          prev_row = first_row;
        } else {
          if constexpr (G >= 8) prev_row = static_cast<int>(group_search_gt<G>(p.rowptr, 1, p.M, c0 - 1)) - 1;
          else prev_row = row_of_nnz(p.rowptr, p.M, c0 - 1);
        }
      }
    }
    const bool is_carry = owner && chunk_valid && c0 > 0 && prev_row == first_row;
    if (owner && chunk_valid) {
      if (g_lane == 0 && blockIdx.y == 0) p.carry_row[c] = is_carry ? first_row : -1;
      if (zrun) {
        // the run between the previous chunk's last row (-1 for chunk 0) and this one's first
        if (SEG && zflags && g_lane == 0) p.zflag[first_row >> kZeroRowsLog] = 0;
        owner_empty_run<CF, G, SEG>(p.C, p.ldc, col0, nv, p.M, prev_row, first_row, p.zflag, zflags, g_lane);
      }
    }
    int cur = first_row;
    float acc[CF];
#pragma unroll
    for (int i = 0; i < CF; ++i) acc[i] = 0.0f;

    // `last` marks the flush at the end of the chunk.
    auto flush = [&](int row, bool last) {
      if (nv <= 0) return;
      float* dst = p.C + static_cast<int64_t>(row) * p.ldc + col0;
      if (owner) {
        if (row == first_row && is_carry) store_cols<CF>(p.carry_val + c * p.N + col0, nv, acc, false);
        else store_cols<CF>(dst, nv, acc, acc_mode);
      } else if (!OWN && p.policy == 3 && !last && row != first_row) {
        store_cols<CF>(dst, nv, acc, acc_mode);
      } else {
        atomic_add_cols<CF>(dst, nv, acc);
      }
    };

    // Metadata of the next tile, loaded one tile ahead: column and value of the
    // lane's nnz, and the rowptr window starting after r_hint (a row no later
    // than any nnz still to come in the chunk).
    int r_hint = first_row;
    int pk = 0, prow = INT32_MAX, prp = INT32_MAX, pslot = 0;
    float pv = 0.0f;
    auto prefetch = [&](int64_t off_n) {
      const bool vn = off_n + g_lane < clen;
      const int64_t en = c0 + off_n + g_lane;
      if constexpr (SEG) {
        prow = vn ? p.index[en] : INT32_MAX;
      } else {
        pk = vn ? p.colind[en] : 0;
        if (vbucket) pslot = vn ? p.vslot[en] : 0;  // read from the bucket when staged
        else pv = vn ? (p.vals ? p.vals[p.vmap ? p.vmap[en] : en] : 1.0f) : 0.0f;
        if (off_n < clen) {
          const int jj = r_hint + 1 + g_lane;
          prp = (jj <= p.M) ? p.rowptr[jj] : INT32_MAX;
        }
      }
    };
    prefetch(0);
    // the round's bucket is in LDS before the first tile is staged (the first tile's
    // metadata loads are already in flight)
    if (vbucket) __syncthreads();

    // Entry (row r, value v) whose dense row is in b: a new row flushes the previous one.
    auto consume = [&](int r, float v, const float (&b)[CF]) {
      if (r != cur) {
        flush(cur, false);
        // rows of the run in partial blocks (segments: stage() flags the whole ones)
        if (zrun) owner_empty_run<CF, G, false>(p.C, p.ldc, col0, nv, p.M, cur, r, nullptr, false, g_lane);
        cur = r;
#pragma unroll
        for (int i = 0; i < CF; ++i) acc[i] = 0.0f;
      }
#pragma unroll
      for (int i = 0; i < CF; ++i) acc[i] += v * b[i];
    };

    // Segments under the owner policy: block flags of the tile's entries, lane-parallel
    // (each lane compares its segment with the previous entry's). A lane whose segment
    // starts a new block stores the block's 0; a run between two entries that covers
    // whole blocks is flagged by the group. The run before the chunk's first entry is
    // the chunk start's.
    int tprev = first_row;
    auto flag_tile = [&](int row, bool valid, int cnt) {
      const int up = __shfl(row, __lane_id() - (g_lane > 0 ? 1 : 0));
      const int rp = g_lane > 0 ? up : tprev;
      const bool fresh = valid && row != rp;
      if (fresh && (row >> kZeroRowsLog) != (rp >> kZeroRowsLog)) p.zflag[row >> kZeroRowsLog] = 0;
      const bool whole = fresh && ((rp + kZeroRows) >> kZeroRowsLog) < (row >> kZeroRowsLog);
      uint64_t m = (__ballot(whole) >> wave_gbase) & (G == 64 ? ~0ull : ((1ull << G) - 1));
      while (m) {
        const int src = wave_gbase + __ffsll(static_cast<unsigned long long>(m)) - 1;
        m &= m - 1;
        const int ra = __shfl(rp, src), rb = __shfl(row, src);
        owner_empty_run<1, G, true>(nullptr, 0, 0, 0, p.M, ra, rb, p.zflag, true, g_lane);
      }
      if (cnt > 0) tprev = __shfl(row, wave_gbase + cnt - 1);
    };

    // Stages the tile at `off`: rows found, (column, value, row) written to the
    // group's LDS slots. Returns the group's entry count in the tile.
    auto stage = [&](int64_t off) -> int {
      const bool valid = off + g_lane < clen;
      const int cnt = static_cast<int>(max<int64_t>(0, min<int64_t>(G, clen - off)));
      int row;
      if constexpr (SEG) {
        row = prow;
        if (zflags) flag_tile(row, valid, cnt);
      } else {
        const float v = vbucket ? s_bkt[pslot] : pv;  // read ahead of the row search
        row = group_row_find<G>(p.rowptr, p.M, r_hint, prp, static_cast<int>(c0 + off + g_lane), valid);
        const int last_row = __shfl(row, wave_gbase + max(cnt, 1) - 1);
        if (cnt > 0) r_hint = last_row;
        s_k[L] = pk;
        s_v[L] = v;
      }
      s_r[L] = valid ? row : INT32_MAX;
      // The triples are read back only by lanes of the same wave: order the LDS
      // writes before the reads without a workgroup barrier.
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      return cnt;
    };

    // A tile that is full for every group of the wave, in a wave whose lanes all
    // hold CF columns: no per-entry predicates, and the staged metadata is read
    // four entries per LDS access.
    auto fast_tile = [&](int64_t off, int cnt) {
      if constexpr (kFastTiles) {
        for (int pp = 0; pp < cnt; pp += U) {
          RawCols<VT, CF> raw[U];
#pragma unroll
          for (int u = 0; u < U; u += 4) {
            int4 kk;
            if constexpr (SEG) {
              const int e0 = static_cast<int>(c0 + off + pp + u);
              kk = make_int4(e0, e0 + 1, e0 + 2, e0 + 3);
            } else {
              kk = *reinterpret_cast<const int4*>(s_k + sbase + pp + u);
            }
            load_row(row_at(kk.x), raw[u]);
            load_row(row_at(kk.y), raw[u + 1]);
            load_row(row_at(kk.z), raw[u + 2]);
            load_row(row_at(kk.w), raw[u + 3]);
          }
          if (pp == 0) prefetch(off + G);
#pragma unroll
          for (int u = 0; u < U; u += 4) {
            const int4 rr = *reinterpret_cast<const int4*>(s_r + sbase + pp + u);
            float4 vv = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            if constexpr (!SEG) vv = *reinterpret_cast<const float4*>(s_v + sbase + pp + u);
            float b[CF];
            widen<VT, CF>(raw[u], b);
            consume(rr.x, vv.x, b);
            widen<VT, CF>(raw[u + 1], b);
            consume(rr.y, vv.y, b);
            widen<VT, CF>(raw[u + 2], b);
            consume(rr.z, vv.z, b);
            widen<VT, CF>(raw[u + 3], b);
            consume(rr.w, vv.w, b);
          }
        }
      }
    };

    // Any tile: per-entry predicates for partial tiles and partial column vectors.
    auto slow_tile = [&](int64_t off, int cnt) {
      for (int pp = 0; pp < cnt; pp += U) {
        // Issue the whole batch of dense-row loads before consuming any; the
        // row and value of each entry are read back from LDS only when consumed.
        RawCols<VT, CF> raw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (full && pp + u < cnt) {
            const int64_t krow = SEG ? (c0 + off + pp + u) : static_cast<int64_t>(s_k[sbase + pp + u]);
            load_row(row_at(krow), raw[u]);
          }
        }
        // The next tile's metadata travels while this batch's rows arrive.
        if (pp == 0) prefetch(off + G);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (pp + u < cnt) {
            const int r = s_r[sbase + pp + u];
            const float v = SEG ? 1.0f : s_v[sbase + pp + u];
            float b[CF];
            if (full) {
              widen<VT, CF>(raw[u], b);
            } else if (nv > 0) {
              const int64_t krow = SEG ? (c0 + off + pp + u) : static_cast<int64_t>(s_k[sbase + pp + u]);
              load_cols_partial<VT, CF>(row_at(krow), nv, b);
            } else {
#pragma unroll
              for (int i = 0; i < CF; ++i) b[i] = 0.0f;
            }
            consume(r, v, b);
          }
        }
      }
    };

    // Full tiles run in the fast loop; the first tile that is not full for every
    // group (only the matrix's last chunk has one) hands over, already staged, to
    // the predicated loop. The two loops are separate regions for the register
    // allocator. Tile counts and exits are wave-uniform.
    const int ntile = static_cast<int>((p.chunk + G - 1) / G);
    int j = 0, cnt = 0;
    bool staged = false;
    if constexpr (kFastTiles) {
      if (lanes_full) {
        for (; j < ntile; ++j) {
          const int64_t off = static_cast<int64_t>(j) * G;
          if (__ballot(off < clen) == 0) {  // no group of this wave has entries left
            j = ntile;
            break;
          }
          cnt = stage(off);
          if (__ballot(cnt != 0 && cnt != G) != 0) {
            staged = true;
            break;
          }
          fast_tile(off, cnt);
          __builtin_amdgcn_wave_barrier();  // LDS slots are rewritten by the next tile
        }
      }
    }
    for (; j < ntile; ++j) {
      const int64_t off = static_cast<int64_t>(j) * G;
      if (!staged) {
        if (__ballot(off < clen) == 0) break;
        cnt = stage(off);
      }
      staged = false;
      slow_tile(off, cnt);
      __builtin_amdgcn_wave_barrier();  // LDS slots are rewritten by the next tile
    }
    if (chunk_valid && clen > 0) {
      flush(cur, true);
      // the last chunk owns the run after the last row with entries
      if (zrun && c == p.nchunks - 1) owner_empty_run<CF, G, SEG>(p.C, p.ldc, col0, nv, p.M, cur, p.M, p.zflag, zflags, g_lane);
    }
  }
}

// SDDMM on A's pattern with the same sequential walk: out[p] = <X[row(p), :],
// Y[colind[p], :]>. A group of G lanes covers G*CF columns (looping over wider
// N), walks a chunk of consecutive nnz, finds rows through its own rowptr
// window, and per batch of U entries issues U loads of X rows (repeated rows hit
// in cache) and U loads of Y rows before reducing each dot product across the
// group (DPP row shifts for G <= 16).
template <typename VT, int CF, int G>
__global__ void __launch_bounds__(256) sddmm_sr_kernel(SddmmParams p) {
  constexpr int U = (G < 8) ? G : 8;
  const int L = threadIdx.x;
  const int g = L / G;
  const int g_lane = L - g * G;
  const int wave_gbase = __lane_id() - g_lane;
  const int groups_per_block = blockDim.x / G;
  const int64_t ggid = static_cast<int64_t>(blockIdx.x) * groups_per_block + g;
  const VT* __restrict__ X = static_cast<const VT*>(p.X);
  const VT* __restrict__ Y = static_cast<const VT*>(p.Y);

  for (int64_t c = ggid;; c += p.total_groups) {
    if (__ballot(c < p.nchunks) == 0) break;
    const bool in_range = c < p.nchunks;
    const int64_t c0 = in_range ? c * p.chunk : 0;
    const int64_t clen = in_range ? min(p.chunk, p.nnz - c0) : 0;
    int r_hint;
    if constexpr (G >= 8) r_hint = static_cast<int>(group_search_gt<G>(p.rowptr, 1, p.M, c0)) - 1;
    else r_hint = row_of_nnz(p.rowptr, p.M, c0);

    // metadata of the next tile: column of the lane's nnz and the rowptr window
    int pk = 0, prp = INT32_MAX;
    auto prefetch = [&](int64_t off_n) {
      const bool vn = off_n + g_lane < clen;
      pk = vn ? p.colind[c0 + off_n + g_lane] : 0;
      if (off_n < clen) {
        const int jj = r_hint + 1 + g_lane;
        prp = (jj <= p.M) ? p.rowptr[jj] : INT32_MAX;
      }
    };
    prefetch(0);
    for (int64_t off = 0;; off += G) {
      if (__ballot(off < clen) == 0) break;
      const bool valid = off + g_lane < clen;
      const int cnt = static_cast<int>(max<int64_t>(0, min<int64_t>(G, clen - off)));
      const int row = group_row_find<G>(p.rowptr, p.M, r_hint, prp, static_cast<int>(c0 + off + g_lane), valid);
      const int k = pk;
      const int last_row = __shfl(row, wave_gbase + max(cnt, 1) - 1);
      if (cnt > 0) r_hint = last_row;
      prefetch(off + G);
      for (int pp = 0; pp < cnt; pp += U) {
        int ru[U], ku[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int src = wave_gbase + min(pp + u, cnt - 1);
          ru[u] = __shfl(row, src);
          ku[u] = __shfl(k, src);
        }
        float part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) part[u] = 0.0f;
        for (int cb = 0; cb < p.N; cb += G * CF) {
          const int col0 = cb + g_lane * CF;
          const int nv = min(CF, p.N - col0);
          RawCols<VT, CF> xr[U], yr[U];
          if (nv == CF) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (pp + u < cnt) {
                load_raw<VT, CF>(X + static_cast<int64_t>(ru[u]) * p.ld_x + col0, xr[u]);
                load_raw<VT, CF>(Y + static_cast<int64_t>(ku[u]) * p.ld_y + col0, yr[u]);
              }
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (pp + u < cnt && nv > 0) {
              float xa[CF], ya[CF];
              if (nv == CF) {
                widen<VT, CF>(xr[u], xa);
                widen<VT, CF>(yr[u], ya);
              } else {
                load_cols_partial<VT, CF>(X + static_cast<int64_t>(ru[u]) * p.ld_x + col0, nv, xa);
                load_cols_partial<VT, CF>(Y + static_cast<int64_t>(ku[u]) * p.ld_y + col0, nv, ya);
              }
#pragma unroll
              for (int i = 0; i < CF; ++i) part[u] += xa[i] * ya[i];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float v[1] = {part[u]};
          group_reduce_to_first<G, 1>(v);
          if (g_lane == 0 && pp + u < cnt) p.out[c0 + off + pp + u] = v[0];
        }
      }
    }
  }
}

// ---- instantiation and dispatch ----------------------------------------------------------

// Staged-value launches (p.vslot set) exist for groups of 8 to 64 lanes; the host plans
// them only there.
// The fixed owner form (owner policy, no accumulation) exists for groups of 8 to 64 lanes too;
// other groups take the general kernel for it.
template <typename VT, int CF, int G>
static hipError_t launch_sr_t(const SrParams& p, const Launch& l, bool seg, bool& ok) {
  const bool own = p.policy == 2 && p.accumulate == 0;
  if (seg) {
    sr_kernel<VT, CF, G, true><<<l.grid, l.block, l.shmem, l.stream>>>(p);
  } else if (p.vslot) {
    if constexpr (G >= 8) {
      if (own) sr_kernel<VT, CF, G, false, true, true><<<l.grid, l.block, l.shmem, l.stream>>>(p);
      else sr_kernel<VT, CF, G, false, true><<<l.grid, l.block, l.shmem, l.stream>>>(p);
    } else {
      ok = false;
      return hipSuccess;
    }
  } else {
    if constexpr (G >= 8) {
      if (own) {
        sr_kernel<VT, CF, G, false, false, true><<<l.grid, l.block, l.shmem, l.stream>>>(p);
        return hipGetLastError();
      }
    }
    sr_kernel<VT, CF, G, false><<<l.grid, l.block, l.shmem, l.stream>>>(p);
  }
  return hipGetLastError();
}

template <typename VT, int CF>
static hipError_t launch_sr_g(int g, const SrParams& p, const Launch& l, bool seg, bool& ok) {
  ok = true;
  switch (g) {
    case 1: return launch_sr_t<VT, CF, 1>(p, l, seg, ok);
    case 2: return launch_sr_t<VT, CF, 2>(p, l, seg, ok);
    case 4: return launch_sr_t<VT, CF, 4>(p, l, seg, ok);
    case 8: return launch_sr_t<VT, CF, 8>(p, l, seg, ok);
    case 16: return launch_sr_t<VT, CF, 16>(p, l, seg, ok);
    case 32: return launch_sr_t<VT, CF, 32>(p, l, seg, ok);
    case 64: return launch_sr_t<VT, CF, 64>(p, l, seg, ok);
    default: ok = false; return hipSuccess;
  }
}

template <typename VT>
static hipError_t launch_sr_cf(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok) {
  switch (cf) {
    case 1: return launch_sr_g<VT, 1>(g, p, l, seg, ok);
    case 2: return launch_sr_g<VT, 2>(g, p, l, seg, ok);
    case 4: return launch_sr_g<VT, 4>(g, p, l, seg, ok);
    case 8: return launch_sr_g<VT, 8>(g, p, l, seg, ok);
    default: ok = false; return hipSuccess;
  }
}

template <typename VT, int CF, int G>
static hipError_t launch_sddmm_t(const SddmmParams& p, const Launch& l) {
  sddmm_sr_kernel<VT, CF, G><<<l.grid, l.block, 0, l.stream>>>(p);
  return hipGetLastError();
}

template <typename VT, int CF>
static hipError_t launch_sddmm_g(int g, const SddmmParams& p, const Launch& l, bool& ok) {
  ok = true;
  switch (g) {
    case 1: return launch_sddmm_t<VT, CF, 1>(p, l);
    case 2: return launch_sddmm_t<VT, CF, 2>(p, l);
    case 4: return launch_sddmm_t<VT, CF, 4>(p, l);
    case 8: return launch_sddmm_t<VT, CF, 8>(p, l);
    case 16: return launch_sddmm_t<VT, CF, 16>(p, l);
    case 32: return launch_sddmm_t<VT, CF, 32>(p, l);
    case 64: return launch_sddmm_t<VT, CF, 64>(p, l);
    default: ok = false; return hipSuccess;
  }
}

template <typename VT>
static hipError_t launch_sddmm_cf(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok) {
  switch (cf) {
    case 1: return launch_sddmm_g<VT, 1>(g, p, l, ok);
    case 2: return launch_sddmm_g<VT, 2>(g, p, l, ok);
    case 4: return launch_sddmm_g<VT, 4>(g, p, l, ok);
    case 8:
      if constexpr (sizeof(VT) == 2) return launch_sddmm_g<VT, 8>(g, p, l, ok);
      [[fallthrough]];
    default: ok = false; return hipSuccess;
  }
}

// Per-type entry points (sr_f32.hip, sr_bf16.hip, sr_f16.hip: one translation unit
// per storage type, compiled in parallel).
hipError_t launch_sr_f32(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok);
hipError_t launch_sr_bf16(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok);
hipError_t launch_sr_f16(int cf, int g, const SrParams& p, const Launch& l, bool seg, bool& ok);
hipError_t launch_sddmm_f32(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok);
hipError_t launch_sddmm_bf16(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok);
hipError_t launch_sddmm_f16(int cf, int g, const SddmmParams& p, const Launch& l, bool& ok);

}  // namespace segscan
