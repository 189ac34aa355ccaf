// Device primitives shared by the SpMM and segment_coo kernels (gfx950, wave64).
//
//  * typed 32..256-bit loads of a lane's CF contiguous columns, converted to fp32;
//  * lane shifts within an aligned group of G lanes: DPP row_shl for G <= 16
//    (one VALU op, no LDS traffic), ds_bpermute (__shfl_down) for G = 32, 64;
//  * the segmented suffix reduction of the parallel-reduction kernels, built on a
//    wave ballot of segment ends instead of shuffling row ids every step;
//  * the wave-cooperative CSR row lookup that replaces a per-lane global binary
//    search over rowptr with one coalesced 64-entry window load per wave.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_params.hpp"

namespace segscan {

constexpr int kWave = 64;

// Storage types of the dense operand. bf16 and f16 are kept as raw bits and
// widened to fp32 in registers (both widenings are exact).
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(bf16_t x) { return __uint_as_float(static_cast<uint32_t>(x.bits) << 16); }
__device__ __forceinline__ float to_f32(f16_t x) { return static_cast<float>(__builtin_bit_cast(_Float16, x.bits)); }

template <int BYTES> struct RawVec;
template <> struct RawVec<2> { using T = uint16_t; };
template <> struct RawVec<4> { using T = uint32_t; };
template <> struct RawVec<8> { using T = uint2; };
template <> struct RawVec<16> { using T = uint4; };

// Loads CF contiguous elements at p (aligned to CF * sizeof(VT) up to 16 bytes)
// as one vector load (two for 32 bytes) and widens them to fp32.
template <typename VT, int CF>
__device__ __forceinline__ void load_cols(const VT* __restrict__ p, float (&out)[CF]) {
  constexpr int BYTES = CF * static_cast<int>(sizeof(VT));
  if constexpr (BYTES <= 16) {
    using R = typename RawVec<BYTES>::T;
    R r = *reinterpret_cast<const R*>(p);
    const VT* e = reinterpret_cast<const VT*>(&r);
#pragma unroll
    for (int i = 0; i < CF; ++i) out[i] = to_f32(e[i]);
  } else {
    static_assert(BYTES == 32, "at most 32 bytes per lane");
    uint4 r0 = reinterpret_cast<const uint4*>(p)[0];
    uint4 r1 = reinterpret_cast<const uint4*>(p)[1];
    const VT* e0 = reinterpret_cast<const VT*>(&r0);
    const VT* e1 = reinterpret_cast<const VT*>(&r1);
#pragma unroll
    for (int i = 0; i < CF / 2; ++i) {
      out[i] = to_f32(e0[i]);
      out[i + CF / 2] = to_f32(e1[i]);
    }
  }
}

// A lane's CF columns kept as raw storage (one or two 16-byte registers groups),
// so that many dense-row loads can be in flight before any is widened.
template <typename VT, int CF>
struct RawCols {
  static constexpr int BYTES = CF * static_cast<int>(sizeof(VT));
  static constexpr int PART = BYTES < 16 ? BYTES : 16;
  static constexpr int NPART = BYTES / PART;
  using P = typename RawVec<PART>::T;
  P part[NPART];
};

template <typename VT, int CF>
__device__ __forceinline__ void load_raw(const VT* __restrict__ p, RawCols<VT, CF>& r) {
  const auto* q = reinterpret_cast<const typename RawCols<VT, CF>::P*>(p);
#pragma unroll
  for (int i = 0; i < RawCols<VT, CF>::NPART; ++i) r.part[i] = q[i];
}

// load_raw with non-temporal loads, for rows read exactly once in order (segment_coo's
// src): they then take no L2 lines from data that is reused. The gathers of the SpMM
// kernels keep load_raw: as non-temporal loads they measured 1.3-1.4x slower.
template <typename VT, int CF>
__device__ __forceinline__ void load_raw_stream(const VT* __restrict__ p, RawCols<VT, CF>& r) {
  using P = typename RawCols<VT, CF>::P;
  constexpr int W = static_cast<int>(sizeof(P)) / 4;
  if constexpr (W >= 2) {
    typedef unsigned int uv __attribute__((ext_vector_type(W)));
#pragma unroll
    for (int i = 0; i < RawCols<VT, CF>::NPART; ++i) {
      const uv v = __builtin_nontemporal_load(reinterpret_cast<const uv*>(p) + i);
      __builtin_memcpy(&r.part[i], &v, sizeof(P));
    }
  } else if constexpr (W == 1) {
#pragma unroll
    for (int i = 0; i < RawCols<VT, CF>::NPART; ++i) {
      const unsigned int v = __builtin_nontemporal_load(reinterpret_cast<const unsigned int*>(p) + i);
      __builtin_memcpy(&r.part[i], &v, sizeof(P));
    }
  } else {
    load_raw<VT, CF>(p, r);
  }
}

template <typename VT, int CF>
__device__ __forceinline__ void widen(const RawCols<VT, CF>& r, float (&out)[CF]) {
  const VT* e = reinterpret_cast<const VT*>(r.part);
#pragma unroll
  for (int i = 0; i < CF; ++i) out[i] = to_f32(e[i]);
}

// Loads the first nv (< CF) of a lane's columns element by element (residue panel).
template <typename VT, int CF>
__device__ __forceinline__ void load_cols_partial(const VT* __restrict__ p, int nv, float (&out)[CF]) {
#pragma unroll
  for (int i = 0; i < CF; ++i) out[i] = (i < nv) ? to_f32(p[i]) : 0.0f;
}

template <typename VT, int CF>
__device__ __forceinline__ void load_cols_any(const VT* __restrict__ p, int nv, float (&out)[CF]) {
  if (nv >= CF) load_cols<VT, CF>(p, out);
  else load_cols_partial<VT, CF>(p, nv, out);
}

// Stores / accumulates CF fp32 columns (vector store when the panel is full).
template <int CF>
__device__ __forceinline__ void store_cols(float* __restrict__ p, int nv, const float (&v)[CF], bool accumulate) {
  if (nv >= CF && !accumulate) {
    if constexpr (CF == 1) {
      p[0] = v[0];
    } else if constexpr (CF == 2) {
      *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else if constexpr (CF == 4) {
      *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
      reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
      reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < CF; ++i) {
    if (i < nv) p[i] = accumulate ? p[i] + v[i] : v[i];
  }
}

template <int CF>
__device__ __forceinline__ void atomic_add_cols(float* __restrict__ p, int nv, const float (&v)[CF]) {
#pragma unroll
  for (int i = 0; i < CF; ++i) {
    if (i < nv) unsafeAtomicAdd(p + i, v[i]);  // coarse-grained memory: hardware fp32 atomic
  }
}

template <int CF>
__device__ __forceinline__ void zero_cols(float* __restrict__ p, int nv) {
  float z[CF];
#pragma unroll
  for (int i = 0; i < CF; ++i) z[i] = 0.0f;
  store_cols<CF>(p, nv, z, false);
}

// ---- lane shifts -------------------------------------------------------------------------

// Lane i receives lane i + D of its 16-lane DPP row (0 past the row end).
template <int D>
__device__ __forceinline__ float dpp_row_shl(float v) {
  static_assert(D >= 1 && D <= 15, "row_shl takes 1..15");
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + D, 0xF, 0xF, true));
}

// Lane i of an aligned group of G lanes receives lane i + D of the same group.
// Lanes whose source falls outside the group get an unspecified value; callers
// mask them out.
template <int G, int D>
__device__ __forceinline__ float group_shift_down(float v) {
  if constexpr (G <= 16) {
    return dpp_row_shl<D>(v);
  } else {
    return __shfl_down(v, D, G);
  }
}

// Segmented suffix sum within aligned groups of G lanes. Rows are non-decreasing
// inside a group. On return the first lane of each run of equal rows (a segment
// head, `is_head`) holds the run's total in c[]. The segment end of every lane
// comes from one wave ballot, so each of the log2(G) steps shifts values only.
template <int G, int CF>
__device__ __forceinline__ bool group_segmented_reduce(float (&c)[CF], int row) {
  const int lane = __lane_id();
  const int g_lane = lane & (G - 1);
  if constexpr (G == 1) {
    return true;
  } else {
    const int row_next = __shfl_down(row, 1, G);
    const bool is_end = (g_lane == G - 1) || (row_next != row);
    const uint64_t ends = __ballot(is_end);
    const bool is_head = (g_lane == 0) || ((ends >> (lane - 1)) & 1ull);
    const int seg_end = lane + __builtin_ctzll(ends >> lane);
#define SEGSCAN_SEG_STEP(D)                                              \
    if constexpr ((D) < G) {                                             \
      _Pragma("unroll") for (int i = 0; i < CF; ++i) {                   \
        const float t = group_shift_down<G, (D)>(c[i]);                  \
        if (lane + (D) <= seg_end) c[i] += t;                            \
      }                                                                  \
    }
    SEGSCAN_SEG_STEP(1)
    SEGSCAN_SEG_STEP(2)
    SEGSCAN_SEG_STEP(4)
    SEGSCAN_SEG_STEP(8)
    SEGSCAN_SEG_STEP(16)
    SEGSCAN_SEG_STEP(32)
#undef SEGSCAN_SEG_STEP
    return is_head;
  }
}

// Plain sum within aligned groups of G lanes; lane 0 of the group holds the total.
template <int G, int CF>
__device__ __forceinline__ void group_reduce_to_first(float (&c)[CF]) {
#define SEGSCAN_RED_STEP(D)                                              \
  if constexpr ((D) < G) {                                               \
    _Pragma("unroll") for (int i = 0; i < CF; ++i) c[i] += group_shift_down<G, (D)>(c[i]); \
  }
  SEGSCAN_RED_STEP(32)
  SEGSCAN_RED_STEP(16)
  SEGSCAN_RED_STEP(8)
  SEGSCAN_RED_STEP(4)
  SEGSCAN_RED_STEP(2)
  SEGSCAN_RED_STEP(1)
#undef SEGSCAN_RED_STEP
}

// One step of a reduce-scatter over an aligned group of G lanes: lanes with bit m clear
// return a + a' and lanes with bit m set b + b' (primes: the partner lane l ^ m). Fed
// the U per-entry partial sums of a batch in pairs, level by level as they are produced
// (a streaming butterfly, at most log2 U pending values), it leaves lane l holding the
// group total of entry sum_h bit(l, G >> (h + 1)) << h.
template <int G>
__device__ __forceinline__ float butterfly(float a, float b, int m) {
  const bool hi = (__lane_id() & m) != 0;
  const float send = hi ? a : b;
  const float keep = hi ? b : a;
  return keep + __shfl_xor(send, m, G);
}

// Entry of a batch whose group total the butterfly leaves in lane gl.
template <int G, int U>
__device__ __forceinline__ int butterfly_entry(int gl) {
  constexpr int H = __builtin_ctz(U);
  int e = 0;
#pragma unroll
  for (int h = 0; h < H; ++h) e |= ((gl & (G >> (h + 1))) ? 1 : 0) << h;
  return e;
}

// ---- CSR row lookup ---------------------------------------------------------------------

// Row containing nnz position e: the r with rowptr[r] <= e < rowptr[r+1], empty
// rows skipped; e >= nnz maps to M - 1. Same result as the reference's
// binary_search_segment_number (test-cg/include/util/check.cuh:88-102).
__device__ __forceinline__ int row_of_nnz(const int32_t* __restrict__ rowptr, int M, int64_t e) {
  int lo = 1, hi = M;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rowptr[mid] <= e) lo = mid + 1;
    else hi = mid;
  }
  return hi - 1;
}

template <int G>
__device__ __forceinline__ uint64_t group_mask() {
  if constexpr (G >= 64) return ~0ull;
  else return (1ull << G) - 1ull;
}

// Wave-cooperative 64-ary search (every lane calls it with the same arguments,
// full EXEC): the first j in [lo, hi] with a[j] > key, given a non-decreasing
// and a[hi] > key. Each round probes 64 evenly spaced entries with one load per
// lane and narrows the range 64-fold: 3 dependent loads for 2^18 entries
// instead of 18 for a binary search.
__device__ __forceinline__ int64_t wave_search_gt(const int32_t* __restrict__ a, int64_t lo, int64_t hi, int64_t key) {
  const int lane = __lane_id();
  while (lo < hi) {
    const int64_t step = (hi - lo + 64) / 64;
    // probes are non-decreasing and the last lane always probes hi
    const int64_t j = (lane == 63) ? hi : min(lo + lane * step, hi);
    const uint64_t gt = __ballot(a[j] > key);
    const int f = __builtin_ctzll(gt);  // a[hi] > key, so gt != 0
    const int64_t new_hi = (f == 63) ? hi : min(lo + f * step, hi);
    // probe f-1 was <= key, so it was not clamped to hi
    const int64_t new_lo = (f == 0) ? lo : lo + (f - 1) * step + 1;
    lo = new_lo;
    hi = new_hi;
  }
  return hi;
}

// Group-cooperative G-ary search (G >= 8; aligned groups, each with its own
// group-uniform lo / hi / key): the first j in [lo, hi] with a[j] > key, given a
// non-decreasing and a[hi] > key. Groups leave the loop independently.
template <int G>
__device__ __forceinline__ int64_t group_search_gt(const int32_t* __restrict__ a, int64_t lo, int64_t hi, int64_t key) {
  static_assert(G >= 8, "use a binary search for narrower groups");
  const int lane = __lane_id();
  const int gbase = lane & ~(G - 1);
  const int gl = lane - gbase;
  while (lo < hi) {
    const int64_t step = (hi - lo + G) / G;
    const int64_t j = (gl == G - 1) ? hi : min(lo + gl * step, hi);
    const uint64_t gt = (__ballot(a[j] > key) >> gbase) & group_mask<G>();
    const int f = __builtin_ctzll(gt);  // the last lane probes hi
    const int64_t new_hi = (f == G - 1) ? hi : min(lo + f * step, hi);
    const int64_t new_lo = (f == 0) ? lo : lo + (f - 1) * step + 1;
    lo = new_lo;
    hi = new_hi;
  }
  return hi;
}

// The first j in [lo, hi] with a[j] > key (a non-decreasing, a[hi] > key; lo / hi / key
// group-uniform): a G-ary search by the group, a binary search for narrower groups.
template <int G>
__device__ __forceinline__ int first_gt(const int32_t* __restrict__ a, int lo, int hi, int key) {
  if constexpr (G >= 8) {
    return static_cast<int>(group_search_gt<G>(a, lo, hi, key));
  } else {
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if (a[mid] > key) hi = mid;
      else lo = mid + 1;
    }
    return hi;
  }
}

// Group-cooperative lookup for an aligned group of G lanes walking consecutive
// nnz of one chunk. `rp` holds rowptr[r_hint + 1 + g_lane] (INT32_MAX past M),
// loaded ahead of time by the caller so that its latency overlaps other work.
// Precondition: rowptr[r_hint] <= e for every active lane, e < rowptr[M].
// Each lane counts the boundaries <= e in the window with a log2(G)-step
// binary search over the group's lanes; lanes past the window step to the next
// G boundaries (group-uniform loop, rare when rows average more than a few nnz),
// or, when those too lie at or below the lowest pending entry (a run of empty
// rows), jump to that entry's row with a search instead of walking the run.
template <int G>
__device__ __forceinline__ int group_row_find(const int32_t* __restrict__ rowptr, int M, int r_hint, int rp, int e,
                                              bool active) {
  const int lane = __lane_id();
  const int gbase = lane & ~(G - 1);
  const int g_lane = lane - gbase;
  int base = r_hint + 1;
  int row = r_hint;
  bool done = !active;
  while (true) {
    int pos = 0;
#pragma unroll
    for (int s = G / 2; s >= 1; s >>= 1) {
      const int v = __shfl(rp, gbase + pos + s - 1);
      if (v <= e) pos += s;
    }
    const int last = __shfl(rp, gbase + G - 1);
    if (pos == G - 1 && last <= e) pos = G;
    if (!done && pos < G) {
      row = base + pos - 1;
      done = true;
    }
    const uint64_t pending = (__ballot(!done) >> gbase) & group_mask<G>();
    if (pending == 0) break;
    // Pending lanes saw rowptr[base + G - 1] <= e < rowptr[M], so base + G <= M. The
    // lowest pending lane holds the smallest pending entry (lanes walk consecutive ones).
    base += G;
    const int e_min = __shfl(e, gbase + __builtin_ctzll(pending));
    if (rowptr[min(base + G - 1, M)] <= e_min) base = first_gt<G>(rowptr, base, M, e_min);
    const int j = base + g_lane;
    rp = (j <= M) ? rowptr[j] : INT32_MAX;
  }
  return row;
}

// Wave-cooperative lookup: every lane of the wave must call it (full EXEC).
// Precondition: rowptr[r0] <= e for every lane and e < rowptr[M] (callers clamp
// padding lanes to a valid position). The wave loads rowptr[r0+1 .. r0+64]
// once (one coalesced 256-byte read) and each lane counts the boundaries <= e
// with a 7-step binary search over the window held in lanes; lanes whose row
// lies beyond the window move on to the next 64 entries together.
__device__ __forceinline__ int wave_row_lookup(const int32_t* __restrict__ rowptr, int M, int r0, int e) {
  const int lane = __lane_id();
  int base = r0 + 1;
  int row = r0;
  bool done = false;
  while (true) {
    const int j = base + lane;
    const int rp = (j <= M) ? rowptr[j] : INT32_MAX;
    int pos = 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
      const int v = __shfl(rp, pos + s - 1);
      if (v <= e) pos += s;
    }
    const int last = __shfl(rp, 63);
    if (pos == 63 && last <= e) pos = 64;
    if (!done && pos < 64) {
      row = base + pos - 1;
      done = true;
    }
    if (__ballot(!done) == 0) break;
    // Pending lanes saw rowptr[base + 63] <= e < rowptr[M]. When the next 64 boundaries
    // also lie at or below every pending entry (a run of empty rows), jump to the row
    // of the smallest one with a 64-ary search instead of walking the run.
    base += 64;
    int e_min = done ? INT32_MAX : e;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) e_min = min(e_min, __shfl_xor(e_min, s));
    if (rowptr[min(base + 63, M)] <= e_min) base = static_cast<int>(wave_search_gt(rowptr, base, M, e_min));
  }
  return row;
}

// Empty rows under the owner write policy. Rows are grouped in blocks of kZeroRows. A
// block whose rows are all empty is zeroed by the fix-up pass after the kernel, with its
// whole grid; the kernel itself zeroes only the rows of an empty run that share a block
// with a row with entries: at most 2 x (kZeroRows - 1) per run, whatever its length.
// The fix-up pass finds entry-free blocks from rowptr (CSR: rowptr[16j] == rowptr[16j + 16])
// or, for segments, from zflag[j], one byte per block, rewritten by every call: 0 when
// block j holds a segment with entries (stored by a chunk that reaches it), 1 when it
// holds none (stored by the owner of the empty run that covers it). Nothing needs
// clearing between calls.
//
// The run of empty rows strictly between rows a and b (a = -1: before the first row;
// b = M: after the last), seen by one group: lanes with nv > 0 zero their CF columns (at
// col0) of the run's partial blocks; with FLAGS and `flags` the group's G lanes also flag
// its whole blocks (only groups of the first column slice pass flags).
template <int CF, int G, bool FLAGS>
__device__ __forceinline__ void owner_empty_run(float* __restrict__ C, int64_t ldc, int col0, int nv, int M, int a,
                                                int b, uint8_t* __restrict__ zflag, bool flags, int g_lane) {
  const int s = a + 1;
  if (s >= b) return;
  const int j0 = (s + kZeroRows - 1) >> kZeroRowsLog;                            // first block starting at s or later
  const int j1 = b >= M ? static_cast<int>(zero_blocks(M)) : (b >> kZeroRowsLog);  // blocks below j1 end by b
  int z0 = b, z1 = b;  // rows [s, z0) and [z1, b) are zeroed here
  if (j0 < j1) {
    z0 = j0 << kZeroRowsLog;
    z1 = min(b, j1 << kZeroRowsLog);
    if constexpr (FLAGS) {
      if (flags) {
        // bytes up to a 4-aligned block, then 4 blocks per 32-bit store, then the rest
        const int w0 = min(j1, (j0 + 3) & ~3), w1 = max(w0, j1 & ~3);
        for (int j = j0 + g_lane; j < w0; j += G) zflag[j] = 1;
        uint32_t* zw = reinterpret_cast<uint32_t*>(zflag);
        for (int w = (w0 >> 2) + g_lane; w < (w1 >> 2); w += G) zw[w] = 0x01010101u;
        for (int j = w1 + g_lane; j < j1; j += G) zflag[j] = 1;
      }
    }
  }
  if (nv > 0) {
    for (int r = s; r < z0; ++r) zero_cols<CF>(C + static_cast<int64_t>(r) * ldc + col0, nv);
    for (int r = z1; r < b; ++r) zero_cols<CF>(C + static_cast<int64_t>(r) * ldc + col0, nv);
  }
}

}  // namespace segscan
