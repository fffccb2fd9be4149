// Destination tiles of the grouped plan.
//
// Reference: the per-order loop of _energy_exchange
// (sparrowpy/classes/RadiosityFast.py:1073-1145), as grouped by
// spr_plan_grouped_sort / spr_plan_grouped_fill.
//
// Neighbouring destination patches (consecutive patch ids on one wall) see
// nearly the same sources through nearly the same outgoing directions, so
// their plan rows read nearly the same source rows prev[i, d, b, :], only
// shifted by slightly different delays.  A tile of JT consecutive rows is
// merged into one list of distinct (group key, source row) entries, each
// carrying one (delay, weight) per slot (row of the tile): a workgroup then
// issues the JT slots' gathers of one source row back to back, so only the
// first misses in the caches and the rest hit the lines it brought in.  On
// the street-canyon plan JT = 4 merges 16.4 M entries into 5.2 M.
//
// Within a key run the merged list is in ascending source row order, the
// order of the entries of each group in the grouped plan, so every slot sums
// its group in exactly the grouped plan's order: the per-group sums are
// bit-identical to spr_exchange_gather_grouped's.
#include "spr_common.hpp"
#include "spr_exchange.hpp"
#include "spr_util.hpp"

namespace spr {
namespace {

// Largest k in [lo, hi) with a[k] <= x (a ascending, a[lo] <= x).
__device__ __forceinline__ int64_t last_le(const int64_t *__restrict__ a,
                                           int64_t lo, int64_t hi,
                                           int64_t x) {
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Number of k in [lo, hi) with a[k] < x (strict) or a[k] <= x (!strict).
__device__ __forceinline__ int64_t count_below(const int64_t *__restrict__ a,
                                               int64_t lo, int64_t hi,
                                               int64_t x, bool strict) {
  const int64_t base = lo;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (strict ? a[mid] < x : a[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - base;
}

// skey[e] = (group key << 32) | source row: ascending within every plan row
// (groups in key order, entries of a group in source order).
__global__ __launch_bounds__(256) void tile_skey_kernel(
    int64_t Vd, int64_t G, const int64_t *__restrict__ grp_start,
    const int32_t *__restrict__ grp_key, const int32_t *__restrict__ rowid,
    int64_t *__restrict__ skey) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Vd) return;
  const int64_t g = last_le(grp_start, 0, G, e);
  skey[e] = ((int64_t)grp_key[g] << 32) | (int64_t)(uint32_t)rowid[e];
}

// Merge of the JT sorted rows of each tile: entry e of row j (slot
// k = j mod JT) goes to position pos = row_ptr[tile start] + (its rank in
// its row) + (entries of the other rows of the tile that sort before it,
// ties ordered by slot).  A tile's positions are the span of its rows.
__global__ __launch_bounds__(256) void tile_merge_kernel(
    int64_t P, int JT, int64_t Vd, const int64_t *__restrict__ row_ptr,
    const int64_t *__restrict__ skey, int64_t *__restrict__ sorted,
    int64_t *__restrict__ payload) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Vd) return;
  const int64_t j = last_le(row_ptr, 0, P, e);
  const int k = (int)(j % JT);
  const int64_t j0 = j - k;
  const int64_t s = skey[e];
  int64_t pos = row_ptr[j0] + (e - row_ptr[j]);
  for (int q = 0; q < JT; ++q) {
    const int64_t jq = j0 + q;
    if (q == k || jq >= P) continue;
    pos += count_below(skey, row_ptr[jq], row_ptr[jq + 1], s, q > k);
  }
  sorted[pos] = s;
  payload[pos] = (e << 4) | k;
}

__device__ __forceinline__ bool run_start(const int64_t *__restrict__ sorted,
                                          int64_t p, int64_t tile_begin) {
  return p == tile_begin || sorted[p] != sorted[p - 1];
}

__device__ __forceinline__ int64_t tile_begin_of(
    int64_t P, int JT, const int64_t *__restrict__ row_ptr, int64_t p) {
  const int64_t j = last_le(row_ptr, 0, P, p);
  return row_ptr[j - j % JT];
}

// first[p] = 1 where a distinct (key, source row) starts in its tile.
__global__ __launch_bounds__(256) void tile_first_kernel(
    int64_t P, int JT, int64_t Vd, const int64_t *__restrict__ row_ptr,
    const int64_t *__restrict__ sorted, int64_t *__restrict__ first) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Vd) return;
  first[p] = run_start(sorted, p, tile_begin_of(P, JT, row_ptr, p)) ? 1 : 0;
}

// tptr[t] = merged index of tile t's first entry (first = exclusive scan).
__global__ void tile_ptr_kernel(int64_t P, int JT, int64_t n_tiles,
                                const int64_t *__restrict__ row_ptr,
                                const int64_t *__restrict__ first,
                                int64_t *__restrict__ tptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  const int64_t j = t * JT < P ? t * JT : P;
  tptr[t] = first[row_ptr[j]];
}

__device__ __forceinline__ int64_t merged_index(
    const int64_t *__restrict__ sorted, const int64_t *__restrict__ first,
    int64_t p, int64_t tile_begin) {
  return run_start(sorted, p, tile_begin) ? first[p] : first[p] - 1;
}

// Merged entry of position p: its source row, and slot k's delay (without
// flag) and per-band weights.  mlag starts at kLagMask, mw at 0 (slots a
// merged entry does not have).
__global__ __launch_bounds__(256) void tile_fill_kernel(
    int64_t P, int JT, int64_t B, int64_t Vd, int64_t M,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ rowid,
    const int32_t *__restrict__ lag, const double *__restrict__ scal,
    const int64_t *__restrict__ sorted, const int64_t *__restrict__ payload,
    const int64_t *__restrict__ first, int32_t *__restrict__ mrow,
    int32_t *__restrict__ mlag, double *__restrict__ mw) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Vd) return;
  const int64_t m =
      merged_index(sorted, first, p, tile_begin_of(P, JT, row_ptr, p));
  const int64_t e = payload[p] >> 4;
  const int k = (int)(payload[p] & 15);
  mrow[m] = rowid[e];
  mlag[m * JT + k] = lag[e] & kLagMask;
  for (int64_t b = 0; b < B; ++b) mw[(b * M + m) * JT + k] = scal[b * Vd + e];
}

// Group ends: the entry flagged last of its group in slot k's row sets the
// flag on slot k of the last merged entry of its key run in the tile, where
// the tile kernel stores slot k's sum and moves to its next group.
__global__ __launch_bounds__(256) void tile_flag_kernel(
    int64_t P, int JT, int64_t Vd, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ lag, const int64_t *__restrict__ sorted,
    const int64_t *__restrict__ payload, const int64_t *__restrict__ first,
    int32_t *__restrict__ mlag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Vd) return;
  const int64_t e = payload[p] >> 4;
  if (!(lag[e] & kLastInGroup)) return;
  const int k = (int)(payload[p] & 15);
  const int64_t j = last_le(row_ptr, 0, P, p);
  const int64_t tb = row_ptr[j - j % JT];
  const int64_t jt_end = j - j % JT + JT < P ? j - j % JT + JT : P;
  const int64_t te = row_ptr[jt_end];
  // last position of the tile whose key equals this entry's key
  const int64_t key_end = ((sorted[p] >> 32) + 1) << 32;
  const int64_t q = p + count_below(sorted, p, te, key_end, true) - 1;
  const int64_t m = merged_index(sorted, first, q, tb);
  mlag[m * JT + k] |= kLastInGroup;
}

// ---------------------------------------------------------------------
// the tiled gather
// ---------------------------------------------------------------------
// Merged entries are walked in blocks of MB: one vector load per field
// brings a block's metadata (source row, JT delay words, JT weights), one
// entry per lane, broadcast with v_readlane.  A block is fully unrolled
// into MB/C chunks of C merged entries; chunk q+1's C*JT*R gathers are in
// flight while chunk q is summed.  A slot without an entry has delay
// kLagMask (a descriptor of size 0: loads give 0) and weight 0.
// Requires T < 2^29.
template <int JW>
struct TileMeta {
  int32_t row;
  int32_t lag[JW];
  double w[JW];
};

// Block of merged entries m0 .. m0+63 (one per lane), slots s0 .. s0+JW-1
// of JT.
template <int JW>
__device__ __forceinline__ TileMeta<JW> load_tile_meta(
    int64_t m0, int64_t m_end, int JT, int s0,
    const int32_t *__restrict__ mrow, const int32_t *__restrict__ mlag,
    const double *__restrict__ w) {
  const int64_t m = m0 + (int64_t)(threadIdx.x & 63);
  const bool in = m < m_end;
  const int64_t mc = in ? m : m_end - 1;
  TileMeta<JW> t;
  t.row = mrow[mc];
#pragma unroll
  for (int k = 0; k < JW; ++k) {
    t.lag[k] = in ? mlag[mc * JT + s0 + k] : kLagMask;
    t.w[k] = in ? w[mc * JT + s0 + k] : 0.0;
  }
  return t;
}

// A workgroup takes one (tile, band, block of 256*R bins) with JT/JW
// groups of 4 waves: group h sums slots h*JW .. h*JW+JW-1 over the same
// bins, so the groups' gathers of one source row meet in the CU's cache.
template <int R, int C, int JT, int JW, int MB>
__global__ __launch_bounds__(256 * (JT / JW)) void tile_gather_kernel(
    int64_t P, int B, int T, int n_tblk, int64_t nblocks, int64_t M,
    const int64_t *__restrict__ tptr, const int64_t *__restrict__ grp_ptr,
    const int32_t *__restrict__ mrow, const int32_t *__restrict__ mlag,
    const double *__restrict__ mw, const float *__restrict__ prev,
    float *__restrict__ gsum) {
  static_assert(MB <= 64 && MB % (2 * C) == 0, "chunks pair up in a block");
  static_assert(JT % JW == 0, "slot groups split the tile");
  constexpr int NQ = MB / C, BS = 256;
  int64_t tile;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, tile, b, tb);
  const int t0 = tb * (BS * R);
  const int t_end = min(T, t0 + BS * R);
  const int64_t m_begin = tptr[tile], m_end = tptr[tile + 1];
  if (m_end <= m_begin) return;
  const int s0 = (int)(threadIdx.x / BS) * JW;  // first slot of this group
  const double *w = mw + (int64_t)b * M * JT;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int tl = t0 + (int)(threadIdx.x % BS);
  int64_t g[JW];
#pragma unroll
  for (int k = 0; k < JW; ++k) {
    const int64_t j = tile * JT + s0 + k;
    g[k] = j < P ? grp_ptr[j] : 0;
  }

  auto fetch = [&](const TileMeta<JW> &m, int q, float (&v)[C][JW][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int idx = q * C + c;
      const int32_t rk = __builtin_amdgcn_readlane(m.row, idx);
      const float *src = prev + ((int64_t)rk * B + b) * T;
#pragma unroll
      for (int k = 0; k < JW; ++k) {
        const int n = min(__builtin_amdgcn_readlane(m.lag[k], idx) & kLagMask,
                          T);
        const __amdgpu_buffer_rsrc_t rs =
            make_rsrc(src, n < t_end ? row_bytes : 0u);
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[c][k][r] = buffer_load(rs, bin_offset(tl + BS * r - n));
      }
    }
  };

  double S[JW][R];
#pragma unroll
  for (int k = 0; k < JW; ++k)
#pragma unroll
    for (int r = 0; r < R; ++r) S[k][r] = 0.0;
  auto consume = [&](const TileMeta<JW> &m, int q,
                     const float (&v)[C][JW][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int idx = q * C + c;
#pragma unroll
      for (int k = 0; k < JW; ++k) {
        const double wk = readlane_f64(m.w[k], idx);
#pragma unroll
        for (int r = 0; r < R; ++r)
          S[k][r] = fma(wk, (double)v[c][k][r], S[k][r]);
      }
#pragma unroll
      for (int k = 0; k < JW; ++k) {
        if (__builtin_amdgcn_readlane(m.lag[k], idx) & kLastInGroup) {
          const __amdgpu_buffer_rsrc_t os =
              make_rsrc(gsum + (g[k] * B + b) * (int64_t)T, row_bytes);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            buffer_store((float)S[k][r], os, (uint32_t)(tl + BS * r) * 4u);
            S[k][r] = 0.0;
          }
          g[k] += 1;
        }
      }
    }
  };

  TileMeta<JW> cur =
      load_tile_meta<JW>(m_begin, m_end, JT, s0, mrow, mlag, w);
  TileMeta<JW> next =
      load_tile_meta<JW>(m_begin + MB, m_end, JT, s0, mrow, mlag, w);
  float va[C][JW][R], vb[C][JW][R];
  fetch(cur, 0, va);
  for (int64_t blk = m_begin; blk < m_end; blk += MB) {
#pragma unroll
    for (int q = 0; q < NQ; q += 2) {
      fetch(cur, q + 1, vb);
      consume(cur, q, va);
      if (q + 2 < NQ)
        fetch(cur, q + 2, va);
      else
        fetch(next, 0, va);  // first chunk of the next block
      consume(cur, q + 1, vb);
    }
    cur = next;
    next = load_tile_meta<JW>(blk + 2 * MB, m_end, JT, s0, mrow, mlag, w);
  }
}

// Launch shapes, measured on the street canyon (T = 6,000): JT = 4 (5.2 M
// merged entries) with R = 4 bins per lane and one merged entry per chunk
// (4 waves per SIMD) takes 172 ms (166 ms before offsets were kept whole,
// see bin_offset); R = 2 with two entries per chunk 211 ms; R = 8, 184 ms.
// JT = 8 (3.3 M merged entries, 38 % of the slots empty) in two groups of
// 4 waves takes 213 ms: its lower traffic does not pay for the empty
// slots' issue slots.  16-byte loads (4 consecutive bins per lane, one load
// per slot) give the same sums in 177.7 ms against 172.8 ms.
template <int R, int C, int JT, int JW, int MB>
int launch_tile_shape(int64_t P, int64_t B, int64_t T, int64_t M,
                      const int64_t *tptr, const int64_t *grp_ptr,
                      const int32_t *mrow, const int32_t *mlag,
                      const double *mw, const float *prev, float *gsum,
                      hipStream_t s) {
  constexpr int BS = 256;
  const int64_t n_tiles = ceil_div(P, JT);
  const int64_t nt = ceil_div(T, (int64_t)BS * R);
  const int64_t nb = n_tiles * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(T < ((int64_t)1 << 29), "histogram too long (T >= 2^29)");
  hipLaunchKernelGGL((tile_gather_kernel<R, C, JT, JW, MB>),
                     dim3((unsigned)nb), dim3(BS * (JT / JW)), 0, s, P,
                     (int)B, (int)T, (int)nt, nb, M, tptr, grp_ptr, mrow,
                     mlag, mw, prev, gsum);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

template <int R, int JT, int NB, int SMAX, bool DS, int DEPTH, bool PIPE>
int launch_tile_lds(int64_t P, int64_t B, int64_t T, int64_t M,
                    const int64_t *tptr, const int64_t *grp_ptr,
                    const int64_t *gdiff, const int32_t *mrow,
                    const int32_t *mlag, const double *mw, const double *mdw,
                    const float *prev, float *gsum, hipStream_t s);

int launch_tile_gather(int64_t P, int JT, int64_t B, int64_t T, int64_t M,
                       const int64_t *tptr, const int64_t *grp_ptr,
                       const int32_t *mrow, const int32_t *mlag,
                       const double *mw, const float *prev, float *gsum,
                       hipStream_t s) {
  if (JT == 4)
    return launch_tile_shape<4, 1, 4, 4, 64>(P, B, T, M, tptr, grp_ptr, mrow,
                                             mlag, mw, prev, gsum, s);
  if (JT == 8)
    return launch_tile_shape<4, 1, 8, 4, 64>(P, B, T, M, tptr, grp_ptr,
                                             mrow, mlag, mw, prev, gsum, s);
  return launch_tile_lds<16, 16, 16, 64, false, 1, false>(
      P, B, T, M, tptr, grp_ptr, nullptr, mrow, mlag, mw, nullptr, prev, gsum,
      s);
}

// ---------------------------------------------------------------------
// the LDS-staged tiled gather (16-row tiles)
// ---------------------------------------------------------------------
// With 16 rows per tile a merged entry serves several slots, but loading
// its source row once per slot through the caches (as tile_gather_kernel
// does for 4 rows) costs one cache read per slot and, across many waves,
// evicts the lines before the last slot reads them.  Here the workgroup
// stages each merged entry's window of the source row in LDS once, and
// every slot (one wave per destination row of the tile) reads its delayed
// view of it from LDS.
//
// A workgroup owns (tile, band, block of TB = 64*R bins).  Merged entries
// are taken in batches of NB; for each, the window of source bins
// [t0 - maxlag, t0 - maxlag + W) with W = TB + SMAX covers every present
// slot whose delay is within SMAX of the entry's largest delay (maxlag); a
// slot with a larger spread reads its bins from global memory instead
// (uniform branch, rare).  Two LDS buffers: the global loads of batch q+1
// are in flight while batch q is summed from LDS; one barrier per batch.
// Each wave sums its slot's entries in merged order, which within a group
// is the grouped plan's entry order: the per-group sums are bit-identical
// to spr_exchange_gather_grouped's.  Requires T < 2^29.
template <int JT>
struct SlotMeta {
  int32_t row;     // source row of the merged entry
  int32_t lag;     // this wave's slot delay word (kLagMask: absent)
  int32_t maxlag;  // largest delay of the entry's present slots
  int32_t minlag;  // smallest
  double w;        // this wave's slot weight in the workgroup's band
  double dw;       // its diffuse weight (diffuse + specular plan)
};

// Merged entries m0 .. m0+63, one per lane, as seen by slot k.
template <int JT>
__device__ __forceinline__ SlotMeta<JT> load_slot_meta(
    int64_t m0, int64_t m_end, int k, const int32_t *__restrict__ mrow,
    const int32_t *__restrict__ mlag, const double *__restrict__ w,
    const double *__restrict__ dw) {
  const int64_t m = m0 + (int64_t)(threadIdx.x & 63);
  const bool in = m < m_end;
  const int64_t mc = in ? m : m_end - 1;
  SlotMeta<JT> t;
  t.row = mrow[mc];
  int32_t mx = 0, mn = kLagMask;
  const int4 *lw = reinterpret_cast<const int4 *>(mlag + mc * JT);
#pragma unroll
  for (int q = 0; q < JT / 4; ++q) {
    const int4 v = lw[q];
    const int32_t a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t l = a[u] & kLagMask;
      if (l != kLagMask) {
        mx = max(mx, l);
        mn = min(mn, l);
      }
    }
  }
  t.maxlag = mx;
  t.minlag = mn == kLagMask ? mx : mn;
  t.lag = in ? mlag[mc * JT + k] : kLagMask;
  t.w = in ? w[mc * JT + k] : 0.0;
  t.dw = in && dw ? dw[mc * JT + k] : 0.0;
  return t;
}

template <int R, int JT, int NB, int SMAX, bool DS, int DEPTH, bool PIPE>
__global__ __launch_bounds__(64 * JT) void tile_lds_gather_kernel(
    int64_t P, int B, int T, int n_tblk, int64_t nblocks, int64_t M,
    const int64_t *__restrict__ tptr, const int64_t *__restrict__ grp_ptr,
    const int64_t *__restrict__ gdiff, const int32_t *__restrict__ mrow,
    const int32_t *__restrict__ mlag, const double *__restrict__ mw,
    const double *__restrict__ mdw, const float *__restrict__ prev,
    float *__restrict__ gsum) {
  constexpr int TB = 64 * R;       // bins of the workgroup
  constexpr int W = TB + SMAX;     // LDS window of one merged entry
  constexpr int NX = W / 64;       // window floats per lane
  constexpr int NE = NB / JT;      // entries each wave stages per batch
  static_assert(SMAX % 64 == 0 && NB % JT == 0 && 64 % NB == 0, "shape");
  static_assert(DEPTH == 1 || DEPTH == 2, "one or two batches in flight");
  static_assert(64 / NB >= DEPTH, "staging stays within the next block");
  __shared__ float lds[2][NB][W];
  int64_t tile;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, tile, b, tb);
  const int64_t m_begin = tptr[tile], m_end = tptr[tile + 1];
  if (m_end <= m_begin) return;  // uniform: before any barrier
  const int t0 = tb * TB;
  // slot = wave; readfirstlane makes it (and everything derived from it:
  // the slot's row, its group counter, its branches) a scalar value, so
  // the per-entry branches are scalar branches, not exec-mask juggling
  const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int64_t j = tile * JT + k;
  const bool active = j < P;
  int64_t g = active ? grp_ptr[j] : 0;
  const double *w = mw + (int64_t)b * M * JT;
  const double *dw = DS ? mdw + (int64_t)b * M * JT : nullptr;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int64_t bstride = (int64_t)B * T;

  // batch staging: wave k loads entries k, k + JT, ... of a batch into
  // registers st[slot]; DEPTH register sets, so with DEPTH = 2 the loads of
  // batch q+2 are issued while batch q+1's are still in flight
  float st[DEPTH][NE][NX];
  auto stage_load = [&](int slot, const SlotMeta<JT> &m, int64_t mb,
                        int64_t blk0) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      const int64_t mm = mb + k + e * JT;
      const int idx = (int)(mm - blk0);
      const bool have = mm < m_end;
      const int32_t row = __builtin_amdgcn_readlane(m.row, have ? idx : 0);
      const int32_t mx = __builtin_amdgcn_readlane(m.maxlag, have ? idx : 0);
      const int32_t mn = __builtin_amdgcn_readlane(m.minlag, have ? idx : 0);
      const int need = have ? min(W, TB + (mx - mn)) : 0;
      const __amdgpu_buffer_rsrc_t rs =
          make_rsrc(prev + (int64_t)row * bstride + (int64_t)b * T,
                    row_bytes);
      const int s0 = t0 - mx + lane;
#pragma unroll
      for (int x = 0; x < NX; ++x)
        st[slot][e][x] = 64 * x < need
                             ? buffer_load(rs, bin_offset(s0 + 64 * x))
                             : 0.0f;
    }
  };
  auto stage_write = [&](int slot, int buf) {
#pragma unroll
    for (int e = 0; e < NE; ++e)
#pragma unroll
      for (int x = 0; x < NX; ++x)
        lds[buf][k + e * JT][lane + 64 * x] = st[slot][e][x];
  };

  double S[R], Sd[DS ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) S[r] = 0.0;
#pragma unroll
  for (int r = 0; r < (DS ? R : 1); ++r) Sd[r] = 0.0;
  auto compute = [&](const SlotMeta<JT> &m, int64_t mb, int64_t blk0,
                     int buf) {
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int64_t mm = mb + e;
      if (mm >= m_end) break;  // uniform
      const int idx = (int)(mm - blk0);
      const int32_t lw = __builtin_amdgcn_readlane(m.lag, idx);
      const int32_t l = lw & kLagMask;
      if (active && l != kLagMask) {
        const double wk = readlane_f64(m.w, idx);
        const int32_t mx = __builtin_amdgcn_readlane(m.maxlag, idx);
        const int off = mx - l;
        float v[R];
        if (off <= SMAX) {
#pragma unroll
          for (int r = 0; r < R; ++r) v[r] = lds[buf][e][off + lane + 64 * r];
        } else {  // delay far from the entry's others: global memory
          const int32_t row = __builtin_amdgcn_readlane(m.row, idx);
          const __amdgpu_buffer_rsrc_t rs =
              make_rsrc(prev + (int64_t)row * bstride + (int64_t)b * T,
                        row_bytes);
#pragma unroll
          for (int r = 0; r < R; ++r)
            v[r] = buffer_load(rs, bin_offset(t0 + lane + 64 * r - l));
        }
        if (DS) {
          const double dk = readlane_f64(m.dw, idx);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double x = (double)v[r];
            S[r] = fma(wk, x, S[r]);
            Sd[DS ? r : 0] = fma(dk, x, Sd[DS ? r : 0]);
          }
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) S[r] = fma(wk, (double)v[r], S[r]);
        }
      }
      if (active && (lw & kLastInGroup)) {
        const __amdgpu_buffer_rsrc_t os =
            make_rsrc(gsum + (g * B + b) * (int64_t)T, row_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          buffer_store((float)S[r], os, (uint32_t)(t0 + lane + 64 * r) * 4u);
          S[r] = 0.0;
        }
        g += 1;
      }
    }
  };

  // The same sums, software-pipelined: the wave walks only the entries of
  // the batch where its slot is present or ends a group (set bits of a
  // ballot over the metadata lanes), and issues the next present entry's
  // LDS reads into the other register set before summing the current one.
  auto read_entry = [&](const SlotMeta<JT> &m, int idx, int e, int buf,
                        float (&v)[R]) {
    const int32_t l = __builtin_amdgcn_readlane(m.lag, idx) & kLagMask;
    const int32_t mx = __builtin_amdgcn_readlane(m.maxlag, idx);
    const int off = mx - l;
    if (off <= SMAX) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = lds[buf][e][off + lane + 64 * r];
    } else {
      const int32_t row = __builtin_amdgcn_readlane(m.row, idx);
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(
          prev + (int64_t)row * bstride + (int64_t)b * T, row_bytes);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[r] = buffer_load(rs, bin_offset(t0 + lane + 64 * r - l));
    }
  };
  auto use_entry = [&](const SlotMeta<JT> &m, int idx, bool present,
                       const float (&v)[R]) {
    const int32_t lw = __builtin_amdgcn_readlane(m.lag, idx);
    if (present) {
      const double wk = readlane_f64(m.w, idx);
      if (DS) {
        const double dk = readlane_f64(m.dw, idx);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double x = (double)v[r];
          S[r] = fma(wk, x, S[r]);
          Sd[DS ? r : 0] = fma(dk, x, Sd[DS ? r : 0]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) S[r] = fma(wk, (double)v[r], S[r]);
      }
    }
    if (lw & kLastInGroup) {
      const __amdgpu_buffer_rsrc_t os =
          make_rsrc(gsum + (g * B + b) * (int64_t)T, row_bytes);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        buffer_store((float)S[r], os, (uint32_t)(t0 + lane + 64 * r) * 4u);
        S[r] = 0.0;
      }
      g += 1;
    }
  };
  auto compute_pipe = [&](const SlotMeta<JT> &m, int64_t mb, int64_t blk0,
                          int buf) {
    if (!active) return;  // uniform per wave
    const int i0 = (int)(mb - blk0);
    const int n = (int)min((int64_t)NB, m_end - mb);
    const uint64_t window = (n >= 64 ? ~0ull : ((1ull << n) - 1ull)) << i0;
    const int32_t lw = m.lag;  // this lane's entry of the block
    const uint64_t pres =
        __ballot((lw & kLagMask) != kLagMask) & window;
    uint64_t todo = (pres | __ballot((lw & kLastInGroup) != 0)) & window;
    if (!todo) return;
    float va[R], vb[R];
    int idx = __builtin_ctzll(todo);
    if ((pres >> idx) & 1) read_entry(m, idx, idx - i0, buf, va);
    while (true) {
      todo &= todo - 1;
      int nxt_idx = todo ? __builtin_ctzll(todo) : -1;
      if (nxt_idx >= 0 && ((pres >> nxt_idx) & 1))
        read_entry(m, nxt_idx, nxt_idx - i0, buf, vb);
      use_entry(m, idx, (pres >> idx) & 1, va);
      if (nxt_idx < 0) break;
      idx = nxt_idx;
      todo &= todo - 1;
      nxt_idx = todo ? __builtin_ctzll(todo) : -1;
      if (nxt_idx >= 0 && ((pres >> nxt_idx) & 1))
        read_entry(m, nxt_idx, nxt_idx - i0, buf, va);
      use_entry(m, idx, (pres >> idx) & 1, vb);
      if (nxt_idx < 0) break;
      idx = nxt_idx;
    }
  };

  // Metadata blocks of 64 merged entries: `cur` holds the block of the
  // batch being summed (first entry blk0), `nxt` the following one; a
  // batch never straddles two blocks (NB divides 64), and the batches being
  // staged are at most one block ahead.
  SlotMeta<JT> cur =
      load_slot_meta<JT>(m_begin, m_end, k, mrow, mlag, w, dw);
  SlotMeta<JT> nxt =
      load_slot_meta<JT>(m_begin + 64, m_end, k, mrow, mlag, w, dw);
  int64_t blk0 = m_begin;
  auto stage = [&](int slot, int64_t mb) {
    if (mb - blk0 >= 64)
      stage_load(slot, nxt, mb, blk0 + 64);
    else
      stage_load(slot, cur, mb, blk0);
  };
  // batch 0 into LDS buffer 0; with DEPTH 2, batch 1's loads in flight
  stage(0, m_begin);
  stage_write(0, 0);
  if (DEPTH == 2 && m_begin + NB < m_end) stage(1 % DEPTH, m_begin + NB);
  __syncthreads();
  // one step: sum batch mb from LDS buffer `buf`; register set `rs_next`
  // holds the loads of batch mb + NB (issued a step earlier with DEPTH 2,
  // or now with DEPTH 1), which go to the other LDS buffer afterwards
  auto step = [&](int64_t mb, int buf, int rs_next) {
    const int64_t mb1 = mb + NB, mb2 = mb + 2 * NB;
    if (DEPTH == 1) {
      if (mb1 < m_end) stage(0, mb1);
    } else {
      if (mb2 < m_end) stage(rs_next ^ 1, mb2);
    }
    if (PIPE)
      compute_pipe(cur, mb, blk0, buf);
    else
      compute(cur, mb, blk0, buf);
    if (mb1 < m_end) stage_write(DEPTH == 1 ? 0 : rs_next, buf ^ 1);
    __syncthreads();
    if (mb1 - blk0 >= 64) {
      cur = nxt;
      blk0 += 64;
      nxt = load_slot_meta<JT>(blk0 + 64, m_end, k, mrow, mlag, w, dw);
    }
  };
  // two steps per iteration, so the buffer and register-set indices are
  // compile-time constants
  for (int64_t mb = m_begin; mb < m_end; mb += 2 * NB) {
    step(mb, 0, 1 % DEPTH);
    if (mb + NB < m_end) step(mb + NB, 1, 0);
  }
  // diffuse + specular plan: the row's diffuse sum over all its entries
  // goes to its diffuse group (the row's last), as ds_gather_kernel stores
  if (DS && active && gdiff[j] >= 0) {
    const __amdgpu_buffer_rsrc_t os =
        make_rsrc(gsum + (gdiff[j] * B + b) * (int64_t)T, row_bytes);
#pragma unroll
    for (int r = 0; r < R; ++r)
      buffer_store((float)Sd[DS ? r : 0], os,
                   (uint32_t)(t0 + lane + 64 * r) * 4u);
  }
}

// 16-row tiles: R = 16 bins per lane (1,024-bin blocks; 8 for the diffuse
// + specular plan, whose second sums take the registers), batches of 16
// merged entries, windows of TB + 64 bins (2 x 16 x 1,088 floats of LDS
// at R = 16, one workgroup of 16 waves per CU).
template <int R, int JT, int NB, int SMAX, bool DS, int DEPTH, bool PIPE>
int launch_tile_lds(int64_t P, int64_t B, int64_t T, int64_t M,
                    const int64_t *tptr, const int64_t *grp_ptr,
                    const int64_t *gdiff, const int32_t *mrow,
                    const int32_t *mlag, const double *mw, const double *mdw,
                    const float *prev, float *gsum, hipStream_t s) {
  const int64_t n_tiles = ceil_div(P, JT);
  const int64_t nt = ceil_div(T, (int64_t)64 * R);
  const int64_t nb = n_tiles * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(T < ((int64_t)1 << 29), "histogram too long (T >= 2^29)");
  hipLaunchKernelGGL(
      (tile_lds_gather_kernel<R, JT, NB, SMAX, DS, DEPTH, PIPE>),
                     dim3((unsigned)nb), dim3(64 * JT), 0, s, P, (int)B,
                     (int)T, (int)nt, nb, M, tptr, grp_ptr, gdiff, mrow,
                     mlag, mw, mdw, prev, gsum);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

unsigned grid_of(int64_t n) { return (unsigned)ceil_div(n > 0 ? n : 1, 256); }

}  // namespace
}  // namespace spr

using namespace spr;

extern "C" int spr_plan_tile_order(int64_t P, int64_t JT, int64_t Vd,
                                   int64_t G, const int64_t *row_ptr,
                                   const int64_t *grp_start,
                                   const int32_t *grp_key,
                                   const int32_t *rowid, int64_t *skey,
                                   int64_t *sorted, int64_t *payload,
                                   int64_t *first, int64_t *tptr,
                                   void *stream) {
  clear_error();
  SPR_REQUIRE(P > 0 && Vd >= 0 && G >= 0 && JT >= 2 && JT <= 16,
              "spr_plan_tile_order: bad sizes (JT must be 2..16)");
  SPR_REQUIRE(Vd < ((int64_t)1 << 40), "spr_plan_tile_order: plan too large");
  SPR_REQUIRE(row_ptr && tptr && first, "spr_plan_tile_order: null pointer");
  const bool given = !grp_start && !grp_key;
  SPR_REQUIRE(Vd == 0 || (skey && sorted && payload &&
                          (given || (grp_start && grp_key && rowid &&
                                     G > 0))),
              "spr_plan_tile_order: null pointer");
  hipStream_t s = as_stream(stream);
  const int64_t n_tiles = ceil_div(P, JT);
  if (Vd > 0) {
    if (!given) {
      hipLaunchKernelGGL(tile_skey_kernel, dim3(grid_of(Vd)), dim3(256), 0,
                         s, Vd, G, grp_start, grp_key, rowid, skey);
      SPR_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(tile_merge_kernel, dim3(grid_of(Vd)), dim3(256), 0, s,
                       P, (int)JT, Vd, row_ptr, skey, sorted, payload);
    SPR_LAUNCH_CHECK();
    hipLaunchKernelGGL(tile_first_kernel, dim3(grid_of(Vd)), dim3(256), 0, s,
                       P, (int)JT, Vd, row_ptr, sorted, first);
    SPR_LAUNCH_CHECK();
  }
  int rc = scan_exclusive(first, Vd, s);
  if (rc != SPR_OK) return rc;
  hipLaunchKernelGGL(tile_ptr_kernel, dim3(grid_of(n_tiles + 1)), dim3(256),
                     0, s, P, (int)JT, n_tiles, row_ptr, first, tptr);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_plan_tile_fill(int64_t P, int64_t JT, int64_t B,
                                  int64_t Vd, int64_t M,
                                  const int64_t *row_ptr,
                                  const int32_t *rowid, const int32_t *lag,
                                  const double *scal, const int64_t *sorted,
                                  const int64_t *payload,
                                  const int64_t *first, int32_t *mrow,
                                  int32_t *mlag, double *mw, void *stream) {
  clear_error();
  SPR_REQUIRE(P > 0 && B > 0 && Vd >= 0 && M >= 0 && M <= Vd && JT >= 2 &&
                  JT <= 16,
              "spr_plan_tile_fill: bad sizes (JT must be 2..16)");
  if (M == 0) return SPR_OK;
  SPR_REQUIRE(row_ptr && rowid && lag && scal && sorted && payload && first &&
                  mrow && mlag && mw,
              "spr_plan_tile_fill: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = fill<int32_t>(mlag, M * JT, kLagMask, s);
  if (rc == SPR_OK) rc = fill<double>(mw, B * M * JT, 0.0, s);
  if (rc != SPR_OK) return rc;
  hipLaunchKernelGGL(tile_fill_kernel, dim3(grid_of(Vd)), dim3(256), 0, s, P,
                     (int)JT, B, Vd, M, row_ptr, rowid, lag, scal, sorted,
                     payload, first, mrow, mlag, mw);
  SPR_LAUNCH_CHECK();
  hipLaunchKernelGGL(tile_flag_kernel, dim3(grid_of(Vd)), dim3(256), 0, s, P,
                     (int)JT, Vd, row_ptr, lag, sorted, payload, first, mlag);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_gather_tiled(int64_t P, int64_t JT, int64_t B,
                                         int64_t T, int64_t M,
                                         const int64_t *tptr,
                                         const int64_t *grp_ptr,
                                         const int32_t *mrow,
                                         const int32_t *mlag,
                                         const double *mw, const float *prev,
                                         float *gsum, void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 &&
                  (JT == 4 || JT == 8 || JT == 16),
              "spr_exchange_gather_tiled: bad sizes (JT must be 4, 8 or 16)");
  SPR_REQUIRE(tptr && grp_ptr && mrow && mlag && mw && prev && gsum,
              "spr_exchange_gather_tiled: null pointer");
  return launch_tile_gather(P, (int)JT, B, T, M, tptr, grp_ptr, mrow, mlag,
                            mw, prev, gsum, as_stream(stream));
}

extern "C" int spr_exchange_gather_tiled_ds(
    int64_t P, int64_t JT, int64_t B, int64_t T, int64_t M,
    const int64_t *tptr, const int64_t *gstart, const int64_t *gdiff,
    const int32_t *mrow, const int32_t *mlag, const double *mw,
    const double *mdw, const float *prev, float *gsum, void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 && JT == 16,
              "spr_exchange_gather_tiled_ds: bad sizes (JT must be 16)");
  SPR_REQUIRE(tptr && gstart && gdiff && mrow && mlag && mw && mdw && prev &&
                  gsum,
              "spr_exchange_gather_tiled_ds: null pointer");
  return launch_tile_lds<8, 16, 16, 64, true, 1, false>(
      P, B, T, M, tptr, gstart, gdiff, mrow, mlag, mw, mdw, prev, gsum,
      as_stream(stream));
}

extern "C" int spr_exchange_gather_tiled_shape(
    int64_t P, int64_t JT, int64_t R, int64_t B, int64_t T, int64_t M,
    const int64_t *tptr, const int64_t *grp_ptr, const int32_t *mrow,
    const int32_t *mlag, const double *mw, const float *prev, float *gsum,
    void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 && JT == 16 &&
                  (R == 8 || R == 16 || R == -8 || R == -16),
              "spr_exchange_gather_tiled_shape: bad sizes (JT 16; R 16, 8, "
              "-8: 8 with two batches in flight, -16: 16 pipelined over "
              "the present entries)");
  SPR_REQUIRE(tptr && grp_ptr && mrow && mlag && mw && prev && gsum,
              "spr_exchange_gather_tiled_shape: null pointer");
  hipStream_t s = as_stream(stream);
  if (R == 8)
    return launch_tile_lds<8, 16, 16, 64, false, 1, false>(
        P, B, T, M, tptr, grp_ptr, nullptr, mrow, mlag, mw, nullptr, prev,
        gsum, s);
  if (R == -16)
    return launch_tile_lds<16, 16, 16, 64, false, 1, true>(
        P, B, T, M, tptr, grp_ptr, nullptr, mrow, mlag, mw, nullptr, prev,
        gsum, s);
  if (R == -8)
    return launch_tile_lds<8, 16, 16, 64, false, 2, false>(
        P, B, T, M, tptr, grp_ptr, nullptr, mrow, mlag, mw, nullptr, prev,
        gsum, s);
  return launch_tile_lds<16, 16, 16, 64, false, 1, false>(
      P, B, T, M, tptr, grp_ptr, nullptr, mrow, mlag, mw, nullptr, prev,
      gsum, s);
}
