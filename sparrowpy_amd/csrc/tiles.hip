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

int launch_tile_gather(int64_t P, int JT, int64_t B, int64_t T, int64_t M,
                       const int64_t *tptr, const int64_t *grp_ptr,
                       const int32_t *mrow, const int32_t *mlag,
                       const double *mw, const float *prev, float *gsum,
                       hipStream_t s) {
  if (JT == 4)
    return launch_tile_shape<4, 1, 4, 4, 64>(P, B, T, M, tptr, grp_ptr, mrow,
                                             mlag, mw, prev, gsum, s);
  return launch_tile_shape<4, 1, 8, 4, 64>(P, B, T, M, tptr, grp_ptr, mrow,
                                           mlag, mw, prev, gsum, s);
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
// are taken in batches of NB = JT; for each, the window of source bins
// [t0 - maxlag, t0 - maxlag + W) with W = TB + SMAX covers every present
// slot whose delay is within SMAX of the entry's largest delay (maxlag); a
// slot with a larger spread reads its bins from global memory instead
// (uniform branch, rare).  The batch's metadata goes to LDS beside the
// windows: per (entry, slot) the byte offset of the slot's view in the
// window buffer with the flags absent / far / group end, and the weights;
// per entry its source row and maxlag.  A wave reads its slot's 16 words
// with one LDS load (one per lane), walks only the entries its slot has or
// ends a group at (ballots over the words), takes each entry's word with
// v_readlane and reads its weight with a broadcast LDS load: a present
// entry costs the wave one v_readlane, one mask and one add for its view's
// address, R + 1 LDS reads, R conversions and R FP64 FMAs (the walk's
// scalar work was as large as the sums' vector work when each entry built
// its view's address from the delay word: 100.4 against 105.0 ms on the
// street canyon); the next entry's reads are issued before the current
// entry is summed.  Two LDS buffers (two
// LDS objects): batch q+1's windows are copied from memory into LDS by the
// loads themselves (no registers) while batch q is summed; each wave waits
// for its copies (vmcnt) only before the one barrier per batch.  Each wave
// sums its slot's entries in merged order, which within a group is the
// grouped plan's entry order: the per-group sums are bit-identical to
// spr_exchange_gather_grouped's.  Requires T < 2^29.
typedef __attribute__((address_space(3))) void lds_void_t;

// One buffer load of a dword per lane into LDS: lane l's dword goes to
// dst + 4 l (+ the immediate offset, which also moves the memory address).
__device__ __forceinline__ void dma_dword(__amdgpu_buffer_rsrc_t rs,
                                          float *dst, uint32_t off) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void_t *)dst, 4, off, 0, 0,
                                           0);
}

// Chunks X .. N-1 of 64 dwords, chunk x at immediate offset 256 x (a
// constant expression, as the instruction field must be).
template <int X, int N>
struct WindowDma {
  __device__ __forceinline__ static void run(__amdgpu_buffer_rsrc_t rs,
                                             float *dst, uint32_t vo) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void_t *)dst, 4, vo, 0,
                                             256 * X, 0);
    WindowDma<X + 1, N>::run(rs, dst, vo);
  }
};
template <int N>
struct WindowDma<N, N> {
  __device__ __forceinline__ static void run(__amdgpu_buffer_rsrc_t,
                                             float *, uint32_t) {}
};

// flags of a staged word (its low 29 bits the view's byte offset)
constexpr int32_t kAbsent = (int32_t)0x80000000u;
constexpr int32_t kFar = 0x20000000;
constexpr int32_t kAddrMask = 0x1fffffff;

template <int NB, int JT, bool DS>
struct BatchMeta {
  // rows padded to JT + 1: a wave reads one slot's column (16 entries) per
  // array, which at a row stride of JT words would fall on a few LDS banks
  int32_t word[NB][JT + 1];       // view byte offset | kAbsent, kFar, end
  double w[NB][JT + 1];           // weights in the workgroup's band
  double dw[DS ? NB : 1][JT + 1]; // diffuse weights (diffuse + specular)
  int32_t row[NB];            // source row
  int32_t maxlag[NB];         // largest delay of the entry's present slots
};

// R = 8 without the diffuse sums: 80 KiB of LDS and at most 64 VGPRs, so
// two workgroups share a CU (8 waves per SIMD): the sums are latency-bound
// (each entry's reads, conversions and FP64 FMAs form one chain per wave),
// and twice the waves hide twice the latency.
template <int R, bool DS>
constexpr int lds_gather_waves_per_eu() {
  return R == 8 && !DS ? 8 : 1;
}

template <int R, int JT, int SMAX, bool DS>
__global__ __launch_bounds__(64 * JT) __attribute__((
    amdgpu_waves_per_eu(lds_gather_waves_per_eu<R, DS>()))) void
tile_lds_gather_kernel(
    int64_t P, int B, int T, int n_tblk, int64_t nblocks, int64_t M,
    const int64_t *__restrict__ tptr, const int64_t *__restrict__ grp_ptr,
    const int64_t *__restrict__ gdiff, const int4 *__restrict__ mspan,
    const int32_t *__restrict__ mlag, const double *__restrict__ mw,
    const double *__restrict__ mdw, const float *__restrict__ prev,
    float *__restrict__ gsum) {
  constexpr int NB = JT;           // entries per batch: one window per wave
  constexpr int TB = 64 * R;       // bins of the workgroup
  constexpr int W = TB + SMAX;     // LDS window of one merged entry
  constexpr int NX = W / 64;       // window floats per lane
  constexpr int NM = NB * JT;      // (entry, slot) pairs of a batch
  static_assert(SMAX % 64 == 0 && NM % 64 == 0 &&
                    (DS ? 3 : 2) * NM <= 64 * JT && NB <= 32 &&
                    (NB & (NB - 1)) == 0,
                "shape: whole waves stage the metadata");
  // the two window buffers are two LDS objects, not one [2][NB][W] array:
  // the compiler waits (vmcnt) for copies into LDS before any LDS read it
  // cannot prove to be of another object, and a read of one buffer would
  // otherwise wait for the other buffer's copies (the next batch's) before
  // the sums could start
  __shared__ float lds_a[NB][W];
  __shared__ float lds_b[NB][W];
  __shared__ BatchMeta<NB, JT, DS> meta[2];
  auto win = [&](int buf) -> float(*)[W] { return buf ? lds_b : lds_a; };
  int64_t tile;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, tile, b, tb);
  const int64_t m_begin = tptr[tile];
  // entries of the tile (< 2^31): batch positions are 32-bit offsets from
  // m_begin, so every loop bound is a scalar 32-bit compare
  const int n_tile = (int)(tptr[tile + 1] - m_begin);
  if (n_tile <= 0) return;  // uniform: before any barrier
  const int t0 = tb * TB;
  // slot = wave; readfirstlane makes it (and everything derived from it:
  // the slot's row, its group counter, its branches) a scalar value, so
  // the per-entry branches are scalar branches, not exec-mask juggling
  const int k = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int64_t j = tile * JT + k;
  const bool active = j < P;
  int64_t g = active ? grp_ptr[j] : 0;
  const double *w = mw + ((int64_t)b * M + m_begin) * JT;
  const double *dw = DS ? mdw + ((int64_t)b * M + m_begin) * JT : nullptr;
  const int32_t *lagw = mlag + m_begin * JT;
  const int4 *span = mspan + m_begin;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int64_t bstride = (int64_t)B * T;
  auto n_at = [&](int mb) { return min(NB, n_tile - mb); };

  // window staging: wave k copies entry mb + k's window of its source row
  // straight from memory into LDS (buffer loads with the LDS flag: lane l
  // of a load writes dword l at the load's LDS address), from the entry's
  // span (source row, maxlag, minlag: a scalar load, one batch ahead).  A
  // window starting at or after the row's first bin (w0 >= 0) takes a
  // descriptor based at bin w0, so the 64-bin steps are immediate offsets
  // (added to both the memory and the LDS address) that cannot wrap; one
  // starting before it keeps each lane's offset whole (bin_offset),
  // negative bins reading 0.  Nothing passes through registers: the loads
  // are waited for (vmcnt) before the barrier that publishes the buffer.
  auto span_of = [&](int mm) {
    return mm < n_tile ? span[mm] : make_int4(0, 0, 0, 0);
  };
  auto stage_load = [&](int buf, int n_in, const int4 &sp) {
    if (k >= n_in) return;  // uniform
    const int need = min(W, TB + (sp.y - sp.z));
    const float *row = prev + (int64_t)sp.x * bstride + (int64_t)b * T;
    const int w0 = t0 - sp.y;
    float *dst = &win(buf)[k][0];
    if (w0 >= 0) {
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(
          row + w0, w0 < T ? (uint32_t)(T - w0) * 4u : 0u);
      const uint32_t vo = (uint32_t)lane * 4u;
      WindowDma<0, (NX < 16 ? NX : 16)>::run(rs, dst, vo);
#pragma unroll
      for (int x = 16; x < NX; ++x)
        if (x < TB / 64 || 64 * x < need)
          dma_dword(rs, dst + 64 * x, vo + 256u * (uint32_t)x);
    } else {
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(row, row_bytes);
#pragma unroll
      for (int x = 0; x < NX; ++x)
        if (x < TB / 64 || 64 * x < need)
          dma_dword(rs, dst + 64 * x, bin_offset(w0 + lane + 64 * x));
    }
    if (lane == 0) {
      meta[buf].row[k] = sp.x;
      meta[buf].maxlag[k] = sp.y;
    }
  };
  // metadata staging: waves 0 .. NM/64-1 take the delay words (pair q =
  // (e, kk) = (q / JT, q % JT)), the next NM/64 waves the weights, and for
  // the diffuse + specular plan the next NM/64 the diffuse weights.  Every
  // thread issues the same loads (entry index clamped into the batch,
  // weight array chosen per wave), without branches: the values are only
  // used after the sums, so nothing waits for them (or, vmcnt being in
  // order, for the window copies issued before them) in between.  Pairs
  // past the batch's end are staged with the last entry's values (their
  // own entry index in the offset): the sums never use them, and the
  // walk's final prefetch, which may read entry NB - 1, stays inside that
  // entry's window.
  const int q_part = k / (NM / 64);
  const int q = (int)threadIdx.x - q_part * NM;
  const int q_e = q / JT, q_k = q % JT;
  const double *wsrc = DS && q_part == 2 ? dw : w;
  int32_t m_word = 0, m_max = 0;
  double m_w = 0.0;
  auto meta_load = [&](int mb, int n_in) {
    const int64_t mm = (int64_t)(mb + min(q_e, n_in - 1));
    m_word = lagw[mm * JT + q_k];
    m_max = reinterpret_cast<const int32_t *>(span)[mm * 4 + 1];
    m_w = wsrc[mm * JT + q_k];
  };
  // the staged word of (entry e, slot kk): the byte offset of the slot's
  // view in the window buffer, (e * W + min(maxlag - delay, SMAX)) * 4,
  // with flags: kAbsent (slot without the entry; the offset then e * W * 4),
  // kFar (delay more than SMAX below maxlag: the view is read from global
  // memory) and kLastInGroup -- so an entry costs the walk one mask and
  // one add to address its view
  auto meta_write = [&](int buf) {
    if (q_part == 0) {
      const int32_t l = m_word & kLagMask;
      const bool absent = l == kLagMask;
      const int32_t off = absent ? 0 : m_max - l;
      meta[buf].word[q_e][q_k] = ((q_e * W + min(off, SMAX)) * 4) |
                                 (absent ? kAbsent : 0) |
                                 (off > SMAX ? kFar : 0) |
                                 (m_word & kLastInGroup);
    } else if (q_part == 1) {
      meta[buf].w[q_e][q_k] = m_w;
    } else if (DS && q_part == 2) {
      meta[buf].dw[DS ? q_e : 0][q_k] = m_w;
    }
  };

  double S[R], Sd[DS ? R : 1];
#pragma unroll
  for (int r = 0; r < R; ++r) S[r] = 0.0;
#pragma unroll
  for (int r = 0; r < (DS ? R : 1); ++r) Sd[r] = 0.0;

  // The sums of one batch, software-pipelined: the wave walks the entries
  // where its slot is present or ends a group (set bits of two ballots
  // over its 16 offset words, one per lane), in order, and issues the next
  // one's LDS reads (values and weights) before summing the current one.
  // The reads are unconditional (an absent entry, or one whose delay is
  // beyond the window, reads an in-window view it does not use), so the
  // sums wait only for the reads of their own entry.
  struct Cur {
    float v[R];
    double w, dw;
    int32_t lw;
  };
  // words: lane e (mod NB) holds entry e's staged word of this slot, taken
  // with v_readlane; the weights are read by broadcast LDS loads (with
  // v_readlane from a per-batch lane load instead: 94.8 against 93.3 ms)
  auto read_entry = [&](int buf, int idx, int32_t words, Cur &c) {
    c.lw = __builtin_amdgcn_readlane(words, idx);
    const float *src =
        reinterpret_cast<const float *>(
            reinterpret_cast<const char *>(&win(buf)[0][0]) +
            (c.lw & kAddrMask)) +
        lane;
#pragma unroll
    for (int r = 0; r < R; ++r) c.v[r] = src[64 * r];
    c.w = meta[buf].w[idx][k];
    if (DS) c.dw = meta[buf].dw[DS ? idx : 0][k];
  };
  auto sum_entry = [&](const float (&v)[R], double wk, double dk) {
    if (DS) {
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
  };
  auto use_entry = [&](int buf, int mb, int idx, const Cur &c) {
    if (!(c.lw & (kAbsent | kFar))) {
      sum_entry(c.v, c.w, c.dw);
    } else if (!(c.lw & kAbsent)) {
      // delay far from the entry's others: its bins from global memory,
      // into registers of their own (the prefetched view stays untouched)
      const int32_t row = __builtin_amdgcn_readfirstlane(meta[buf].row[idx]);
      const int32_t l = lagw[(int64_t)(mb + idx) * JT + k] & kLagMask;
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(
          prev + (int64_t)row * bstride + (int64_t)b * T, row_bytes);
      float v[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[r] = buffer_load(rs, bin_offset(t0 + lane + 64 * r - l));
      // drain here (vmcnt(0); expcnt and lgkmcnt left alone): with loads
      // of this path still pending at the join, the compiler would make
      // later reads into the same registers wait for vmcnt, and with it
      // for the next batch's window copies
      __builtin_amdgcn_s_waitcnt(0x0f70);
      sum_entry(v, c.w, c.dw);
    }
    if (c.lw & kLastInGroup) {
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
  auto compute = [&](int buf, int mb, int n_in) {
    const int32_t words = meta[buf].word[lane & (NB - 1)][k];
    const uint32_t in_mask = (uint32_t)((1ull << n_in) - 1ull);
    const uint32_t pres =
        (uint32_t)__ballot((words & kAbsent) == 0) & in_mask;
    const uint32_t ends =
        (uint32_t)__ballot((words & kLastInGroup) != 0) & in_mask;
    uint32_t todo = __builtin_amdgcn_readfirstlane(pres | ends);
    const int n_todo = __builtin_popcount(todo);
    if (n_todo == 0) return;
    // i-th entry of the walk in ca (i even) or cb (i odd); the next one's
    // reads are issued before the current one is summed (after the last,
    // a harmless read of entry NB - 1, whose staged word always addresses
    // its own window)
    constexpr uint32_t kStop = 1u << (NB - 1);
    Cur ca, cb;
    int idx = __builtin_ctz(todo);
    read_entry(buf, idx, words, ca);
#pragma unroll
    for (int i = 0; i < NB; i += 2) {
      if (i >= n_todo) break;  // uniform
      todo &= todo - 1u;
      int nx = __builtin_ctz(todo | kStop);
      read_entry(buf, nx, words, cb);
      use_entry(buf, mb, idx, ca);
      idx = nx;
      if (i + 1 >= n_todo) break;
      todo &= todo - 1u;
      nx = __builtin_ctz(todo | kStop);
      read_entry(buf, nx, words, ca);
      use_entry(buf, mb, idx, cb);
      idx = nx;
    }
  };

  // batch 0 into buffer 0; sp holds the span of this wave's entry of the
  // next batch to stage
  int4 sp = span_of(k);
  stage_load(0, n_at(0), sp);
  meta_load(0, n_at(0));
  meta_write(0);
  sp = span_of(NB + k);
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's window copies
  __syncthreads();
  // one step: sum batch mb from LDS buffer `buf` while batch mb + NB's
  // loads are in flight; they go to the other buffer afterwards
  auto step = [&](int mb, int buf) {
    const int n1 = n_at(mb + NB);
    if (n1 > 0) {
      stage_load(buf ^ 1, n1, sp);
      meta_load(mb + NB, n1);
      sp = span_of(mb + 2 * NB + k);
    }
    if (active) compute(buf, mb, n_at(mb));
    if (n1 > 0) meta_write(buf ^ 1);
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's window copies
    __syncthreads();
  };
  // two steps per iteration, so the buffer indices are compile-time
  // constants
  for (int mb = 0; mb < n_tile; mb += 2 * NB) {
    step(mb, 0);
    if (mb + NB < n_tile) step(mb + NB, 1);
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

// 16-row tiles: batches of 16 merged entries, windows of TB + 64 bins.
// R = 8 (512-bin blocks): 2 x 16 x 576 floats of LDS and 6.5 KiB of batch
// metadata, two workgroups of 16 waves per CU; R = 16 (1,024-bin blocks):
// 2 x 16 x 1,088 floats, one workgroup per CU.  Measured on the street
// canyon (T = 6,000, 4 x 4 blocks, 1.71 M merged entries): R = 8 98 ms,
// R = 16 108 ms; with the metadata broadcast by v_readlane from per-lane
// loads instead of staged in LDS, and the windows staged through
// registers, R = 16 took 131 ms.  The diffuse + specular plan runs R = 8
// with one workgroup per CU (its second sums need 95 VGPRs).
template <int R, int JT, int SMAX, bool DS>
int launch_tile_lds(int64_t P, int64_t B, int64_t T, int64_t M,
                    const int64_t *tptr, const int64_t *grp_ptr,
                    const int64_t *gdiff, const int32_t *mspan,
                    const int32_t *mlag, const double *mw, const double *mdw,
                    const float *prev, float *gsum, hipStream_t s) {
  const int64_t n_tiles = ceil_div(P, JT);
  const int64_t nt = ceil_div(T, (int64_t)64 * R);
  const int64_t nb = n_tiles * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(T < ((int64_t)1 << 29), "histogram too long (T >= 2^29)");
  hipLaunchKernelGGL((tile_lds_gather_kernel<R, JT, SMAX, DS>),
                     dim3((unsigned)nb), dim3(64 * JT), 0, s, P, (int)B,
                     (int)T, (int)nt, nb, M, tptr, grp_ptr, gdiff,
                     reinterpret_cast<const int4 *>(mspan), mlag, mw, mdw,
                     prev, gsum);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

// mspan[m] = {source row, largest and smallest delay of the merged entry's
// present slots, 0}.
__global__ __launch_bounds__(256) void tile_span_kernel(
    int64_t M, int JT, const int32_t *__restrict__ mrow,
    const int32_t *__restrict__ mlag, int4 *__restrict__ mspan) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  int32_t mx = 0, mn = kLagMask;
  for (int kk = 0; kk < JT; ++kk) {
    const int32_t l = mlag[m * JT + kk] & kLagMask;
    if (l != kLagMask) {
      mx = max(mx, l);
      mn = min(mn, l);
    }
  }
  mspan[m] = make_int4(mrow[m], mx, mn == kLagMask ? mx : mn, 0);
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

extern "C" int spr_plan_tile_span(int64_t M, int64_t JT, const int32_t *mrow,
                                  const int32_t *mlag, int32_t *mspan,
                                  void *stream) {
  clear_error();
  SPR_REQUIRE(M >= 0 && JT >= 2 && JT <= 16,
              "spr_plan_tile_span: bad sizes (JT must be 2..16)");
  if (M == 0) return SPR_OK;
  SPR_REQUIRE(mrow && mlag && mspan, "spr_plan_tile_span: null pointer");
  hipLaunchKernelGGL(tile_span_kernel, dim3(grid_of(M)), dim3(256), 0,
                     as_stream(stream), M, (int)JT, mrow, mlag,
                     reinterpret_cast<int4 *>(mspan));
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
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 && (JT == 4 || JT == 8),
              "spr_exchange_gather_tiled: bad sizes (JT must be 4 or 8; "
              "16-row tiles: spr_exchange_gather_tiled_shape)");
  SPR_REQUIRE(tptr && grp_ptr && mrow && mlag && mw && prev && gsum,
              "spr_exchange_gather_tiled: null pointer");
  return launch_tile_gather(P, (int)JT, B, T, M, tptr, grp_ptr, mrow, mlag,
                            mw, prev, gsum, as_stream(stream));
}

extern "C" int spr_exchange_gather_tiled_ds(
    int64_t P, int64_t JT, int64_t B, int64_t T, int64_t M,
    const int64_t *tptr, const int64_t *gstart, const int64_t *gdiff,
    const int32_t *mspan, const int32_t *mlag, const double *mw,
    const double *mdw, const float *prev, float *gsum, void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 && JT == 16,
              "spr_exchange_gather_tiled_ds: bad sizes (JT must be 16)");
  SPR_REQUIRE(tptr && gstart && gdiff && mspan && mlag && mw && mdw && prev &&
                  gsum,
              "spr_exchange_gather_tiled_ds: null pointer");
  return launch_tile_lds<8, 16, 64, true>(P, B, T, M, tptr, gstart, gdiff,
                                          mspan, mlag, mw, mdw, prev, gsum,
                                          as_stream(stream));
}

extern "C" int spr_exchange_gather_tiled_shape(
    int64_t P, int64_t JT, int64_t R, int64_t B, int64_t T, int64_t M,
    const int64_t *tptr, const int64_t *grp_ptr, const int32_t *mspan,
    const int32_t *mlag, const double *mw, const float *prev, float *gsum,
    void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && M >= 0 && JT == 16 &&
                  (R == 8 || R == 16),
              "spr_exchange_gather_tiled_shape: bad sizes (JT 16, R 16 or 8)");
  SPR_REQUIRE(tptr && grp_ptr && mspan && mlag && mw && prev && gsum,
              "spr_exchange_gather_tiled_shape: null pointer");
  hipStream_t s = as_stream(stream);
  if (R == 8)
    return launch_tile_lds<8, 16, 64, false>(P, B, T, M, tptr, grp_ptr,
                                             nullptr, mspan, mlag, mw,
                                             nullptr, prev, gsum, s);
  return launch_tile_lds<16, 16, 64, false>(P, B, T, M, tptr, grp_ptr,
                                            nullptr, mspan, mlag, mw, nullptr,
                                            prev, gsum, s);
}
