// Energy exchange: destination-major plan construction, order-0 injection
// and the per-order gather kernel.
//
// Reference: _energy_exchange (sparrowpy/classes/RadiosityFast.py:1073-1145)
// and _energy_exchange_init_energy (:1037-1070).
//
// Data layout in HBM (float32, C order): every state buffer is [P, D, B, T]
// like the reference's E_matrix, so a (patch, direction, band) histogram is
// one contiguous row of T bins.  One reflection order reads, for each
// destination j and band b, the rows prev[i, dir_ij, b, :] of all patches
// i that see j, shifted by their delays, and writes the D rows
// cur[j, :, b, :].
//
// Kernel mapping of the per-pair plan: a workgroup owns (j, b, a block of
// time bins).  With 16 or more (padded) directions the order is a matrix
// product on the FP64 matrix cores (exchange_mfma_kernel); below that each
// thread owns R bins (strided by BS so every load instruction is one
// contiguous 256-byte wavefront access) and accumulates all DP directions
// in float64 registers, the coefficient rows coef[e, b, :] staged in LDS.
// The grouped plans (group sums, then a spread by each group's BRDF row)
// are below, with their construction from the geometry or from a dense F~.
#include <mutex>

#include "spr_common.hpp"
#include "spr_fft.hpp"
#include "spr_util.hpp"
#include "spr_exchange.hpp"

namespace spr {
namespace {

// ---------------------------------------------------------------------
// plan rows: for destination j, every i != j visible in either orientation
// ---------------------------------------------------------------------
__device__ __forceinline__ bool sym_vis(const uint8_t *vis, int64_t P,
                                        int64_t i, int64_t j) {
  if (i == j) return false;
  return (i < j ? vis[i * P + j] : vis[j * P + i]) != 0;
}

// One wavefront per destination: ballot over 64 candidate sources at a time.
__global__ __launch_bounds__(256) void plan_rows_kernel(
    int64_t P, const uint8_t *__restrict__ vis, int64_t *__restrict__ row_ptr,
    int32_t *__restrict__ src, int fill) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= P) return;
  int64_t pos = fill ? row_ptr[j] : 0;
  for (int64_t base = 0; base < P; base += 64) {
    int64_t i = base + lane;
    bool f = (i < P) && sym_vis(vis, P, i, j);
    uint64_t m = __ballot(f);
    if (fill && f) {
      uint64_t lt = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
      src[pos + __popcll(lt)] = (int32_t)i;
    }
    pos += __popcll(m);
  }
  if (!fill && lane == 0) row_ptr[j] = pos;
}

// Pair list -> dense upper-triangular visibility (for spr_energy_exchange).
__global__ void pairs_to_vis_kernel(int64_t P, int64_t V,
                                    const int32_t *__restrict__ pairs,
                                    uint8_t *__restrict__ vis) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  int64_t a = pairs[2 * v], b = pairs[2 * v + 1];
  if (a < 0 || b < 0 || a >= P || b >= P) return;
  if (a < b)
    vis[a * P + b] = 1;
  else if (b < a)
    vis[b * P + a] = 1;
}

// ---------------------------------------------------------------------
// plan delays / directions / coefficients
// ---------------------------------------------------------------------
__device__ __forceinline__ int32_t clamp_lag(int64_t n) {
  // a delay >= T contributes nothing; clamp so it fits int32
  return n > 0x3fffffff ? 0x3fffffff : (int32_t)n;
}

// Outgoing direction of a visible pair.  The direction map holds a valid
// index for every visible pair; an out-of-range entry (invalid input) would
// make the exchange read outside the patch's histograms, so it is clamped
// to direction 0 to keep every access in bounds.
__device__ __forceinline__ int32_t checked_dir(int64_t d, int64_t D) {
  return (d < 0 || d >= D) ? 0 : (int32_t)d;
}

__global__ __launch_bounds__(256) void plan_from_dense_kernel(
    int64_t P, int64_t D, int64_t B, int64_t DP, double c, double dt,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const double *__restrict__ dist, const int64_t *__restrict__ p2o,
    const double *__restrict__ fft, int32_t *__restrict__ lag,
    int32_t *__restrict__ dir, double *__restrict__ coef) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x) {
    int64_t i = src[e];
    lag[e] = clamp_lag(delay_floor(dist[i * P + j], c, dt));
    dir[e] = checked_dir(p2o[i * P + j], D);
    const double *f = fft + (i * P + j) * D * B;
    for (int64_t b = 0; b < B; ++b)
      for (int64_t d = 0; d < DP; ++d)
        coef[(e * B + b) * DP + d] = d < D ? f[d * B + b] : 0.0;
  }
}

__global__ __launch_bounds__(256) void plan_from_geometry_kernel(
    int64_t P, int64_t D, int64_t B, int64_t DP, double c, double dt,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const double *__restrict__ centers, const double *__restrict__ areas,
    const double *__restrict__ ff, const int64_t *__restrict__ p2o,
    const double *__restrict__ air, const int64_t *__restrict__ wall_ids,
    int64_t S, const double *__restrict__ brdf,
    const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs, int32_t *__restrict__ lag,
    int32_t *__restrict__ dir, double *__restrict__ coef) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x) {
    int64_t i = src[e];
    // RadiosityFast.py:540-543: |c_i - c_j|
    double dij = norm(load3(centers + 3 * i) - load3(centers + 3 * j));
    lag[e] = clamp_lag(delay_floor(dij, c, dt));
    dir[e] = checked_dir(p2o[i * P + j], D);
    PairFactor pf = pair_factor(i, j, P, centers, areas, ff, wall_ids, S, D,
                                B, brdf, brdf_index, in_dirs);
    for (int64_t b = 0; b < B; ++b) {
      double ab = air ? exp(-air[b] * pf.distance) : 1.0;
      for (int64_t d = 0; d < DP; ++d)
        coef[(e * B + b) * DP + d] =
            d < D ? pair_coef(pf, ab, air != nullptr, d, b, B) : 0.0;
    }
  }
}

// ---------------------------------------------------------------------
// order 0 (RadiosityFast.py:1066-1069)
// ---------------------------------------------------------------------
__global__ void init_kernel(int64_t P, int64_t D, int64_t B, int64_t T,
                            const double *__restrict__ e0dir,
                            const double *__restrict__ d0, double c,
                            double dt, float *__restrict__ prev,
                            float *__restrict__ tot_hi,
                            float *__restrict__ tot_lo, int *err) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P * D * B) return;
  int64_t i = q / (D * B);
  int64_t n = delay_floor(d0[i], c, dt);
  if (n < 0 || n >= T) {
    atomicOr(err, 1);
    return;
  }
  double v = e0dir[q];
  float hi = (float)v;
  int64_t idx = q * T + n;
  prev[idx] = hi;
  tot_hi[idx] = hi;
  tot_lo[idx] = (float)(v - (double)hi);
}

// ---------------------------------------------------------------------
// one reflection order
// ---------------------------------------------------------------------

// Gather of C consecutive plan entries starting at e (row ends at e_end):
// the prev histogram row `row(e)` shifted by the entry's delay.  Metadata
// loads are unconditional (the index is clamped into the row) so they can
// be issued together; entries past the row end or whose delay reaches
// beyond the tile are inactive and load nothing.
template <int R, int BS, int C, typename RowFn>
__device__ __forceinline__ void gather_chunk(
    int64_t e, int64_t e_end, int B, int T, int b, int t0, int t_end,
    RowFn row_of, const int32_t *__restrict__ lag,
    const float *__restrict__ prev, float (&buf)[C][R], int (&n)[C],
    bool (&act)[C]) {
  const int64_t dstride = (int64_t)B * T;
  int64_t row[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int64_t ee = e + c < e_end ? e + c : e_end - 1;
    n[c] = lag[ee];
    row[c] = row_of(ee);
    act[c] = (e + c < e_end) && ((n[c] & kLagMask) < t_end);
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float *s = prev + row[c] * dstride + (int64_t)b * T;
    const int nc = n[c] & kLagMask;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int t = t0 + (int)threadIdx.x + BS * r;
      const int ts = t - nc;
      buf[c][r] = (act[c] && t < T && ts >= 0) ? s[ts] : 0.f;
    }
  }
}

// Write one order's tile (cur) and add it to the compensated total
// (tot_hi + tot_lo, updated in float64 and split back into two floats).
template <int DP, int R, int BS>
__device__ __forceinline__ void store_tile(const double (&acc)[DP][R], int D,
                                           int B, int T, int64_t j, int b,
                                           int t0, float *__restrict__ cur,
                                           float *__restrict__ tot_hi,
                                           float *__restrict__ tot_lo) {
  const int64_t dstride = (int64_t)B * T;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = t0 + (int)threadIdx.x + BS * r;
    if (t >= T) continue;
#pragma unroll
    for (int d = 0; d < DP; ++d) {
      if (d >= D) continue;  // padded direction (compile-time d)
      const int64_t idx = (j * D + d) * dstride + (int64_t)b * T + t;
      const double a = acc[d][r];
      cur[idx] = (float)a;
      const double sum = ((double)tot_hi[idx] + (double)tot_lo[idx]) + a;
      const float h = (float)sum;
      tot_hi[idx] = h;
      tot_lo[idx] = (float)(sum - (double)h);
    }
  }
}

// Generic plan: every entry carries its own coefficient row coef[e, b, :].
// The next chunk's gathers are issued before the current chunk's FMAs so
// that C * R loads per lane are in flight while the VALU works.  The
// coefficient rows (each used once, so a scalar load per entry would wait
// for L2 every time) come through LDS: the workgroup loads the rows of the
// next NBE entries with coalesced vector loads while the current NBE are
// summed, and every wave reads the current rows by broadcast.
template <int DP, int R, int BS, int C>
__global__ __launch_bounds__(BS) void exchange_step_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const int32_t *__restrict__ lag, const int32_t *__restrict__ dir,
    const double *__restrict__ coef, const float *__restrict__ prev,
    float *__restrict__ cur, float *__restrict__ tot_hi,
    float *__restrict__ tot_lo) {
  constexpr int NBE = 16;                      // entries per row batch
  constexpr int NW = NBE * DP;                 // doubles per batch
  constexpr int NLD = (NW + BS - 1) / BS;      // of them per thread
  static_assert(NBE % C == 0, "chunks tile a batch");
  __shared__ double cfs[2][NW];
  const int64_t tile = xcd_tile(blockIdx.x, nblocks);
  const int tb = (int)(tile % n_tblk);
  const int64_t jb = tile / n_tblk;
  const int b = (int)(jb % B);
  const int64_t j = jb / B;
  const int t0 = tb * (BS * R);
  const int t_end = min(T, t0 + BS * R);

  double acc[DP][R];
#pragma unroll
  for (int d = 0; d < DP; ++d)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[d][r] = 0.0;

  const int64_t e_begin = row_ptr[j], e_end = row_ptr[j + 1];
  if (e_end > e_begin) {  // uniform: before any barrier
    auto row_of = [&](int64_t ee) {
      return (int64_t)src[ee] * D + dir[ee];
    };
    // rows of entries eb .. eb+NBE-1 (band b) into registers, then LDS
    double pre[NLD];
    auto rows_load = [&](int64_t eb) {
#pragma unroll
      for (int k = 0; k < NLD; ++k) {
        const int i = (int)threadIdx.x + k * BS;
        const int q = i / DP, d = i % DP;
        pre[k] = (i < NW && eb + q < e_end)
                     ? coef[((eb + q) * B + b) * DP + d]
                     : 0.0;
      }
    };
    auto rows_store = [&](int buf) {
#pragma unroll
      for (int k = 0; k < NLD; ++k) {
        const int i = (int)threadIdx.x + k * BS;
        if (i < NW) cfs[buf][i] = pre[k];
      }
    };
    rows_load(e_begin);
    rows_store(0);
    __syncthreads();
    float buf[C][R];
    int n[C];
    bool act[C];
    gather_chunk<R, BS, C>(e_begin, e_end, B, T, b, t0, t_end, row_of, lag,
                           prev, buf, n, act);
    int cb = 0;
    for (int64_t eb = e_begin; eb < e_end; eb += NBE) {
      const bool more = eb + NBE < e_end;  // uniform
      if (more) rows_load(eb + NBE);
#pragma unroll
      for (int q0 = 0; q0 < NBE; q0 += C) {
        const int64_t e = eb + q0;
        if (e >= e_end) break;  // uniform
        float nbuf[C][R];
        int nn[C];
        bool nact[C];
        gather_chunk<R, BS, C>(e + C, e_end, B, T, b, t0, t_end, row_of,
                               lag, prev, nbuf, nn, nact);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          if (!act[c]) continue;
          const double *cf = &cfs[cb][(q0 + c) * DP];
          double ev[R];
#pragma unroll
          for (int r = 0; r < R; ++r) ev[r] = (double)buf[c][r];
#pragma unroll
          for (int d = 0; d < DP; ++d) {
            const double cd = cf[d];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[d][r] = fma(cd, ev[r], acc[d][r]);
          }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
          act[c] = nact[c];
#pragma unroll
          for (int r = 0; r < R; ++r) buf[c][r] = nbuf[c][r];
        }
      }
      // one barrier per batch: the next rows are in the other buffer
      // before any wave reads them, and every wave has finished reading
      // this buffer before it is written again (after the next batch)
      if (more) rows_store(cb ^ 1);
      __syncthreads();
      cb ^= 1;
    }
  }
  store_tile<DP, R, BS>(acc, D, B, T, j, b, t0, cur, tot_hi, tot_lo);
}

// Per-pair plan on the matrix cores (DP >= 16).  For destination j and band
// b one order is the product
//   cur[j, :, b, t] = sum_e coef[e, b, :] x_e[t],
//   x_e[t] = prev[src_e * D + dir_e, b, t - lag_e]  (0 for t < lag_e),
// of the [DP x E_j] coefficient rows and the [E_j x T] shifted source
// histograms, formed with v_mfma_f64_16x16x4f64: K = 4 entries per step;
// lane l holds A[l & 15][l >> 4] (coefficient of direction l & 15 of entry
// l >> 4), B[l >> 4][l & 15] (bin l & 15 of that entry) and, in register
// r, D[(l >> 4) + 4r][l & 15].  A wave owns MT = DP / 16 direction tiles
// by NT tiles of 16 bins (MT * NT = 8 accumulators of 4 doubles); each
// step loads, per lane, its entry's MT coefficients and NT source bins,
// the next step's while the current one is multiplied.  Entries past the
// row end are zero.  FP64 throughout; the sums' order differs from the
// vector kernel's (float64 rounding).
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

template <int DP, int NT>
__global__ __launch_bounds__(256) void exchange_mfma_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const int32_t *__restrict__ lag, const int32_t *__restrict__ dir,
    const double *__restrict__ coef, const float *__restrict__ prev,
    float *__restrict__ cur, float *__restrict__ tot_hi,
    float *__restrict__ tot_lo) {
  constexpr int MT = DP / 16;
  constexpr int WBINS = 16 * NT;  // bins per wave
  static_assert(MT * NT == 8, "eight accumulator tiles per wave");
  const int64_t tile = xcd_tile(blockIdx.x, nblocks);
  const int tb = (int)(tile % n_tblk);
  const int64_t jb = tile / n_tblk;
  const int b = (int)(jb % B);
  const int64_t j = jb / B;
  const int lane = (int)threadIdx.x & 63;
  const int kk = lane >> 4, cc = lane & 15;
  const int tw = (tb * 4 + ((int)threadIdx.x >> 6)) * WBINS;
  const int64_t BT = (int64_t)B * T;
  const int64_t e_begin = row_ptr[j], e_end = row_ptr[j + 1];

  mfma_d4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma_d4{0.0, 0.0, 0.0, 0.0};

  // Software pipeline, all loads branch-free: the metadata of step s + 2
  // and the operands of step s + 1 are loaded while step s is multiplied,
  // and a load's value is first used one step after it was issued.  An
  // entry past the row end reads the row's last entry, a bin outside the
  // source row (t < lag, or t - lag >= T) reads the row's first bin, and
  // selects at the use zero both.  (Each 16-lane group reads its own source
  // row, so a buffer descriptor would not be wave-uniform.)
  struct Meta {
    int64_t e;  // entry, clamped to the row
    bool in;
    int32_t src, dir, lag;
  };
  struct Ops {
    double a[MT];
    float x[NT];
    bool in, ok[NT];
  };
  auto meta = [&](int64_t e0) {
    Meta m;
    const int64_t e = e0 + kk;
    m.in = e < e_end;
    m.e = m.in ? e : e_end - 1;
    m.src = src[m.e];
    m.dir = dir[m.e];
    m.lag = lag[m.e];
    return m;
  };
  auto issue = [&](const Meta &m) {
    Ops o;
    o.in = m.in;
    const double *cf = coef + (m.e * B + b) * DP + cc;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) o.a[mt] = cf[mt * 16];
    const float *row =
        prev + ((int64_t)m.src * D + m.dir) * BT + (int64_t)b * T;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int ts = tw + nt * 16 + cc - m.lag;
      o.ok[nt] = m.in && (unsigned)ts < (unsigned)T;
      o.x[nt] = row[o.ok[nt] ? ts : 0];
    }
    return o;
  };

  auto mult = [&](const Ops &o) {
    double a[MT], x[NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) a[mt] = o.in ? o.a[mt] : 0.0;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) x[nt] = o.ok[nt] ? (double)o.x[nt] : 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
            a[mt], x[nt], acc[mt][nt], 0, 0, 0);
  };
  if (e_end > e_begin) {  // uniform
    // two operand sets in turn; the scheduling barriers keep each step's
    // loads ahead of the previous step's products
    Meta m1 = meta(e_begin + 4);
    Ops oa = issue(meta(e_begin));
    for (int64_t e0 = e_begin; e0 < e_end; e0 += 8) {
      const Meta m2 = meta(e0 + 8);
      const Ops ob = issue(m1);
      __builtin_amdgcn_sched_barrier(0);
      mult(oa);
      m1 = meta(e0 + 12);
      oa = issue(m2);
      __builtin_amdgcn_sched_barrier(0);
      mult(ob);
    }
  }

  // cur and the compensated total (as store_tile)
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int d = mt * 16 + kk + 4 * r;
        const int t = tw + nt * 16 + cc;
        if (d >= D || t >= T) continue;
        const int64_t idx = (j * D + d) * BT + (int64_t)b * T + t;
        const double v = acc[mt][nt][r];
        cur[idx] = (float)v;
        const double sum = ((double)tot_hi[idx] + (double)tot_lo[idx]) + v;
        const float h = (float)sum;
        tot_hi[idx] = h;
        tot_lo[idx] = (float)(sum - (double)h);
      }
}

// Grouped plan, pass 1 of an order: per group g of destination j's sources
// (entries sharing one BRDF row), the weighted shifted sum
//   gsum[g, b, t] = sum_{e in g} scal[b, e] * prev[rowid[e], b, t - lag[e]]
// in float64, stored as float32.
//
// The destination's entries are walked in blocks of 64: one coalesced
// vector load brings a block's metadata (source row, delay word, weight),
// one entry per lane, and each entry's values are broadcast with
// v_readlane.  A block is fully unrolled into 64/C chunks of C entries;
// chunk q+1's C*R gathers are issued into one register set while chunk q
// is summed from the other, so no register holding an in-flight load is
// ever copied and the compiler can wait for exactly the chunk it needs.
// Each lane owns R bins.  A source bin before the histogram start
// (t < lag) is a negative bin whose wrapped byte offset (bin_offset) lies
// beyond the row and reads as 0; so does every bin of an entry past the
// row end or whose delay reaches beyond the tile (descriptor of size 0).
// Requires T < 2^29, so that t - lag >= -T maps to an offset >= 2^31.
constexpr int kMetaBlock = 64;

struct EntryMeta {
  int32_t row, lag;
  double w;
};

__device__ __forceinline__ EntryMeta load_meta(
    int64_t e0, int64_t e_end, const int32_t *__restrict__ rowid,
    const int32_t *__restrict__ lag, const double *__restrict__ w) {
  const int64_t e = e0 + (int64_t)(threadIdx.x & 63);
  const bool in = e < e_end;
  const int64_t ec = in ? e : e_end - 1;
  EntryMeta m;
  m.row = rowid[ec];
  m.lag = in ? lag[ec] : kLagMask;  // past the row end: no bins, no flag
  m.w = in ? w[ec] : 0.0;
  return m;
}

template <int R, int BS, int C>
__global__ __launch_bounds__(BS) void group_gather_kernel(
    int B, int T, int n_tblk, int64_t nblocks, int64_t Vd,
    const int64_t *__restrict__ row_ptr, const int64_t *__restrict__ grp_ptr,
    const int32_t *__restrict__ rowid, const int32_t *__restrict__ lag,
    const double *__restrict__ scal, const float *__restrict__ prev,
    float *__restrict__ gsum) {
  static_assert(kMetaBlock % (2 * C) == 0, "chunks pair up in a block");
  constexpr int NQ = kMetaBlock / C;  // chunks per block
  int64_t j;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, j, b, tb);
  const int t0 = tb * (BS * R);
  const int t_end = min(T, t0 + BS * R);
  const int64_t e_begin = row_ptr[j], e_end = row_ptr[j + 1];
  if (e_end <= e_begin) return;
  const double *w = scal + (int64_t)b * Vd;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int tl = t0 + (int)threadIdx.x;
  int64_t g = grp_ptr[j];

  // gathers of chunk q (entries q*C .. q*C+C-1 of the block of `m`)
  auto fetch = [&](const EntryMeta &m, int q, float (&v)[C][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = q * C + c;
      const int32_t rk = __builtin_amdgcn_readlane(m.row, k);
      const int32_t lk = __builtin_amdgcn_readlane(m.lag, k);
      const int n = min(lk & kLagMask, T);
      const float *src = prev + ((int64_t)rk * B + b) * T;
      const __amdgpu_buffer_rsrc_t rs =
          make_rsrc(src, n < t_end ? row_bytes : 0u);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[c][r] = buffer_load(rs, bin_offset(tl + BS * r - n));
    }
  };

  double S[R];
#pragma unroll
  for (int r = 0; r < R; ++r) S[r] = 0.0;
  // sum chunk q; store a group's sums at its last entry
  auto consume = [&](const EntryMeta &m, int q, const float (&v)[C][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = q * C + c;
      const double wk = readlane_f64(m.w, k);
#pragma unroll
      for (int r = 0; r < R; ++r) S[r] = fma(wk, (double)v[c][r], S[r]);
      if (__builtin_amdgcn_readlane(m.lag, k) & kLastInGroup) {
        const __amdgpu_buffer_rsrc_t os =
            make_rsrc(gsum + (g * B + b) * (int64_t)T, row_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          buffer_store((float)S[r], os, (uint32_t)(tl + BS * r) * 4u);
          S[r] = 0.0;
        }
        g += 1;
      }
    }
  };

  EntryMeta cur = load_meta(e_begin, e_end, rowid, lag, w);
  EntryMeta next = load_meta(e_begin + kMetaBlock, e_end, rowid, lag, w);
  float va[C][R], vb[C][R];
  fetch(cur, 0, va);
  for (int64_t blk = e_begin; blk < e_end; blk += kMetaBlock) {
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
    next = load_meta(blk + 2 * kMetaBlock, e_end, rowid, lag, w);
  }
}

// Gather pass of the diffuse + specular plan: as group_gather_kernel, with
// a second weight per entry (dscal) summed over the whole row into the
// row's diffuse group (its last group slot).  Written as a runtime loop over
// pairs of C-entry chunks with scalar metadata loads: with two weights the
// fully unrolled 64-entry blocks of group_gather_kernel exhaust the
// register file.
template <int R, int BS, int C>
__global__ __launch_bounds__(BS) void ds_gather_kernel(
    int B, int T, int n_tblk, int64_t nblocks, int64_t Vd,
    const int64_t *__restrict__ row_ptr, const int64_t *__restrict__ grp_ptr,
    const int32_t *__restrict__ rowid, const int32_t *__restrict__ lag,
    const double *__restrict__ scal, const double *__restrict__ dscal,
    const float *__restrict__ prev, float *__restrict__ gsum) {
  int64_t j;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, j, b, tb);
  const int t0 = tb * (BS * R);
  const int t_end = min(T, t0 + BS * R);
  const int64_t e_begin = row_ptr[j], e_end = row_ptr[j + 1];
  if (e_end <= e_begin) return;
  const double *w = scal + (int64_t)b * Vd;
  const double *w2 = dscal + (int64_t)b * Vd;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int tl = t0 + (int)threadIdx.x;
  int64_t g = grp_ptr[j];

  auto fetch = [&](int64_t e0, float (&v)[C][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const bool in_row = e0 + c < e_end;
      const int64_t ee = in_row ? e0 + c : e_end - 1;
      const int n = min(lag[ee] & kLagMask, T);
      const float *src = prev + ((int64_t)rowid[ee] * B + b) * T;
      const __amdgpu_buffer_rsrc_t rs =
          make_rsrc(src, in_row && n < t_end ? row_bytes : 0u);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[c][r] = buffer_load(rs, bin_offset(tl + BS * r - n));
    }
  };
  double S[R], Sd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) S[r] = Sd[r] = 0.0;
  auto consume = [&](int64_t e0, const float (&v)[C][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int64_t e = e0 + c;
      if (e >= e_end) break;
      const double wk = w[e], dk = w2[e];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double x = (double)v[c][r];
        S[r] = fma(wk, x, S[r]);
        Sd[r] = fma(dk, x, Sd[r]);
      }
      if (lag[e] & kLastInGroup) {
        const __amdgpu_buffer_rsrc_t os =
            make_rsrc(gsum + (g * B + b) * (int64_t)T, row_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          buffer_store((float)S[r], os, (uint32_t)(tl + BS * r) * 4u);
          S[r] = 0.0;
        }
        g += 1;
      }
    }
  };
  float va[C][R], vb[C][R];
  fetch(e_begin, va);
  for (int64_t e = e_begin; e < e_end; e += 2 * C) {
    fetch(e + C, vb);
    consume(e, va);
    fetch(e + 2 * C, va);
    consume(e + C, vb);
  }
  const __amdgpu_buffer_rsrc_t os = make_rsrc(
      gsum + ((grp_ptr[j + 1] - 1) * B + b) * (int64_t)T, row_bytes);
#pragma unroll
  for (int r = 0; r < R; ++r)
    buffer_store((float)Sd[r], os, (uint32_t)(tl + BS * r) * 4u);
}

// Grouped plan, pass 2: spread the group sums over the D directions with
// their BRDF rows and update the order histogram and the running total:
//   cur[j, :, b, t] = sum_{g of j} rtab[key_g, b, :] * gsum[g, b, t].
// A workgroup of 4 waves takes one (destination, band, time block); each
// lane keeps DW (up to 32) directions of R bins as float64 sums (DW * R =
// 32 below 32 directions, 64 at 32); beyond 32 directions the waves split
// them NWD = DP / DW ways and the bins NWT = 4 / NWD ways.
// The running total (hi + lo, updated in float64) is read K (d, r) pairs
// at a time through buffer descriptors whose out-of-range offsets stand
// for padded directions and bins past T (loads give 0, stores are
// dropped).  Requires D * B * T < 2^29 (byte offsets below 2^31).
template <int DP>
struct SpreadShape {
  static constexpr int DW = DP < 32 ? DP : 32;   // directions per lane
  static constexpr int NWD = DP / DW;            // wave groups along d
  static constexpr int NWT = 4 / NWD;            // wave groups along t
  // 2 bins per lane at 32 directions: 21.1 / 45.2 ms (cur only / paired
  // total) on the street canyon against 22.0 / 47.1 ms with 1
  static constexpr int R = DW >= 32 ? 2 : 32 / DW > 8 ? 8 : 32 / DW;
  static constexpr int TILE = NWT * 64 * R;      // bins per workgroup
};

// MODE: 0 writes cur only; 1 adds cur to the total; 2 adds prev + cur (two
// orders at once, so the total is read and written every other order).
template <int DP, int MODE>
__global__ __launch_bounds__(256) void group_spread_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks,
    const int64_t *__restrict__ grp_ptr, const int32_t *__restrict__ grp_key,
    const double *__restrict__ rtab, const float *__restrict__ gsum,
    const float *__restrict__ prev, float *__restrict__ cur,
    float *__restrict__ tot_hi, float *__restrict__ tot_lo) {
  using Sh = SpreadShape<DP>;
  constexpr int DW = Sh::DW, R = Sh::R, K = 8;
  // time-major tile order: the blocks one XCD runs at a time write long
  // runs of each direction row (no reuse across destinations to keep here;
  // 41.6 ms instead of 47.1 ms with the destination-major order on the
  // street canyon)
  const int64_t tile = xcd_tile(blockIdx.x, nblocks);
  const int64_t P = nblocks / (B * n_tblk);
  const int tb = (int)(tile % n_tblk);
  const int64_t j = (tile / n_tblk) % P;
  const int b = (int)(tile / (n_tblk * P));
  // the wave index as a scalar (readfirstlane): the BRDF-row reads below
  // then stay scalar loads shared by the wave, not per-lane loads
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  const int wd = wave % Sh::NWD, wt = wave / Sh::NWD;
  const int d0 = wd * DW;  // first direction of this wave
  const int tl = tb * Sh::TILE + wt * (64 * R) + lane;  // bin of r = 0
  const uint32_t row_bytes = (uint32_t)T * 4u;

  double acc[DW][R];
#pragma unroll
  for (int d = 0; d < DW; ++d)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[d][r] = 0.0;
  // groups are taken GC at a time, the next GC rows' loads in flight while
  // the current ones are summed; rows past the destination's last group
  // read as 0 (descriptor of size 0) and add nothing
  constexpr int GC = DW >= 32 ? 2 : 4;
  const int64_t g0 = grp_ptr[j], g1 = grp_ptr[j + 1];
  auto load = [&](int64_t g, float (&v)[GC][R]) {
#pragma unroll
    for (int c = 0; c < GC; ++c) {
      const int64_t gg = g + c;
      const __amdgpu_buffer_rsrc_t rs =
          make_rsrc(gsum + (min(gg, g1 - 1) * B + b) * (int64_t)T,
                    gg < g1 ? row_bytes : 0u);
#pragma unroll
      for (int r = 0; r < R; ++r)
        v[c][r] = buffer_load(rs, (uint32_t)(tl + 64 * r) * 4u);
    }
  };
  auto step = [&](int64_t g, const float (&v)[GC][R]) {
#pragma unroll
    for (int c = 0; c < GC; ++c) {
      const int64_t gg = min(g + c, g1 - 1);
      const double *rt = rtab + ((int64_t)grp_key[gg] * B + b) * DP + d0;
#pragma unroll
      for (int d = 0; d < DW; ++d) {
        const double cd = rt[d];
#pragma unroll
        for (int r = 0; r < R; ++r)
          acc[d][r] = fma(cd, (double)v[c][r], acc[d][r]);
      }
    }
  };
  if (g1 > g0) {
    float va[GC][R], vb[GC][R];
    load(g0, va);
    for (int64_t g = g0; g < g1; g += 2 * GC) {
      // the scheduling barriers keep each prefetch ahead of the sums
      load(g + GC, vb);
      __builtin_amdgcn_sched_barrier(0);
      step(g, va);
      load(g + 2 * GC, va);
      __builtin_amdgcn_sched_barrier(0);
      step(g + GC, vb);
    }
  }

  // histogram and running total: element (j, d, b, t) at
  // ((j * D + d) * B + b) * T + t; one descriptor spans d = 0 .. D-1
  const int64_t dstride = (int64_t)B * T;
  const int64_t base = j * D * dstride + (int64_t)b * T;
  const uint32_t span = (uint32_t)(((int64_t)(D - 1) * dstride + T) * 4);
  const __amdgpu_buffer_rsrc_t rc = make_rsrc(cur + base, span);
  constexpr int NP = DW * R;  // (d, r) pairs of this lane
  if (MODE == 0) {  // order histogram only (no running total)
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int d = d0 + p / R, r = p % R;
      const int t = tl + 64 * r;
      const uint32_t off = (d < D && t < T)
                               ? (uint32_t)((int64_t)d * dstride + t) * 4u
                               : kNoOffset;
      buffer_store((float)acc[p / R][r], rc, off);
    }
    return;
  }
  const __amdgpu_buffer_rsrc_t rh = make_rsrc(tot_hi + base, span);
  const __amdgpu_buffer_rsrc_t rl = make_rsrc(tot_lo + base, span);
  const __amdgpu_buffer_rsrc_t rp =
      make_rsrc(MODE == 2 ? prev + base : cur + base, MODE == 2 ? span : 0u);
#pragma unroll
  for (int p0 = 0; p0 < NP; p0 += K) {
    uint32_t off[K];
    float h[K], l[K], e[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int p = p0 + q;
      if (p >= NP) break;
      const int d = d0 + p / R, r = p % R;
      const int t = tl + 64 * r;
      off[q] = (d < D && t < T) ? (uint32_t)((int64_t)d * dstride + t) * 4u
                                : kNoOffset;
      h[q] = buffer_load(rh, off[q]);
      l[q] = buffer_load(rl, off[q]);
      if (MODE == 2) e[q] = buffer_load(rp, off[q]);
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int p = p0 + q;
      if (p >= NP) break;
      const double a = acc[p / R][p % R];
      double sum = (double)h[q] + (double)l[q];
      if (MODE == 2) sum += (double)e[q];
      sum += a;
      const float hs = (float)sum;
      buffer_store((float)a, rc, off[q]);
      buffer_store(hs, rh, off[q]);
      buffer_store((float)(sum - (double)hs), rl, off[q]);
    }
  }
}

// Spread of the diffuse + specular plan.  Its BRDF rows are the unit
// vector e_m (a specular group, key m < D) and the all-ones row (the
// destination's last group, its diffuse sum), so
//   cur[j, d, b, t] = gsum[diffuse] + gsum[group with key d]
// (just the diffuse sum where no group has key d): the float64 sum that
// group_spread_kernel forms with multiply-adds by ones and zeros, bit for
// bit, without the D multiply-adds per group.  A workgroup takes one
// (destination, band, 256 bins); each lane one bin and all D directions,
// every direction's group-sum load issued before any is used.  MODE as in
// group_spread_kernel.
template <int DP, int MODE>
__global__ __launch_bounds__(256) void ds_spread_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks,
    const int64_t *__restrict__ grp_ptr, const int32_t *__restrict__ grp_key,
    const float *__restrict__ gsum, const float *__restrict__ prev,
    float *__restrict__ cur, float *__restrict__ tot_hi,
    float *__restrict__ tot_lo) {
  constexpr int K = 8;
  __shared__ int32_t gmap[DP];  // direction -> specular group offset, or -1
  const int64_t tile = xcd_tile(blockIdx.x, nblocks);
  const int64_t P = nblocks / (B * n_tblk);
  const int tb = (int)(tile % n_tblk);
  const int64_t j = (tile / n_tblk) % P;
  const int b = (int)(tile / (n_tblk * P));
  const int t = tb * 256 + (int)threadIdx.x;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const int64_t g0 = grp_ptr[j], g1 = grp_ptr[j + 1];
  for (int d = (int)threadIdx.x; d < DP; d += 256) gmap[d] = -1;
  __syncthreads();
  for (int64_t g = g0 + (int64_t)threadIdx.x; g < g1 - 1; g += 256) {
    const int32_t k = grp_key[g];
    if (k >= 0 && k < D) gmap[k] = (int32_t)(g - g0);
  }
  __syncthreads();
  float v[DP];
#pragma unroll
  for (int d = 0; d < DP; ++d) {
    const int32_t gi = __builtin_amdgcn_readfirstlane(gmap[d]);
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(gsum + ((g0 + (gi > 0 ? gi : 0)) * B + b) * (int64_t)T,
                  gi >= 0 ? row_bytes : 0u);
    v[d] = buffer_load(rs, (uint32_t)t * 4u);
  }
  const __amdgpu_buffer_rsrc_t rd =
      make_rsrc(gsum + ((g1 > g0 ? g1 - 1 : g0) * B + b) * (int64_t)T,
                g1 > g0 ? row_bytes : 0u);
  const double gd = (double)buffer_load(rd, (uint32_t)t * 4u);

  const int64_t dstride = (int64_t)B * T;
  const int64_t base = j * D * dstride + (int64_t)b * T;
  const uint32_t span = (uint32_t)(((int64_t)(D - 1) * dstride + T) * 4);
  const __amdgpu_buffer_rsrc_t rc = make_rsrc(cur + base, span);
  auto offset = [&](int d) {
    return (d < D && t < T) ? (uint32_t)((int64_t)d * dstride + t) * 4u
                            : kNoOffset;
  };
  if (MODE == 0) {
#pragma unroll
    for (int d = 0; d < DP; ++d)
      buffer_store((float)((double)v[d] + gd), rc, offset(d));
    return;
  }
  const __amdgpu_buffer_rsrc_t rh = make_rsrc(tot_hi + base, span);
  const __amdgpu_buffer_rsrc_t rl = make_rsrc(tot_lo + base, span);
  const __amdgpu_buffer_rsrc_t rp =
      make_rsrc(MODE == 2 ? prev + base : cur + base, MODE == 2 ? span : 0u);
#pragma unroll
  for (int d0 = 0; d0 < DP; d0 += K) {
    uint32_t off[K];
    float h[K], l[K], e[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (d0 + q >= DP) break;
      off[q] = offset(d0 + q);
      h[q] = buffer_load(rh, off[q]);
      l[q] = buffer_load(rl, off[q]);
      if (MODE == 2) e[q] = buffer_load(rp, off[q]);
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (d0 + q >= DP) break;
      const double a = (double)v[d0 + q] + gd;
      double sum = (double)h[q] + (double)l[q];
      if (MODE == 2) sum += (double)e[q];
      sum += a;
      const float hs = (float)sum;
      buffer_store((float)a, rc, off[q]);
      buffer_store(hs, rh, off[q]);
      buffer_store((float)(sum - (double)hs), rl, off[q]);
    }
  }
}

template <int DP>
int launch_ds_spread(int64_t P, int64_t D, int64_t B, int64_t T,
                     const int64_t *grp_ptr, const int32_t *grp_key,
                     const float *gsum, const float *prev, float *cur,
                     float *tot_hi, float *tot_lo, hipStream_t s) {
  const int64_t nt = ceil_div(T, (int64_t)256);
  const int64_t nb = P * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(D * B * T < ((int64_t)1 << 29),
              "histogram too large for one patch (D*B*T >= 2^29)");
#define SPR_DS_MODE(M)                                                        \
  hipLaunchKernelGGL((ds_spread_kernel<DP, M>), dim3((unsigned)nb),           \
                     dim3(256), 0, s, (int)D, (int)B, (int)T, (int)nt, nb,    \
                     grp_ptr, grp_key, gsum, prev, cur, tot_hi, tot_lo)
  if (!tot_hi)
    SPR_DS_MODE(0);
  else if (!prev)
    SPR_DS_MODE(1);
  else
    SPR_DS_MODE(2);
#undef SPR_DS_MODE
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

template <int DP, int R>
int launch_step(int64_t P, int64_t D, int64_t B, int64_t T,
                const int64_t *row_ptr, const int32_t *src,
                const int32_t *lag, const int32_t *dir, const double *coef,
                const float *prev, float *cur, float *tot_hi, float *tot_lo,
                hipStream_t s) {
  if constexpr (DP >= 16) {
    constexpr int NT = 8 / (DP / 16);
    const int64_t n_tblk = ceil_div(T, (int64_t)4 * 16 * NT);
    const int64_t nblocks = P * B * n_tblk;
    SPR_REQUIRE(nblocks < (int64_t)0x7fffffff, "exchange grid too large");
    hipLaunchKernelGGL((exchange_mfma_kernel<DP, NT>),
                       dim3((unsigned)nblocks), dim3(256), 0, s, (int)D,
                       (int)B, (int)T, (int)n_tblk, nblocks, row_ptr, src,
                       lag, dir, coef, prev, cur, tot_hi, tot_lo);
    SPR_LAUNCH_CHECK();
    return SPR_OK;
  }
  constexpr int BS = 256;
  const int64_t n_tblk = ceil_div(T, (int64_t)BS * R);
  const int64_t nblocks = P * B * n_tblk;
  SPR_REQUIRE(nblocks < (int64_t)0x7fffffff, "exchange grid too large");
  hipLaunchKernelGGL((exchange_step_kernel<DP, R, BS, 4>),
                     dim3((unsigned)nblocks), dim3(BS), 0, s, (int)D, (int)B,
                     (int)T, (int)n_tblk, nblocks, row_ptr, src, lag, dir,
                     coef, prev, cur, tot_hi, tot_lo);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

__global__ void finalize_kernel(int64_t n, const float *__restrict__ hi,
                                const float *__restrict__ lo,
                                float *__restrict__ out) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += stride)
    out[k] = (float)((double)hi[k] + (double)lo[k]);
}

// The grouped step: gather-sum then spread.
constexpr int kGatherR = 2;
constexpr int kGatherC = 8;
// The diffuse + specular gather: 8 bins per lane, 4 entries per chunk.
// Directional shoebox (1,072 patches, D = 64, B = 6, T = 6,000), one order:
// 10.5 ms; 4 bins x 4 entries 11.4 ms, 8 x 2 11.8 ms, 4 x 8 13.0 ms,
// 2 x 8 19.3 ms, 1 x 16 39.7 ms (the same sums bit for bit).
constexpr int kDsGatherR = 8;
constexpr int kDsGatherC = 4;

int launch_gather(int64_t P, int64_t B, int64_t T, int64_t Vd,
                  const int64_t *row_ptr, const int64_t *grp_ptr,
                  const int32_t *rowid, const int32_t *lag,
                  const double *scal, const double *dscal, const float *prev,
                  float *gsum, hipStream_t s) {
  constexpr int BS = 256;
  const int64_t nt =
      ceil_div(T, (int64_t)BS * (dscal ? kDsGatherR : kGatherR));
  const int64_t nb = P * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(T < ((int64_t)1 << 29), "histogram too long (T >= 2^29)");
  if (dscal)
    hipLaunchKernelGGL((ds_gather_kernel<kDsGatherR, BS, kDsGatherC>),
                       dim3((unsigned)nb), dim3(BS), 0, s, (int)B, (int)T,
                       (int)nt, nb, Vd, row_ptr, grp_ptr, rowid, lag, scal,
                       dscal, prev, gsum);
  else
    hipLaunchKernelGGL((group_gather_kernel<kGatherR, BS, kGatherC>),
                       dim3((unsigned)nb), dim3(BS), 0, s, (int)B, (int)T,
                       (int)nt, nb, Vd, row_ptr, grp_ptr, rowid, lag, scal,
                       prev, gsum);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

template <int DP>
int launch_spread(int64_t P, int64_t D, int64_t B, int64_t T,
                  const int64_t *grp_ptr, const int32_t *grp_key,
                  const double *rtab, const float *gsum, const float *prev,
                  float *cur, float *tot_hi, float *tot_lo, hipStream_t s) {
  const int64_t nt = ceil_div(T, (int64_t)SpreadShape<DP>::TILE);
  const int64_t nb = P * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "exchange grid too large");
  SPR_REQUIRE(D * B * T < ((int64_t)1 << 29),
              "histogram too large for one patch (D*B*T >= 2^29)");
#define SPR_SPREAD_MODE(M)                                                    \
  hipLaunchKernelGGL((group_spread_kernel<DP, M>), dim3((unsigned)nb),        \
                     dim3(256), 0, s, (int)D, (int)B, (int)T, (int)nt, nb,    \
                     grp_ptr, grp_key, rtab, gsum, prev, cur, tot_hi, tot_lo)
  if (!tot_hi)
    SPR_SPREAD_MODE(0);
  else if (!prev)
    SPR_SPREAD_MODE(1);
  else
    SPR_SPREAD_MODE(2);
#undef SPR_SPREAD_MODE
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

// tot_hi + tot_lo += x (compensated, in float64).
__global__ void total_add_kernel(int64_t n, const float *__restrict__ x,
                                 float *__restrict__ hi,
                                 float *__restrict__ lo) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += stride) {
    const double sum = ((double)hi[k] + (double)lo[k]) + (double)x[k];
    const float hs = (float)sum;
    hi[k] = hs;
    lo[k] = (float)(sum - (double)hs);
  }
}

// ---------------------------------------------------------------------
// grouped plan construction
// ---------------------------------------------------------------------
constexpr int kSortMax = 8192;  // longest destination row the LDS sort takes

// Key of entry e: its BRDF row, or with spec_dir (diffuse + specular plan)
// the specular direction of that row.
__global__ __launch_bounds__(256) void group_key_kernel(
    int64_t P, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ src, const double *__restrict__ centers,
    const int64_t *__restrict__ wall_ids, int64_t S,
    const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs,
    const int32_t *__restrict__ spec_dir, int32_t *__restrict__ key) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x) {
    const int64_t k = entry_brdf_row(src[e], j, centers, wall_ids, S,
                                     brdf_index, in_dirs);
    key[e] = spec_dir ? spec_dir[k] : (int32_t)k;
  }
}

// Ascending bitonic sort of k[0, np) in LDS by the whole workgroup; np is a
// power of two, the same for every thread, and k is filled and
// synchronised before the call.
__device__ void lds_bitonic_sort(uint64_t *k, int np) {
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = threadIdx.x; q < np; q += blockDim.x) {
        int p = q ^ stride;
        if (p > q) {
          bool up = (q & size) == 0;
          uint64_t a = k[q], c = k[p];
          if ((a > c) == up) {
            k[q] = c;
            k[p] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Sort each destination row by (key, position) with a bitonic sort in LDS
// (stable: ties keep ascending source order) and count its groups.
// Entries whose key is none_key form no group; a non-empty row gets
// `extra` more groups (the diffuse group of the diffuse + specular plan).
__global__ __launch_bounds__(1024) void group_sort_kernel(
    int64_t P, const int64_t *__restrict__ row_ptr, int32_t *__restrict__ src,
    int32_t *__restrict__ key, int64_t *__restrict__ grp_count, int npad,
    int32_t none_key, int extra) {
  // dynamic LDS only (8-byte aligned base): keys, source ids, group count
  extern __shared__ uint64_t smem64[];
  uint64_t *k = smem64;
  int32_t *sv = reinterpret_cast<int32_t *>(k + npad);
  int &groups = sv[npad];
  const int64_t j = blockIdx.x;
  const int64_t e0 = row_ptr[j];
  const int n = (int)(row_ptr[j + 1] - e0);
  for (int q = threadIdx.x; q < npad; q += blockDim.x) {
    if (q < n) {
      k[q] = ((uint64_t)(uint32_t)key[e0 + q] << 32) | (uint32_t)q;
      sv[q] = src[e0 + q];
    } else {
      k[q] = ~0ull;
    }
  }
  if (threadIdx.x == 0) groups = 0;
  __syncthreads();
  lds_bitonic_sort(k, npad);
  int local = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const uint32_t kq = (uint32_t)(k[q] >> 32);
    src[e0 + q] = sv[(uint32_t)(k[q] & 0xffffffffu)];
    key[e0 + q] = (int32_t)kq;
    if ((int32_t)kq != none_key &&
        (q == 0 || (uint32_t)(k[q - 1] >> 32) != kq))
      local += 1;
  }
  atomicAdd(&groups, local);
  __syncthreads();
  if (threadIdx.x == 0) grp_count[j] = groups + (n > 0 ? extra : 0);
}

// Group starts and keys, one wavefront per destination row.
// With diffuse_key >= 0 a non-empty row's last group is its diffuse group
// (no entries of its own: start = row end, key = diffuse_key).
__global__ __launch_bounds__(256) void group_fill_kernel(
    int64_t P, int64_t Vd, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ key, const int64_t *__restrict__ grp_ptr,
    int64_t *__restrict__ grp_start, int32_t *__restrict__ grp_key,
    int32_t none_key, int32_t diffuse_key) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= P) return;
  const int64_t e0 = row_ptr[j], e1 = row_ptr[j + 1];
  int64_t pos = grp_ptr[j];
  for (int64_t base = e0; base < e1; base += 64) {
    const int64_t e = base + lane;
    const bool f = e < e1 && key[e] != none_key &&
                   (e == e0 || key[e] != key[e - 1]);
    const uint64_t m = __ballot(f);
    if (f) {
      const uint64_t lt = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
      const int64_t g = pos + __popcll(lt);
      grp_start[g] = e;
      grp_key[g] = key[e];
    }
    pos += __popcll(m);
  }
  if (diffuse_key >= 0 && e1 > e0 && lane == 0) {
    grp_start[grp_ptr[j + 1] - 1] = e1;
    grp_key[grp_ptr[j + 1] - 1] = diffuse_key;
  }
  if (j == P - 1 && lane == 0) grp_start[grp_ptr[P]] = Vd;
}

// Per-entry delay (with the last-in-group flag), source row and per-band
// scalar F~_ij * exp(-air_b |u|), stored band-major: scal[b * Vd + e].
__global__ __launch_bounds__(256) void group_entry_kernel(
    int64_t P, int64_t D, int64_t B, double c, double dt, int64_t Vd,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const int32_t *__restrict__ key, const double *__restrict__ centers,
    const double *__restrict__ areas, const double *__restrict__ ff,
    const int64_t *__restrict__ p2o, const double *__restrict__ air,
    int32_t *__restrict__ rowid, int32_t *__restrict__ lag,
    double *__restrict__ scal) {
  const int64_t j = blockIdx.x;
  const int64_t e1 = row_ptr[j + 1];
  for (int64_t e = row_ptr[j] + threadIdx.x; e < e1; e += blockDim.x) {
    const int64_t i = src[e];
    const double dij = norm(load3(centers + 3 * i) - load3(centers + 3 * j));
    const bool last = e + 1 == e1 || key[e + 1] != key[e];
    lag[e] = clamp_lag(delay_floor(dij, c, dt)) | (last ? kLastInGroup : 0);
    rowid[e] = (int32_t)(i * D + checked_dir(p2o[i * P + j], D));
    PairFactor pf = pair_factor(i, j, P, centers, areas, ff, nullptr, 1, 1, B,
                                nullptr, nullptr, nullptr);
    for (int64_t b = 0; b < B; ++b) {
      double v = pf.f;
      if (air) v = v * exp(-air[b] * pf.distance);
      scal[b * Vd + e] = v;
    }
  }
}

// rtab[(bi*S + s), b, d] = brdf[bi, s, d, b] (zero-padded to DP).
__global__ void rtab_kernel(int64_t NK, int64_t D, int64_t B, int64_t DP,
                            const double *__restrict__ brdf,
                            double *__restrict__ rtab) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NK * B * DP) return;
  const int64_t d = q % DP, b = (q / DP) % B, k = q / (DP * B);
  double v = 0.0;
  if (d < D) v = brdf ? brdf[(k * D + d) * B + b] : 1.0;
  rtab[q] = v;
}

// ---------------------------------------------------------------------
// grouped plan from the dense F~ (spr_energy_exchange)
// ---------------------------------------------------------------------
// The F~ rows of a destination's entries that share a BRDF row are, band by
// band, multiples of one vector: F~[i, j, :, b] = s_ijb * rho_b.  The rows
// are matched by a hash of their band-wise normalised values (each band
// divided by its first element of largest magnitude, the pivot, and
// rounded to multiples of 2^-30), sorted by it within the destination, and
// every entry is then checked against its group's first entry:
//   |F~_e[d, b] - s_eb F~_first[d, b]| <= 2^-28 max_d |F~_e[d, b]|
// with s_eb = F~_e[piv, b] / F~_first[piv, b] at the first entry's pivot.
// A failed check (a hash collision, or values that are not numbers) sends
// the whole call to the per-pair plan.
constexpr double kRowQuantum = 1073741824.0;                  // 2^30
constexpr double kRowTol = 3.7252902984619140625e-09;         // 2^-28
constexpr uint64_t kHashPosMask = (uint64_t)kSortMax - 1;     // 13 bits
static_assert((kSortMax & (kSortMax - 1)) == 0, "kSortMax: a power of 2");

// splitmix64's finaliser (Steele, Lea and Flood, OOPSLA 2014) on h + x.
__device__ __forceinline__ uint64_t mix64(uint64_t h, uint64_t x) {
  uint64_t z = h + x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Index of the first element of largest magnitude of f[0], f[stride], ...,
// f[(n-1) stride]; -1 when none is nonzero (all 0, or not numbers).
__device__ __forceinline__ int row_pivot(const double *__restrict__ f,
                                         int64_t n, int64_t stride) {
  int piv = -1;
  double m = 0.0;
  for (int64_t d = 0; d < n; ++d) {
    const double a = fabs(f[d * stride]);
    if (a > m) {
      m = a;
      piv = (int)d;
    }
  }
  return piv;
}

// hkey[e] = hash of entry e's normalised F~ row (all bands).
__global__ __launch_bounds__(256) void dense_row_hash_kernel(
    int64_t P, int64_t D, int64_t B, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ src, const double *__restrict__ fft,
    uint64_t *__restrict__ hkey) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x) {
    const double *f = fft + ((int64_t)src[e] * P + j) * D * B;
    uint64_t h = 0x243f6a8885a308d3ull;
    for (int64_t b = 0; b < B; ++b) {
      const int piv = row_pivot(f + b, D, B);
      h = mix64(h, (uint64_t)(int64_t)piv);
      if (piv < 0) continue;
      const double p = f[piv * B + b];
      for (int64_t d = 0; d < D; ++d)
        h = mix64(h, (uint64_t)__double2ll_rn(f[d * B + b] / p * kRowQuantum));
    }
    hkey[e] = h;
  }
}

// Sort each destination row by (hash, position) in LDS -- the position in
// the hash's low 13 bits, so ties keep ascending source order -- and mark
// the group starts: start[e] = 1 where the rest of the hash changes.
// A row longer than kSortMax sets *bad.  LDS: npad >= every row's length
// (capped at kSortMax) keys, source ids and a group count.
__global__ __launch_bounds__(1024) void dense_hash_sort_kernel(
    int64_t P, const int64_t *__restrict__ row_ptr, int32_t *__restrict__ src,
    const uint64_t *__restrict__ hkey, int32_t *__restrict__ start,
    int64_t *__restrict__ grp_count, int32_t *__restrict__ bad) {
  extern __shared__ uint64_t smem64[];
  const int64_t j = blockIdx.x;
  const int64_t e0 = row_ptr[j];
  const int64_t len = row_ptr[j + 1] - e0;
  if (len > kSortMax) {  // uniform
    if (threadIdx.x == 0) {
      grp_count[j] = 0;
      atomicOr(bad, 1);
    }
    return;
  }
  const int n = (int)len;
  int np = 1;
  while (np < n) np <<= 1;
  uint64_t *k = smem64;
  int32_t *sv = reinterpret_cast<int32_t *>(k + np);
  int &groups = sv[np];
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    if (q < n) {
      k[q] = (hkey[e0 + q] & ~kHashPosMask) | (uint64_t)q;
      sv[q] = src[e0 + q];
    } else {
      k[q] = ~0ull;
    }
  }
  if (threadIdx.x == 0) groups = 0;
  __syncthreads();
  lds_bitonic_sort(k, np);
  int local = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    src[e0 + q] = sv[k[q] & kHashPosMask];
    const bool st =
        q == 0 || (k[q - 1] & ~kHashPosMask) != (k[q] & ~kHashPosMask);
    start[e0 + q] = st ? 1 : 0;
    local += st ? 1 : 0;
  }
  atomicAdd(&groups, local);
  __syncthreads();
  if (threadIdx.x == 0) grp_count[j] = groups;
}

// The grouped plan on the sorted rows, one wavefront per destination row:
// group keys (the entry index of each group's first entry, whose F~ row in
// coef is the group's table row), source rows, last-in-group flags and
// the per-band scalars scal[b * Vd + e] = s_eb; an entry that is not s_e
// times its group's first row sets *bad.
__global__ __launch_bounds__(256) void dense_group_fill_kernel(
    int64_t P, int64_t D, int64_t B, int64_t DP, int64_t Vd,
    const int64_t *__restrict__ row_ptr, const int64_t *__restrict__ grp_ptr,
    const int32_t *__restrict__ src, const int32_t *__restrict__ start,
    const int32_t *__restrict__ dir, const double *__restrict__ coef,
    int32_t *__restrict__ grp_key, int32_t *__restrict__ rowid,
    int32_t *__restrict__ lag, double *__restrict__ scal,
    int32_t *__restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= P) return;
  const int64_t e0 = row_ptr[j], e1 = row_ptr[j + 1];
  int64_t g = grp_ptr[j] - 1;  // group of the last entry before the chunk
  int64_t first = e0;          // and its first entry
  bool ok = true;
  const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
  for (int64_t base = e0; base < e1; base += 64) {
    const int64_t e = base + lane;
    const bool in = e < e1;
    const bool st = in && start[e] != 0;
    const uint64_t m = __ballot(st);
    const uint64_t ml = m & upto;
    const int64_t ge = g + __popcll(ml);
    const int64_t fe = ml ? base + 63 - __clzll(ml) : first;
    if (m) first = base + 63 - __clzll(m);
    g += __popcll(m);
    if (!in) continue;
    if (st) grp_key[ge] = (int32_t)e;
    const bool last = e + 1 == e1 || start[e + 1] != 0;
    lag[e] |= last ? kLastInGroup : 0;
    rowid[e] = (int32_t)((int64_t)src[e] * D + dir[e]);
    for (int64_t b = 0; b < B; ++b) {
      const double *r = coef + (fe * B + b) * DP;
      const double *f = coef + (e * B + b) * DP;
      const int piv = row_pivot(r, D, 1);
      const double s = piv >= 0 ? f[piv] / r[piv] : 0.0;
      double mx = 0.0;
      for (int64_t d = 0; d < D; ++d) mx = fmax(mx, fabs(f[d]));
      for (int64_t d = 0; d < D; ++d)
        if (!(fabs(f[d] - s * r[d]) <= kRowTol * mx)) ok = false;
      scal[b * Vd + e] = s;
    }
  }
  if (!ok) atomicOr(bad, 1);
}

// Diffuse + specular decomposition of BRDF rows: row k (brdf[k, :, b]) is
// a_kb on every direction except at most one, spec_dir[k] = m, where it is
// a_kb + c_kb (the form create_from_scattering produces: s/pi everywhere
// plus the specular lobe in one direction, times 1 - alpha).  m is the
// first maximum; rows whose specular part vanishes in every band get
// m = D (no specular group).  *ok is cleared if any row has another form.
__global__ void ds_table_kernel(int64_t NK, int64_t D, int64_t B,
                                const double *__restrict__ brdf,
                                double *__restrict__ a, double *__restrict__ c,
                                int32_t *__restrict__ spec_dir,
                                int32_t *__restrict__ ok) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NK) return;
  if (!brdf) {
    for (int64_t b = 0; b < B; ++b) {
      a[k * B + b] = 1.0;
      c[k * B + b] = 0.0;
    }
    spec_dir[k] = (int32_t)D;
    return;
  }
  const double *row = brdf + k * D * B;
  int64_t m = D;
  bool good = true;
  for (int64_t b = 0; b < B; ++b) {
    int64_t mb = 0;
    for (int64_t d = 1; d < D; ++d)
      if (row[d * B + b] > row[mb * B + b]) mb = d;
    const double base = row[(mb == 0 && D > 1 ? 1 : 0) * B + b];
    for (int64_t d = 0; d < D; ++d)
      if (d != mb && row[d * B + b] != base) good = false;
    const double spec = D > 1 ? row[mb * B + b] - base : 0.0;
    a[k * B + b] = D > 1 ? base : row[b];
    c[k * B + b] = spec;
    if (spec != 0.0) {
      if (m != D && m != mb) good = false;
      m = mb;
    }
  }
  spec_dir[k] = (int32_t)m;
  if (!good) atomicMin(ok, 0);
}

// Entries of the diffuse + specular plan: delay (flagged at the end of a
// specular group), source row, and per band the specular weight
// scal = F~_ij exp(-air_b |u|) c_kb and the diffuse weight ... a_kb.
__global__ __launch_bounds__(256) void ds_entry_kernel(
    int64_t P, int64_t D, int64_t B, double c, double dt, int64_t Vd,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const int32_t *__restrict__ key, const double *__restrict__ centers,
    const double *__restrict__ areas, const double *__restrict__ ff,
    const int64_t *__restrict__ p2o, const double *__restrict__ air,
    const int64_t *__restrict__ wall_ids, int64_t S,
    const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs, const double *__restrict__ ta,
    const double *__restrict__ tc, int32_t *__restrict__ rowid,
    int32_t *__restrict__ lag, double *__restrict__ scal,
    double *__restrict__ dscal, int32_t *__restrict__ krow,
    double *__restrict__ wgt) {
  const int64_t j = blockIdx.x;
  const int64_t e1 = row_ptr[j + 1];
  for (int64_t e = row_ptr[j] + threadIdx.x; e < e1; e += blockDim.x) {
    const int64_t i = src[e];
    const double dij = norm(load3(centers + 3 * i) - load3(centers + 3 * j));
    const bool last = key[e] != (int32_t)D &&
                      (e + 1 == e1 || key[e + 1] != key[e]);
    lag[e] = clamp_lag(delay_floor(dij, c, dt)) | (last ? kLastInGroup : 0);
    rowid[e] = (int32_t)(i * D + checked_dir(p2o[i * P + j], D));
    const int64_t k = entry_brdf_row(i, j, centers, wall_ids, S, brdf_index,
                                     in_dirs);
    PairFactor pf = pair_factor(i, j, P, centers, areas, ff, nullptr, 1, 1, B,
                                nullptr, nullptr, nullptr);
    if (krow) krow[e] = (int32_t)k;
    for (int64_t b = 0; b < B; ++b) {
      double v = pf.f;
      if (air) v = v * exp(-air[b] * pf.distance);
      scal[b * Vd + e] = v * tc[k * B + b];
      dscal[b * Vd + e] = v * ta[k * B + b];
      if (wgt) wgt[b * Vd + e] = v;
    }
  }
}

// rtab of the diffuse + specular plan: row m < D the unit vector e_m,
// row D zeros, row D + 1 ones (padding zero).
__global__ void ds_rtab_kernel(int64_t D, int64_t B, int64_t DP,
                               double *__restrict__ rtab) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (D + 2) * B * DP) return;
  const int64_t d = q % DP, k = q / (DP * B);
  double v = 0.0;
  if (d < D) v = (k == D + 1) ? 1.0 : (k == d ? 1.0 : 0.0);
  rtab[q] = v;
}

// ---------------------------------------------------------------------
// transposed (source-major) view of a diffuse + specular plan, for the
// adjoint: row i lists the entries e = (i -> j) sorted by (direction of
// the source row, destination)
// ---------------------------------------------------------------------
constexpr int kTransposeDstBits = 25;  // destination index bits of a key

__global__ __launch_bounds__(256) void t_count_kernel(
    int64_t P, int64_t D, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ rowid, unsigned long long *__restrict__ cnt) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x)
    atomicAdd(&cnt[rowid[e] / D], 1ull);
}

__global__ __launch_bounds__(256) void t_scatter_kernel(
    int64_t P, int64_t D, const int64_t *__restrict__ row_ptr,
    const int32_t *__restrict__ rowid, const int64_t *__restrict__ trow_ptr,
    unsigned long long *__restrict__ cursor, int32_t *__restrict__ te,
    int32_t *__restrict__ tkey) {
  const int64_t j = blockIdx.x;
  for (int64_t e = row_ptr[j] + threadIdx.x; e < row_ptr[j + 1];
       e += blockDim.x) {
    const int64_t i = rowid[e] / D, d = rowid[e] % D;
    const int64_t q = trow_ptr[i] + (int64_t)atomicAdd(&cursor[i], 1ull);
    te[q] = (int32_t)e;
    tkey[q] = (int32_t)((d << kTransposeDstBits) | j);
  }
}

// After the per-row sort by key: destination, delay (bit 30 set on the last
// entry of each direction group), specular direction (0 with a zero
// weight when the entry has none), and the group pointers
// tgptr[i*D + d] = first transposed entry of source i with direction >= d.
__global__ __launch_bounds__(256) void t_fill_kernel(
    int64_t P, int64_t D, const int64_t *__restrict__ trow_ptr,
    const int32_t *__restrict__ te, const int32_t *__restrict__ tkey,
    const int32_t *__restrict__ lag, const int32_t *__restrict__ key,
    int32_t *__restrict__ tj, int32_t *__restrict__ tlag,
    int32_t *__restrict__ tm, int64_t *__restrict__ tgptr) {
  const int64_t i = blockIdx.x;
  const int64_t q0 = trow_ptr[i], q1 = trow_ptr[i + 1];
  const int32_t dmask = (1 << kTransposeDstBits) - 1;
  for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
    const int32_t k = tkey[q];
    const int32_t d = k >> kTransposeDstBits;
    const bool last = q + 1 == q1 || (tkey[q + 1] >> kTransposeDstBits) != d;
    const int32_t e = te[q];
    tj[q] = k & dmask;
    tlag[q] = (lag[e] & kLagMask) | (last ? kLastInGroup : 0);
    tm[q] = key[e] < (int32_t)D ? key[e] : 0;
  }
  for (int64_t d = threadIdx.x; d <= D; d += blockDim.x) {
    if (d == D) {
      if (i == P - 1) tgptr[P * D] = q1;
      continue;
    }
    int64_t lo = q0, hi = q1;  // first q with direction >= d
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if ((tkey[mid] >> kTransposeDstBits) < d)
        lo = mid + 1;
      else
        hi = mid;
    }
    tgptr[i * D + d] = lo;
  }
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace spr

using namespace spr;

extern "C" int64_t spr_padded_directions(int64_t D) {
  if (D < 1) return 0;
  int64_t p = 1;
  while (p < D) p <<= 1;
  return p <= 64 ? p : 0;
}

extern "C" int spr_plan_rows(int64_t P, const uint8_t *vis, int64_t *row_ptr,
                             int32_t *src, void *stream) {
  SPR_REQUIRE(P > 0 && vis && row_ptr, "spr_plan_rows: bad arguments");
  SPR_REQUIRE(P < 0x7fffffff, "spr_plan_rows: too many patches");
  hipStream_t s = as_stream(stream);
  unsigned blocks = (unsigned)ceil_div(P, 4);
  if (!src) {
    hipLaunchKernelGGL(plan_rows_kernel, dim3(blocks), dim3(256), 0, s, P,
                       vis, row_ptr, (int32_t *)nullptr, 0);
    SPR_LAUNCH_CHECK();
    return scan_exclusive(row_ptr, P, s);
  }
  hipLaunchKernelGGL(plan_rows_kernel, dim3(blocks), dim3(256), 0, s, P, vis,
                     row_ptr, src, 1);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_plan_from_dense(int64_t P, int64_t D, int64_t B, double c,
                                   double dt, int64_t Vd,
                                   const int64_t *row_ptr, const int32_t *src,
                                   const double *dist, const int64_t *p2o,
                                   const double *fft, int32_t *lag,
                                   int32_t *dir, double *coef, void *stream) {
  int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_plan_from_dense: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && c > 0 && dt > 0,
              "spr_plan_from_dense: bad sizes");
  SPR_REQUIRE(row_ptr && src && dist && p2o && fft && lag && dir && coef,
              "spr_plan_from_dense: null pointer");
  if (Vd == 0) return SPR_OK;
  hipLaunchKernelGGL(plan_from_dense_kernel, dim3((unsigned)P), dim3(256), 0,
                     as_stream(stream), P, D, B, DP, c, dt, row_ptr, src,
                     dist, p2o, fft, lag, dir, coef);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_plan_from_geometry(
    int64_t P, int64_t D, int64_t B, double c, double dt, int64_t Vd,
    const int64_t *row_ptr, const int32_t *src, const double *centers,
    const double *areas, const double *ff, const int64_t *p2o,
    const double *air, const int64_t *wall_ids, int64_t S,
    const double *brdf, const int64_t *brdf_index, const double *in_dirs,
    int32_t *lag, int32_t *dir, double *coef, void *stream) {
  int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_plan_from_geometry: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && c > 0 && dt > 0,
              "spr_plan_from_geometry: bad sizes");
  SPR_REQUIRE(row_ptr && src && centers && areas && ff && p2o && wall_ids &&
                  lag && dir && coef,
              "spr_plan_from_geometry: null pointer");
  SPR_REQUIRE(!brdf || (brdf_index && in_dirs && S > 0),
              "spr_plan_from_geometry: brdf needs brdf_index and in_dirs");
  SPR_REQUIRE(brdf || D == 1, "spr_plan_from_geometry: D must be 1 "
                              "without a BRDF");
  if (Vd == 0) return SPR_OK;
  hipLaunchKernelGGL(plan_from_geometry_kernel, dim3((unsigned)P), dim3(256),
                     0, as_stream(stream), P, D, B, DP, c, dt, row_ptr, src,
                     centers, areas, ff, p2o, air, wall_ids, S, brdf,
                     brdf_index, in_dirs, lag, dir, coef);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_init(int64_t P, int64_t D, int64_t B, int64_t T,
                                 const double *e0dir, const double *d0,
                                 double c, double dt, float *prev,
                                 float *tot_hi, float *tot_lo, int *err_flag,
                                 void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0 && T > 0 && c > 0 && dt > 0,
              "spr_exchange_init: bad sizes");
  SPR_REQUIRE(e0dir && d0 && prev && tot_hi && tot_lo && err_flag,
              "spr_exchange_init: null pointer");
  hipStream_t s = as_stream(stream);
  const int64_t n = P * D * B * T;
  int rc;
  if ((rc = fill<float>(prev, n, 0.f, s))) return rc;
  if ((rc = fill<float>(tot_hi, n, 0.f, s))) return rc;
  if ((rc = fill<float>(tot_lo, n, 0.f, s))) return rc;
  const int64_t q = P * D * B;
  hipLaunchKernelGGL(init_kernel, dim3((unsigned)ceil_div(q, 256)), dim3(256),
                     0, s, P, D, B, T, e0dir, d0, c, dt, prev, tot_hi, tot_lo,
                     err_flag);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_step(int64_t P, int64_t D, int64_t B, int64_t T,
                                 const int64_t *row_ptr, const int32_t *src,
                                 const int32_t *lag, const int32_t *dir,
                                 const double *coef, const float *prev,
                                 float *cur, float *tot_hi, float *tot_lo,
                                 void *stream) {
  int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_exchange_step: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && T < 0x3fffffff,
              "spr_exchange_step: bad sizes");
  SPR_REQUIRE(row_ptr && src && lag && dir && coef && prev && cur && tot_hi &&
                  tot_lo,
              "spr_exchange_step: null pointer");
  SPR_REQUIRE(prev != cur, "spr_exchange_step: prev and cur must differ");
  hipStream_t s = as_stream(stream);
#define SPR_STEP(DPV, RV)                                                   \
  return launch_step<DPV, RV>(P, D, B, T, row_ptr, src, lag, dir, coef,     \
                              prev, cur, tot_hi, tot_lo, s)
  switch (DP) {
    case 1: SPR_STEP(1, 8);
    case 2: SPR_STEP(2, 8);
    case 4: SPR_STEP(4, 8);
    case 8: SPR_STEP(8, 8);
    case 16: SPR_STEP(16, 4);
    case 32: SPR_STEP(32, 2);
    case 64: SPR_STEP(64, 1);
  }
#undef SPR_STEP
  return fail(SPR_ERR_UNSUPPORTED, "spr_exchange_step: D unsupported");
}

extern "C" int spr_exchange_finalize(int64_t n, const float *tot_hi,
                                     const float *tot_lo, float *etc,
                                     void *stream) {
  SPR_REQUIRE(n >= 0 && tot_hi && tot_lo && etc,
              "spr_exchange_finalize: bad arguments");
  if (n == 0) return SPR_OK;
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     as_stream(stream), n, tot_hi, tot_lo, etc);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

namespace spr {
namespace {
// Keys, per-row stable sort by key and exclusive scan of the group counts.
int sort_rows(int64_t P, int64_t max_row, const int64_t *row_ptr,
              int32_t *src, const double *centers, const int64_t *wall_ids,
              int64_t S, const int64_t *brdf_index, const double *in_dirs,
              const int32_t *spec_dir, int32_t none_key, int extra,
              int32_t *key, int64_t *grp_ptr, hipStream_t s) {
  hipLaunchKernelGGL(group_key_kernel, dim3((unsigned)P), dim3(256), 0, s, P,
                     row_ptr, src, centers, wall_ids, S, brdf_index, in_dirs,
                     spec_dir, key);
  SPR_LAUNCH_CHECK();
  int npad = 1;
  while (npad < max_row) npad <<= 1;
  const size_t lds =
      (size_t)npad * (sizeof(uint64_t) + sizeof(int32_t)) + sizeof(int32_t);
  SPR_HIP(hipFuncSetAttribute((const void *)group_sort_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  hipLaunchKernelGGL(group_sort_kernel, dim3((unsigned)P), dim3(1024), lds, s,
                     P, row_ptr, src, key, grp_ptr, npad, none_key, extra);
  SPR_LAUNCH_CHECK();
  return scan_exclusive(grp_ptr, P, s);
}
}  // namespace
}  // namespace spr

extern "C" int spr_plan_grouped_sort(
    int64_t P, int64_t max_row, const int64_t *row_ptr, int32_t *src,
    const double *centers, const int64_t *wall_ids, int64_t S,
    const int64_t *brdf_index, const double *in_dirs, int32_t *key,
    int64_t *grp_ptr, void *stream) {
  SPR_REQUIRE(P > 0 && max_row >= 0, "spr_plan_grouped_sort: bad sizes");
  SPR_REQUIRE(row_ptr && src && centers && key && grp_ptr,
              "spr_plan_grouped_sort: null pointer");
  SPR_REQUIRE(!in_dirs || (wall_ids && brdf_index && S > 0),
              "spr_plan_grouped_sort: in_dirs needs wall_ids, brdf_index");
  if (max_row > kSortMax)
    return fail(SPR_ERR_UNSUPPORTED,
                "spr_plan_grouped_sort: a destination sees %lld patches, "
                "the grouped plan supports %d",
                (long long)max_row, kSortMax);
  hipStream_t s = as_stream(stream);
  return sort_rows(P, max_row, row_ptr, src, centers, wall_ids, S,
                   brdf_index, in_dirs, nullptr, -1, 0, key, grp_ptr, s);
}

extern "C" int spr_plan_grouped_fill(
    int64_t P, int64_t D, int64_t B, double c, double dt, int64_t Vd,
    const int64_t *row_ptr, const int32_t *src, const int32_t *key,
    const int64_t *grp_ptr, const double *centers, const double *areas,
    const double *ff, const int64_t *p2o, const double *air, int64_t NK,
    const double *brdf, int64_t *grp_start, int32_t *grp_key, int32_t *rowid,
    int32_t *lag, double *scal, double *rtab, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_plan_grouped_fill: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && NK > 0 && c > 0 && dt > 0 && Vd >= 0,
              "spr_plan_grouped_fill: bad sizes");
  SPR_REQUIRE(row_ptr && src && key && grp_ptr && centers && areas && ff &&
                  p2o && grp_start && grp_key && rowid && lag && scal && rtab,
              "spr_plan_grouped_fill: null pointer");
  SPR_REQUIRE(brdf || (D == 1 && NK == 1),
              "spr_plan_grouped_fill: without a BRDF D and NK must be 1");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(group_fill_kernel, dim3((unsigned)ceil_div(P, 4)),
                     dim3(256), 0, s, P, Vd, row_ptr, key, grp_ptr, grp_start,
                     grp_key, (int32_t)-1, (int32_t)-1);
  SPR_LAUNCH_CHECK();
  hipLaunchKernelGGL(group_entry_kernel, dim3((unsigned)P), dim3(256), 0, s,
                     P, D, B, c, dt, Vd, row_ptr, src, key, centers, areas, ff,
                     p2o, air, rowid, lag, scal);
  SPR_LAUNCH_CHECK();
  const int64_t nr = NK * B * DP;
  hipLaunchKernelGGL(rtab_kernel, dim3((unsigned)ceil_div(nr, 256)),
                     dim3(256), 0, s, NK, D, B, DP, brdf, rtab);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_plan_dense_hash(int64_t P, int64_t D, int64_t B,
                                   const int64_t *row_ptr, const int32_t *src,
                                   const double *fft, uint64_t *hkey,
                                   void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0, "spr_plan_dense_hash: bad sizes");
  SPR_REQUIRE(row_ptr && src && fft && hkey,
              "spr_plan_dense_hash: null pointer");
  hipLaunchKernelGGL(dense_row_hash_kernel, dim3((unsigned)P), dim3(256), 0,
                     as_stream(stream), P, D, B, row_ptr, src, fft, hkey);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_brdf_diffuse_specular(int64_t NK, int64_t D, int64_t B,
                                         const double *brdf, double *a,
                                         double *c, int32_t *spec_dir,
                                         int32_t *ok, void *stream) {
  SPR_REQUIRE(NK > 0 && D > 0 && B > 0, "spr_brdf_diffuse_specular: bad sizes");
  SPR_REQUIRE(a && c && spec_dir && ok,
              "spr_brdf_diffuse_specular: null pointer");
  SPR_REQUIRE(brdf || D == 1,
              "spr_brdf_diffuse_specular: D must be 1 without a BRDF");
  hipStream_t s = as_stream(stream);
  int rc = fill<int32_t>(ok, 1, 1, s);
  if (rc) return rc;
  hipLaunchKernelGGL(ds_table_kernel, dim3((unsigned)ceil_div(NK, 256)),
                     dim3(256), 0, s, NK, D, B, brdf, a, c, spec_dir, ok);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_plan_ds_sort(int64_t P, int64_t D, int64_t max_row,
                                const int64_t *row_ptr, int32_t *src,
                                const double *centers,
                                const int64_t *wall_ids, int64_t S,
                                const int64_t *brdf_index,
                                const double *in_dirs,
                                const int32_t *spec_dir, int32_t *key,
                                int64_t *grp_ptr, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && max_row >= 0, "spr_plan_ds_sort: bad sizes");
  SPR_REQUIRE(row_ptr && src && centers && key && grp_ptr && spec_dir,
              "spr_plan_ds_sort: null pointer");
  SPR_REQUIRE(!in_dirs || (wall_ids && brdf_index && S > 0),
              "spr_plan_ds_sort: in_dirs needs wall_ids, brdf_index");
  if (max_row > kSortMax)
    return fail(SPR_ERR_UNSUPPORTED,
                "spr_plan_ds_sort: a destination sees %lld patches, "
                "the plan supports %d",
                (long long)max_row, kSortMax);
  return sort_rows(P, max_row, row_ptr, src, centers, wall_ids, S,
                   brdf_index, in_dirs, spec_dir, (int32_t)D, 1, key, grp_ptr,
                   as_stream(stream));
}

extern "C" int spr_plan_ds_fill(
    int64_t P, int64_t D, int64_t B, double c, double dt, int64_t Vd,
    const int64_t *row_ptr, const int32_t *src, const int32_t *key,
    const int64_t *grp_ptr, const double *centers, const double *areas,
    const double *ff, const int64_t *p2o, const double *air,
    const int64_t *wall_ids, int64_t S, const int64_t *brdf_index,
    const double *in_dirs, const double *ta, const double *tc,
    int64_t *grp_start, int32_t *grp_key, int32_t *rowid, int32_t *lag,
    double *scal, double *dscal, double *rtab, int32_t *krow, double *wgt,
    void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_plan_ds_fill: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && c > 0 && dt > 0 && Vd >= 0,
              "spr_plan_ds_fill: bad sizes");
  SPR_REQUIRE(row_ptr && src && key && grp_ptr && centers && areas && ff &&
                  p2o && ta && tc && grp_start && grp_key && rowid && lag &&
                  scal && dscal && rtab,
              "spr_plan_ds_fill: null pointer");
  SPR_REQUIRE(!in_dirs || (wall_ids && brdf_index && S > 0),
              "spr_plan_ds_fill: in_dirs needs wall_ids, brdf_index");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(group_fill_kernel, dim3((unsigned)ceil_div(P, 4)),
                     dim3(256), 0, s, P, Vd, row_ptr, key, grp_ptr, grp_start,
                     grp_key, (int32_t)D, (int32_t)(D + 1));
  SPR_LAUNCH_CHECK();
  if (Vd > 0) {
    hipLaunchKernelGGL(ds_entry_kernel, dim3((unsigned)P), dim3(256), 0, s, P,
                       D, B, c, dt, Vd, row_ptr, src, key, centers, areas, ff,
                       p2o, air, wall_ids, S, brdf_index, in_dirs, ta, tc,
                       rowid, lag, scal, dscal, krow, wgt);
    SPR_LAUNCH_CHECK();
  }
  const int64_t nr = (D + 2) * B * DP;
  hipLaunchKernelGGL(ds_rtab_kernel, dim3((unsigned)ceil_div(nr, 256)),
                     dim3(256), 0, s, D, B, DP, rtab);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_gather_ds(
    int64_t P, int64_t B, int64_t T, int64_t Vd, const int64_t *row_ptr,
    const int64_t *grp_ptr, const int32_t *rowid, const int32_t *lag,
    const double *scal, const double *dscal, const float *prev, float *gsum,
    void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && Vd >= 0,
              "spr_exchange_gather_ds: bad sizes");
  SPR_REQUIRE(row_ptr && grp_ptr && rowid && lag && scal && dscal && prev &&
                  gsum,
              "spr_exchange_gather_ds: null pointer");
  return launch_gather(P, B, T, Vd, row_ptr, grp_ptr, rowid, lag, scal,
                       dscal, prev, gsum, as_stream(stream));
}

extern "C" int spr_plan_ds_transpose_rows(int64_t P, int64_t D,
                                          const int64_t *row_ptr,
                                          const int32_t *rowid,
                                          int64_t *trow_ptr, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0, "spr_plan_ds_transpose_rows: bad sizes");
  SPR_REQUIRE(row_ptr && rowid && trow_ptr,
              "spr_plan_ds_transpose_rows: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = fill<int64_t>(trow_ptr, P + 1, 0, s);
  if (rc) return rc;
  hipLaunchKernelGGL(t_count_kernel, dim3((unsigned)P), dim3(256), 0, s, P, D,
                     row_ptr, rowid,
                     reinterpret_cast<unsigned long long *>(trow_ptr));
  SPR_LAUNCH_CHECK();
  return scan_exclusive(trow_ptr, P, s);
}

extern "C" int spr_plan_ds_transpose_fill(
    int64_t P, int64_t D, int64_t Vd, int64_t max_trow,
    const int64_t *row_ptr, const int32_t *rowid, const int32_t *lag,
    const int32_t *key, const int64_t *trow_ptr, int64_t *cursor,
    int32_t *te, int32_t *tkey, int32_t *tj, int32_t *tlag, int32_t *tm,
    int64_t *tgptr, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && Vd >= 0 && max_trow >= 0,
              "spr_plan_ds_transpose_fill: bad sizes");
  SPR_REQUIRE(P < ((int64_t)1 << kTransposeDstBits) && D <= 64,
              "spr_plan_ds_transpose_fill: too many patches or directions");
  SPR_REQUIRE(row_ptr && rowid && lag && key && trow_ptr && cursor && te &&
                  tkey && tj && tlag && tm && tgptr,
              "spr_plan_ds_transpose_fill: null pointer");
  if (max_trow > kSortMax)
    return fail(SPR_ERR_UNSUPPORTED,
                "spr_plan_ds_transpose_fill: a source sees %lld patches, "
                "the plan supports %d",
                (long long)max_trow, kSortMax);
  hipStream_t s = as_stream(stream);
  int rc = fill<int64_t>(cursor, P, 0, s);
  if (rc) return rc;
  if (Vd > 0) {
    hipLaunchKernelGGL(t_scatter_kernel, dim3((unsigned)P), dim3(256), 0, s,
                       P, D, row_ptr, rowid, trow_ptr,
                       reinterpret_cast<unsigned long long *>(cursor), te,
                       tkey);
    SPR_LAUNCH_CHECK();
    int npad = 1;
    while (npad < max_trow) npad <<= 1;
    const size_t lds =
        (size_t)npad * (sizeof(uint64_t) + sizeof(int32_t)) + sizeof(int32_t);
    SPR_HIP(hipFuncSetAttribute((const void *)group_sort_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    // sorts te by tkey within each transposed row (cursor: group counts,
    // not used)
    hipLaunchKernelGGL(group_sort_kernel, dim3((unsigned)P), dim3(1024), lds,
                       s, P, trow_ptr, te, tkey, cursor, npad, (int32_t)-1, 0);
    SPR_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(t_fill_kernel, dim3((unsigned)P), dim3(256), 0, s, P, D,
                     trow_ptr, te, tkey, lag, key, tj, tlag, tm, tgptr);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_gather_grouped(
    int64_t P, int64_t B, int64_t T, int64_t Vd, const int64_t *row_ptr,
    const int64_t *grp_ptr, const int32_t *rowid, const int32_t *lag,
    const double *scal, const float *prev, float *gsum, void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && T < 0x3fffffff && Vd >= 0,
              "spr_exchange_gather_grouped: bad sizes");
  SPR_REQUIRE(row_ptr && grp_ptr && rowid && lag && scal && prev && gsum,
              "spr_exchange_gather_grouped: null pointer");
  return launch_gather(P, B, T, Vd, row_ptr, grp_ptr, rowid, lag, scal,
                       nullptr, prev, gsum, as_stream(stream));
}

namespace {
int spread_grouped(const char *name, int64_t P, int64_t D, int64_t B,
                   int64_t T, const int64_t *grp_ptr, const int32_t *grp_key,
                   const double *rtab, const float *gsum, const float *prev,
                   float *cur, float *tot_hi, float *tot_lo, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "%s: D=%lld unsupported", name, (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && T < 0x3fffffff, "%s: bad sizes",
              name);
  SPR_REQUIRE(grp_ptr && grp_key && rtab && gsum && cur,
              "%s: null pointer", name);
  SPR_REQUIRE(!tot_hi == !tot_lo, "%s: tot_hi and tot_lo go together",
              name);
  SPR_REQUIRE(!prev || (tot_hi && prev != cur),
              "%s: prev needs the total and a separate buffer", name);
  hipStream_t s = as_stream(stream);
#define SPR_SPREAD(DPV)                                                       \
  return launch_spread<DPV>(P, D, B, T, grp_ptr, grp_key, rtab, gsum, prev,   \
                            cur, tot_hi, tot_lo, s)
  switch (DP) {
    case 1: SPR_SPREAD(1);
    case 2: SPR_SPREAD(2);
    case 4: SPR_SPREAD(4);
    case 8: SPR_SPREAD(8);
    case 16: SPR_SPREAD(16);
    case 32: SPR_SPREAD(32);
    case 64: SPR_SPREAD(64);
  }
#undef SPR_SPREAD
  return fail(SPR_ERR_UNSUPPORTED, "%s: bad D", name);
}
}  // namespace

extern "C" int spr_exchange_spread_grouped(
    int64_t P, int64_t D, int64_t B, int64_t T, const int64_t *grp_ptr,
    const int32_t *grp_key, const double *rtab, const float *gsum,
    float *cur, float *tot_hi, float *tot_lo, void *stream) {
  return spread_grouped("spr_exchange_spread_grouped", P, D, B, T, grp_ptr,
                        grp_key, rtab, gsum, nullptr, cur, tot_hi, tot_lo,
                        stream);
}

extern "C" int spr_exchange_spread_grouped_pair(
    int64_t P, int64_t D, int64_t B, int64_t T, const int64_t *grp_ptr,
    const int32_t *grp_key, const double *rtab, const float *gsum,
    const float *prev, float *cur, float *tot_hi, float *tot_lo,
    void *stream) {
  SPR_REQUIRE(prev && tot_hi && tot_lo,
              "spr_exchange_spread_grouped_pair: null pointer");
  return spread_grouped("spr_exchange_spread_grouped_pair", P, D, B, T,
                        grp_ptr, grp_key, rtab, gsum, prev, cur, tot_hi,
                        tot_lo, stream);
}

extern "C" int spr_exchange_spread_ds(int64_t P, int64_t D, int64_t B,
                                      int64_t T, const int64_t *grp_ptr,
                                      const int32_t *grp_key,
                                      const float *gsum, const float *prev,
                                      float *cur, float *tot_hi,
                                      float *tot_lo, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(DP > 0, "spr_exchange_spread_ds: D=%lld unsupported",
              (long long)D);
  SPR_REQUIRE(P > 0 && B > 0 && T > 0 && T < 0x3fffffff,
              "spr_exchange_spread_ds: bad sizes");
  SPR_REQUIRE(grp_ptr && grp_key && gsum && cur,
              "spr_exchange_spread_ds: null pointer");
  SPR_REQUIRE(!tot_hi == !tot_lo,
              "spr_exchange_spread_ds: tot_hi and tot_lo go together");
  SPR_REQUIRE(!prev || (tot_hi && prev != cur),
              "spr_exchange_spread_ds: prev needs the total and a separate "
              "buffer");
  hipStream_t s = as_stream(stream);
#define SPR_DSS(DPV)                                                          \
  return launch_ds_spread<DPV>(P, D, B, T, grp_ptr, grp_key, gsum, prev, cur, \
                               tot_hi, tot_lo, s)
  switch (DP) {
    case 1: SPR_DSS(1);
    case 2: SPR_DSS(2);
    case 4: SPR_DSS(4);
    case 8: SPR_DSS(8);
    case 16: SPR_DSS(16);
    case 32: SPR_DSS(32);
    case 64: SPR_DSS(64);
  }
#undef SPR_DSS
  return fail(SPR_ERR_UNSUPPORTED, "spr_exchange_spread_ds: bad D");
}

extern "C" int spr_exchange_total_add(int64_t n, const float *x,
                                      float *tot_hi, float *tot_lo,
                                      void *stream) {
  SPR_REQUIRE(n >= 0, "spr_exchange_total_add: bad size");
  if (n == 0) return SPR_OK;
  SPR_REQUIRE(x && tot_hi && tot_lo, "spr_exchange_total_add: null pointer");
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(total_add_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     as_stream(stream), n, x, tot_hi, tot_lo);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_step_grouped(
    int64_t P, int64_t D, int64_t B, int64_t T, int64_t Vd,
    const int64_t *row_ptr, const int64_t *grp_ptr, const int32_t *grp_key,
    const int32_t *rowid, const int32_t *lag, const double *scal,
    const double *rtab, float *gsum, const float *prev, float *cur,
    float *tot_hi, float *tot_lo, void *stream) {
  SPR_REQUIRE(prev != cur, "spr_exchange_step_grouped: prev == cur");
  int rc = spr_exchange_gather_grouped(P, B, T, Vd, row_ptr, grp_ptr, rowid,
                                       lag, scal, prev, gsum, stream);
  if (rc) return rc;
  return spr_exchange_spread_grouped(P, D, B, T, grp_ptr, grp_key, rtab,
                                     gsum, cur, tot_hi, tot_lo, stream);
}

namespace {
// Workspace of spr_energy_exchange.  Besides the per-pair plan and the
// three state buffers it holds the grouped plan built from the dense F~
// (hashes alias coef, which is filled after the sort) with group sums for
// at most gcap = min(ceil(Vd / 4), 2 P D) groups: beyond a quarter of the
// entries the per-pair plan is used, and the group sums never take more
// than two state buffers.
struct DenseLayout {
  size_t vis, row_ptr, src, lag, dir, coef, prev, cur, lo, err, start, rowid,
      grp_ptr, grp_key, scal, gsum, bad, total;
  int64_t gcap;
};

DenseLayout dense_layout(int64_t P, int64_t D, int64_t B, int64_t T,
                         int64_t V) {
  DenseLayout L;
  int64_t DP = spr_padded_directions(D);
  int64_t Vd = 2 * V;
  L.gcap = std::min((Vd + 3) / 4, 2 * P * D);
  size_t n_state = (size_t)P * D * B * T;
  size_t off = 0;
  auto take = [&](size_t &at, size_t bytes) {
    at = off;
    off = align256(off + bytes);
  };
  take(L.vis, (size_t)P * P);
  take(L.row_ptr, sizeof(int64_t) * (P + 1));
  take(L.src, sizeof(int32_t) * Vd);
  take(L.lag, sizeof(int32_t) * Vd);
  take(L.dir, sizeof(int32_t) * Vd);
  take(L.coef, sizeof(double) * Vd * B * (DP > 0 ? DP : 1));
  take(L.prev, sizeof(float) * n_state);
  take(L.cur, sizeof(float) * n_state);
  take(L.lo, sizeof(float) * n_state);
  take(L.err, sizeof(int));
  take(L.start, sizeof(int32_t) * Vd);
  take(L.rowid, sizeof(int32_t) * Vd);
  take(L.grp_ptr, sizeof(int64_t) * (P + 1));
  take(L.grp_key, sizeof(int32_t) * L.gcap);
  take(L.scal, sizeof(double) * B * Vd);
  take(L.gsum, sizeof(float) * L.gcap * B * T);
  take(L.bad, sizeof(int32_t));
  L.total = off;
  return L;
}

// Groups of the last spr_energy_exchange call on this thread (0: the
// per-pair plan ran).
thread_local int64_t last_dense_groups = 0;
}  // namespace

extern "C" int64_t spr_energy_exchange_workspace(int64_t P, int64_t D,
                                                 int64_t B, int64_t T,
                                                 int64_t V) {
  if (P <= 0 || D <= 0 || B <= 0 || T <= 0 || V < 0) return -1;
  return (int64_t)dense_layout(P, D, B, T, V).total;
}

extern "C" int64_t spr_energy_exchange_groups(void) {
  return last_dense_groups;
}

namespace {
// The grouped plan from the dense F~ on the destination rows (row_ptr,
// src) already in the workspace: *groups = G on success, 0 (with src
// possibly reordered) when the rows do not factorise into at most gcap
// groups, a row is longer than the LDS sort takes, or the sizes are beyond
// the grouped kernels' limits.
int dense_grouped_plan(int64_t T, int64_t P, int64_t D, int64_t B,
                       const double *dist, const double *fft,
                       const int64_t *p2o, double c, double dt, int64_t V,
                       char *ws, const DenseLayout &L, hipStream_t s,
                       int64_t *groups) {
  *groups = 0;
  const int64_t Vd = 2 * V, DP = spr_padded_directions(D);
  if (L.gcap < 1 || T >= ((int64_t)1 << 29) ||
      D * B * T >= ((int64_t)1 << 29) || P * D >= 0x7fffffffll ||
      Vd >= 0x7fffffffll)
    return SPR_OK;
  auto *row_ptr = reinterpret_cast<int64_t *>(ws + L.row_ptr);
  auto *src = reinterpret_cast<int32_t *>(ws + L.src);
  auto *lag = reinterpret_cast<int32_t *>(ws + L.lag);
  auto *dir = reinterpret_cast<int32_t *>(ws + L.dir);
  auto *coef = reinterpret_cast<double *>(ws + L.coef);
  auto *hkey = reinterpret_cast<uint64_t *>(ws + L.coef);
  auto *start = reinterpret_cast<int32_t *>(ws + L.start);
  auto *rowid = reinterpret_cast<int32_t *>(ws + L.rowid);
  auto *grp_ptr = reinterpret_cast<int64_t *>(ws + L.grp_ptr);
  auto *grp_key = reinterpret_cast<int32_t *>(ws + L.grp_key);
  auto *scal = reinterpret_cast<double *>(ws + L.scal);
  auto *bad = reinterpret_cast<int32_t *>(ws + L.bad);
  int rc;
  if ((rc = fill<int32_t>(bad, 1, 0, s))) return rc;
  hipLaunchKernelGGL(dense_row_hash_kernel, dim3((unsigned)P), dim3(256), 0,
                     s, P, D, B, row_ptr, src, fft, hkey);
  SPR_LAUNCH_CHECK();
  int npad = 1;
  while (npad < std::min<int64_t>(P, kSortMax)) npad <<= 1;  // rows <= P
  const size_t lds =
      (size_t)npad * (sizeof(uint64_t) + sizeof(int32_t)) + sizeof(int32_t);
  SPR_HIP(hipFuncSetAttribute((const void *)dense_hash_sort_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  hipLaunchKernelGGL(dense_hash_sort_kernel, dim3((unsigned)P), dim3(1024),
                     lds, s, P, row_ptr, src, hkey, start, grp_ptr, bad);
  SPR_LAUNCH_CHECK();
  if ((rc = scan_exclusive(grp_ptr, P, s))) return rc;
  int64_t h_groups = 0;
  int32_t h_bad = 0;
  SPR_HIP(hipMemcpyAsync(&h_groups, grp_ptr + P, sizeof(int64_t),
                         hipMemcpyDeviceToHost, s));
  SPR_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost,
                         s));
  SPR_HIP(hipStreamSynchronize(s));
  if (h_bad || h_groups < 1 || h_groups > L.gcap) return SPR_OK;
  if ((rc = spr_plan_from_dense(P, D, B, c, dt, Vd, row_ptr, src, dist, p2o,
                                fft, lag, dir, coef, s)))
    return rc;
  hipLaunchKernelGGL(dense_group_fill_kernel, dim3((unsigned)ceil_div(P, 4)),
                     dim3(256), 0, s, P, D, B, DP, Vd, row_ptr, grp_ptr, src,
                     start, dir, coef, grp_key, rowid, lag, scal, bad);
  SPR_LAUNCH_CHECK();
  SPR_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int32_t), hipMemcpyDeviceToHost,
                         s));
  SPR_HIP(hipStreamSynchronize(s));
  if (!h_bad) *groups = h_groups;
  return SPR_OK;
}

// The body of spr_energy_exchange on a workspace `ws` of the dense layout
// (sizes already checked by the caller): the grouped two-pass exchange
// when the F~ rows factorise, else the per-pair plan.
int energy_exchange_on(const char *name, int64_t T, int64_t P, int64_t D,
                       int64_t B, const double *e0dir, const double *d0,
                       const double *dist, const double *fft,
                       const int64_t *p2o, double c, double dt, int64_t K,
                       int64_t V, const int32_t *pairs, float *etc, char *ws,
                       const DenseLayout &L, void *stream) {
  hipStream_t s = as_stream(stream);
  auto *vis = reinterpret_cast<uint8_t *>(ws + L.vis);
  auto *row_ptr = reinterpret_cast<int64_t *>(ws + L.row_ptr);
  auto *src = reinterpret_cast<int32_t *>(ws + L.src);
  auto *lag = reinterpret_cast<int32_t *>(ws + L.lag);
  auto *dir = reinterpret_cast<int32_t *>(ws + L.dir);
  auto *coef = reinterpret_cast<double *>(ws + L.coef);
  auto *prev = reinterpret_cast<float *>(ws + L.prev);
  auto *cur = reinterpret_cast<float *>(ws + L.cur);
  auto *lo = reinterpret_cast<float *>(ws + L.lo);
  auto *err = reinterpret_cast<int *>(ws + L.err);
  last_dense_groups = 0;
  int rc;
  if ((rc = fill<int32_t>(err, 1, 0, s))) return rc;
  if ((rc = spr_exchange_init(P, D, B, T, e0dir, d0, c, dt, prev, etc, lo,
                              err, stream)))
    return rc;
  if (K >= 1 && V > 0) {
    const int64_t Vd = 2 * V;
    if ((rc = fill<uint8_t>(vis, P * P, 0, s))) return rc;
    hipLaunchKernelGGL(pairs_to_vis_kernel, dim3((unsigned)ceil_div(V, 256)),
                       dim3(256), 0, s, P, V, pairs, vis);
    SPR_LAUNCH_CHECK();
    if ((rc = spr_plan_rows(P, vis, row_ptr, nullptr, stream))) return rc;
    if ((rc = spr_plan_rows(P, vis, row_ptr, src, stream))) return rc;
    int64_t G = 0;
    if ((rc = dense_grouped_plan(T, P, D, B, dist, fft, p2o, c, dt, V, ws, L,
                                 s, &G)))
      return rc;
    if (G > 0) {
      // two passes per order; the total is updated every other order with
      // two orders at once (spread mode 2), an odd last order on its own
      auto *grp_ptr = reinterpret_cast<const int64_t *>(ws + L.grp_ptr);
      auto *grp_key = reinterpret_cast<const int32_t *>(ws + L.grp_key);
      auto *rowid = reinterpret_cast<const int32_t *>(ws + L.rowid);
      auto *scal = reinterpret_cast<const double *>(ws + L.scal);
      auto *gsum = reinterpret_cast<float *>(ws + L.gsum);
      for (int64_t k = 0; k < K; ++k) {
        if ((rc = launch_gather(P, B, T, Vd, row_ptr, grp_ptr, rowid, lag,
                                scal, nullptr, prev, gsum, s)))
          return rc;
        const bool pair = k % 2 == 1;
        if ((rc = spread_grouped(name, P, D, B, T, grp_ptr, grp_key, coef,
                                 gsum, pair ? prev : nullptr, cur,
                                 pair ? etc : nullptr, pair ? lo : nullptr,
                                 stream)))
          return rc;
        float *t = prev;
        prev = cur;
        cur = t;
      }
      if (K % 2 == 1 &&
          (rc = spr_exchange_total_add(P * D * B * T, prev, etc, lo, stream)))
        return rc;
      last_dense_groups = G;
    } else {
      // the sort may have reordered the rows: rebuild them in source order
      if ((rc = spr_plan_rows(P, vis, row_ptr, src, stream))) return rc;
      if ((rc = spr_plan_from_dense(P, D, B, c, dt, Vd, row_ptr, src, dist,
                                    p2o, fft, lag, dir, coef, stream)))
        return rc;
      for (int64_t k = 0; k < K; ++k) {
        if ((rc = spr_exchange_step(P, D, B, T, row_ptr, src, lag, dir, coef,
                                    prev, cur, etc, lo, stream)))
          return rc;
        float *t = prev;
        prev = cur;
        cur = t;
      }
    }
  }
  if ((rc = spr_exchange_finalize(P * D * B * T, etc, lo, etc, stream)))
    return rc;
  int h_err = 0;
  SPR_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, s));
  SPR_HIP(hipStreamSynchronize(s));
  if (h_err)
    return fail(SPR_ERR_RANGE,
                "%s: a source delay is beyond the histogram length", name);
  return SPR_OK;
}

int check_exchange_args(const char *name, int64_t T, int64_t P, int64_t D,
                        int64_t B, const double *e0dir, const double *d0,
                        const double *dist, const double *fft,
                        const int64_t *p2o, int64_t K, int64_t V,
                        const int32_t *pairs, const float *etc) {
  SPR_REQUIRE(T > 0 && P > 0 && D > 0 && B > 0 && V >= 0,
              "%s: bad sizes", name);
  SPR_REQUIRE(spr_padded_directions(D) > 0,
              "%s: D=%lld unsupported (max 64)", name, (long long)D);
  SPR_REQUIRE(e0dir && d0 && etc, "%s: null pointer", name);
  SPR_REQUIRE(K < 1 || (dist && fft && p2o && (pairs || V == 0)),
              "%s: null pointer", name);
  return SPR_OK;
}

// Scratch of spr_energy_exchange_noalloc: one copy per GPU, held until the
// process ends.  Host code never uses `scratch` itself (on the host the name
// stands for a host-side shadow, not device memory): the device address of
// the current GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept in noalloc_ptr.
static __device__ __align__(256) unsigned char
    scratch[SPR_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex noalloc_mutex;
void *noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(void **out) {
  int dev = -1;
  SPR_HIP(hipGetDevice(&dev));
  SPR_REQUIRE(dev >= 0 && dev < kMaxDevices,
              "spr_energy_exchange_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(noalloc_mutex);
  if (!noalloc_ptr[dev]) {
    void *p = nullptr;
    SPR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SPR_REQUIRE(p != nullptr,
                "spr_energy_exchange_noalloc: no scratch on device %d", dev);
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return SPR_OK;
}
}  // namespace

extern "C" int spr_energy_exchange(int64_t T, int64_t P, int64_t D, int64_t B,
                                   const double *e0dir, const double *d0,
                                   const double *dist, const double *fft,
                                   const int64_t *p2o, double c, double dt,
                                   int64_t K, int64_t V,
                                   const int32_t *pairs, float *etc,
                                   void *workspace, int64_t workspace_bytes,
                                   void *stream) {
  int rc = check_exchange_args("spr_energy_exchange", T, P, D, B, e0dir, d0,
                               dist, fft, p2o, K, V, pairs, etc);
  if (rc) return rc;
  SPR_REQUIRE(workspace, "spr_energy_exchange: null pointer");
  DenseLayout L = dense_layout(P, D, B, T, V);
  if ((int64_t)L.total > workspace_bytes)
    return fail(SPR_ERR_WORKSPACE,
                "spr_energy_exchange: workspace of %lld bytes, need %lld",
                (long long)workspace_bytes, (long long)L.total);
  return energy_exchange_on("spr_energy_exchange", T, P, D, B, e0dir, d0,
                            dist, fft, p2o, c, dt, K, V, pairs, etc,
                            static_cast<char *>(workspace), L, stream);
}

extern "C" int spr_energy_exchange_noalloc(
    int64_t T, int64_t P, int64_t D, int64_t B, const double *e0dir,
    const double *d0, const double *dist, const double *fft,
    const int64_t *p2o, double c, double dt, int64_t K, int64_t V,
    const int32_t *pairs, float *etc, void *stream) {
  clear_error();
  int rc = check_exchange_args("spr_energy_exchange_noalloc", T, P, D, B,
                               e0dir, d0, dist, fft, p2o, K, V, pairs, etc);
  if (rc) return rc;
  DenseLayout L = dense_layout(P, D, B, T, V);
  if ((int64_t)L.total > (int64_t)SPR_NOALLOC_SCRATCH_BYTES)
    return fail(SPR_ERR_WORKSPACE,
                "spr_energy_exchange_noalloc: the scene needs %lld bytes of "
                "scratch, more than the %lld bytes built in; use "
                "spr_energy_exchange with a workspace",
                (long long)L.total, (long long)SPR_NOALLOC_SCRATCH_BYTES);
  void *ws = nullptr;
  if ((rc = noalloc_scratch(&ws))) return rc;
  return energy_exchange_on("spr_energy_exchange_noalloc", T, P, D, B, e0dir,
                            d0, dist, fft, p2o, c, dt, K, V, pairs, etc,
                            static_cast<char *>(ws), L, stream);
}
