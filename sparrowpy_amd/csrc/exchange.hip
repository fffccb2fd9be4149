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
// Kernel mapping: a workgroup owns (j, b, a block of BS*R time bins); each
// thread owns R bins (strided by BS so every load instruction is one
// contiguous 256-byte wavefront access) and accumulates all DP directions
// in float64 registers.  The per-pair coefficient row coef[e, b, :] is
// uniform across the workgroup and is read through the scalar unit.
#include "spr_common.hpp"
#include "spr_fft.hpp"
#include "spr_util.hpp"

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
template <int DP, int R, int BS>
__global__ __launch_bounds__(BS) void exchange_step_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks,
    const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ src,
    const int32_t *__restrict__ lag, const int32_t *__restrict__ dir,
    const double *__restrict__ coef, const float *__restrict__ prev,
    float *__restrict__ cur, float *__restrict__ tot_hi,
    float *__restrict__ tot_lo) {
  const int64_t tile = xcd_tile(blockIdx.x, nblocks);
  const int tb = (int)(tile % n_tblk);
  const int64_t jb = tile / n_tblk;
  const int b = (int)(jb % B);
  const int64_t j = jb / B;
  const int t0 = tb * (BS * R);
  const int t_end = min(T, t0 + BS * R);
  const int64_t dstride = (int64_t)B * T;  // between directions of a patch

  double acc[DP][R];
#pragma unroll
  for (int d = 0; d < DP; ++d)
#pragma unroll
    for (int r = 0; r < R; ++r) acc[d][r] = 0.0;

  const int64_t e_begin = row_ptr[j], e_end = row_ptr[j + 1];
  for (int64_t e = e_begin; e < e_end; ++e) {
    const int n = lag[e];
    if (n >= t_end) continue;
    const float *s =
        prev + ((int64_t)src[e] * D + dir[e]) * dstride + (int64_t)b * T;
    double ev[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int t = t0 + (int)threadIdx.x + BS * r;
      const int ts = t - n;
      ev[r] = (t < T && ts >= 0) ? (double)s[ts] : 0.0;
    }
    const double *cf = coef + (e * B + b) * DP;
#pragma unroll
    for (int d = 0; d < DP; ++d) {
      const double cd = cf[d];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[d][r] = fma(cd, ev[r], acc[d][r]);
    }
  }

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

template <int DP, int R>
int launch_step(int64_t P, int64_t D, int64_t B, int64_t T,
                const int64_t *row_ptr, const int32_t *src,
                const int32_t *lag, const int32_t *dir, const double *coef,
                const float *prev, float *cur, float *tot_hi, float *tot_lo,
                hipStream_t s) {
  constexpr int BS = 256;
  const int64_t n_tblk = ceil_div(T, (int64_t)BS * R);
  const int64_t nblocks = P * B * n_tblk;
  SPR_REQUIRE(nblocks < (int64_t)0x7fffffff, "exchange grid too large");
  hipLaunchKernelGGL((exchange_step_kernel<DP, R, BS>),
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

namespace {
struct DenseLayout {
  size_t vis, row_ptr, src, lag, dir, coef, prev, cur, lo, err, total;
};

DenseLayout dense_layout(int64_t P, int64_t D, int64_t B, int64_t T,
                         int64_t V) {
  DenseLayout L;
  int64_t DP = spr_padded_directions(D);
  int64_t Vd = 2 * V;
  size_t n_state = (size_t)P * D * B * T;
  size_t off = 0;
  L.vis = off;
  off = align256(off + (size_t)P * P);
  L.row_ptr = off;
  off = align256(off + sizeof(int64_t) * (P + 1));
  L.src = off;
  off = align256(off + sizeof(int32_t) * Vd);
  L.lag = off;
  off = align256(off + sizeof(int32_t) * Vd);
  L.dir = off;
  off = align256(off + sizeof(int32_t) * Vd);
  L.coef = off;
  off = align256(off + sizeof(double) * Vd * B * (DP > 0 ? DP : 1));
  L.prev = off;
  off = align256(off + sizeof(float) * n_state);
  L.cur = off;
  off = align256(off + sizeof(float) * n_state);
  L.lo = off;
  off = align256(off + sizeof(float) * n_state);
  L.err = off;
  off = align256(off + sizeof(int));
  L.total = off;
  return L;
}
}  // namespace

extern "C" int64_t spr_energy_exchange_workspace(int64_t P, int64_t D,
                                                 int64_t B, int64_t T,
                                                 int64_t V) {
  if (P <= 0 || D <= 0 || B <= 0 || T <= 0 || V < 0) return -1;
  return (int64_t)dense_layout(P, D, B, T, V).total;
}

extern "C" int spr_energy_exchange(int64_t T, int64_t P, int64_t D, int64_t B,
                                   const double *e0dir, const double *d0,
                                   const double *dist, const double *fft,
                                   const int64_t *p2o, double c, double dt,
                                   int64_t K, int64_t V,
                                   const int32_t *pairs, float *etc,
                                   void *workspace, int64_t workspace_bytes,
                                   void *stream) {
  SPR_REQUIRE(T > 0 && P > 0 && D > 0 && B > 0 && V >= 0,
              "spr_energy_exchange: bad sizes");
  SPR_REQUIRE(spr_padded_directions(D) > 0,
              "spr_energy_exchange: D=%lld unsupported (max 64)",
              (long long)D);
  SPR_REQUIRE(e0dir && d0 && etc && workspace,
              "spr_energy_exchange: null pointer");
  SPR_REQUIRE(K < 1 || (dist && fft && p2o && (pairs || V == 0)),
              "spr_energy_exchange: null pointer");
  DenseLayout L = dense_layout(P, D, B, T, V);
  if ((int64_t)L.total > workspace_bytes)
    return fail(SPR_ERR_WORKSPACE,
                "spr_energy_exchange: workspace of %lld bytes, need %lld",
                (long long)workspace_bytes, (long long)L.total);
  hipStream_t s = as_stream(stream);
  char *ws = static_cast<char *>(workspace);
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
  int rc;
  if ((rc = fill<int32_t>(err, 1, 0, s))) return rc;
  if ((rc = spr_exchange_init(P, D, B, T, e0dir, d0, c, dt, prev, etc, lo,
                              err, stream)))
    return rc;
  if (K >= 1 && V > 0) {
    if ((rc = fill<uint8_t>(vis, P * P, 0, s))) return rc;
    hipLaunchKernelGGL(pairs_to_vis_kernel, dim3((unsigned)ceil_div(V, 256)),
                       dim3(256), 0, s, P, V, pairs, vis);
    SPR_LAUNCH_CHECK();
    if ((rc = spr_plan_rows(P, vis, row_ptr, nullptr, stream))) return rc;
    if ((rc = spr_plan_rows(P, vis, row_ptr, src, stream))) return rc;
    if ((rc = spr_plan_from_dense(P, D, B, c, dt, 2 * V, row_ptr, src, dist,
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
  if ((rc = spr_exchange_finalize(P * D * B * T, etc, lo, etc, stream)))
    return rc;
  int h_err = 0;
  SPR_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, s));
  SPR_HIP(hipStreamSynchronize(s));
  if (h_err)
    return fail(SPR_ERR_RANGE,
                "spr_energy_exchange: a source delay is beyond the "
                "histogram length");
  return SPR_OK;
}
