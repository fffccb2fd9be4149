// Receiver side: patch-to-receiver factors and the histogram collection.
//
// Reference: _patch2receiver_energy_universal (universal.py:149-160),
// _collect_energy_patches (RadiosityFast.py:686-752),
// _collect_receiver_energy (RadiosityFast.py:1148-1185) and the patch sum
// of collect_energy_receiver_mono (RadiosityFast.py:591-592).
//
// Mapping: a workgroup owns (receiver r, band b, a block of time bins) and
// walks all patches k; per patch the direction index, factor, delay and air
// factor are uniform across the workgroup and the histogram row
// etc[k, idx, b, :] is read contiguously (circularly shifted by the delay).
#include "spr_common.hpp"
#include "spr_geom.hpp"
#include "spr_util.hpp"

namespace spr {
namespace {

__global__ void receiver_energy_kernel(int64_t NQ,
                                       const double *__restrict__ points,
                                       int64_t P,
                                       const double *__restrict__ pts,
                                       const uint8_t *__restrict__ vis,
                                       double *__restrict__ factor) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (k >= P) return;
  factor[q * P + k] = vis[q * P + k]
                          ? pt_solution(load3(points + 3 * q), pts + 12 * k,
                                        true)
                          : 0.0;
}

constexpr int kCollectBS = 256;
constexpr int kCollectR = 4;  // bins per thread

// out[r, b, t] = sum_k (etc[k, idx_rk, b, (t - n_rk) mod T] * g_rk) * a_rkb
__global__ __launch_bounds__(kCollectBS) void collect_mono_kernel(
    int64_t NR, int64_t P, int64_t D, int64_t B, int64_t T, int n_tblk,
    const float *__restrict__ etc, const double *__restrict__ points,
    const double *__restrict__ centers, const int64_t *__restrict__ idx,
    const double *__restrict__ factor, const double *__restrict__ air,
    double c, double dt, float *__restrict__ out) {
  const int tb = blockIdx.x % n_tblk;
  const int64_t rb = blockIdx.x / n_tblk;
  const int64_t b = rb % B, r = rb / B;
  const int64_t t0 = (int64_t)tb * kCollectBS * kCollectR;
  const d3 rp = load3(points + 3 * r);
  double acc[kCollectR];
#pragma unroll
  for (int q = 0; q < kCollectR; ++q) acc[q] = 0.0;
  for (int64_t k = 0; k < P; ++k) {
    const double g = factor[r * P + k];
    if (g == 0.0) continue;  // contributes exactly +0 (etc >= 0)
    const double dist = norm(load3(centers + 3 * k) - rp);
    const int64_t n = delay_ceil(dist, c, dt);
    const int64_t sh = ((n % T) + T) % T;
    const double a = exp(-air[b] * dist);
    const float *row = etc + ((k * D + idx[r * P + k]) * B + b) * T;
#pragma unroll
    for (int q = 0; q < kCollectR; ++q) {
      const int64_t t = t0 + threadIdx.x + (int64_t)kCollectBS * q;
      if (t < T) {
        int64_t ts = t - sh;
        if (ts < 0) ts += T;
        acc[q] += ((double)row[ts] * g) * a;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kCollectR; ++q) {
    const int64_t t = t0 + threadIdx.x + (int64_t)kCollectBS * q;
    if (t < T) out[(r * B + b) * T + t] = (float)acc[q];
  }
}

// patchwise: out[r, k, b, t], thread per bin
__global__ __launch_bounds__(256) void collect_patchwise_kernel(
    int64_t NR, int64_t P, int64_t D, int64_t B, int64_t T, int n_tblk,
    const float *__restrict__ etc, const double *__restrict__ points,
    const double *__restrict__ centers, const int64_t *__restrict__ idx,
    const double *__restrict__ factor, const double *__restrict__ air,
    double c, double dt, float *__restrict__ out) {
  const int tb = blockIdx.x % n_tblk;
  const int64_t rkb = blockIdx.x / n_tblk;
  const int64_t b = rkb % B;
  const int64_t k = (rkb / B) % P;
  const int64_t r = rkb / (B * P);
  const int64_t t = (int64_t)tb * 256 + threadIdx.x;
  if (t >= T) return;
  const double g = factor[r * P + k];
  const double dist = norm(load3(centers + 3 * k) - load3(points + 3 * r));
  const int64_t n = delay_ceil(dist, c, dt);
  const int64_t sh = ((n % T) + T) % T;
  const double a = exp(-air[b] * dist);
  int64_t ts = t - sh;
  if (ts < 0) ts += T;
  const float *row = etc + ((k * D + idx[r * P + k]) * B + b) * T;
  out[((r * P + k) * B + b) * T + t] = (float)(((double)row[ts] * g) * a);
}

// Adjoint of the mono collection: lam[k, idx_rk, b, s] += g_rk a_rkb
// grad[r, b, (s + n_rk) mod T] for every receiver r (lam zero-filled
// before); one thread per (patch, band, bin), receivers in order.
__global__ __launch_bounds__(256) void collect_adjoint_kernel(
    int64_t NR, int64_t P, int64_t D, int64_t B, int64_t T,
    const float *__restrict__ grad, const double *__restrict__ points,
    const double *__restrict__ centers, const int64_t *__restrict__ idx,
    const double *__restrict__ factor, const double *__restrict__ air,
    double c, double dt, float *__restrict__ lam) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P * B * T) return;
  const int64_t s = q % T, b = (q / T) % B, k = q / (B * T);
  const d3 ck = load3(centers + 3 * k);
  for (int64_t r = 0; r < NR; ++r) {
    const double g = factor[r * P + k];
    if (g == 0.0) continue;
    const double dist = norm(ck - load3(points + 3 * r));
    const int64_t n = delay_ceil(dist, c, dt);
    const int64_t sh = ((n % T) + T) % T;
    int64_t t = s + sh;
    if (t >= T) t -= T;
    const double a = exp(-air[b] * dist);
    float *row = lam + ((k * D + idx[r * P + k]) * B + b) * T;
    row[s] = (float)((double)row[s] + (g * a) * (double)grad[(r * B + b) * T + t]);
  }
}

}  // namespace
}  // namespace spr

using namespace spr;

extern "C" int spr_receiver_energy(int64_t NQ, const double *points,
                                   int64_t P, const double *patch_points,
                                   const uint8_t *vis, double *factor,
                                   void *stream) {
  SPR_REQUIRE(NQ > 0 && P > 0 && NQ < 65536, "spr_receiver_energy: sizes");
  SPR_REQUIRE(points && patch_points && vis && factor,
              "spr_receiver_energy: null pointer");
  hipLaunchKernelGGL(receiver_energy_kernel,
                     dim3((unsigned)ceil_div(P, 128), (unsigned)NQ),
                     dim3(128), 0, as_stream(stream), NQ, points, P,
                     patch_points, vis, factor);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_collect(int64_t NR, int64_t P, int64_t D, int64_t B,
                           int64_t T, const float *etc, const double *points,
                           const double *centers, const int64_t *idx,
                           const double *factor, const double *air, double c,
                           double dt, int patchwise, float *out,
                           void *stream) {
  SPR_REQUIRE(NR > 0 && P > 0 && D > 0 && B > 0 && T > 0 && c > 0 && dt > 0,
              "spr_collect: bad sizes");
  SPR_REQUIRE(etc && points && centers && idx && factor && air && out,
              "spr_collect: null pointer");
  hipStream_t s = as_stream(stream);
  if (!patchwise) {
    const int64_t n_tblk = ceil_div(T, (int64_t)kCollectBS * kCollectR);
    const int64_t nb = NR * B * n_tblk;
    SPR_REQUIRE(nb < 0x7fffffff, "spr_collect: grid too large");
    hipLaunchKernelGGL(collect_mono_kernel, dim3((unsigned)nb),
                       dim3(kCollectBS), 0, s, NR, P, D, B, T, (int)n_tblk,
                       etc, points, centers, idx, factor, air, c, dt, out);
  } else {
    const int64_t n_tblk = ceil_div(T, 256);
    const int64_t nb = NR * P * B * n_tblk;
    SPR_REQUIRE(nb < 0x7fffffff, "spr_collect: grid too large");
    hipLaunchKernelGGL(collect_patchwise_kernel, dim3((unsigned)nb),
                       dim3(256), 0, s, NR, P, D, B, T, (int)n_tblk, etc,
                       points, centers, idx, factor, air, c, dt, out);
  }
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_collect_adjoint(int64_t NR, int64_t P, int64_t D,
                                   int64_t B, int64_t T, const float *grad,
                                   const double *points,
                                   const double *centers, const int64_t *idx,
                                   const double *factor, const double *air,
                                   double c, double dt, float *lam,
                                   void *stream) {
  SPR_REQUIRE(NR > 0 && P > 0 && D > 0 && B > 0 && T > 0 && c > 0 && dt > 0,
              "spr_collect_adjoint: bad sizes");
  SPR_REQUIRE(grad && points && centers && idx && factor && air && lam,
              "spr_collect_adjoint: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = fill<float>(lam, P * D * B * T, 0.0f, s);
  if (rc) return rc;
  hipLaunchKernelGGL(collect_adjoint_kernel,
                     dim3((unsigned)ceil_div(P * B * T, 256)), dim3(256), 0,
                     s, NR, P, D, B, T, grad, points, centers, idx, factor,
                     air, c, dt, lam);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}
