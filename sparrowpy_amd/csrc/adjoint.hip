// Adjoint of the energy exchange on a diffuse + specular plan.
//
// Forward (RadiosityFast.py:1073-1145): E_{k+1}[j,d,b,t] = sum over entries
// e = (i -> j) of F~_e[d,b] * E_k[i, dir_e, b, t - n_e], total = sum_k E_k,
// with F~_e[d,b] = w_eb (A[k_e,b] + C[k_e,b] [d = m_e]) (A: diffuse
// constant, C: specular part of BRDF row k_e; w_eb the pair factor).
// For a loss L(total) with lambda = dL/dtotal:
//   mu_K = lambda,  mu_k = lambda + A^T mu_{k+1},
//   (A^T mu)[i,d',b,t] = sum_{e: i -> j, dir_e = d'}
//        w_eb (A_b sum_d mu[j,d,b,t+n_e] + C_b mu[j,m_e,b,t+n_e]),
//   dL/d(w A)_eb += sum_t (sum_d mu_{k+1}[j,d,b,t]) E_k[i,d',b,t-n_e],
//   dL/d(w C)_eb += sum_t mu_{k+1}[j,m_e,b,t] E_k[i,d',b,t-n_e],
// and dL/dE_0 = mu_0, which reaches the tables through the injection
// E_0dir[i,:,b] = E0[i,b] (A + C e_m)[k0_i, b] (RadiosityFast.py:988-1034).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "spr_common.hpp"
#include "spr_exchange.hpp"
#include "spr_util.hpp"

namespace spr {
namespace {

// scal[b,e] = wgt[b,e] C[krow_e, b], dscal[b,e] = wgt[b,e] A[krow_e, b]
__global__ void ds_weights_kernel(int64_t B, int64_t Vd,
                                  const int32_t *__restrict__ krow,
                                  const double *__restrict__ wgt,
                                  const double *__restrict__ ta,
                                  const double *__restrict__ tc,
                                  double *__restrict__ scal,
                                  double *__restrict__ dscal) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B * Vd) return;
  const int64_t b = q / Vd, e = q % Vd;
  const int64_t k = krow[e];
  scal[q] = wgt[q] * tc[k * B + b];
  dscal[q] = wgt[q] * ta[k * B + b];
}

// Transposed weights: twa[b,q] = dscal[b, te[q]], twc[b,q] = scal[b, te[q]].
__global__ void t_weights_kernel(int64_t B, int64_t Vd,
                                 const int32_t *__restrict__ te,
                                 const double *__restrict__ scal,
                                 const double *__restrict__ dscal,
                                 double *__restrict__ twa,
                                 double *__restrict__ twc) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B * Vd) return;
  const int64_t b = q / Vd, p = q % Vd;
  const int64_t e = te[p];
  twa[q] = dscal[b * Vd + e];
  twc[q] = scal[b * Vd + e];
}

// Injection from the tables: k0[p] = BRDF row of patch p for the direction
// of the source, e0dir[p,d,b] = energy0[p,b] (A[k0,b] + C[k0,b] [d = m]).
__global__ void ds_injection_kernel(
    int64_t P, int64_t D, int64_t B, const double *__restrict__ source,
    const double *__restrict__ centers, const int64_t *__restrict__ wall_ids,
    int64_t S, const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs, const double *__restrict__ energy0,
    const double *__restrict__ ta, const double *__restrict__ tc,
    const int32_t *__restrict__ spec_dir, int32_t *__restrict__ k0,
    double *__restrict__ e0dir) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int64_t k = 0;
  if (in_dirs) {
    d3 u = load3(source) - load3(centers + 3 * p);
    u = div(u, norm(u));
    const int64_t w = wall_ids[p];
    k = brdf_index[w] * S + nearest_direction(in_dirs + w * S * 3, S, u);
  }
  k0[p] = (int32_t)k;
  const int64_t m = spec_dir[k];
  for (int64_t b = 0; b < B; ++b) {
    const double e = energy0[p * B + b];
    for (int64_t d = 0; d < D; ++d)
      e0dir[(p * D + d) * B + b] =
          e * (d == m ? ta[k * B + b] + tc[k * B + b] : ta[k * B + b]);
  }
}

// musum[j,b,t] = sum_d mu[j,d,b,t] (float64 sum in direction order,
// float32 result); 8 directions' loads are issued before they are added.
__global__ void direction_sum_kernel(int64_t P, int64_t D, int64_t B,
                                     int64_t T, const float *__restrict__ mu,
                                     float *__restrict__ musum) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P * B * T) return;
  const int64_t t = q % T, b = (q / T) % B, j = q / (B * T);
  const float *p = mu + (j * D * B + b) * T + t;  // element (j, 0, b, t)
  const int64_t stride = B * T;                  // next direction
  double s = 0.0;
  int64_t d = 0;
  for (; d + 8 <= D; d += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p[(d + k) * stride];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += (double)v[k];
  }
  for (; d < D; ++d) s += (double)p[d * stride];
  musum[q] = (float)s;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One backward order: mu_out += A^T mu_next (mu_out holds lambda on entry),
// and the entries' table gradients ga[0,b,q] += sum_t musum[j,b,t+n]
// ek[i,d',b,t], ga[1,b,q] += sum_t mu_next[j,m,b,t+n] ek[i,d',b,t]
// (q: transposed entry, d' its source direction).  A workgroup takes one
// (source i, band, block of 256*R bins), sources fastest, and streams all
// entries of source i -- sorted by direction, a flag on the last entry of
// each direction -- in chunks of C, the next chunk's loads (mu-sum row,
// specular mu row, the source's own order-k row) in flight while the
// current one is summed.  At a direction's last entry the sum is added to
// its mu_out row with float atomics (each element receives exactly one
// add, so the result does not depend on their order).  A bin t + n beyond
// the histogram is an out-of-range buffer offset and reads 0.  The
// per-entry dot products are reduced over each wave and added (float64
// atomics) by lanes 0 .. 2C-1.  Requires T < 2^29.
template <int R, int C>
__global__ __launch_bounds__(256) void adjoint_ds_kernel(
    int D, int B, int T, int n_tblk, int64_t nblocks, int64_t Vd,
    const int64_t *__restrict__ tgptr, const int32_t *__restrict__ tkey,
    const int32_t *__restrict__ tlag, const int32_t *__restrict__ tm,
    const double *__restrict__ twa, const double *__restrict__ twc,
    const float *__restrict__ musum, const float *__restrict__ munext,
    const float *__restrict__ ek, float *__restrict__ muout,
    double *__restrict__ ga) {
  constexpr int BS = 256;
  constexpr int32_t kDstMask = (1 << 25) - 1;
  int64_t i;
  int b, tb;
  dest_major_tile(xcd_tile(blockIdx.x, nblocks), nblocks / (B * n_tblk),
                  n_tblk, i, b, tb);
  const int64_t q0 = tgptr[i * D], q1 = tgptr[i * D + D];
  if (q1 == q0) return;
  const int tl = tb * (BS * R) + (int)threadIdx.x;
  const int lane = (int)threadIdx.x & 63;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  const double *wa = twa + (int64_t)b * Vd, *wc = twc + (int64_t)b * Vd;
  double *gda = ga + (int64_t)b * Vd, *gdc = ga + ((int64_t)B + b) * Vd;
  const int64_t srow = i * D;  // row (i, d') = srow + d'

  auto fetch = [&](int64_t qq, float (&x)[C][R], float (&y)[C][R],
                   float (&e)[C][R]) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const bool in = qq + c < q1;
      const int64_t qc = in ? qq + c : q1 - 1;
      const int32_t key = tkey[qc];
      const int64_t j = key & kDstMask, d = key >> 25, m = tm[qc];
      const int n = min(tlag[qc] & kLagMask, T);  // >= T: reads 0
      const uint32_t sz = in ? row_bytes : 0u;
      const __amdgpu_buffer_rsrc_t rx =
          make_rsrc(musum + (j * B + b) * (int64_t)T, sz);
      const __amdgpu_buffer_rsrc_t ry =
          make_rsrc(munext + ((j * D + m) * B + b) * (int64_t)T, sz);
      const __amdgpu_buffer_rsrc_t re =
          make_rsrc(ek + ((srow + d) * B + b) * (int64_t)T, sz);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int t = tl + BS * r;
        x[c][r] = buffer_load(rx, (uint32_t)(t + n) * 4u);
        y[c][r] = buffer_load(ry, (uint32_t)(t + n) * 4u);
        e[c][r] = buffer_load(re, (uint32_t)t * 4u);
      }
    }
  };
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  auto consume = [&](int64_t qq, const float (&x)[C][R],
                     const float (&y)[C][R], const float (&e)[C][R]) {
    double pa[C], pc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      pa[c] = pc[c] = 0.0;
      if (qq + c < q1) {
        const double a = wa[qq + c], s = wc[qq + c];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double xd = (double)x[c][r], yd = (double)y[c][r];
          const double ed = (double)e[c][r];
          acc[r] = fma(a, xd, acc[r]);
          acc[r] = fma(s, yd, acc[r]);
          pa[c] = fma(ed, xd, pa[c]);
          pc[c] = fma(ed, yd, pc[c]);
        }
        if (tlag[qq + c] & kLastInGroup) {
          const int64_t d = tkey[qq + c] >> 25;
          float *o = muout + ((srow + d) * B + b) * (int64_t)T;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int t = tl + BS * r;
            if (t < T) unsafeAtomicAdd(o + t, (float)acc[r]);
            acc[r] = 0.0;
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      pa[c] = wave_sum(pa[c]);
      pc[c] = wave_sum(pc[c]);
    }
    if (lane < 2 * C) {
      const int c = lane % C;
      double v = 0.0;
#pragma unroll
      for (int cc = 0; cc < C; ++cc)
        if (cc == c) v = lane < C ? pa[cc] : pc[cc];
      if (qq + c < q1) unsafeAtomicAdd((lane < C ? gda : gdc) + qq + c, v);
    }
  };
  float xa[C][R], ya[C][R], ea[C][R], xb[C][R], yb[C][R], eb[C][R];
  fetch(q0, xa, ya, ea);
  for (int64_t q = q0; q < q1; q += 2 * C) {
    fetch(q + C, xb, yb, eb);
    consume(q, xa, ya, ea);
    fetch(q + 2 * C, xa, ya, ea);
    consume(q + C, xb, yb, eb);
  }
}

// Table gradients from the entries' accumulated dot products:
// gA[k,b] += wgt[b,e] ga[0,b,q], gC[k,b] += wgt[b,e] ga[1,b,q] with
// e = te[q], k = krow[e].
__global__ void table_grad_kernel(int64_t B, int64_t Vd,
                                  const int32_t *__restrict__ te,
                                  const int32_t *__restrict__ krow,
                                  const double *__restrict__ wgt,
                                  const double *__restrict__ ga,
                                  double *__restrict__ gA,
                                  double *__restrict__ gC) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B * Vd) return;
  const int64_t b = q / Vd, p = q % Vd;
  const int64_t e = te[p];
  const int64_t k = krow[e];
  const double w = wgt[b * Vd + e];
  unsafeAtomicAdd(gA + k * B + b, w * ga[q]);
  unsafeAtomicAdd(gC + k * B + b, w * ga[B * Vd + q]);
}

// Injection gradient: dL/de0dir[p,d,b] = mu0[p,d,b,n_p] with n_p the
// order-0 bin of patch p; gA[k0_p,b] += energy0[p,b] sum_d mu0[p,d,b,n_p],
// gC[k0_p,b] += energy0[p,b] mu0[p,m,b,n_p].
__global__ void injection_grad_kernel(
    int64_t P, int64_t D, int64_t B, int64_t T,
    const double *__restrict__ energy0, const double *__restrict__ d0,
    double c, double dt, const int32_t *__restrict__ k0,
    const int32_t *__restrict__ spec_dir, const float *__restrict__ mu0,
    double *__restrict__ gA, double *__restrict__ gC) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P * B) return;
  const int64_t p = q / B, b = q % B;
  const int64_t n = delay_floor(d0[p], c, dt);
  if (n < 0 || n >= T) return;
  const int64_t k = k0[p], m = spec_dir[k];
  double s = 0.0, sm = 0.0;
  for (int64_t d = 0; d < D; ++d) {
    const double v = mu0[((p * D + d) * B + b) * T + n];
    s += v;
    if (d == m) sm = v;
  }
  const double e = energy0[p * B + b];
  unsafeAtomicAdd(gA + k * B + b, e * s);
  if (m < D) unsafeAtomicAdd(gC + k * B + b, e * sm);
}

// ---------------------------------------------------------------------
// Adjoint of the grouped plan (any BRDF rows)
// ---------------------------------------------------------------------
// One forward order on the grouped plan (exchange.hip, RadiosityFast.py:
// 1114-1143 grouped by BRDF row):
//   gsum[g,b,t] = sum_{e in g} scal[b,e] E_k[row_e, b, t - n_e],
//   E_{k+1}[j,d,b,t] = sum_{g of j} rtab[key_g, b, d] gsum[g,b,t],
// with rtab[k,b,d] = table[k,d,b] (BRDF row k = brdf * S + s).  For
// mu = dL/dE_{k+1}:
//   nu[g,b,t] = sum_d rtab[key_g,b,d] mu[j_g,d,b,t]              (spread^T)
//   dL/dE_k[r,b,t] += sum_{e: row_e = r} scal[b,e] nu[g_e,b,t+n_e] (gather^T)
//   dL/dtable[key_g,d,b] += sum_t mu[j_g,d,b,t] gsum[g,b,t].

// nu for every group of destination j with up to 8 (padded) directions: a
// workgroup per (j, band, 256 bins), one bin per lane; the lane keeps
// mu[j,:,b,t] as DPT doubles (converted once), and the workgroup stages its
// groups' BRDF rows in LDS, 16 at a time, read by broadcast.  (With 16 or
// more directions grouped_spread_adjoint_mfma_kernel forms nu; this kernel
// took 12.2 ms per call on the directional shoebox.)
template <int DPT>
__global__ __launch_bounds__(256) void grouped_spread_adjoint_kernel(
    int D, int B, int T, int DP, int n_tblk,
    const int64_t *__restrict__ grp_ptr, const int32_t *__restrict__ grp_key,
    const double *__restrict__ rtab, const float *__restrict__ mu,
    float *__restrict__ nu) {
  constexpr int NG = 16;
  __shared__ double rts[NG][DPT];
  const int64_t blk = blockIdx.x;
  const int tb = (int)(blk % n_tblk);
  const int b = (int)((blk / n_tblk) % B);
  const int64_t j = blk / ((int64_t)n_tblk * B);
  const int64_t g0 = grp_ptr[j], g1 = grp_ptr[j + 1];
  if (g0 == g1) return;  // uniform
  const int t = tb * 256 + (int)threadIdx.x;
  const bool in = t < T;
  const int64_t dstride = (int64_t)B * T;
  const float *m = mu + (j * D * B + b) * (int64_t)T + (in ? t : 0);
  double v[DPT];
#pragma unroll
  for (int d = 0; d < DPT; ++d)
    v[d] = (d < D && in) ? (double)m[d * dstride] : 0.0;
  for (int64_t gc = g0; gc < g1; gc += NG) {
    const int ng = (int)min((int64_t)NG, g1 - gc);
    for (int idx = (int)threadIdx.x; idx < NG * DPT; idx += 256) {
      const int gi = idx / DPT, d = idx % DPT;
      rts[gi][d] = (gi < ng && d < DP)
                       ? rtab[((int64_t)grp_key[gc + gi] * B + b) * DP + d]
                       : 0.0;
    }
    __syncthreads();
    for (int gi = 0; gi < ng; ++gi) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < DPT; ++d) s = fma(rts[gi][d], v[d], s);
      if (in) nu[((gc + gi) * B + b) * (int64_t)T + t] = (float)s;
    }
    __syncthreads();
  }
}

// gg[g,b,d] += sum_t mu[j_g,d,b,t] gsum[g,b,t] for the groups g of
// destination j: per (j, band) the product of the group sums [G_j x T]
// with mu[j,:,b,:]^T [T x D], on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: lane l holds A[l & 15][l >> 4], B[l >> 4][l & 15]
// and D[(l >> 4) + 4r][l & 15]).  Tiles of 16 NQ groups x 16 NR directions
// (NQ from the groups left, NR from D); the bins in chunks of 32 staged in
// LDS as float32, both operands read from there; wave q takes group tile
// q against every direction tile (NR accumulators of 16 x 16).  Each
// (group, direction) is summed by one lane in float64 and owned by it: no
// atomics, and the result does not depend on the launch's timing.  The
// directional shoebox's launch takes 5.5 ms (12.0 ms on the vector unit,
// each lane summing NQ x NR pairs; 10.7 ms here without the prefetch of
// the next chunk; 17.8 ms with chunks of 64 bins).
constexpr int kTgChunk = 32;
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

template <int NQ, int NR>
__device__ __forceinline__ void table_grad_tile(
    int D, int B, int T, int DP, int b, int64_t gt, int64_t g1,
    const float *__restrict__ mrow, const float *__restrict__ gsum,
    double *__restrict__ gg, float (&sg)[64][kTgChunk + 1],
    float (&sm)[64][kTgChunk + 1]) {
  constexpr int TC = kTgChunk;
  const int tid = (int)threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int kk = lane >> 4, cc = lane & 15;
  const int64_t dstride = (int64_t)B * T;
  mfma_d4 acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  // 256 lanes stage 8 rows of 32 bins per step; the next chunk's values
  // are loaded into registers while the current chunk is multiplied
  const int col = tid % TC;
  float pg[2 * NQ], pm[2 * NR];
  auto fetch = [&](int t0) {
    const int t = t0 + col;
#pragma unroll
    for (int s = 0; s < 2 * NQ; ++s) {
      const int64_t g = gt + tid / TC + 8 * s;
      pg[s] = (g < g1 && t < T) ? gsum[(g * B + b) * (int64_t)T + t] : 0.f;
    }
#pragma unroll
    for (int s = 0; s < 2 * NR; ++s) {
      const int d = tid / TC + 8 * s;
      pm[s] = (d < D && t < T) ? mrow[d * dstride + t] : 0.f;
    }
  };
  fetch(0);
  for (int t0 = 0; t0 < T; t0 += TC) {
#pragma unroll
    for (int s = 0; s < 2 * NQ; ++s) sg[tid / TC + 8 * s][col] = pg[s];
#pragma unroll
    for (int s = 0; s < 2 * NR; ++s) sm[tid / TC + 8 * s][col] = pm[s];
    __syncthreads();
    if (t0 + TC < T) fetch(t0 + TC);  // uniform
    if (wave < NQ) {  // uniform per wave
#pragma unroll 4
      for (int c = 0; c < TC; c += 4) {
        const double a = (double)sg[16 * wave + cc][c + kk];
#pragma unroll
        for (int r = 0; r < NR; ++r)
          acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a, (double)sm[16 * r + cc][c + kk], acc[r], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (wave >= NQ) return;
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t g = gt + 16 * wave + kk + 4 * i;
      const int d = 16 * r + cc;
      if (g < g1 && d < D) gg[(g * B + b) * DP + d] += acc[r][i];
    }
}

template <int NR>
__global__ __launch_bounds__(256) void grouped_table_grad_kernel(
    int D, int B, int T, int DP, const int64_t *__restrict__ grp_ptr,
    const float *__restrict__ mu, const float *__restrict__ gsum,
    double *__restrict__ gg) {
  __shared__ float sg[64][kTgChunk + 1];
  __shared__ float sm[64][kTgChunk + 1];
  const int b = (int)(blockIdx.x % B);
  const int64_t j = blockIdx.x / B;
  const int64_t g0 = grp_ptr[j], g1 = grp_ptr[j + 1];
  const float *mrow = mu + (j * D * B + b) * (int64_t)T;  // (j, 0, b, 0)
  for (int64_t gt = g0; gt < g1; gt += 64) {
    const int64_t left = g1 - gt;  // uniform
    if (left > 48)
      table_grad_tile<4, NR>(D, B, T, DP, b, gt, g1, mrow, gsum, gg, sg, sm);
    else if (left > 32)
      table_grad_tile<3, NR>(D, B, T, DP, b, gt, g1, mrow, gsum, gg, sg, sm);
    else if (left > 16)
      table_grad_tile<2, NR>(D, B, T, DP, b, gt, g1, mrow, gsum, gg, sg, sm);
    else
      table_grad_tile<1, NR>(D, B, T, DP, b, gt, g1, mrow, gsum, gg, sg, sm);
  }
}

// nu on the FP64 matrix cores (16, 32 or 64 padded directions): per
// destination j and band b the product of the groups' BRDF rows
// [G_j x DP] with mu[j,:,b,:] [DP x T], K = DP in steps of 4.  A workgroup
// per (j, band, 4 x 16 NT bins), each wave 16 NT bins (NT = 2 at 64
// directions, so that the operands fit four waves per SIMD, else 4): the
// wave's mu operands (lane
// l: direction 4s + (l >> 4), bin l & 15 of each 16-bin tile) are loaded
// once into registers, then every tile of 16 groups is multiplied against
// them, its BRDF rows staged in LDS.  Lane layout of
// v_mfma_f64_16x16x4f64 as in grouped_table_grad_kernel.
template <int DP>
struct SpreadAdjMfma {
  static constexpr int NT = DP == 64 ? 2 : 4;  // 16-bin tiles per wave
  static constexpr int BINS = 4 * 16 * NT;      // bins per workgroup
};

template <int DP>
__global__ __launch_bounds__(256) void grouped_spread_adjoint_mfma_kernel(
    int D, int B, int T, int n_tblk, const int64_t *__restrict__ grp_ptr,
    const int32_t *__restrict__ grp_key, const double *__restrict__ rtab,
    const float *__restrict__ mu, float *__restrict__ nu) {
  constexpr int KS = DP / 4, NT = SpreadAdjMfma<DP>::NT;
  __shared__ double rts[16][DP + 1];
  const int64_t blk = blockIdx.x;
  const int tb = (int)(blk % n_tblk);
  const int b = (int)((blk / n_tblk) % B);
  const int64_t j = blk / ((int64_t)n_tblk * B);
  const int64_t g0 = grp_ptr[j], g1 = grp_ptr[j + 1];
  if (g0 == g1) return;  // uniform
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63, kk = lane >> 4, cc = lane & 15;
  const int tw = (tb * 4 + (tid >> 6)) * 16 * NT;
  const int64_t dstride = (int64_t)B * T;
  const float *m = mu + (j * D * B + b) * (int64_t)T;  // (j, 0, b, 0)
  float mv[KS][NT];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int d = 4 * s + kk, t = tw + 16 * nt + cc;
      const bool ok = d < D && t < T;
      const float v = m[ok ? d * dstride + t : 0];
      mv[s][nt] = ok ? v : 0.f;
    }
  for (int64_t gt = g0; gt < g1; gt += 16) {
    for (int idx = tid; idx < 16 * DP; idx += 256) {
      const int gi = idx / DP, d = idx % DP;
      rts[gi][d] = (gt + gi < g1 && d < D)
                       ? rtab[((int64_t)grp_key[gt + gi] * B + b) * DP + d]
                       : 0.0;
    }
    __syncthreads();
    mfma_d4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const double a = rts[cc][4 * s + kk];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
            a, (double)mv[s][nt], acc[nt], 0, 0, 0);
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t g = gt + kk + 4 * r;
        const int t = tw + 16 * nt + cc;
        if (g < g1 && t < T) nu[(g * B + b) * (int64_t)T + t] = (float)acc[nt][r];
      }
    __syncthreads();
  }
}

// muout[r,b,t] = lam[r,b,t] + sum_{q in T(r)} tscal[b,q] nu[tgrp_q,b,t+n_q]
// (float64 sum in transposed-entry order; a bin t + n beyond the histogram
// reads 0 through the row's buffer descriptor and adds nothing).  A
// workgroup per (source row r, band, 1,024 bins), 4 bins per lane, rows
// taken direction-major (r = i*D + d for d slow, i fast: neighbouring
// patches sending in one direction feed the same groups, whose nu rows
// they then share through L2); the entries are taken C at a time, every
// load of the chunk issued before any is summed.
__global__ __launch_bounds__(256) void grouped_gather_adjoint_kernel(
    int64_t P, int D, int B, int T, int n_tblk, int64_t Vd,
    const int64_t *__restrict__ trow_ptr,
    const int32_t *__restrict__ tgrp, const int32_t *__restrict__ tlag,
    const double *__restrict__ tscal, const float *__restrict__ nu,
    const float *__restrict__ lam, float *__restrict__ muout) {
  constexpr int R = 4, C = 4;
  const int64_t blk = blockIdx.x;
  const int tb = (int)(blk % n_tblk);
  const int b = (int)((blk / n_tblk) % B);
  const int64_t rd = blk / ((int64_t)n_tblk * B);  // d * P + i
  const int64_t r = (rd % P) * D + rd / P;
  const int64_t q0 = trow_ptr[r], q1 = trow_ptr[r + 1];
  const int64_t base = (r * B + b) * (int64_t)T;
  const uint32_t row_bytes = (uint32_t)T * 4u;
  int t[R];
  double s[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    t[i] = tb * 256 * R + (int)threadIdx.x + 256 * i;
    s[i] = t[i] < T ? (double)lam[base + t[i]] : 0.0;
  }
  for (int64_t q = q0; q < q1; q += C) {
    float v[C][R];
    double w[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const bool ok = q + c < q1;  // uniform
      const int64_t qq = ok ? q + c : q0;
      const int n = tlag[qq];
      w[c] = ok ? tscal[b * Vd + qq] : 0.0;
      const __amdgpu_buffer_rsrc_t rs = make_rsrc(
          nu + ((int64_t)tgrp[qq] * B + b) * (int64_t)T, ok ? row_bytes : 0u);
#pragma unroll
      for (int i = 0; i < R; ++i) v[c][i] = buffer_load(rs, bin_offset(t[i] + n));
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (q + c >= q1) break;  // uniform
#pragma unroll
      for (int i = 0; i < R; ++i) s[i] = fma(w[c], (double)v[c][i], s[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
    if (t[i] < T) muout[base + t[i]] = (float)s[i];
}

// gt[key_g,d,b] += gg[g,b,d] (float64 atomics: the groups of one BRDF row
// add in no fixed order).
__global__ void grouped_key_grad_kernel(int64_t G, int D, int B, int DP,
                                        const int32_t *__restrict__ grp_key,
                                        const double *__restrict__ gg,
                                        double *__restrict__ gt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= G * B * D) return;
  const int d = (int)(q % D);
  const int b = (int)((q / D) % B);
  const int64_t g = q / ((int64_t)D * B);
  unsafeAtomicAdd(gt + ((int64_t)grp_key[g] * D + d) * B + b,
                  gg[(g * B + b) * DP + d]);
}

// Injection part: E_0dir[p,d,b] = energy0[p,b] table[k0_p,d,b], placed at
// bin n_p, so gt[k0_p,d,b] += energy0[p,b] mu0[p,d,b,n_p].
__global__ void injection_table_grad_kernel(
    int64_t P, int64_t D, int64_t B, int64_t T,
    const double *__restrict__ energy0, const double *__restrict__ d0,
    double c, double dt, const int32_t *__restrict__ k0,
    const float *__restrict__ mu0, double *__restrict__ gt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P * D * B) return;
  const int64_t b = q % B, d = (q / B) % D, p = q / (D * B);
  const int64_t n = delay_floor(d0[p], c, dt);
  if (n < 0 || n >= T) return;
  unsafeAtomicAdd(gt + ((int64_t)k0[p] * D + d) * B + b,
                  energy0[p * B + b] * (double)mu0[q * T + n]);
}

}  // namespace
}  // namespace spr

using namespace spr;

extern "C" int64_t spr_padded_directions(int64_t D);  // exchange.hip

extern "C" int spr_ds_weights(int64_t B, int64_t Vd, const int32_t *krow,
                              const double *wgt, const double *a,
                              const double *cs, double *scal, double *dscal,
                              void *stream) {
  SPR_REQUIRE(B > 0 && Vd >= 0, "spr_ds_weights: bad sizes");
  SPR_REQUIRE(krow && wgt && a && cs && scal && dscal,
              "spr_ds_weights: null pointer");
  if (Vd == 0) return SPR_OK;
  hipLaunchKernelGGL(ds_weights_kernel, dim3((unsigned)ceil_div(B * Vd, 256)),
                     dim3(256), 0, as_stream(stream), B, Vd, krow, wgt, a, cs,
                     scal, dscal);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_ds_transposed_weights(int64_t B, int64_t Vd,
                                         const int32_t *te,
                                         const double *scal,
                                         const double *dscal, double *twa,
                                         double *twc, void *stream) {
  SPR_REQUIRE(B > 0 && Vd >= 0, "spr_ds_transposed_weights: bad sizes");
  SPR_REQUIRE(te && scal && dscal && twa && twc,
              "spr_ds_transposed_weights: null pointer");
  if (Vd == 0) return SPR_OK;
  hipLaunchKernelGGL(t_weights_kernel, dim3((unsigned)ceil_div(B * Vd, 256)),
                     dim3(256), 0, as_stream(stream), B, Vd, te, scal, dscal,
                     twa, twc);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_ds_injection(int64_t P, int64_t D, int64_t B,
                                const double *source, const double *centers,
                                const int64_t *wall_ids, int64_t S,
                                const int64_t *brdf_index,
                                const double *in_dirs, const double *energy0,
                                const double *a, const double *cs,
                                const int32_t *spec_dir, int32_t *k0,
                                double *e0dir, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0, "spr_ds_injection: bad sizes");
  SPR_REQUIRE(source && centers && energy0 && a && cs && spec_dir && k0 &&
                  e0dir,
              "spr_ds_injection: null pointer");
  SPR_REQUIRE(!in_dirs || (wall_ids && brdf_index && S > 0),
              "spr_ds_injection: in_dirs needs wall_ids, brdf_index");
  hipLaunchKernelGGL(ds_injection_kernel, dim3((unsigned)ceil_div(P, 128)),
                     dim3(128), 0, as_stream(stream), P, D, B, source, centers,
                     wall_ids, S, brdf_index, in_dirs, energy0, a, cs,
                     spec_dir, k0, e0dir);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_direction_sum(int64_t P, int64_t D, int64_t B, int64_t T,
                                 const float *mu, float *musum,
                                 void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0 && T > 0,
              "spr_direction_sum: bad sizes");
  SPR_REQUIRE(mu && musum, "spr_direction_sum: null pointer");
  hipLaunchKernelGGL(direction_sum_kernel,
                     dim3((unsigned)ceil_div(P * B * T, 256)), dim3(256), 0,
                     as_stream(stream), P, D, B, T, mu, musum);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_exchange_adjoint_ds(
    int64_t P, int64_t D, int64_t B, int64_t T, int64_t Vd,
    const int64_t *tgptr, const int32_t *tkey, const int32_t *tlag,
    const int32_t *tm, const double *twa, const double *twc,
    const float *lam, const float *musum, const float *munext,
    const float *ek, float *muout, double *ga, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && D <= 64 && B > 0 && T > 0 && Vd >= 0,
              "spr_exchange_adjoint_ds: bad sizes");
  SPR_REQUIRE(T < ((int64_t)1 << 29) && P < ((int64_t)1 << 25),
              "spr_exchange_adjoint_ds: histogram or scene too large");
  SPR_REQUIRE(tgptr && tkey && tlag && tm && twa && twc && lam && musum &&
                  munext && ek && muout && ga,
              "spr_exchange_adjoint_ds: null pointer");
  SPR_REQUIRE(muout != munext && muout != lam && muout != ek,
              "spr_exchange_adjoint_ds: muout aliases an input");
  // 8 bins per lane, 2 entries per chunk: each entry's two wave
  // reductions (its table-gradient dot products) then cover 512 bins.
  // Directional shoebox, order 30: backward 1.49 s with this shape; 1.66 s
  // with R = 4, C = 4, 1.76 s with R = 8, C = 1, 1.78 s with R = 16, C = 1
  // or R = 4, C = 2, 1.95 s with R = 16, C = 2, 2.68 s with R = 2, C = 4.
  constexpr int R = 8, C = 2;
  const int64_t nt = ceil_div(T, (int64_t)256 * R);
  const int64_t nb = P * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff, "spr_exchange_adjoint_ds: grid too "
                                        "large");
  hipStream_t s = as_stream(stream);
  SPR_HIP(hipMemcpyAsync(muout, lam, sizeof(float) * P * D * B * T,
                         hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL((adjoint_ds_kernel<R, C>), dim3((unsigned)nb), dim3(256),
                     0, s, (int)D, (int)B, (int)T, (int)nt, nb, Vd, tgptr,
                     tkey, tlag, tm, twa, twc, musum, munext, ek, muout, ga);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_ds_table_grad(int64_t B, int64_t Vd, const int32_t *te,
                                 const int32_t *krow, const double *wgt,
                                 const double *ga, double *gA, double *gC,
                                 void *stream) {
  SPR_REQUIRE(B > 0 && Vd >= 0, "spr_ds_table_grad: bad sizes");
  SPR_REQUIRE(te && krow && wgt && ga && gA && gC,
              "spr_ds_table_grad: null pointer");
  if (Vd == 0) return SPR_OK;
  hipLaunchKernelGGL(table_grad_kernel, dim3((unsigned)ceil_div(B * Vd, 256)),
                     dim3(256), 0, as_stream(stream), B, Vd, te, krow, wgt, ga,
                     gA, gC);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_ds_injection_grad(int64_t P, int64_t D, int64_t B,
                                     int64_t T, const double *energy0,
                                     const double *d0, double c, double dt,
                                     const int32_t *k0,
                                     const int32_t *spec_dir,
                                     const float *mu0, double *gA,
                                     double *gC, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0 && T > 0 && c > 0 && dt > 0,
              "spr_ds_injection_grad: bad sizes");
  SPR_REQUIRE(energy0 && d0 && k0 && spec_dir && mu0 && gA && gC,
              "spr_ds_injection_grad: null pointer");
  hipLaunchKernelGGL(injection_grad_kernel,
                     dim3((unsigned)ceil_div(P * B, 256)), dim3(256), 0,
                     as_stream(stream), P, D, B, T, energy0, d0, c, dt, k0,
                     spec_dir, mu0, gA, gC);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_grouped_spread_adjoint(int64_t P, int64_t D, int64_t B,
                                          int64_t T, const int64_t *grp_ptr,
                                          const int32_t *grp_key,
                                          const double *rtab, const float *mu,
                                          float *nu, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(P > 0 && DP > 0 && B > 0 && T > 0 && T < ((int64_t)1 << 29),
              "spr_grouped_spread_adjoint: bad sizes");
  SPR_REQUIRE(grp_ptr && grp_key && rtab && mu && nu,
              "spr_grouped_spread_adjoint: null pointer");
  const int64_t nt = ceil_div(T, (int64_t)256);
  const int64_t nb = P * B * nt;
  SPR_REQUIRE(2 * nb < (int64_t)0x7fffffff,  // 128-bin blocks at D = 64
              "spr_grouped_spread_adjoint: grid too large");
  hipStream_t s = as_stream(stream);
#define SPR_NU_MFMA(DPV)                                                      \
  hipLaunchKernelGGL(                                                         \
      (grouped_spread_adjoint_mfma_kernel<DPV>),                              \
      dim3((unsigned)(P * B * ceil_div(T, (int64_t)SpreadAdjMfma<DPV>::BINS))), \
      dim3(256), 0, s, (int)D, (int)B, (int)T,                                \
      (int)ceil_div(T, (int64_t)SpreadAdjMfma<DPV>::BINS), grp_ptr, grp_key,  \
      rtab, mu, nu)
  if (DP == 16)
    SPR_NU_MFMA(16);
  else if (DP == 32)
    SPR_NU_MFMA(32);
  else if (DP == 64)
    SPR_NU_MFMA(64);
  else  // up to 8 directions: the vector kernel
    hipLaunchKernelGGL((grouped_spread_adjoint_kernel<8>), dim3((unsigned)nb),
                       dim3(256), 0, s, (int)D, (int)B, (int)T, (int)DP,
                       (int)nt, grp_ptr, grp_key, rtab, mu, nu);
#undef SPR_NU_MFMA
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_grouped_table_grad(int64_t P, int64_t D, int64_t B,
                                      int64_t T, const int64_t *grp_ptr,
                                      const float *mu, const float *gsum,
                                      double *gg, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(P > 0 && DP > 0 && B > 0 && T > 0 && T < ((int64_t)1 << 29),
              "spr_grouped_table_grad: bad sizes");
  SPR_REQUIRE(grp_ptr && mu && gsum && gg,
              "spr_grouped_table_grad: null pointer");
  SPR_REQUIRE(P * B < (int64_t)0x7fffffff,
              "spr_grouped_table_grad: grid too large");
  hipStream_t s = as_stream(stream);
#define SPR_TG(NRV)                                                           \
  hipLaunchKernelGGL((grouped_table_grad_kernel<NRV>),                        \
                     dim3((unsigned)(P * B)), dim3(256), 0, s, (int)D,        \
                     (int)B, (int)T, (int)DP, grp_ptr, mu, gsum, gg)
  if (D <= 16)
    SPR_TG(1);
  else if (D <= 32)
    SPR_TG(2);
  else if (D <= 48)
    SPR_TG(3);
  else
    SPR_TG(4);
#undef SPR_TG
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_grouped_gather_adjoint(
    int64_t P, int64_t D, int64_t B, int64_t T, int64_t Vd,
    const int64_t *trow_ptr, const int32_t *tgrp, const int32_t *tlag,
    const double *tscal, const float *nu, const float *lam, float *muout,
    void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0 && T > 0 && Vd >= 0 &&
                  T < ((int64_t)1 << 29),
              "spr_grouped_gather_adjoint: bad sizes");
  SPR_REQUIRE(trow_ptr && tgrp && tlag && tscal && nu && lam && muout,
              "spr_grouped_gather_adjoint: null pointer");
  SPR_REQUIRE(muout != lam, "spr_grouped_gather_adjoint: muout aliases lam");
  const int64_t nt = ceil_div(T, (int64_t)1024);
  const int64_t nb = P * D * B * nt;
  SPR_REQUIRE(nb < (int64_t)0x7fffffff,
              "spr_grouped_gather_adjoint: grid too large");
  hipLaunchKernelGGL(grouped_gather_adjoint_kernel, dim3((unsigned)nb),
                     dim3(256), 0, as_stream(stream), P, (int)D, (int)B,
                     (int)T, (int)nt, Vd, trow_ptr, tgrp, tlag, tscal, nu,
                     lam, muout);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_grouped_key_grad(int64_t G, int64_t D, int64_t B,
                                    const int32_t *grp_key, const double *gg,
                                    double *gt, void *stream) {
  const int64_t DP = spr_padded_directions(D);
  SPR_REQUIRE(G >= 0 && DP > 0 && B > 0, "spr_grouped_key_grad: bad sizes");
  SPR_REQUIRE(grp_key && gg && gt, "spr_grouped_key_grad: null pointer");
  if (G == 0) return SPR_OK;
  hipLaunchKernelGGL(grouped_key_grad_kernel,
                     dim3((unsigned)ceil_div(G * B * D, 256)), dim3(256), 0,
                     as_stream(stream), G, (int)D, (int)B, (int)DP, grp_key,
                     gg, gt);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_injection_table_grad(int64_t P, int64_t D, int64_t B,
                                        int64_t T, const double *energy0,
                                        const double *d0, double c, double dt,
                                        const int32_t *k0, const float *mu0,
                                        double *gt, void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && B > 0 && T > 0 && c > 0 && dt > 0,
              "spr_injection_table_grad: bad sizes");
  SPR_REQUIRE(energy0 && d0 && k0 && mu0 && gt,
              "spr_injection_table_grad: null pointer");
  hipLaunchKernelGGL(injection_table_grad_kernel,
                     dim3((unsigned)ceil_div(P * D * B, 256)), dim3(256), 0,
                     as_stream(stream), P, D, B, T, energy0, d0, c, dt, k0,
                     mu0, gt);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}
