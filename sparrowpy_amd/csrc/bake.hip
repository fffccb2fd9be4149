// Bake kernels: visibility, visible-pair list, universal form factors,
// outgoing-direction maps, the dense directional form factor and source
// injection.  All geometry is float64 (see spr_common.hpp on rounding).
//
// Reference: sparrowpy/geometry.py (visibility), sparrowpy/form_factor/
// universal.py and integration.py (form factors, point-patch factors),
// sparrowpy/classes/RadiosityFast.py (bake_geometry, init_source_energy).
#include "spr_common.hpp"
#include "spr_fft.hpp"
#include "spr_geom.hpp"
#include "spr_util.hpp"

namespace spr {
namespace {

// ---------------------------------------------------------------------
// visibility
// ---------------------------------------------------------------------
__global__ void blocker_kernel(int64_t NS, const double *__restrict__ normals,
                               const double *__restrict__ pts,
                               BlockerRec *__restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NS) return;
  BlockerRec r;
  make_blocker(pts + 12 * k, normals + 3 * k, r);
  out[k] = r;
}

// Thread per pair (i, j) of a 16x16 tile.  The workgroup first culls the
// blockers for the whole tile (may_block_tile over the boxes of its 16 row
// and 16 column patch centres) into a candidate list in LDS, 256 blockers
// at a time; every lane then walks the candidates (a uniform loop, so
// records arrive through scalar loads) until its pair is blocked.  The cull
// only drops blockers that cannot block any pair of the tile, so the flags
// are those of the full scan (canyon, 5,122 patches as blockers: ~1,500
// candidates per tile, 105 ms instead of 277 ms).
__global__ __launch_bounds__(256) void p2p_visibility_kernel(
    int64_t P, const double *__restrict__ centers, int64_t NS,
    const BlockerRec *__restrict__ recs, uint8_t *__restrict__ vis) {
  __shared__ int32_t cand[256];
  __shared__ int32_t wave_count[4];
  __shared__ double box[12];
  const int64_t ti = blockIdx.y, tj = blockIdx.x;
  const int64_t i = ti * 16 + (threadIdx.x >> 4);
  const int64_t j = tj * 16 + (threadIdx.x & 15);
  if (ti > tj) {  // tile entirely below the diagonal
    if (i < P && j < P) vis[i * P + j] = 0;
    return;
  }
  const bool valid = i < P && j < P;
  bool alive = valid && i < j;
  d3 a{0, 0, 0}, b{0, 0, 0};
  if (alive) {
    a = load3(centers + 3 * i);
    b = load3(centers + 3 * j);
  }
  if (threadIdx.x < 2) {
    // bounding box of the tile's row (i) or column (j) patch centres
    const int64_t t = threadIdx.x == 0 ? ti : tj;
    d3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (int q = 0; q < 16; ++q) {
      const int64_t k = t * 16 + q;
      if (k >= P) break;
      const d3 c = load3(centers + 3 * k);
      lo = d3{fmin(lo.x, c.x), fmin(lo.y, c.y), fmin(lo.z, c.z)};
      hi = d3{fmax(hi.x, c.x), fmax(hi.y, c.y), fmax(hi.z, c.z)};
    }
    double *o = box + 6 * threadIdx.x;
    o[0] = lo.x, o[1] = lo.y, o[2] = lo.z, o[3] = hi.x, o[4] = hi.y,
    o[5] = hi.z;
  }
  __syncthreads();
  const d3 ilo{box[0], box[1], box[2]}, ihi{box[3], box[4], box[5]};
  const d3 jlo{box[6], box[7], box[8]}, jhi{box[9], box[10], box[11]};
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  for (int64_t k0 = 0; k0 < NS; k0 += 256) {
    const int64_t kk = k0 + (int64_t)threadIdx.x;
    const bool keep = kk < NS && may_block_tile(ilo, ihi, jlo, jhi, recs[kk]);
    const uint64_t m = __ballot(keep);
    if (lane == 0) wave_count[wave] = __popcll(m);
    __syncthreads();
    int base = 0, n_cand = 0;
    for (int w = 0; w < 4; ++w) {
      base += w < wave ? wave_count[w] : 0;
      n_cand += wave_count[w];
    }
    if (keep)
      cand[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)kk;
    __syncthreads();
    for (int c = 0; c < n_cand; ++c) {
      if (!__any(alive)) break;
      const int k = __builtin_amdgcn_readfirstlane(cand[c]);
      if (alive) alive = basic_visibility(a, b, recs[k]);
    }
    if (!__syncthreads_or(alive)) break;
  }
  if (valid) vis[i * P + j] = (uint8_t)alive;
}

// Thread per (query point q, patch k); walls as blockers.
__global__ __launch_bounds__(256) void point_visibility_kernel(
    int64_t NQ, const double *__restrict__ points, int64_t P,
    const double *__restrict__ centers, int64_t NS,
    const BlockerRec *__restrict__ recs, uint8_t *__restrict__ vis) {
  const int64_t q = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool alive = k < P;
  d3 a = load3(points + 3 * q), b{0, 0, 0};
  if (alive) b = load3(centers + 3 * k);
  for (int64_t s = 0; s < NS; ++s) {
    if (!__any(alive)) break;
    if (alive) alive = basic_visibility(a, b, recs[s]);
  }
  if (k < P) vis[q * P + k] = (uint8_t)alive;
}

// ---------------------------------------------------------------------
// visible-pair list (RadiosityFast.py:377-387), one wave per row i
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void visible_pairs_kernel(
    int64_t P, const uint8_t *__restrict__ vis, int64_t *__restrict__ offs,
    int32_t *__restrict__ pairs, int fill) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= P) return;
  int64_t pos = fill ? offs[i] : 0;
  for (int64_t base = i + 1; base < P; base += 64) {
    int64_t j = base + lane;
    bool f = j < P && vis[i * P + j] != 0;
    uint64_t m = __ballot(f);
    if (fill && f) {
      uint64_t lt = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
      int64_t p = pos + __popcll(lt);
      pairs[2 * p] = (int32_t)i;
      pairs[2 * p + 1] = (int32_t)j;
    }
    pos += __popcll(m);
  }
  if (!fill && lane == 0) offs[i] = pos;
}

// ---------------------------------------------------------------------
// universal form factor (universal.py:54-96)
// ---------------------------------------------------------------------

// geometry.py:719-748 _coincidence_check(receiver, source), threshold 1e-6
__device__ inline bool coincident(const double *p0, const double *p1) {
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      if (norm(load3(p0 + 3 * a) - load3(p1 + 3 * b)) < 1e-6) return true;
  return false;
}

// integration.py:608-651 _sample_boundary_regular, point q of a 4-corner
// patch sampled with n_div segments per side.
__device__ inline d3 boundary_point(const double *el, int q, int n_div) {
  int i = q / n_div, ii = q % n_div;
  d3 a = load3(el + 3 * i), b = load3(el + 3 * ((i + 1) & 3));
  double t = (double)ii, nd = (double)n_div;
  return d3{a.x + (t * (b.x - a.x)) / nd, a.y + (t * (b.y - a.y)) / nd,
            a.z + (t * (b.z - a.z)) / nd};
}

__device__ inline double coord(d3 p, int dim) {
  return dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
}

// integration.py:462-486 _newton_cotes_4th (Boole's rule)
__device__ inline double boole(const double *x, const double *y) {
  double h = x[1] - x[0];
  return ((2.0 * h) / 45.0) *
         ((((7.0 * y[0] + 32.0 * y[1]) + 12.0 * y[2]) + 32.0 * y[3]) +
          7.0 * y[4]);
}

// integration.py:38-114 stokes_integration, one wavefront per pair:
// lanes fill the 16x16 ln|r| table, 48 lanes integrate over patch j's
// boundary (one per sample point of i and coordinate), lane 0 integrates
// over patch i's boundary in the reference's (dim, segment) order.
__global__ __launch_bounds__(64) void stokes_kernel(
    int64_t P, const double *__restrict__ pts, const double *__restrict__ areas,
    int64_t V, const int32_t *__restrict__ pairs, double *__restrict__ ff) {
  __shared__ double form[16][16];
  __shared__ double inner[16][3];
  const int64_t v = blockIdx.x;
  const int64_t i = pairs[2 * v], j = pairs[2 * v + 1];
  if (i < 0 || j < 0 || i >= P || j >= P) return;  // invalid pair: skip
  const double *pi = pts + 12 * i, *pj = pts + 12 * j;
  if (coincident(pj, pi)) return;  // handled by nusselt_kernel
  const int lane = threadIdx.x;
  for (int q = lane; q < 256; q += 64) {
    int a = q >> 4, b = q & 15;
    form[a][b] = log(norm(boundary_point(pi, a, 4) - boundary_point(pj, b, 4)));
  }
  __syncthreads();
  if (lane < 48) {
    int a = lane / 3, dim = lane % 3;
    double acc = 0.0, x[5], y[5];
    for (int s = 0; s < 4; ++s) {
      for (int k = 0; k < 5; ++k) {
        int q = (4 * s + k) & 15;
        x[k] = coord(boundary_point(pj, q, 4), dim);
      }
      if (fabs(x[4] - x[0]) > 1e-3) {
        for (int k = 0; k < 5; ++k) y[k] = form[a][(4 * s + k) & 15];
        acc += boole(x, y);
      }
    }
    inner[a][dim] = acc;
  }
  __syncthreads();
  if (lane == 0) {
    double outer = 0.0, x[5], y[5];
    for (int dim = 0; dim < 3; ++dim)
      for (int s = 0; s < 4; ++s) {
        for (int k = 0; k < 5; ++k)
          x[k] = coord(boundary_point(pi, (4 * s + k) & 15, 4), dim);
        if (fabs(x[4] - x[0]) > 1e-3) {
          for (int k = 0; k < 5; ++k) y[k] = inner[(4 * s + k) & 15][dim];
          outer += boole(x, y);
        }
      }
    ff[i * P + j] = fabs(outer / ((2.0 * M_PI) * areas[i]));
  }
}

// 3x3 inverse by elimination with partial pivoting, column-wise solve of
// A X = I (np.linalg.inv's LU route).
__device__ inline void inv3(const double (&a_in)[3][3], double (&out)[3][3]) {
  double a[3][3];
  int piv[3] = {0, 1, 2};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = a_in[r][c];
  for (int k = 0; k < 3; ++k) {
    int p = k;
    double best = fabs(a[k][k]);
    for (int r = k + 1; r < 3; ++r)
      if (fabs(a[r][k]) > best) {
        best = fabs(a[r][k]);
        p = r;
      }
    if (p != k) {
      for (int c = 0; c < 3; ++c) {
        double t = a[k][c];
        a[k][c] = a[p][c];
        a[p][c] = t;
      }
      int t = piv[k];
      piv[k] = piv[p];
      piv[p] = t;
    }
    for (int r = k + 1; r < 3; ++r) {
      a[r][k] = a[r][k] / a[k][k];
      for (int c = k + 1; c < 3; ++c) a[r][c] = a[r][c] - a[r][k] * a[k][c];
    }
  }
  for (int col = 0; col < 3; ++col) {
    double y[3];
    for (int r = 0; r < 3; ++r) {
      double s = (piv[r] == col) ? 1.0 : 0.0;
      for (int c = 0; c < r; ++c) s = s - a[r][c] * y[c];
      y[r] = s;
    }
    for (int r = 2; r >= 0; --r) {
      double s = y[r];
      for (int c = r + 1; c < 3; ++c) s = s - a[r][c] * out[c][col];
      out[r][col] = s / a[r][r];
    }
  }
}

// integration.py:416-460 _area_under_curve (order 2) with
// _poly_estimation_Lagrange (:349-382) and _poly_integration (:385-412).
__device__ inline double area_under_curve(d2 p0, d2 p1, d2 p2) {
  d2 f = p2 - p0;
  double nf = norm(f);
  d2 r0{f.x / nf, f.y / nf};
  d2 r1{-f.y / nf, f.x / nf};
  double x[3] = {0.0, 0.0, 0.0}, y[3] = {0.0, 0.0, 0.0};
  d2 c1 = p1 - p0, c2 = p2 - p0;
  x[1] = dot(r0, c1);
  y[1] = dot(r1, c1);
  x[2] = dot(r0, c2);
  y[2] = dot(r1, c2);
  double co[3];
  if (fabs(x[2] - x[0]) < 1e-6) {
    co[0] = co[1] = co[2] = 0.0;
  } else {
    double xm[3][3], inv[3][3];
    for (int r = 0; r < 3; ++r)
      for (int o = 0; o < 3; ++o) xm[r][2 - o] = pow(x[r], (double)o);
    inv3(xm, inv);
    for (int r = 0; r < 3; ++r)
      co[r] = (inv[r][0] * y[0] + inv[r][1] * y[1]) + inv[r][2] * y[2];
  }
  double out = 0.0;
  for (int r = 0; r < 3; ++r) {
    double p = (double)(3 - r);
    out += (co[r] * pow(x[2], p)) / p;
    out -= (co[r] * pow(x[0], p)) / p;
  }
  return out;
}

__device__ inline d2 project2(const double *m0, const double *m1, d3 v) {
  return d2{dot(load3(m0), v), dot(load3(m1), v)};
}

// integration.py:116-230 nusselt_analog
__device__ double nusselt_analog(d3 origin, const double *m0,
                                 const double *m1, const double *pj,
                                 const double *nj) {
  d3 sph[8];
  d2 pln[8];
  for (int q = 0; q < 8; ++q) {
    d3 d = boundary_point(pj, q, 2) - origin;
    sph[q] = div(d, norm(d));
    pln[q] = project2(m0, m1, sph[q]);
  }
  d3 e1 = load3(pj + 3) - load3(pj), e2 = load3(pj + 6) - load3(pj + 3);
  double hand = sgn(dot(cross(e1, e2), load3(nj)));
  // _polygon_area of the projected corners (samples 0, 2, 4, 6; z = 0)
  double big_poly = 0.0;
  {
    d3 q0{pln[0].x, pln[0].y, 0.0};
    for (int t = 0; t < 2; ++t) {
      d3 qa{pln[2 * (t + 1)].x, pln[2 * (t + 1)].y, 0.0};
      d3 qb{pln[2 * (t + 2)].x, pln[2 * (t + 2)].y, 0.0};
      big_poly += 0.5 * norm(cross(qa - q0, qb - q0));
    }
  }
  double curved = 0.0;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    int s0 = 2 * jj, s1 = 2 * jj + 1, s2 = (2 * jj + 2) & 7;
    d3 pa{pln[s2].x, pln[s2].y, 0.0}, pb{pln[s0].x, pln[s0].y, 0.0};
    if (!(norm(cross(pa, pb)) > 1e-6)) continue;
    if (dot(pln[s2], pln[s0]) >= 1e-6) {
      curved += area_under_curve(pln[s0], pln[s1], pln[s2]);
    } else {
      d3 mpoint{sph[s0].x + (sph[s2].x - sph[s0].x) / 2.0,
                sph[s0].y + (sph[s2].y - sph[s0].y) / 2.0,
                sph[s0].z + (sph[s2].z - sph[s0].z) / 2.0};
      d3 marc = div(mpoint, norm(mpoint));
      d3 a{sph[s0].x + (marc.x - sph[s0].x) / 2.0,
           sph[s0].y + (marc.y - sph[s0].y) / 2.0,
           sph[s0].z + (marc.z - sph[s0].z) / 2.0};
      d3 b{marc.x + (sph[s2].x - marc.x) / 2.0,
           marc.y + (sph[s2].y - marc.y) / 2.0,
           marc.z + (sph[s2].z - marc.z) / 2.0};
      d2 mp2 = project2(m0, m1, mpoint);
      d2 marc2 = project2(m0, m1, marc);
      d2 a2 = project2(m0, m1, div(a, norm(a)));
      d2 b2 = project2(m0, m1, div(b, norm(b)));
      double lin = (norm(pln[s2] - pln[s0]) * norm(mp2 - marc2)) / 2.0;
      double left = area_under_curve(pln[s0], a2, marc2);
      double right = area_under_curve(marc2, b2, pln[s2]);
      curved += ((lin * sgn(left)) + left) + right;
    }
  }
  return big_poly + hand * curved;
}

// numpy.linspace(0, stop, num)[k]
__device__ inline double linspace0(double stop, int num, int k) {
  if (num == 1) return 0.0;
  if (k == num - 1) return stop;
  return (double)k * (stop / (double)(num - 1)) + 0.0;
}

// integration.py:232-289 nusselt_integration over the regular grid of
// integration.py:533-605; one wavefront per pair, one lane per sample,
// samples summed in the reference's order by lane 0.
__global__ __launch_bounds__(64) void nusselt_kernel(
    int64_t P, const double *__restrict__ pts,
    const double *__restrict__ normals, int64_t V,
    const int32_t *__restrict__ pairs, double *__restrict__ ff) {
  __shared__ double vals[64];
  __shared__ double m[6];
  const int64_t v = blockIdx.x;
  const int64_t i = pairs[2 * v], j = pairs[2 * v + 1];
  if (i < 0 || j < 0 || i >= P || j >= P) return;  // invalid pair: skip
  const double *pi = pts + 12 * i, *pj = pts + 12 * j;
  if (!coincident(pj, pi)) return;  // handled by stokes_kernel
  const int lane = threadIdx.x;
  if (lane == 0) rotation_rows(normals + 3 * i, m, m + 3);
  d3 u = load3(pi + 3) - load3(pi), w = load3(pi + 9) - load3(pi);
  double nu = norm(u), nw = norm(w);
  double sq = sqrt(1.0 * 64.0);
  int npx = (int)rint(nu / nw * sq);
  int npz = (int)rint(nw / nu * sq);
  if (npz == 0) npz = 1;
  if (npx == 0) npx = 1;
  const int n = npx * npz;
  const double stx = 1.0 - 1.0 / (double)npx, stz = 1.0 - 1.0 / (double)npz;
  const double ssx = 1.0 / (double)(npx * 2), ssz = 1.0 / (double)(npz * 2);
  __syncthreads();
  double out = 0.0;
  for (int base = 0; base < n; base += 64) {
    int q = base + lane;
    if (q < n) {
      int a = q / npz, b = q % npz;
      double s = linspace0(stx, npx, a) + ssx;
      double t = linspace0(stz, npz, b) + ssz;
      d3 p0{(s * u.x + t * w.x) + pi[0], (s * u.y + t * w.y) + pi[1],
            (s * u.z + t * w.z) + pi[2]};
      vals[lane] = nusselt_analog(p0, m, m + 3, pj, normals + 3 * j);
    }
    __syncthreads();
    if (lane == 0) {
      int cnt = n - base < 64 ? n - base : 64;
      for (int k = 0; k < cnt; ++k) out += vals[k];
    }
    __syncthreads();
  }
  if (lane == 0) {
    out *= 1.0 / (M_PI * (double)n);
    ff[i * P + j] = out;
  }
}

// ---------------------------------------------------------------------
// direction maps
// ---------------------------------------------------------------------
__global__ void direction_map_kernel(int64_t P, const uint8_t *__restrict__ vis,
                                     const double *__restrict__ centers,
                                     int64_t D,
                                     const double *__restrict__ out_dirs,
                                     const int64_t *__restrict__ wall_ids,
                                     int64_t *__restrict__ p2o) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j >= P) return;
  int64_t r = D;
  if (i != j && (i < j ? vis[i * P + j] : vis[j * P + i])) {
    d3 d = load3(centers + 3 * j) - load3(centers + 3 * i);
    d = div(d, norm(d));
    r = nearest_direction(out_dirs + wall_ids[i] * D * 3, D, d);
  }
  p2o[i * P + j] = r;
}

__global__ void point_direction_kernel(int64_t NQ,
                                       const double *__restrict__ points,
                                       int64_t P,
                                       const double *__restrict__ centers,
                                       int64_t D,
                                       const double *__restrict__ out_dirs,
                                       const int64_t *__restrict__ wall_ids,
                                       int64_t *__restrict__ idx) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (k >= P) return;
  d3 d = load3(points + 3 * q) - load3(centers + 3 * k);
  d = div(d, norm(d));
  idx[q * P + k] = nearest_direction(out_dirs + wall_ids[k] * D * 3, D, d);
}

// ---------------------------------------------------------------------
// dense directional form factor (RadiosityFast.py:1234-1272)
// ---------------------------------------------------------------------
__global__ void fft_dense_kernel(int64_t P, const uint8_t *__restrict__ vis,
                                 const double *__restrict__ ff, int64_t B,
                                 const double *__restrict__ centers,
                                 const double *__restrict__ areas,
                                 const double *__restrict__ air,
                                 const int64_t *__restrict__ wall_ids,
                                 int64_t S, int64_t D,
                                 const double *__restrict__ brdf,
                                 const int64_t *__restrict__ brdf_index,
                                 const double *__restrict__ in_dirs,
                                 double *__restrict__ fft) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j >= P) return;
  double *o = fft + (i * P + j) * D * B;
  bool visible = i != j && (i < j ? vis[i * P + j] : vis[j * P + i]);
  if (!visible) {
    for (int64_t k = 0; k < D * B; ++k) o[k] = 0.0;
    return;
  }
  PairFactor pf = pair_factor(i, j, P, centers, areas, ff, wall_ids, S, D, B,
                              brdf, brdf_index, in_dirs);
  for (int64_t b = 0; b < B; ++b) {
    double ab = air ? exp(-air[b] * pf.distance) : 1.0;
    for (int64_t d = 0; d < D; ++d)
      o[d * B + b] = pair_coef(pf, ab, air != nullptr, d, b, B);
  }
}

// ---------------------------------------------------------------------
// source injection (universal.py:98-147, RadiosityFast.py:988-1034)
// ---------------------------------------------------------------------
__global__ void source_energy_kernel(
    const double *__restrict__ source, int64_t P,
    const double *__restrict__ centers, const double *__restrict__ pts,
    const uint8_t *__restrict__ vis, const double *__restrict__ air,
    int64_t B, const int64_t *__restrict__ wall_ids, int64_t S, int64_t D,
    const double *__restrict__ brdf, const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs, double *__restrict__ energy0,
    double *__restrict__ e0dir, double *__restrict__ distance) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const d3 src = load3(source);
  const d3 ck = load3(centers + 3 * k);
  double dist = 0.0, pt = 0.0;
  const bool v = vis[k] != 0;
  if (v) {
    dist = norm(src - ck);
    pt = pt_solution(src, pts + 12 * k, false);
  }
  distance[k] = dist;
  // get_scattering_data_source(source, c_k): nearest incoming direction of
  // wall w_k to unit(source - c_k)
  d3 u = src - ck;
  u = div(u, norm(u));
  const int64_t w = wall_ids[k];
  const int64_t s = nearest_direction(in_dirs + w * S * 3, S, u);
  const double *row = brdf + (brdf_index[w] * S + s) * D * B;
  for (int64_t b = 0; b < B; ++b) {
    double e = 0.0;
    if (v) e = air ? exp(-air[b] * dist) * pt : pt;
    energy0[k * B + b] = e;
    for (int64_t d = 0; d < D; ++d)
      e0dir[(k * D + d) * B + b] = e * row[d * B + b];
  }
}

}  // namespace
}  // namespace spr

using namespace spr;

extern "C" int64_t spr_visibility_workspace(int64_t NS) {
  return NS < 0 ? -1 : (int64_t)sizeof(BlockerRec) * (NS > 0 ? NS : 1);
}

static int build_blockers(int64_t NS, const double *normals,
                          const double *pts, void *ws, int64_t ws_bytes,
                          hipStream_t s) {
  if (ws_bytes < spr_visibility_workspace(NS))
    return fail(SPR_ERR_WORKSPACE, "visibility workspace of %lld bytes, "
                "need %lld", (long long)ws_bytes,
                (long long)spr_visibility_workspace(NS));
  if (NS == 0) return SPR_OK;
  hipLaunchKernelGGL(blocker_kernel, dim3((unsigned)ceil_div(NS, 128)),
                     dim3(128), 0, s, NS, normals, pts,
                     static_cast<BlockerRec *>(ws));
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_p2p_visibility(int64_t P, const double *centers,
                                  int64_t NS, const double *surf_normals,
                                  const double *surf_points, uint8_t *vis,
                                  void *workspace, int64_t workspace_bytes,
                                  void *stream) {
  SPR_REQUIRE(P > 0 && NS >= 0 && P < (int64_t)1 << 20,
              "spr_p2p_visibility: bad sizes");
  SPR_REQUIRE(centers && vis && workspace && (NS == 0 || (surf_normals &&
                                                          surf_points)),
              "spr_p2p_visibility: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = build_blockers(NS, surf_normals, surf_points, workspace,
                          workspace_bytes, s);
  if (rc) return rc;
  unsigned nt = (unsigned)ceil_div(P, 16);
  hipLaunchKernelGGL(p2p_visibility_kernel, dim3(nt, nt), dim3(256), 0, s, P,
                     centers, NS, static_cast<const BlockerRec *>(workspace),
                     vis);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_point_visibility(int64_t NQ, const double *points,
                                    int64_t P, const double *centers,
                                    int64_t NS, const double *surf_normals,
                                    const double *surf_points, uint8_t *vis,
                                    void *workspace, int64_t workspace_bytes,
                                    void *stream) {
  SPR_REQUIRE(NQ > 0 && P > 0 && NS >= 0 && NQ < 65536,
              "spr_point_visibility: bad sizes");
  SPR_REQUIRE(points && centers && vis && workspace &&
                  (NS == 0 || (surf_normals && surf_points)),
              "spr_point_visibility: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = build_blockers(NS, surf_normals, surf_points, workspace,
                          workspace_bytes, s);
  if (rc) return rc;
  hipLaunchKernelGGL(point_visibility_kernel,
                     dim3((unsigned)ceil_div(P, 256), (unsigned)NQ), dim3(256),
                     0, s, NQ, points, P, centers, NS,
                     static_cast<const BlockerRec *>(workspace), vis);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_visible_pairs(int64_t P, const uint8_t *vis,
                                 int64_t *row_offsets, int32_t *pairs,
                                 void *stream) {
  SPR_REQUIRE(P > 0 && vis && row_offsets, "spr_visible_pairs: bad args");
  hipStream_t s = as_stream(stream);
  unsigned blocks = (unsigned)ceil_div(P, 4);
  if (!pairs) {
    hipLaunchKernelGGL(visible_pairs_kernel, dim3(blocks), dim3(256), 0, s, P,
                       vis, row_offsets, (int32_t *)nullptr, 0);
    SPR_LAUNCH_CHECK();
    return scan_exclusive(row_offsets, P, s);
  }
  hipLaunchKernelGGL(visible_pairs_kernel, dim3(blocks), dim3(256), 0, s, P,
                     vis, row_offsets, pairs, 1);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_form_factors(int64_t P, const double *points,
                                const double *normals, const double *areas,
                                int64_t V, const int32_t *pairs, double *ff,
                                void *stream) {
  SPR_REQUIRE(P > 0 && V >= 0 && V < 0x7fffffff,
              "spr_form_factors: bad sizes");
  SPR_REQUIRE(points && normals && areas && ff && (V == 0 || pairs),
              "spr_form_factors: null pointer");
  hipStream_t s = as_stream(stream);
  int rc = fill<double>(ff, P * P, 0.0, s);
  if (rc || V == 0) return rc;
  hipLaunchKernelGGL(stokes_kernel, dim3((unsigned)V), dim3(64), 0, s, P,
                     points, areas, V, pairs, ff);
  SPR_LAUNCH_CHECK();
  hipLaunchKernelGGL(nusselt_kernel, dim3((unsigned)V), dim3(64), 0, s, P,
                     points, normals, V, pairs, ff);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_direction_map(int64_t P, const uint8_t *vis,
                                 const double *centers, int64_t D,
                                 const double *out_dirs,
                                 const int64_t *wall_ids, int64_t *p2o,
                                 void *stream) {
  SPR_REQUIRE(P > 0 && D > 0 && P < 65536, "spr_direction_map: bad sizes");
  SPR_REQUIRE(vis && centers && out_dirs && wall_ids && p2o,
              "spr_direction_map: null pointer");
  hipLaunchKernelGGL(direction_map_kernel,
                     dim3((unsigned)ceil_div(P, 256), (unsigned)P), dim3(256),
                     0, as_stream(stream), P, vis, centers, D, out_dirs,
                     wall_ids, p2o);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_point_direction_index(int64_t NQ, const double *points,
                                         int64_t P, const double *centers,
                                         int64_t D, const double *out_dirs,
                                         const int64_t *wall_ids,
                                         int64_t *idx, void *stream) {
  SPR_REQUIRE(NQ > 0 && P > 0 && D > 0 && NQ < 65536,
              "spr_point_direction_index: bad sizes");
  SPR_REQUIRE(points && centers && out_dirs && wall_ids && idx,
              "spr_point_direction_index: null pointer");
  hipLaunchKernelGGL(point_direction_kernel,
                     dim3((unsigned)ceil_div(P, 256), (unsigned)NQ),
                     dim3(256), 0, as_stream(stream), NQ, points, P, centers,
                     D, out_dirs, wall_ids, idx);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_fft_dense(int64_t P, const uint8_t *vis, const double *ff,
                             int64_t B, const double *centers,
                             const double *areas, const double *air,
                             const int64_t *wall_ids, int64_t S, int64_t D,
                             const double *brdf, const int64_t *brdf_index,
                             const double *in_dirs, double *fft,
                             void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && D > 0 && P < 65536,
              "spr_fft_dense: bad sizes");
  SPR_REQUIRE(vis && ff && centers && areas && wall_ids && fft,
              "spr_fft_dense: null pointer");
  SPR_REQUIRE(!brdf || (brdf_index && in_dirs && S > 0),
              "spr_fft_dense: brdf needs brdf_index and in_dirs");
  SPR_REQUIRE(brdf || D == 1, "spr_fft_dense: D must be 1 without a BRDF");
  hipLaunchKernelGGL(fft_dense_kernel,
                     dim3((unsigned)ceil_div(P, 128), (unsigned)P), dim3(128),
                     0, as_stream(stream), P, vis, ff, B, centers, areas, air,
                     wall_ids, S, D, brdf, brdf_index, in_dirs, fft);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

extern "C" int spr_source_energy(const double *source, int64_t P,
                                 const double *centers, const double *points,
                                 const uint8_t *vis, const double *air,
                                 int64_t B, const int64_t *wall_ids,
                                 int64_t S, int64_t D, const double *brdf,
                                 const int64_t *brdf_index,
                                 const double *in_dirs, double *energy0,
                                 double *e0dir, double *distance,
                                 void *stream) {
  SPR_REQUIRE(P > 0 && B > 0 && S > 0 && D > 0,
              "spr_source_energy: bad sizes");
  SPR_REQUIRE(source && centers && points && vis && wall_ids && brdf &&
                  brdf_index && in_dirs && energy0 && e0dir && distance,
              "spr_source_energy: null pointer");
  hipLaunchKernelGGL(source_energy_kernel, dim3((unsigned)ceil_div(P, 128)),
                     dim3(128), 0, as_stream(stream), source, P, centers,
                     points, vis, air, B, wall_ids, S, D, brdf, brdf_index,
                     in_dirs, energy0, e0dir, distance);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}
