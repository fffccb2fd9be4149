/*
 * FP64 CPU oracle of sparrowpy's DirectionalRadiosityFast hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the package smoke check and the
 * benchmark's CPU-baseline leg, as the values the GPU results are compared
 * against; never by the product.
 *
 * This is a statement-by-statement restatement of the reference numerics
 * (sparrowpy 1.0.1).  Operation order follows the Python source: a.b for
 * 3-vectors is (a0*b0 + a1*b1) + a2*b2, norms are sqrt of that, products
 * and sums are separate roundings (build with -ffp-contract=off).  Where the
 * reference defers to BLAS/LAPACK (np.dot, np.linalg.norm, np.linalg.inv)
 * this file uses the plain sequential formula; results can differ from a
 * numba run in the last bit there.
 */
#include "sparrow_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXPTS 16

/* ------------------------------------------------------------------ */
/* small vector helpers                                                 */
/* ------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline double dot2(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1];
}
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }
static inline double norm2(const double *a) { return sqrt(dot2(a, a)); }
static inline void sub3(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static inline void sub2(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
}
static inline void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double sign(double x) {
  return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0);
}

int64_t orc_delay_floor(double d, double c, double dt) {
  return (int64_t)(d / c / dt);
}
int64_t orc_delay_ceil(double d, double c, double dt) {
  return (int64_t)ceil(d / c / dt);
}

/* geometry.py:498-559 _rotation_matrix(n_in) with n_out = (0, 0, 1) */
static void rotation_matrix(const double *n_in, double m[3][3]) {
  const double n_out[3] = {0.0, 0.0, 1.0};
  int counter = 0;
  for (int i = 0; i < 3; ++i)
    if (n_in[i] == n_out[i]) counter += 1;
  if (counter == 3) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = (i == j) ? 1.0 : 0.0;
    return;
  }
  double a[3], b[3];
  double na = norm3(n_in);
  for (int i = 0; i < 3; ++i) a[i] = n_in[i] / na;
  double nb = norm3(n_out);
  for (int i = 0; i < 3; ++i) b[i] = n_out[i] / nb;
  double c = dot3(a, b);
  if (c != -1.0) {
    double v[3];
    cross3(a, b, v);
    double s = norm3(v);
    double k[3][3] = {{0.0, -v[2], v[1]}, {v[2], 0.0, -v[0]},
                      {-v[1], v[0], 0.0}};
    double f = (1.0 - c) / (s * s);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double kk = (k[i][0] * k[0][j] + k[i][1] * k[1][j]) + k[i][2] * k[2][j];
        double e = (i == j) ? 1.0 : 0.0;
        m[i][j] = (e + k[i][j]) + kk * f;
      }
  } else {
    double r[3][3] = {{-1., 0., 0.}, {0., 1., 0.}, {0., 0., -1.}};
    memcpy(m, r, sizeof(r));
  }
}

/* geometry.py:474-496 _matrix_vector_product, first two components */
static inline void rot2(double m[3][3], const double *v, double *o) {
  o[0] = dot3(m[0], v);
  o[1] = dot3(m[1], v);
}

/* geometry.py:561-612 _project_to_plane (3-D) */
static int project_to_plane3(const double *origin, const double *point,
                             const double *plane_pt, const double *n,
                             double eps, int check_normal, double *out) {
  double v[3];
  sub3(point, origin, v);
  double dotprod = dot3(v, n);
  int cond = check_normal ? (dotprod < -eps) : (fabs(dotprod) > eps);
  if (!cond) return 0;
  double w[3];
  sub3(point, plane_pt, w);
  double fac = -(dot3(n, w) / dotprod);
  for (int i = 0; i < 3; ++i) out[i] = (w[i] + plane_pt[i]) + fac * v[i];
  return 1;
}

/* geometry.py:614-686 _point_in_polygon (winding count, eta = 1e-6) */
int orc_point_in_polygon(const double *point3d, const double *polygon3d,
                         int npts, const double *plane_normal) {
  const double eta = 1e-6;
  double d[3];
  sub3(point3d, polygon3d, d);
  if (fabs(dot3(d, plane_normal)) > eta) return 0;
  double m[3][3];
  rotation_matrix(plane_normal, m);
  double pt[2];
  rot2(m, point3d, pt);
  double poly[MAXPTS][2];
  for (int i = 0; i < npts; ++i) rot2(m, polygon3d + 3 * i, poly[i]);
  int count = 0;
  for (int i = 0; i < npts; ++i) {
    const double *a1 = poly[(i + 1) % npts];
    const double *a0 = poly[i % npts];
    double side[2];
    sub2(a1, a0, side);
    double ns = norm2(side);
    double nl[2] = {-side[1] / ns, side[0] / ns};
    /* _project_to_plane(origin=pt, point=pt+[1,0], plane_pt=a1,
       plane_normal=nl, check_normal=False) in 2-D */
    double ptp[2] = {pt[0] + 1.0, pt[1] + 0.0};
    double v[2];
    sub2(ptp, pt, v);
    double dotprod = dot2(v, nl);
    if (!(fabs(dotprod) > 1e-6)) continue;
    double w[2];
    sub2(ptp, a1, w);
    double fac = -(dot2(nl, w) / dotprod);
    double b[2] = {(w[0] + a1[0]) + fac * v[0], (w[1] + a1[1]) + fac * v[1]};
    if (b[0] > pt[0]) {
      double ba0[2], ba1[2], a10[2];
      sub2(b, a0, ba0);
      sub2(b, a1, ba1);
      sub2(a1, a0, a10);
      if (fabs((norm2(ba0) + norm2(ba1)) - norm2(a10)) <= eta) {
        double bp[2];
        sub2(b, pt, bp);
        double dd = dot2(bp, nl);
        if (dd > 0)
          count += 1;
        else if (dd < 0)
          count -= 1;
      }
    }
  }
  return count != 0;
}

/* geometry.py:841-909 _basic_visibility */
int orc_basic_visibility(const double *vis_point, const double *eval_point,
                         const double *surf_points, int npts,
                         const double *surf_normal) {
  const double eta = 1e-6;
  int is_visible = 1;
  int vin = orc_point_in_polygon(vis_point, surf_points, npts, surf_normal);
  int ein = orc_point_in_polygon(eval_point, surf_points, npts, surf_normal);
  if (!vin && !ein) {
    double pt[3];
    if (project_to_plane3(vis_point, eval_point, surf_points, surf_normal,
                          1e-6, 0, pt)) {
      if (orc_point_in_polygon(pt, surf_points, npts, surf_normal)) {
        double a[3], b[3];
        sub3(pt, vis_point, a);
        sub3(pt, eval_point, b);
        if (dot3(a, b) < 0) is_visible = 0;
      }
    }
  } else {
    double ev[3], ve[3], v0[3], e0[3];
    sub3(eval_point, vis_point, ev);
    sub3(vis_point, eval_point, ve);
    if (vin && !ein && dot3(surf_normal, ev) < 0) {
      is_visible = 0;
    } else if (!vin && ein && dot3(surf_normal, ve) < 0) {
      is_visible = 0;
    } else {
      sub3(vis_point, surf_points, v0);
      sub3(eval_point, surf_points, e0);
      if (fabs(dot3(v0, surf_normal)) < eta &&
          fabs(dot3(e0, surf_normal)) < eta && (vin || ein))
        is_visible = 0;
    }
  }
  return is_visible;
}

/* geometry.py:750-797 _check_patch2patch_visibility */
void orc_p2p_visibility(int64_t P, const double *centers, int64_t NS,
                        const double *surf_normals, const double *surf_pts,
                        int npts, uint8_t *out, int nthreads) {
  memset(out, 0, (size_t)(P * P));
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t i = 0; i < P; ++i) {
    for (int64_t j = i + 1; j < P; ++j) {
      int vis = 1;
      int64_t k = 0;
      while (vis && k != NS) {
        vis = orc_basic_visibility(centers + 3 * i, centers + 3 * j,
                                   surf_pts + (size_t)k * npts * 3, npts,
                                   surf_normals + 3 * k);
        k += 1;
      }
      out[i * P + j] = (uint8_t)vis;
    }
  }
}

/* geometry.py:750-797 _check_patch2patch_visibility, the loop body for a
 * list of pairs (i < j each, int64 [n,2]): the same AND over all surfaces
 * as orc_p2p_visibility, for sampled rows of large scenes. */
void orc_p2p_visibility_pairs(const double *centers, int64_t NS,
                              const double *surf_normals,
                              const double *surf_pts, int npts, int64_t n,
                              const int64_t *pairs, uint8_t *out,
                              int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
  for (int64_t q = 0; q < n; ++q) {
    int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
    int vis = 1;
    int64_t k = 0;
    while (vis && k != NS) {
      vis = orc_basic_visibility(centers + 3 * i, centers + 3 * j,
                                 surf_pts + (size_t)k * npts * 3, npts,
                                 surf_normals + 3 * k);
      k += 1;
    }
    out[q] = (uint8_t)vis;
  }
}

/* geometry.py:799-839 _check_point2patch_visibility */
void orc_point_visibility(const double *point, int64_t P,
                          const double *centers, int64_t NS,
                          const double *surf_normals, const double *surf_pts,
                          int npts, uint8_t *out) {
  for (int64_t i = 0; i < P; ++i) {
    int vis = 1;
    int64_t k = 0;
    while (vis && k != NS) {
      vis = orc_basic_visibility(point, centers + 3 * i,
                                 surf_pts + (size_t)k * npts * 3, npts,
                                 surf_normals + 3 * k);
      k += 1;
    }
    out[i] = (uint8_t)vis;
  }
}

/* ------------------------------------------------------------------ */
/* form factors                                                         */
/* ------------------------------------------------------------------ */

/* integration.py:608-651 _sample_boundary_regular for a 4-corner patch */
static void sample_boundary(const double *el, int npoints, double *pts,
                            int conn[4][8]) {
  const int n = 4;
  int n_div = npoints - 1;
  for (int i = 0; i < n; ++i) {
    conn[i][0] = (i * n_div) % (n_div * n);
    conn[i][npoints - 1] = (i * n_div + n_div) % (n_div * n);
    const double *a = el + 3 * i;
    const double *b = el + 3 * ((i + 1) % n);
    for (int ii = 0; ii < n_div; ++ii) {
      for (int c = 0; c < 3; ++c)
        pts[3 * (i * n_div + ii) + c] =
            a[c] + ((double)ii * (b[c] - a[c])) / (double)n_div;
      conn[i][ii] = (i * n_div + ii) % (n_div * n);
    }
  }
}

/* integration.py:462-486 _newton_cotes_4th (Boole's rule) */
static inline double newton_cotes_4th(const double *x, const double *y) {
  double h = x[1] - x[0];
  return ((2.0 * h) / 45.0) *
         ((((7.0 * y[0] + 32.0 * y[1]) + 12.0 * y[2]) + 32.0 * y[3]) +
          7.0 * y[4]);
}

/* integration.py:38-114 stokes_integration */
double orc_stokes(const double *patch_i, const double *patch_j,
                  double patch_i_area) {
  double ib[16 * 3], jb[16 * 3];
  int ic[4][8], jc[4][8];
  sample_boundary(patch_i, 5, ib, ic);
  sample_boundary(patch_j, 5, jb, jc);
  double form[16][16];
  for (int a = 0; a < 16; ++a)
    for (int b = 0; b < 16; ++b) {
      double d[3];
      sub3(ib + 3 * a, jb + 3 * b, d);
      form[a][b] = log(norm3(d));
    }
  double inner[16][3];
  memset(inner, 0, sizeof(inner));
  double outer = 0.0;
  double x[5], y[5];
  for (int dim = 0; dim < 3; ++dim) {
    for (int a = 0; a < 16; ++a) {
      for (int s = 0; s < 4; ++s) {
        for (int k = 0; k < 5; ++k) x[k] = jb[3 * jc[s][k] + dim];
        if (fabs(x[4] - x[0]) > 1e-3) {
          for (int k = 0; k < 5; ++k) y[k] = form[a][jc[s][k]];
          inner[a][dim] += newton_cotes_4th(x, y);
        }
      }
    }
    for (int s = 0; s < 4; ++s) {
      for (int k = 0; k < 5; ++k) x[k] = ib[3 * ic[s][k] + dim];
      if (fabs(x[4] - x[0]) > 1e-3) {
        for (int k = 0; k < 5; ++k) y[k] = inner[ic[s][k]][dim];
        outer += newton_cotes_4th(x, y);
      }
    }
  }
  return fabs(outer / ((2.0 * M_PI) * patch_i_area));
}

/* geometry.py:420-440 _polygon_area */
static double polygon_area(const double *pts, int npts) {
  double area = 0.0;
  for (int t = 0; t < npts - 2; ++t) {
    double a[3], b[3], c[3];
    sub3(pts + 3 * (t + 1), pts, a);
    sub3(pts + 3 * (t + 2), pts, b);
    cross3(a, b, c);
    area += 0.5 * norm3(c);
  }
  return area;
}

/* 3x3 inverse by Gaussian elimination with partial pivoting, solving
 * A X = I column by column (the LU/solve route of np.linalg.inv). */
static void inv3(double a_in[3][3], double out[3][3]) {
  double a[3][3];
  int piv[3] = {0, 1, 2};
  memcpy(a, a_in, sizeof(a));
  for (int k = 0; k < 3; ++k) {
    int p = k;
    double best = fabs(a[k][k]);
    for (int i = k + 1; i < 3; ++i)
      if (fabs(a[i][k]) > best) {
        best = fabs(a[i][k]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < 3; ++j) {
        double t = a[k][j];
        a[k][j] = a[p][j];
        a[p][j] = t;
      }
      int t = piv[k];
      piv[k] = piv[p];
      piv[p] = t;
    }
    for (int i = k + 1; i < 3; ++i) {
      a[i][k] = a[i][k] / a[k][k];
      for (int j = k + 1; j < 3; ++j) a[i][j] = a[i][j] - a[i][k] * a[k][j];
    }
  }
  for (int col = 0; col < 3; ++col) {
    double y[3];
    for (int i = 0; i < 3; ++i) {
      double s = (piv[i] == col) ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) s = s - a[i][j] * y[j];
      y[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < 3; ++j) s = s - a[i][j] * out[j][col];
      out[i][col] = s / a[i][i];
    }
  }
}

/* integration.py:349-382 _poly_estimation_Lagrange (3 samples) */
static void poly_estimation_lagrange(const double *x, const double *y,
                                     double *b) {
  if (fabs(x[2] - x[0]) < 1e-6) {
    b[0] = b[1] = b[2] = 0.0;
    return;
  }
  double xm[3][3], inv[3][3];
  for (int i = 0; i < 3; ++i)
    for (int o = 0; o < 3; ++o) xm[i][2 - o] = pow(x[i], (double)o);
  inv3(xm, inv);
  for (int i = 0; i < 3; ++i) b[i] = dot3(inv[i], y);
}

/* integration.py:385-412 _poly_integration */
static double poly_integration(const double *c, const double *x) {
  double out = 0.0;
  for (int i = 0; i < 3; ++i) {
    double p = (double)(3 - i);
    out += (c[i] * pow(x[2], p)) / p;
    out -= (c[i] * pow(x[0], p)) / p;
  }
  return out;
}

/* integration.py:416-460 _area_under_curve, order 2, three 2-D points */
static double area_under_curve(const double *p0, const double *p1,
                               const double *p2) {
  const double *ps[3] = {p0, p1, p2};
  double f[2];
  sub2(p2, p0, f);
  double nf = norm2(f);
  double r0[2] = {f[0] / nf, f[1] / nf};
  double r1[2] = {-f[1] / nf, f[0] / nf};
  double x[3] = {0, 0, 0}, y[3] = {0, 0, 0};
  for (int k = 1; k < 3; ++k) {
    double c[2];
    sub2(ps[k], p0, c);
    x[k] = dot2(r0, c);
    y[k] = dot2(r1, c);
  }
  double co[3];
  poly_estimation_lagrange(x, y, co);
  return poly_integration(co, x);
}

/* integration.py:116-230 nusselt_analog */
static double nusselt_analog(const double *origin, const double *surf_normal,
                             const double *patch_points,
                             const double *patch_normal) {
  double bp[8 * 3];
  int conn[4][8];
  sample_boundary(patch_points, 3, bp, conn);
  double e1[3], e2[3], cr[3];
  sub3(patch_points + 3, patch_points, e1);
  sub3(patch_points + 6, patch_points + 3, e2);
  cross3(e1, e2, cr);
  double hand = sign(dot3(cr, patch_normal));
  double curved = 0.0;
  double sph[8][3], pln[8][2], proj[8][3];
  for (int i = 0; i < 8; ++i) {
    double d[3];
    sub3(bp + 3 * i, origin, d);
    double nd = norm3(d);
    for (int c = 0; c < 3; ++c) sph[i][c] = d[c] / nd;
  }
  double m[3][3];
  rotation_matrix(surf_normal, m);
  for (int i = 0; i < 8; ++i) {
    rot2(m, sph[i], pln[i]);
    proj[i][0] = pln[i][0];
    proj[i][1] = pln[i][1];
    proj[i][2] = 0.0;
  }
  double quad[4 * 3];
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) quad[3 * i + c] = proj[2 * i][c];
  double big_poly = polygon_area(quad, 4);
  for (int jj = 0; jj < 4; ++jj) {
    int s0 = conn[jj][0], s1 = conn[jj][1], s2 = conn[jj][2];
    double cz[3];
    cross3(proj[s2], proj[s0], cz);
    if (!(norm3(cz) > 1e-6)) continue;
    if (dot2(pln[s2], pln[s0]) >= 1e-6) {
      curved += area_under_curve(pln[s0], pln[s1], pln[s2]);
    } else {
      double mpoint[3], marc[3], a[3], b[3];
      for (int c = 0; c < 3; ++c)
        mpoint[c] = sph[s0][c] + (sph[s2][c] - sph[s0][c]) / 2.0;
      double nm = norm3(mpoint);
      for (int c = 0; c < 3; ++c) marc[c] = mpoint[c] / nm;
      for (int c = 0; c < 3; ++c) a[c] = sph[s0][c] + (marc[c] - sph[s0][c]) / 2.0;
      for (int c = 0; c < 3; ++c) b[c] = marc[c] + (sph[s2][c] - marc[c]) / 2.0;
      double mp2[2], marc2[2], a2[2], b2[2];
      rot2(m, mpoint, mp2);
      rot2(m, marc, marc2);
      double na = norm3(a);
      for (int c = 0; c < 3; ++c) a[c] = a[c] / na;
      rot2(m, a, a2);
      double nb = norm3(b);
      for (int c = 0; c < 3; ++c) b[c] = b[c] / nb;
      rot2(m, b, b2);
      double d1[2], d2[2];
      sub2(pln[s2], pln[s0], d1);
      sub2(mp2, marc2, d2);
      double lin_area = (norm2(d1) * norm2(d2)) / 2.0;
      double left = area_under_curve(pln[s0], a2, marc2);
      double right = area_under_curve(marc2, b2, pln[s2]);
      curved += ((lin_area * sign(left)) + left) + right;
    }
  }
  return big_poly + hand * curved;
}

/* numpy.linspace(0, stop, num) as numpy evaluates it */
static void linspace0(double stop, int num, double *y) {
  if (num == 1) {
    y[0] = 0.0;
    return;
  }
  double step = stop / (double)(num - 1);
  for (int k = 0; k < num; ++k) y[k] = (double)k * step + 0.0;
  y[num - 1] = stop;
}

/* integration.py:232-289 nusselt_integration with
 * integration.py:533-605 _surf_sample_regulargrid (4-corner patch) */
double orc_nusselt(const double *patch_i, const double *patch_j,
                   const double *normal_i, const double *normal_j,
                   int nsamples) {
  double u[3], v[3];
  sub3(patch_i + 3, patch_i, u);
  sub3(patch_i + 9, patch_i, v);
  double nu = norm3(u), nv = norm3(v);
  double sq = sqrt(1.0 * (double)nsamples);
  int npx = (int)nearbyint(nu / nv * sq);
  int npz = (int)nearbyint(nv / nu * sq);
  if (npz == 0) npz = 1;
  if (npx == 0) npx = 1;
  double *tt = (double *)malloc(sizeof(double) * (size_t)(npx + npz));
  double *tz = tt + npx;
  linspace0(1.0 - 1.0 / (double)npx, npx, tt);
  double sstep = 1.0 / (double)(npx * 2);
  for (int k = 0; k < npx; ++k) tt[k] += sstep;
  linspace0(1.0 - 1.0 / (double)npz, npz, tz);
  double sstepz = 1.0 / (double)(npz * 2);
  for (int k = 0; k < npz; ++k) tz[k] += sstepz;
  double out = 0.0;
  for (int a = 0; a < npx; ++a) {
    for (int b = 0; b < npz; ++b) {
      double p0[3];
      for (int c = 0; c < 3; ++c)
        p0[c] = (tt[a] * u[c] + tz[b] * v[c]) + patch_i[c];
      out += nusselt_analog(p0, normal_i, patch_j, normal_j);
    }
  }
  int n = npx * npz;
  free(tt);
  out *= 1.0 / (M_PI * (double)n);
  return out;
}

/* geometry.py:719-748 _coincidence_check (threshold 1e-6) */
static int coincidence_check(const double *p0, const double *p1) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double d[3];
      sub3(p0 + 3 * i, p1 + 3 * j, d);
      if (norm3(d) < 1e-6) return 1;
    }
  return 0;
}

/* universal.py:54-96 universal_form_factor */
double orc_universal_ff(const double *src_pts, const double *src_normal,
                        double src_area, const double *rcv_pts,
                        const double *rcv_normal) {
  if (coincidence_check(rcv_pts, src_pts))
    return orc_nusselt(src_pts, rcv_pts, src_normal, rcv_normal, 64);
  return orc_stokes(src_pts, rcv_pts, src_area);
}

/* universal.py:12-52 patch2patch_ff_universal */
void orc_form_factors(int64_t P, const double *pts, const double *normals,
                      const double *areas, int64_t V, const int32_t *pairs,
                      double *out, int nthreads) {
  memset(out, 0, sizeof(double) * (size_t)(P * P));
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
  for (int64_t v = 0; v < V; ++v) {
    int64_t i = pairs[2 * v], j = pairs[2 * v + 1];
    out[i * P + j] = orc_universal_ff(pts + 12 * i, normals + 3 * i, areas[i],
                                      pts + 12 * j, normals + 3 * j);
  }
}

/* geometry.py:688-715 _sphere_tangent_vector */
static void sphere_tangent(const double *v0, const double *v1, double *out) {
  if (fabs(dot3(v0, v1)) > 1e-10) {
    double d[3];
    sub3(v1, v0, d);
    double s = dot3(d, v0) / dot3(v0, v0);
    for (int c = 0; c < 3; ++c) out[c] = d[c] - s * v0[c];
    double n = norm3(out);
    for (int c = 0; c < 3; ++c) out[c] = out[c] / n;
  } else {
    double n = norm3(v1);
    for (int c = 0; c < 3; ++c) out[c] = v1[c] / n;
  }
}

/* integration.py:295-344 pt_solution */
double orc_pt_solution(const double *point, const double *patch_points,
                       int npts, int mode_receiver) {
  double area = mode_receiver ? polygon_area(patch_points, npts) : 4.0;
  double on[MAXPTS][3];
  for (int i = 0; i < npts; ++i) {
    double d[3];
    sub3(patch_points + 3 * i, point, d);
    double n = norm3(d);
    for (int c = 0; c < 3; ++c) on[i][c] = d[c] / n;
  }
  double sum = 0.0;
  for (int i = 0; i < npts; ++i) {
    double t0[3], t1[3];
    sphere_tangent(on[i], on[(i - 1 + npts) % npts], t0);
    sphere_tangent(on[i], on[(i + 1) % npts], t1);
    sum += acos(dot3(t0, t1));
  }
  double factor = sum - (double)(npts - 2) * M_PI;
  return factor / (M_PI * area);
}

/* universal.py:98-147 _source2patch_energy_universal */
void orc_source_energy(const double *source, int64_t P,
                       const double *centers, const double *pts,
                       const uint8_t *vis, const double *air, int64_t B,
                       double *energy, double *distance) {
  for (int64_t j = 0; j < P; ++j) {
    distance[j] = 0.0;
    for (int64_t b = 0; b < B; ++b) energy[j * B + b] = 0.0;
    if (!vis[j]) continue;
    double d[3];
    sub3(source, centers + 3 * j, d);
    distance[j] = norm3(d);
    double pt = orc_pt_solution(source, pts + 12 * j, 4, 0);
    for (int64_t b = 0; b < B; ++b)
      energy[j * B + b] = air ? exp(-air[b] * distance[j]) * pt : pt;
  }
}

/* universal.py:149-160 _patch2receiver_energy_universal */
void orc_receiver_energy(const double *receiver, int64_t P,
                         const double *pts, const uint8_t *vis, double *out) {
  for (int64_t i = 0; i < P; ++i)
    out[i] = vis[i] ? orc_pt_solution(receiver, pts + 12 * i, 4, 1) : 0.0;
}

/* argmin_k sum_c (dirs[k, c] - d[c])**2, first minimum wins */
static int64_t nearest_direction(const double *dirs, int64_t n,
                                 const double *d) {
  int64_t best = 0;
  double bv = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    const double *r = dirs + 3 * k;
    double e0 = r[0] - d[0], e1 = r[1] - d[1], e2 = r[2] - d[2];
    double s = (e0 * e0 + e1 * e1) + e2 * e2;
    if (k == 0 || s < bv) {
      bv = s;
      best = k;
    }
  }
  return best;
}

/* RadiosityFast.py:1277-1312 get_scattering_data_receiver_index */
void orc_receiver_index(int64_t n, const double *pos_i, const double *pos_j,
                        int64_t D, const double *receivers,
                        const int64_t *wall_ids, int64_t *out) {
  for (int64_t k = 0; k < n; ++k) {
    double d[3];
    sub3(pos_j, pos_i + 3 * k, d);
    double nd = norm3(d);
    for (int c = 0; c < 3; ++c) d[c] = d[c] / nd;
    out[k] = nearest_direction(receivers + (size_t)wall_ids[k] * D * 3, D, d);
  }
}

/* RadiosityFast.py:1359-1390 get_scattering_data_source (index part) */
int64_t orc_source_index(const double *pos_h, const double *pos_i, int64_t S,
                         const double *sources, int64_t wall) {
  double d[3];
  sub3(pos_h, pos_i, d);
  double nd = norm3(d);
  for (int c = 0; c < 3; ++c) d[c] = d[c] / nd;
  return nearest_direction(sources + (size_t)wall * S * 3, S, d);
}

/* RadiosityFast.py:1250-1270, the body of _form_factors_with_directivity_dim
 * for one directed visible pair i -> j: f = ff[i,j] (i < j) or
 * ff[j,i] A_j / A_i (given), times exp(-air_b |unit(c_i - c_j)|), times
 * brdf[bi(wall_i), s*, :, :].  o: [D, B]. */
static void fft_pair(int64_t i, int64_t j, double f, int64_t B,
                     const double *centers, const double *air,
                     const int64_t *wall_ids, int64_t S, int64_t D,
                     const double *brdf, const int64_t *brdf_index,
                     const double *sources, double *o) {
  size_t blk = (size_t)D * B;
  double diff[3];
  sub3(centers + 3 * i, centers + 3 * j, diff);
  double nd = norm3(diff);
  for (int c = 0; c < 3; ++c) diff[c] = diff[c] / nd;
  double distance = norm3(diff);
  for (size_t k = 0; k < blk; ++k) o[k] = f;
  if (air)
    for (int64_t d = 0; d < D; ++d)
      for (int64_t b = 0; b < B; ++b)
        o[d * B + b] = o[d * B + b] * exp(-air[b] * distance);
  if (brdf) {
    int64_t w = wall_ids[i];
    int64_t s = orc_source_index(centers + 3 * i, centers + 3 * j, S,
                                 sources, w);
    const double *r = brdf + ((size_t)brdf_index[w] * S + s) * blk;
    for (size_t k = 0; k < blk; ++k) o[k] = o[k] * r[k];
  }
}

/* RadiosityFast.py:1234-1272 _form_factors_with_directivity_dim */
void orc_fft(int64_t P, const uint8_t *vis, const double *ff, int64_t B,
             const double *centers, const double *areas, const double *air,
             const int64_t *wall_ids, int64_t S, int64_t D,
             const double *brdf, const int64_t *brdf_index,
             const double *sources, double *out, int nthreads) {
  size_t blk = (size_t)D * B;
  memset(out, 0, sizeof(double) * (size_t)P * P * blk);
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
  for (int64_t ii = 0; ii < P * P; ++ii) {
    int64_t i = ii % P;
    int64_t j = (ii / P) % P;
    int visible = (i < j) ? vis[i * P + j] : vis[j * P + i];
    if (!visible) continue;
    double f = (i < j) ? ff[i * P + j] : (ff[j * P + i] * areas[j]) / areas[i];
    fft_pair(i, j, f, B, centers, air, wall_ids, S, D, brdf, brdf_index,
             sources, out + ((size_t)i * P + j) * blk);
  }
}

/* _form_factors_with_directivity_dim (RadiosityFast.py:1234-1272) for a list
 * of directed visible pairs (i -> j, int64 [n,2]); ff_upper[q] = ff[min,
 * max] of pair q (the i < j form factor).  out: [n, D, B]. */
void orc_fft_pairs(int64_t n, const int64_t *pairs, const double *ff_upper,
                   int64_t B, const double *centers, const double *areas,
                   const double *air, const int64_t *wall_ids, int64_t S,
                   int64_t D, const double *brdf, const int64_t *brdf_index,
                   const double *sources, double *out) {
  size_t blk = (size_t)D * B;
  for (int64_t q = 0; q < n; ++q) {
    int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
    double f = (i < j) ? ff_upper[q] : (ff_upper[q] * areas[j]) / areas[i];
    fft_pair(i, j, f, B, centers, air, wall_ids, S, D, brdf, brdf_index,
             sources, out + (size_t)q * blk);
  }
}

/* RadiosityFast.py:988-1034 _add_directional */
void orc_add_directional(int64_t P, int64_t B, const double *energy0,
                         const double *source, const double *centers,
                         const int64_t *wall_ids, int64_t S, int64_t D,
                         const double *sources, const double *brdf,
                         const int64_t *brdf_index, double *out) {
  for (int64_t i = 0; i < P; ++i) {
    int64_t w = wall_ids[i];
    int64_t s = orc_source_index(source, centers + 3 * i, S, sources, w);
    const double *r = brdf + ((size_t)brdf_index[w] * S + s) * D * B;
    for (int64_t d = 0; d < D; ++d)
      for (int64_t b = 0; b < B; ++b)
        out[(i * D + d) * B + b] = energy0[i * B + b] * r[d * B + b];
  }
}

/* RadiosityFast.py:1037-1070 _energy_exchange_init_energy */
int orc_init_energy(int64_t T, int64_t P, int64_t D, int64_t B,
                    const double *e0dir, const double *d0, double c,
                    double dt, double *out) {
  memset(out, 0, sizeof(double) * (size_t)P * D * B * T);
  for (int64_t i = 0; i < P; ++i) {
    int64_t n = orc_delay_floor(d0[i], c, dt);
    if (n < 0 || n >= T) return -1;
    for (int64_t k = 0; k < D * B; ++k)
      out[((size_t)i * D * B + k) * T + n] += e0dir[i * D * B + k];
  }
  return 0;
}

/* accumulate one directed pair into a destination (RadiosityFast.py:
 * 1133-1143): dst[d, b, t] += f[d, b] * src[b, t - n], src the source's
 * [B, T] rows of its outgoing direction, dst the destination's [D, B, T]. */
static inline void row_update(int64_t T, int64_t D, int64_t B,
                              const double *src_rows, double *dst_rows,
                              int64_t n, const double *f) {
  for (int64_t d = 0; d < D; ++d)
    for (int64_t b = 0; b < B; ++b) {
      double fv = f[d * B + b];
      const double *src = src_rows + (size_t)b * T;
      double *dst = dst_rows + ((size_t)d * B + b) * T;
      if (n > 0) {
        for (int64_t t = n; t < T; ++t) dst[t] = dst[t] + fv * src[t - n];
      } else {
        for (int64_t t = 0; t < T; ++t) dst[t] = dst[t] + fv * src[t];
      }
    }
}

/* accumulate one directed pair i -> j into cur (RadiosityFast.py:1133-1143) */
static inline void pair_update(int64_t T, int64_t D, int64_t B,
                               const double *prev, double *cur, int64_t i,
                               int64_t j, int64_t dir, int64_t n,
                               const double *f) {
  row_update(T, D, B, prev + (((size_t)i * D + dir) * B) * T,
             cur + ((size_t)j * D * B) * T, n, f);
}

/* RadiosityFast.py:1073-1145 _energy_exchange */
int orc_exchange(int64_t T, int64_t P, int64_t D, int64_t B,
                 const double *e0dir, const double *d0, const double *dist,
                 const double *fft, const int64_t *p2o, double c, double dt,
                 int64_t K, int64_t V, const int32_t *pairs, double *out,
                 double *work, int nthreads) {
  size_t n_state = (size_t)P * D * B * T;
  size_t blk = (size_t)D * B;
  if (orc_init_energy(T, P, D, B, e0dir, d0, c, dt, out)) return -1;
  double *E[2] = {work, work + n_state};
  memcpy(E[0], out, sizeof(double) * n_state);
  if (K == 0) return 0;
  int64_t *row_ptr = NULL;
  int32_t *row_src = NULL;
  if (nthreads > 1) {
    /* per-destination lists in the order the pair list visits them */
    row_ptr = (int64_t *)calloc((size_t)P + 1, sizeof(int64_t));
    row_src = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * V + 1));
    int64_t *fill = (int64_t *)calloc((size_t)P, sizeof(int64_t));
    for (int64_t v = 0; v < V; ++v) {
      row_ptr[pairs[2 * v + 1] + 1] += 1;
      row_ptr[pairs[2 * v] + 1] += 1;
    }
    for (int64_t j = 0; j < P; ++j) row_ptr[j + 1] += row_ptr[j];
    for (int64_t v = 0; v < V; ++v) {
      int64_t a = pairs[2 * v], b = pairs[2 * v + 1];
      row_src[row_ptr[b] + fill[b]++] = (int32_t)a;
      row_src[row_ptr[a] + fill[a]++] = (int32_t)b;
    }
    free(fill);
  }
  for (int64_t k = 0; k < K; ++k) {
    int cur = (int)((1 + k) % 2);
    int prv = 1 - cur;
    memset(E[cur], 0, sizeof(double) * n_state);
    if (nthreads <= 1) {
      for (int64_t ii = 0; ii < V; ++ii) {
        for (int jj = 0; jj < 2; ++jj) {
          int64_t i = jj == 0 ? pairs[2 * ii] : pairs[2 * ii + 1];
          int64_t j = jj == 0 ? pairs[2 * ii + 1] : pairs[2 * ii];
          int64_t dir = p2o[i * P + j];
          int64_t n = orc_delay_floor(dist[i * P + j], c, dt);
          pair_update(T, D, B, E[prv], E[cur], i, j, dir, n,
                      fft + ((size_t)i * P + j) * blk);
        }
      }
    } else {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
      for (int64_t j = 0; j < P; ++j) {
        for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e) {
          int64_t i = row_src[e];
          int64_t dir = p2o[i * P + j];
          int64_t n = orc_delay_floor(dist[i * P + j], c, dt);
          pair_update(T, D, B, E[prv], E[cur], i, j, dir, n,
                      fft + ((size_t)i * P + j) * blk);
        }
      }
    }
    for (size_t q = 0; q < n_state; ++q) out[q] += E[cur][q];
  }
  free(row_ptr);
  free(row_src);
  return 0;
}

/* One order of _energy_exchange (RadiosityFast.py:1121-1143) on the dense
 * arrays: cur = the order after prev (overwritten), per destination in
 * ascending source order (the threaded order of orc_exchange). */
void orc_exchange_step(int64_t T, int64_t P, int64_t D, int64_t B,
                       const double *prev, double *cur, const double *dist,
                       const double *fft, const int64_t *p2o, double c,
                       double dt, const uint8_t *vis, int nthreads) {
  size_t blk = (size_t)D * B;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t j = 0; j < P; ++j) {
    memset(cur + (size_t)j * blk * T, 0, sizeof(double) * blk * T);
    for (int64_t i = 0; i < P; ++i) {
      if (i == j || !(i < j ? vis[i * P + j] : vis[j * P + i])) continue;
      int64_t n = orc_delay_floor(dist[i * P + j], c, dt);
      if (n >= T) continue;
      pair_update(T, D, B, prev, cur, i, j, p2o[i * P + j], n,
                  fft + ((size_t)i * P + j) * blk);
    }
  }
}

/* Adjoint of one order of _energy_exchange (the transpose of
 * RadiosityFast.py:1133-1143; no counterpart in the reference):
 * mu_out = lam + A^T mu_next, i.e. for every directed visible pair i -> j
 * with delay n < T and outgoing direction d = p2o[i,j]:
 *   mu_out[i,d,b,t] += sum_e fft[i,j,e,b] mu_next[j,e,b,t+n]   (t < T-n)
 *   dfft[i,j,e,b]   += sum_t mu_next[j,e,b,t+n] ek[i,d,b,t]
 * (dfft accumulates over calls).  Parallel over sources i, so every write
 * has one owner. */
void orc_exchange_adjoint_step(int64_t T, int64_t P, int64_t D, int64_t B,
                               const double *lam, const double *mu_next,
                               const double *ek, const double *dist,
                               const double *fft, const int64_t *p2o,
                               double c, double dt, const uint8_t *vis,
                               double *mu_out, double *dfft, int nthreads) {
  size_t blk = (size_t)D * B;
  memcpy(mu_out, lam, sizeof(double) * (size_t)P * blk * T);
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t i = 0; i < P; ++i) {
    for (int64_t j = 0; j < P; ++j) {
      if (i == j || !(i < j ? vis[i * P + j] : vis[j * P + i])) continue;
      int64_t n = orc_delay_floor(dist[i * P + j], c, dt);
      if (n >= T) continue;
      int64_t d = p2o[i * P + j];
      const double *f = fft + ((size_t)i * P + j) * blk;
      double *g = dfft + ((size_t)i * P + j) * blk;
      for (int64_t b = 0; b < B; ++b) {
        double *mo = mu_out + (((size_t)i * D + d) * B + b) * T;
        const double *e = ek + (((size_t)i * D + d) * B + b) * T;
        for (int64_t q = 0; q < D; ++q) {
          const double *mn = mu_next + (((size_t)j * D + q) * B + b) * T + n;
          double fv = f[q * B + b], acc = 0.0;
          for (int64_t t = 0; t < T - n; ++t) {
            mo[t] = mo[t] + fv * mn[t];
            acc = acc + mn[t] * e[t];
          }
          g[q * B + b] = g[q * B + b] + acc;
        }
      }
    }
  }
}

/* get_scattering_data_source's index (RadiosityFast.py:1359-1390) for a list
 * of directed pairs (i -> j): the incoming direction of wall_i nearest to
 * unit(c_i - c_j), as orc_fft picks it.  out[n]. */
void orc_source_index_pairs(int64_t n, const int64_t *pairs,
                            const double *centers, int64_t S,
                            const double *sources, const int64_t *wall_ids,
                            int64_t *out) {
  for (int64_t q = 0; q < n; ++q) {
    int64_t i = pairs[2 * q], j = pairs[2 * q + 1];
    out[q] = orc_source_index(centers + 3 * i, centers + 3 * j, S, sources,
                              wall_ids[i]);
  }
}

void orc_exchange_order(int64_t T, int64_t P, int64_t D, int64_t B,
                        const double *prev, double *cur,
                        const int64_t *row_ptr, const int32_t *row_src,
                        const int32_t *row_lag, const int32_t *row_dir,
                        const double *fft, int64_t j_begin, int64_t j_end,
                        int nthreads) {
  size_t blk = (size_t)D * B;
  (void)P;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t j = j_begin; j < j_end; ++j) {
    memset(cur + (size_t)j * blk * T, 0, sizeof(double) * blk * T);
    for (int64_t e = row_ptr[j]; e < row_ptr[j + 1]; ++e)
      pair_update(T, D, B, prev, cur, row_src[e], j, row_dir[e], row_lag[e],
                  fft + (size_t)e * blk);
  }
}

/* One order of _energy_exchange (RadiosityFast.py:1121-1143) for n_dest
 * destinations whose source rows are given compactly: entry e of
 * destination q (row_ptr[q] <= e < row_ptr[q+1], in accumulation order)
 * reads prev_rows[row_src[e], :, :] ([n_rows, B, T], the source patch's
 * histogram of the outgoing direction towards q) delayed by row_lag[e],
 * with coefficients coef[e, D, B].  cur: [n_dest, D, B, T], overwritten. */
void orc_exchange_rows(int64_t T, int64_t D, int64_t B, int64_t n_dest,
                       const int64_t *row_ptr, const int64_t *row_src,
                       const int64_t *row_lag, const double *coef,
                       const double *prev_rows, double *cur, int nthreads) {
  size_t blk = (size_t)D * B;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t q = 0; q < n_dest; ++q) {
    double *dst = cur + (size_t)q * blk * T;
    memset(dst, 0, sizeof(double) * blk * T);
    for (int64_t e = row_ptr[q]; e < row_ptr[q + 1]; ++e)
      row_update(T, D, B, prev_rows + (size_t)row_src[e] * B * T, dst,
                 row_lag[e], coef + (size_t)e * blk);
  }
}

/* RadiosityFast.py:1148-1185 _collect_receiver_energy */
void orc_collect_receiver_energy(int64_t P, int64_t B, int64_t T,
                                 const double *E, const double *dist,
                                 double c, double dt, const double *air,
                                 double *out) {
  for (int64_t i = 0; i < P; ++i) {
    int64_t n = orc_delay_ceil(dist[i], c, dt);
    int64_t sh = ((n % T) + T) % T;
    for (int64_t b = 0; b < B; ++b) {
      double g = exp(-air[b] * dist[i]);
      const double *src = E + ((size_t)i * B + b) * T;
      double *dst = out + ((size_t)i * B + b) * T;
      for (int64_t t = 0; t < T; ++t) dst[(t + sh) % T] = src[t] * g;
    }
  }
}
