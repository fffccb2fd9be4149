// Device geometry of the bake: the visibility predicate and the
// point-to-patch solid angle, in float64.
//
// The predicate reproduces geometry._basic_visibility /
// _point_in_polygon / _project_to_plane / _rotation_matrix
// (sparrowpy/geometry.py:498-686, 841-909) operation for operation, so the
// boolean results are the reference's bits.  Three things differ in
// organisation only, never in result:
//  * Everything that depends on the blocking polygon alone (rotation rows,
//    rotated corners, edge normals and lengths, the rotated bounding box)
//    is computed once per polygon into a BlockerRec instead of once per
//    query.
//  * Work whose outcome is already decided is skipped: the winding test is
//    not run for a point more than 1e-3 outside the polygon's bounding box
//    in the rotated frame (the reference's edge test only accepts crossing
//    points within ~1e-6 of an edge, so its count is 0 there), and the
//    "segment crosses the polygon" test evaluates the cheap
//    (pt - a).(pt - b) < 0 condition before the winding test (both are pure
//    and are combined with AND).
//  * Blockers that cannot block any segment between two boxes of points
//    (may_block_tile) are not evaluated for those segments.
#pragma once

#include "spr_common.hpp"

namespace spr {

constexpr double kEta = 1e-6;        // geometry.py:616, 845
constexpr double kBoxMargin = 1e-3;  // >> kEta, see header comment

struct BlockerRec {
  double n[3];      // plane normal
  double p0[3];     // first corner
  double m0[3];     // rotation matrix row 0 (normal -> +z)
  double m1[3];     // rotation matrix row 1
  double px[4];     // rotated corners, x
  double py[4];     // rotated corners, y
  double nlx[4];    // edge normals (-side.y, side.x) / |side|
  double nly[4];
  double elen[4];   // |a1 - a0|
  double box[4];    // xmin, xmax, ymin, ymax of the rotated corners
};
static_assert(sizeof(BlockerRec) == 36 * sizeof(double), "BlockerRec size");

// geometry.py:498-559 _rotation_matrix(n_in) with n_out = (0, 0, 1);
// only rows 0 and 1 are used (projection to the plane's 2-D frame).
__host__ __device__ inline void rotation_rows(const double *n_in,
                                              double *m0, double *m1) {
  const double n_out[3] = {0.0, 0.0, 1.0};
  int counter = 0;
  for (int i = 0; i < 3; ++i)
    if (n_in[i] == n_out[i]) counter += 1;
  double m[3][3];
  if (counter == 3) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = (i == j) ? 1.0 : 0.0;
  } else {
    d3 nv{n_in[0], n_in[1], n_in[2]};
    double na = norm(nv);
    d3 a{n_in[0] / na, n_in[1] / na, n_in[2] / na};
    d3 ov{n_out[0], n_out[1], n_out[2]};
    double nb = norm(ov);
    d3 b{n_out[0] / nb, n_out[1] / nb, n_out[2] / nb};
    double c = dot(a, b);
    if (c != -1.0) {
      d3 v = cross(a, b);
      double s = norm(v);
      double k[3][3] = {{0.0, -v.z, v.y}, {v.z, 0.0, -v.x}, {-v.y, v.x, 0.0}};
      double f = (1.0 - c) / (s * s);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double kk =
              (k[i][0] * k[0][j] + k[i][1] * k[1][j]) + k[i][2] * k[2][j];
          double e = (i == j) ? 1.0 : 0.0;
          m[i][j] = (e + k[i][j]) + kk * f;
        }
    } else {
      const double r[3][3] = {{-1., 0., 0.}, {0., 1., 0.}, {0., 0., -1.}};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = r[i][j];
    }
  }
  for (int j = 0; j < 3; ++j) {
    m0[j] = m[0][j];
    m1[j] = m[1][j];
  }
}

__host__ __device__ inline void make_blocker(const double *pts,
                                             const double *normal,
                                             BlockerRec &r) {
  for (int c = 0; c < 3; ++c) {
    r.n[c] = normal[c];
    r.p0[c] = pts[c];
  }
  rotation_rows(normal, r.m0, r.m1);
  for (int i = 0; i < 4; ++i) {
    d3 p = load3(pts + 3 * i);
    r.px[i] = dot(load3(r.m0), p);
    r.py[i] = dot(load3(r.m1), p);
  }
  r.box[0] = r.box[1] = r.px[0];
  r.box[2] = r.box[3] = r.py[0];
  for (int i = 0; i < 4; ++i) {
    d2 a1{r.px[(i + 1) % 4], r.py[(i + 1) % 4]};
    d2 a0{r.px[i], r.py[i]};
    d2 side = a1 - a0;
    double ns = norm(side);
    r.nlx[i] = -side.y / ns;
    r.nly[i] = side.x / ns;
    r.elen[i] = ns;
    r.box[0] = fmin(r.box[0], r.px[i]);
    r.box[1] = fmax(r.box[1], r.px[i]);
    r.box[2] = fmin(r.box[2], r.py[i]);
    r.box[3] = fmax(r.box[3], r.py[i]);
  }
}

// geometry.py:614-686 _point_in_polygon for a 4-corner blocker.
__device__ inline bool point_in_blocker(d3 p, const BlockerRec &r) {
  if (fabs(dot(p - load3(r.p0), load3(r.n))) > kEta) return false;
  d2 pt{dot(load3(r.m0), p), dot(load3(r.m1), p)};
  // Bounding-box early out (see header comment): count stays 0.
  if (pt.x > r.box[1] + kBoxMargin || pt.y < r.box[2] - kBoxMargin ||
      pt.y > r.box[3] + kBoxMargin)
    return false;
  int count = 0;
  const d2 ptp{pt.x + 1.0, pt.y + 0.0};
  const d2 v = ptp - pt;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const d2 a1{r.px[(i + 1) & 3], r.py[(i + 1) & 3]};
    const d2 a0{r.px[i], r.py[i]};
    const d2 nl{r.nlx[i], r.nly[i]};
    double dotprod = dot(v, nl);
    if (!(fabs(dotprod) > 1e-6)) continue;
    d2 w = ptp - a1;
    double fac = -(dot(nl, w) / dotprod);
    d2 b{(w.x + a1.x) + fac * v.x, (w.y + a1.y) + fac * v.y};
    if (b.x > pt.x) {
      if (fabs((norm(b - a0) + norm(b - a1)) - r.elen[i]) <= kEta) {
        double dd = dot(b - pt, nl);
        if (dd > 0)
          count += 1;
        else if (dd < 0)
          count -= 1;
      }
    }
  }
  return count != 0;
}

// Culling that keeps basic_visibility's result.  The predicate can only
// return "blocked" for a segment a-b if a point of it (an endpoint, or the
// crossing point, which lies on the segment up to rounding) passes
// point_in_blocker: within kEta of the plane, at rotated x <= box[1] +
// kBoxMargin and rotated y within box[2..3] +- kBoxMargin.  Such points lie
// on the part of the segment within kCullPlane of the plane, and the
// rotated x and y are linear along it, so a blocker whose accepted x / y
// range that part misses by more than kCullBox leaves the pair visible --
// as does a blocker whose plane the segment never comes within kCullPlane
// of.  The margins exceed the thresholds by far more than any rounding of
// points on the segment.
constexpr double kCullPlane = 1e-3;  // >> kEta
constexpr double kCullBox = 2e-3;    // > kBoxMargin

// Interval bounds of the linear form c.p over the axis-aligned box [lo, hi].
__device__ inline double form_min(d3 c, d3 lo, d3 hi) {
  return (fmin(c.x * lo.x, c.x * hi.x) + fmin(c.y * lo.y, c.y * hi.y)) +
         fmin(c.z * lo.z, c.z * hi.z);
}
__device__ inline double form_max(d3 c, d3 lo, d3 hi) {
  return (fmax(c.x * lo.x, c.x * hi.x) + fmax(c.y * lo.y, c.y * hi.y)) +
         fmax(c.z * lo.z, c.z * hi.z);
}

// For every segment from a point of box I to a point of box J.  If one box
// lies within kCullPlane of the plane and the other beyond it (a patch's
// own wall seen from off the wall), the part of each segment within
// kCullPlane of the plane stays within tau * L of the first box, with
// tau = 2 kCullPlane / (gap - kCullPlane) and L the longest segment, so
// only that neighbourhood of the box is tested; otherwise the whole union
// box is.
__device__ inline bool may_block_tile(d3 ilo, d3 ihi, d3 jlo, d3 jhi,
                                      const BlockerRec &r) {
  const d3 n = load3(r.n), m0 = load3(r.m0), m1 = load3(r.m1);
  const double np0 = dot(n, load3(r.p0));
  const double di0 = form_min(n, ilo, ihi) - np0;
  const double di1 = form_max(n, ilo, ihi) - np0;
  const double dj0 = form_min(n, jlo, jhi) - np0;
  const double dj1 = form_max(n, jlo, jhi) - np0;
  if ((di0 > kCullPlane && dj0 > kCullPlane) ||
      (di1 < -kCullPlane && dj1 < -kCullPlane))
    return false;
  const d3 ulo{fmin(ilo.x, jlo.x), fmin(ilo.y, jlo.y), fmin(ilo.z, jlo.z)};
  const d3 uhi{fmax(ihi.x, jhi.x), fmax(ihi.y, jhi.y), fmax(ihi.z, jhi.z)};
  d3 lo = ulo, hi = uhi;
  double grow = 0.0;
  const bool i_on = di0 >= -kCullPlane && di1 <= kCullPlane;
  const bool j_on = dj0 >= -kCullPlane && dj1 <= kCullPlane;
  const double gap_j = fmax(dj0, -dj1), gap_i = fmax(di0, -di1);
  if (i_on && !j_on && gap_j > 2 * kCullPlane) {
    lo = ilo, hi = ihi;
    grow = 2 * kCullPlane / (gap_j - kCullPlane) * norm(uhi - ulo);
  } else if (j_on && !i_on && gap_i > 2 * kCullPlane) {
    lo = jlo, hi = jhi;
    grow = 2 * kCullPlane / (gap_i - kCullPlane) * norm(uhi - ulo);
  }
  if (form_min(m0, lo, hi) - grow > r.box[1] + kCullBox) return false;
  return !(form_max(m1, lo, hi) + grow < r.box[2] - kCullBox ||
           form_min(m1, lo, hi) - grow > r.box[3] + kCullBox);
}

// geometry.py:841-909 _basic_visibility(vis_point=a, eval_point=b).
__device__ inline bool basic_visibility(d3 a, d3 b, const BlockerRec &r) {
  const bool vin = point_in_blocker(a, r);
  const bool ein = point_in_blocker(b, r);
  const d3 n = load3(r.n);
  if (!vin && !ein) {
    // _project_to_plane(origin=a, point=b, plane_pt=p0, normal=n,
    //                   check_normal=False)
    d3 v = b - a;
    double dotprod = dot(v, n);
    if (fabs(dotprod) > 1e-6) {
      const d3 p0 = load3(r.p0);
      d3 w = b - p0;
      double fac = -(dot(n, w) / dotprod);
      d3 pt{(w.x + p0.x) + fac * v.x, (w.y + p0.y) + fac * v.y,
            (w.z + p0.z) + fac * v.z};
      if (dot(pt - a, pt - b) < 0 && point_in_blocker(pt, r)) return false;
    }
    return true;
  }
  if (vin && !ein && dot(n, b - a) < 0) return false;
  if (!vin && ein && dot(n, a - b) < 0) return false;
  const d3 p0 = load3(r.p0);
  if (fabs(dot(a - p0, n)) < kEta && fabs(dot(b - p0, n)) < kEta &&
      (vin || ein))
    return false;
  return true;
}

// geometry.py:688-715 _sphere_tangent_vector
__device__ inline d3 sphere_tangent(d3 v0, d3 v1) {
  if (fabs(dot(v0, v1)) > 1e-10) {
    d3 d = v1 - v0;
    double s = dot(d, v0) / dot(v0, v0);
    d3 o{d.x - s * v0.x, d.y - s * v0.y, d.z - s * v0.z};
    return div(o, norm(o));
  }
  return div(v1, norm(v1));
}

// geometry.py:420-440 _polygon_area (4 corners)
__device__ inline double quad_area(const double *pts) {
  double area = 0.0;
  d3 p0 = load3(pts);
  for (int t = 0; t < 2; ++t) {
    d3 c = cross(load3(pts + 3 * (t + 1)) - p0, load3(pts + 3 * (t + 2)) - p0);
    area += 0.5 * norm(c);
  }
  return area;
}

// integration.py:295-344 pt_solution for a 4-corner patch;
// receiver: divide by pi*area, source: by pi*4.
__device__ inline double pt_solution(d3 point, const double *pts,
                                     bool receiver) {
  double area = receiver ? quad_area(pts) : 4.0;
  d3 on[4];
  for (int i = 0; i < 4; ++i) {
    d3 d = load3(pts + 3 * i) - point;
    on[i] = div(d, norm(d));
  }
  double sum = 0.0;
  for (int i = 0; i < 4; ++i) {
    d3 t0 = sphere_tangent(on[i], on[(i + 3) & 3]);
    d3 t1 = sphere_tangent(on[i], on[(i + 1) & 3]);
    sum += acos(dot(t0, t1));
  }
  double factor = sum - 2.0 * M_PI;
  return factor / (M_PI * area);
}

}  // namespace spr
