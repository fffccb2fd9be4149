// Shared helpers of the sparrow_hip library: error reporting, launch
// helpers and the float64 vector arithmetic the bake kernels use.
//
// The library is compiled with -ffp-contract=off: every product and sum in
// the geometry code is its own IEEE rounding, in the order the reference's
// numpy expressions evaluate them, so that integer and boolean outputs
// (visibility, delays, direction indices) are bit-exact.  Fused multiply-add
// appears only where it is written out (fma) in the energy kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sparrow_hip.h"

namespace spr {

int fail(int code, const char *fmt, ...);
void clear_error();

inline hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

#define SPR_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return ::spr::fail(SPR_ERR_HIP, "%s failed: %s", #call,              \
                         hipGetErrorString(e_));                           \
  } while (0)

#define SPR_LAUNCH_CHECK()                                                 \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess)                                                  \
      return ::spr::fail(SPR_ERR_HIP, "kernel launch failed: %s",          \
                         hipGetErrorString(e_));                           \
  } while (0)

#define SPR_REQUIRE(cond, ...)                                             \
  do {                                                                     \
    if (!(cond)) return ::spr::fail(SPR_ERR_ARG, __VA_ARGS__);             \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X (block id
// modulo 8 picks the XCD).  xcd_tile maps block ids so that each XCD works
// through one contiguous range of logical tiles: neighbouring tiles (same
// destination patch, neighbouring time blocks) then share that XCD's L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t nblocks) {
  const int64_t n_xcd = 8;
  int64_t per = nblocks / n_xcd, rem = nblocks % n_xcd;
  int64_t x = bid % n_xcd, k = bid / n_xcd;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

// ---------------------------------------------------------------------
// float64 vector arithmetic, evaluated like numpy on 3-vectors:
// a.b = (a0*b0 + a1*b1) + a2*b2 and |a| = sqrt(a.a).
// ---------------------------------------------------------------------
struct d3 {
  double x, y, z;
};
struct d2 {
  double x, y;
};

__host__ __device__ __forceinline__ d3 load3(const double *p) {
  return d3{p[0], p[1], p[2]};
}
__host__ __device__ __forceinline__ d3 operator-(d3 a, d3 b) {
  return d3{a.x - b.x, a.y - b.y, a.z - b.z};
}
__host__ __device__ __forceinline__ d2 operator-(d2 a, d2 b) {
  return d2{a.x - b.x, a.y - b.y};
}
__host__ __device__ __forceinline__ double dot(d3 a, d3 b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__host__ __device__ __forceinline__ double dot(d2 a, d2 b) {
  return a.x * b.x + a.y * b.y;
}
__host__ __device__ __forceinline__ double norm(d3 a) { return sqrt(dot(a, a)); }
__host__ __device__ __forceinline__ double norm(d2 a) { return sqrt(dot(a, a)); }
__host__ __device__ __forceinline__ d3 cross(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ d3 div(d3 a, double s) {
  return d3{a.x / s, a.y / s, a.z / s};
}
__host__ __device__ __forceinline__ double sgn(double x) {
  return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0);
}

// int(d / c / dt) and int(ceil(d / c / dt)) as the reference computes
// histogram delays (RadiosityFast.py:1067-1068, 1135-1136, 1178-1179).
__host__ __device__ __forceinline__ int64_t delay_floor(double d, double c,
                                                        double dt) {
  return (int64_t)(d / c / dt);
}
__host__ __device__ __forceinline__ int64_t delay_ceil(double d, double c,
                                                       double dt) {
  return (int64_t)ceil(d / c / dt);
}

// argmin_k sum_c (dirs[k,c] - u[c])^2, first minimum wins
// (np.argmin(np.sum((dirs - u)**2, axis=-1)), RadiosityFast.py:1308-1310).
__device__ __forceinline__ int64_t nearest_direction(const double *dirs,
                                                     int64_t n, d3 u) {
  int64_t best = 0;
  double bv = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    double e0 = dirs[3 * k] - u.x, e1 = dirs[3 * k + 1] - u.y,
           e2 = dirs[3 * k + 2] - u.z;
    double s = (e0 * e0 + e1 * e1) + e2 * e2;
    if (k == 0 || s < bv) {
      bv = s;
      best = k;
    }
  }
  return best;
}

}  // namespace spr
