// Error reporting, version, and small shared kernels (scan, fills).
#include <cstdarg>
#include <cstdio>

#include "spr_common.hpp"
#include "spr_util.hpp"

namespace spr {

namespace {
thread_local char g_last_error[512] = "";
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_last_error[0] = '\0'; }

// In-place exclusive scan of v[0..n-1] by one block of 1024 threads;
// v[n] receives the total.  Each thread scans a contiguous chunk.
__global__ __launch_bounds__(1024) void scan_exclusive_kernel(int64_t *v,
                                                              int64_t n) {
  __shared__ int64_t part[1024];
  const int tid = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t lo = tid * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  int64_t s = 0;
  for (int64_t k = lo; k < hi; ++k) s += v[k];
  part[tid] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the 1024 partial sums
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t add = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  int64_t run = part[tid] - s;  // exclusive prefix of this chunk
  for (int64_t k = lo; k < hi; ++k) {
    int64_t c = v[k];
    v[k] = run;
    run += c;
  }
  if (tid == 1023) v[n] = part[1023];
}

int scan_exclusive(int64_t *v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(scan_exclusive_kernel, dim3(1), dim3(1024), 0, s, v, n);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

template <typename T>
__global__ void fill_kernel(T *p, int64_t n, T value) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += stride)
    p[k] = value;
}

template <typename T>
int fill(T *p, int64_t n, T value, hipStream_t s) {
  if (n <= 0) return SPR_OK;
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s,
                     p, n, value);
  SPR_LAUNCH_CHECK();
  return SPR_OK;
}

template int fill<float>(float *, int64_t, float, hipStream_t);
template int fill<double>(double *, int64_t, double, hipStream_t);
template int fill<int64_t>(int64_t *, int64_t, int64_t, hipStream_t);
template int fill<int32_t>(int32_t *, int64_t, int32_t, hipStream_t);
template int fill<uint8_t>(uint8_t *, int64_t, uint8_t, hipStream_t);

}  // namespace spr

extern "C" const char *spr_last_error(void) { return spr::g_last_error; }

extern "C" const char *spr_version(void) {
  return "sparrow_hip 0.1 (gfx950)";
}
