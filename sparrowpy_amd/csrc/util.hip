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

// In-place exclusive scan of v[0..n-1], v[n] receiving the total, over
// the whole device and without workspace.  Pass 1: each workgroup scans
// its tile of kScanTile values in place (each thread kScanItems
// consecutive values, the threads' sums scanned in LDS) and parks the
// tile's total in the tile's first slot, whose local prefix is 0.  Pass 2:
// one workgroup scans those totals (stride kScanTile) in place into the
// tiles' offsets and writes the grand total to v[n].  Pass 3: each tile
// adds its offset to its other values.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = (int64_t)kScanThreads * kScanItems;

// Exclusive scan of one value per thread across the workgroup (LDS);
// returns this thread's prefix, *total the workgroup's sum.
template <int NT>
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t x,
                                                        int64_t *total) {
  __shared__ int64_t part[NT];
  const int tid = (int)threadIdx.x;
  part[tid] = x;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    const int64_t add = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  *total = part[NT - 1];
  const int64_t incl = part[tid];
  __syncthreads();  // part is reused by the next call
  return incl - x;
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(int64_t *v,
                                                                 int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int64_t lo = base + (int64_t)threadIdx.x * kScanItems;
  int64_t x[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    x[i] = lo + i < n ? v[lo + i] : 0;
    s += x[i];
  }
  int64_t total;
  int64_t run = block_exclusive_scan<kScanThreads>(s, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    // the tile's first slot gets the tile's total (its prefix is 0)
    if (lo + i < n) v[lo + i] = lo + i == base ? total : run;
    run += x[i];
  }
}

__global__ __launch_bounds__(1024) void scan_totals_kernel(int64_t *v,
                                                           int64_t n,
                                                           int64_t ntiles) {
  // ntiles totals at stride kScanTile; each thread a contiguous run of
  // them (few: one per 4,096 values)
  const int tid = (int)threadIdx.x;
  const int64_t chunk = (ntiles + 1023) / 1024;
  const int64_t lo = tid * chunk;
  const int64_t hi = lo + chunk < ntiles ? lo + chunk : ntiles;
  int64_t s = 0;
  for (int64_t k = lo; k < hi; ++k) s += v[k * kScanTile];
  int64_t total;
  int64_t run = block_exclusive_scan<1024>(s, &total);
  for (int64_t k = lo; k < hi; ++k) {
    const int64_t c = v[k * kScanTile];
    v[k * kScanTile] = run;
    run += c;
  }
  if (tid == 0) v[n] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(int64_t *v,
                                                               int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int64_t off = v[base];  // the tile's offset (pass 2)
  const int64_t lo = base + (int64_t)threadIdx.x * kScanItems;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (lo + i < n && lo + i != base) v[lo + i] += off;
}

int scan_exclusive(int64_t *v, int64_t n, hipStream_t s) {
  if (n <= 0) {
    hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(1024), 0, s, v,
                       (int64_t)0, (int64_t)0);
    SPR_LAUNCH_CHECK();
    return SPR_OK;
  }
  const int64_t ntiles = ceil_div(n, kScanTile);
  SPR_REQUIRE(ntiles < (int64_t)0x7fffffff, "scan too long");
  hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)ntiles),
                     dim3(kScanThreads), 0, s, v, n);
  SPR_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(1024), 0, s, v, n,
                     ntiles);
  SPR_LAUNCH_CHECK();
  if (ntiles > 1) {
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)ntiles),
                       dim3(kScanThreads), 0, s, v, n);
    SPR_LAUNCH_CHECK();
  }
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
