// Declarations of the shared helper kernels in util.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spr {

// In-place exclusive scan of v[0..n-1]; v[n] receives the total.
int scan_exclusive(int64_t *v, int64_t n, hipStream_t s);

// p[0..n-1] = value, by a kernel (safe inside captured graphs).
template <typename T>
int fill(T *p, int64_t n, T value, hipStream_t s);

}  // namespace spr
