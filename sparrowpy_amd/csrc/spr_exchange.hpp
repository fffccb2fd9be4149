// Exchange plan conventions and buffer-descriptor helpers shared by the
// forward (exchange.hip) and adjoint (adjoint.hip) passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "spr_common.hpp"

namespace spr {

// Entry metadata of the grouped plan: the delay word carries, in bit 30,
// the flag "last entry of its group" (delays are clamped below 2^30).
constexpr int32_t kLastInGroup = 0x40000000;
constexpr int32_t kLagMask = 0x3fffffff;

// Tile order of the grouped passes: destination fastest, then time block,
// then band.  With the XCD-contiguous block mapping, the blocks resident on
// one XCD at a time are neighbouring destinations (neighbouring patches of
// a wall) in the same band and time window, which read nearly the same
// source histogram segments -- so those reads are shared through L2/MALL.
__device__ __forceinline__ void dest_major_tile(int64_t tile, int64_t P,
                                                int n_tblk, int64_t &j,
                                                int &b, int &tb) {
  j = tile % P;
  const int64_t rest = tile / P;
  tb = (int)(rest % n_tblk);
  b = (int)(rest / n_tblk);
}

// Raw buffer access with a gfx9 buffer descriptor: a load whose byte
// offset lies outside [0, bytes) returns 0 and a store there is dropped,
// with no branch and no memory access.  This lets every lane issue its
// loads unconditionally, so the compiler can count them and keep the next
// chunk's loads in flight while the current chunk is summed.
constexpr int kBufferWord3 = 0x00020000;   // gfx9 descriptor word 3
constexpr uint32_t kNoOffset = 0x80000000u;  // beyond every range used here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p,
                                                            uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0,
                                           (int)bytes, kBufferWord3);
}

// Byte offset 4k of bin k of a row, kept whole.  For k < 0 the wrapped
// 32-bit value lies beyond every buffer used here and the load reads 0 --
// but only if the instruction receives it as computed: the compiler would
// otherwise move constant parts of the offset into the instruction's
// immediate field, and the hardware does not range-check the wrapped
// register offset plus that immediate as the 32-bit sum it stands for
// (loads of a row's first bins through such a pair return 0).  The empty
// asm makes the value opaque, so it is never split.
__device__ __forceinline__ uint32_t bin_offset(int k) {
  uint32_t off = (uint32_t)k * 4u;
  asm("" : "+v"(off));
  return off;
}

__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r,
                                             uint32_t off) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

__device__ __forceinline__ void buffer_store(float v,
                                             __amdgpu_buffer_rsrc_t r,
                                             uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r,
                                        (int)off, 0, 0);
}

__device__ __forceinline__ double readlane_f64(double v, int k) {
  const int64_t u = __builtin_bit_cast(int64_t, v);
  const int32_t lo = __builtin_amdgcn_readlane((int32_t)u, k);
  const int32_t hi = __builtin_amdgcn_readlane((int32_t)(u >> 32), k);
  return __builtin_bit_cast(double,
                            ((int64_t)hi << 32) | (int64_t)(uint32_t)lo);
}

// BRDF row of entry e (source i -> destination j): brdf_index[w_i] * S +
// s_ij (0 without a BRDF).
__device__ __forceinline__ int64_t entry_brdf_row(
    int64_t i, int64_t j, const double *__restrict__ centers,
    const int64_t *__restrict__ wall_ids, int64_t S,
    const int64_t *__restrict__ brdf_index,
    const double *__restrict__ in_dirs) {
  if (!in_dirs) return 0;
  d3 u = load3(centers + 3 * i) - load3(centers + 3 * j);
  u = div(u, norm(u));
  const int64_t w = wall_ids[i];
  return brdf_index[w] * S + nearest_direction(in_dirs + w * S * 3, S, u);
}

}  // namespace spr
