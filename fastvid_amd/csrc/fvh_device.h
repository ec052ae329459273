// fvh_device.h -- wave64 primitives and geometry helpers for gfx950 kernels.
#pragma once

#include "fvh_common.h"

namespace fvh {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ uint64_t ballot(bool predicate) {
  return __ballot(predicate);
}

__device__ __forceinline__ int popc64(uint64_t v) { return __popcll(v); }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, 64); }

template <typename T>
__device__ __forceinline__ T shfl_up(T v, int delta) { return __shfl_up(v, delta, 64); }

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) { return __shfl_xor(v, mask, 64); }

// Cross-lane moves in the VALU (DPP, gfx9 controls) instead of LDS permutes.
// A lane whose source lies outside the move's range (before its 16-lane row
// for row_shr, past the wave for the wave shifts) reads 0, or keeps `old`
// when kZero is false; rows outside kRowMask keep `old`.
template <int kCtrl, int kRowMask = 0xf, bool kZero = true>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, kCtrl, kRowMask, 0xf, kZero);
}
constexpr int kDppXor1 = 0xb1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4e;        // quad_perm [2,3,0,1]
constexpr int kDppRowShr = 0x110;     // + n: row_shr:n
constexpr int kDppWaveShl1 = 0x130;   // lane i <- lane i + 1
constexpr int kDppWaveShr1 = 0x138;   // lane i <- lane i - 1
constexpr int kDppMirror = 0x140;     // lane i of a row <- lane 15 - i
constexpr int kDppHalfMirror = 0x141; // lane i of a half row <- lane 7 - i
constexpr int kDppBcast15 = 0x142;    // lane 15 of each row -> the next row
constexpr int kDppBcast31 = 0x143;    // lane 31 -> rows 2 and 3

// The previous lane's value (0 on lane 0) and the next lane's (0 on lane 63).
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t v) { return dpp_mov<kDppWaveShr1>(0u, v); }
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v) { return dpp_mov<kDppWaveShl1>(0u, v); }
__device__ __forceinline__ uint64_t from_prev_lane(uint64_t v) {
  return ((uint64_t)from_prev_lane((uint32_t)(v >> 32)) << 32) | from_prev_lane((uint32_t)v);
}
__device__ __forceinline__ uint64_t from_next_lane(uint64_t v) {
  return ((uint64_t)from_next_lane((uint32_t)(v >> 32)) << 32) | from_next_lane((uint32_t)v);
}

// OR over each 16-lane row, in every lane of the row.
__device__ __forceinline__ uint32_t row_or(uint32_t v) {
  v |= dpp_mov<kDppXor1>(0u, v);
  v |= dpp_mov<kDppXor2>(0u, v);
  v |= dpp_mov<kDppHalfMirror>(0u, v);
  v |= dpp_mov<kDppMirror>(0u, v);
  return v;
}

// Reduction of a commutative, associative op over the 64 lanes, returned
// uniform: quads, half rows and rows combine in place, then the row
// broadcasts carry rows 0..2 into lane 63.  `identity` fills rows the
// broadcasts do not write.  The whole wave must be active.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v, uint32_t identity, Op op) {
  v = op(v, dpp_mov<kDppXor1>(identity, v));
  v = op(v, dpp_mov<kDppXor2>(identity, v));
  v = op(v, dpp_mov<kDppHalfMirror>(identity, v));
  v = op(v, dpp_mov<kDppMirror>(identity, v));
  v = op(v, dpp_mov<kDppBcast15, 0xa>(identity, v));
  v = op(v, dpp_mov<kDppBcast31, 0xc>(identity, v));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  return wave_reduce(v, 0u, [](uint32_t a, uint32_t b) { return a | b; });
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  return wave_reduce(v, ~0u, [](uint32_t a, uint32_t b) { return min(a, b); });
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
  return wave_reduce(v, 0u, [](uint32_t a, uint32_t b) { return max(a, b); });
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  if constexpr (sizeof(T) == 4) {
    return (T)wave_reduce((uint32_t)v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
  } else {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += shfl_xor(v, m);
    return v;
  }
}

// Inclusive prefix sum over the 64 lanes (whole wave active).  32-bit
// values: Hillis-Steele within rows by row shifts, then the row broadcasts.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  if constexpr (sizeof(T) == 4) {
    uint32_t x = (uint32_t)v;
    x += dpp_mov<kDppRowShr + 1>(0u, x);
    x += dpp_mov<kDppRowShr + 2>(0u, x);
    x += dpp_mov<kDppRowShr + 4>(0u, x);
    x += dpp_mov<kDppRowShr + 8>(0u, x);
    x += dpp_mov<kDppBcast15, 0xa>(0u, x);
    x += dpp_mov<kDppBcast31, 0xc>(0u, x);
    return (T)x;
  } else {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const T o = shfl_up(v, d);
      if (lane >= d) v += o;
    }
    return v;
  }
}

__device__ __forceinline__ int bit_length(uint32_t v) {
  return v == 0 ? 0 : 32 - __builtin_clz(v);
}

__device__ __forceinline__ int varint_len(uint32_t v) {
  return v < (1u << 7) ? 1 : v < (1u << 14) ? 2 : v < (1u << 21) ? 3 : v < (1u << 28) ? 4 : 5;
}

// Median of three integers; equals the LOCO-I MED predictor for
// (left, above, left+above-upper_left) (encode.cu:169-177).
__device__ __forceinline__ int median3(int a, int b, int c) {
  return max(min(a, b), min(max(a, b), c));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

__device__ __forceinline__ void set_status(FrameState* st, int code) {
  atomicCAS(&st->status, 0, code);
}

// Exact floor(n / d) for n < 2^17 and 2 <= d < 2^15 as mul_hi(n, ceil(2^32/d)):
// with e = m*d - 2^32 < d the error term n*e/2^32 stays below 1/d.
__device__ __forceinline__ uint32_t quant_magic(int32_t d) {
  return (uint32_t)((0x100000000ull + (uint64_t)d - 1) / (uint64_t)d);
}

}  // namespace fvh

namespace fvh {

// Pointers read out of descriptors are generic (flat) to the compiler, which
// tracks flat accesses with conservative counters and drains every
// outstanding load before the next one.
// Re-derive `p` from `gbase`, a pointer kernel argument: same address, but the
// compiler now knows the access is to global memory (kernel arguments are in
// the global address space) and keeps loads in flight with per-load counters.
template <typename T>
__device__ __forceinline__ T* as_global(char* gbase, T* p) {
  return (T*)(gbase + ((const char*)p - (const char*)gbase));
}

}  // namespace fvh
