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

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v |= shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = min(v, (uint32_t)shfl_xor(v, m));
  return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = max(v, (uint32_t)shfl_xor(v, m));
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += shfl_xor(v, m);
  return v;
}

// Inclusive prefix sum over the 64 lanes.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = shfl_up(v, d);
    if (lane >= d) v += o;
  }
  return v;
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
