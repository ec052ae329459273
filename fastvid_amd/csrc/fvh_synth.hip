// fvh_synth.hip -- seeded synthetic planes generated on the device, for the
// serving workload: each request's frame is produced where it is encoded, in
// one launch per plane, from a few dozen parameters the host draws from the
// request's seed.  Not part of the codec (no counterpart in the reference).
//
// Formula (8-bit units, then scaled to the bit depth): a linear ramp along
// a random direction, three sinusoids 20*sin(kx*x + ky*y + phase), Gaussian
// noise, up to 20 axis-aligned rectangles of constant value (drawn over the
// noise, or under the offset and noise when `rects_under_noise` is set, the
// RGB form), an additive offset; then rint(v * max / 255) clipped to
// [0, max].  The noise is a counter-based normal deviate per sample (a
// 64-bit mix of seed and sample index, Box-Muller), so a plane depends only
// on its parameters.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fastvid_hip.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void k_synth_plane(fvh_synth_params p, uint16_t* __restrict__ out) {
  const int64_t n = p.width * p.height;
  const float maxv = (float)((1 << p.bit_depth) - 1);
  const float to_depth = maxv / 255.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t y = i / p.width;
    const int64_t x = i - y * p.width;
    const float fx = (float)x, fy = (float)y;
    float v = (p.ramp_cy * fy / (float)p.height - p.ramp_lo) * p.ramp_scale +
              p.ramp_cx * fx / (float)p.width * p.ramp_scale;
    for (int s = 0; s < 3; ++s) v += 20.0f * sinf(p.kx[s] * fx + p.ky[s] * fy + p.phase[s]);
    int rect = -1;
    for (int r = 0; r < p.rect_count; ++r) {
      const int32_t* q = p.rects[r];
      if (x >= q[0] && x < q[0] + q[2] && y >= q[1] && y < q[1] + q[3]) rect = r;
    }
    float noise = 0.0f;
    if (p.noise_sigma > 0.0f) {
      const uint64_t h = mix64(p.noise_seed ^ mix64((uint64_t)i));
      const float u1 = ((float)(uint32_t)(h >> 40) + 1.0f) * (1.0f / 16777217.0f);  // (0, 1]
      const float u2 = (float)(uint32_t)(h & 0xffffffu) * (1.0f / 16777216.0f);      // [0, 1)
      noise = sqrtf(-2.0f * logf(u1)) * cosf(6.28318530718f * u2) * p.noise_sigma;
    }
    if (p.rects_under_noise) {
      if (rect >= 0) v = (float)p.rects[rect][4];
      v += p.offset + noise;
    } else {
      v += p.offset + noise;
      if (rect >= 0) v = (float)p.rects[rect][4];
    }
    const float q = rintf(v * to_depth);
    out[y * p.pitch + x] = (uint16_t)fminf(fmaxf(q, 0.0f), maxv);
  }
}

}  // namespace

extern "C" int fvh_synth_plane(const fvh_synth_params* params, void* out, fvh_stream_t stream) {
  if (!params || !out || params->width <= 0 || params->height <= 0 ||
      params->pitch < params->width || params->rect_count < 0 || params->rect_count > 20 ||
      !(params->bit_depth == 8 || params->bit_depth == 10 || params->bit_depth == 12 ||
        params->bit_depth == 16)) {
    return FVH_EARG;
  }
  const int64_t n = params->width * params->height;
  const int64_t blocks = (n + 255) / 256;
  const dim3 grid((unsigned)(blocks < 4096 ? blocks : 4096));
  hipLaunchKernelGGL(k_synth_plane, grid, dim3(256), 0, (hipStream_t)stream, *params,
                     (uint16_t*)out);
  return hipGetLastError() == hipSuccess ? FVH_OK : FVH_EHIP;
}
