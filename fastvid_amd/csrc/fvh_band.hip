// fvh_band.hip -- throughput wavefront for the closed-loop predictor
// (encoder) and the reconstruction (decoder) on gfx950.
//
// Mapping: a wave64 is four DPP rows of 16 lanes; each row owns one tile.
// Inside a tile, rows are taken in bands of 16: lane i of the row owns tile
// row 16*b + i of band b and runs 8 steps behind lane i-1.  Because the lag
// is 8, every lane's column satisfies x == step (mod 8), so each lane moves 8
// samples (16 bytes) per 8 steps with one vector load and one vector store,
// and inside an 8-step group every sample sits at a compile-time register
// position.  The "above" neighbour of lane i is lane i-1's reconstruction
// from 8 steps earlier: one DPP row_shr:1 per sample, no LDS and no barrier.
// Lane 0 of a row takes the previous band's last row from a per-row LDS ring
// written by lane 15 one band period (>= 128 steps) earlier.  A tile of
// W x H costs ceil(H/16) * max(W,128) + 120 steps for 16*W*H/16 samples per
// lane, i.e. >= 94% lane utilisation on 256x128 tiles, against the
// reference's W+H-1 barrier-separated antidiagonals with neighbours re-read
// from global memory (cuda/csrc/encode.cu:150-192, cuda/csrc/decode.cu:572-605).
//
// Eligible frames (host flag FrameDesc::band): 16-byte aligned planes, every
// tile width, tile origin, pitch and residual base a multiple of 8 samples,
// tile width <= kBandRingMax.  Other frames use k_wavefront.
#include <type_traits>

#include "fvh_device.h"
#include "fvh_kernels.h"
#include "fvh_timing.h"

namespace fvh {

constexpr int kBandRows = 16;
constexpr int kBandLag = 8;
constexpr int kBandSkew = kBandLag * (kBandRows - 1);  // 120

__device__ __forceinline__ int dpp_row_shr1(int v) {
  // row_shr:1 (0x111): lane i of each 16-lane row receives lane i-1; lane 0
  // receives 0 (bound_ctrl) and is overridden by the caller.
  return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
}

template <typename FT>
struct Vec8;  // eight residuals / samples of one lane

template <>
struct Vec8<uint16_t> {
  uint4 v;
  __device__ __forceinline__ uint32_t get(int j) const {
    const uint32_t w = j < 2 ? v.x : j < 4 ? v.y : j < 6 ? v.z : v.w;
    return (j & 1) ? (w >> 16) : (w & 0xffffu);
  }
  __device__ __forceinline__ void load(const uint16_t* p) { v = *(const uint4*)p; }
  __device__ __forceinline__ void store(uint16_t* p) const { *(uint4*)p = v; }
  __device__ __forceinline__ void zero() { v = make_uint4(0, 0, 0, 0); }
  __device__ __forceinline__ void set(int j, uint32_t x) {
    uint32_t& w = j < 2 ? v.x : j < 4 ? v.y : j < 6 ? v.z : v.w;
    if (j & 1) w = (w & 0xffffu) | (x << 16);
    else w = x & 0xffffu;
  }
};

template <>
struct Vec8<uint32_t> {
  uint4 a, b;
  __device__ __forceinline__ uint32_t get(int j) const {
    switch (j) {
      case 0: return a.x; case 1: return a.y; case 2: return a.z; case 3: return a.w;
      case 4: return b.x; case 5: return b.y; case 6: return b.z; default: return b.w;
    }
  }
  __device__ __forceinline__ void load(const uint32_t* p) {
    a = *(const uint4*)p;
    b = *(const uint4*)(p + 4);
  }
  __device__ __forceinline__ void store(uint32_t* p) const {
    *(uint4*)p = a;
    *(uint4*)(p + 4) = b;
  }
  __device__ __forceinline__ void zero() {
    a = make_uint4(0, 0, 0, 0);
    b = a;
  }
  __device__ __forceinline__ void set(int j, uint32_t x) {
    switch (j) {
      case 0: a.x = x; break; case 1: a.y = x; break; case 2: a.z = x; break; case 3: a.w = x; break;
      case 4: b.x = x; break; case 5: b.y = x; break; case 6: b.z = x; break; default: b.w = x; break;
    }
  }
};

// Exact closed-loop quantiser q = sign(r) * floor((|r| + step/2) / step)
// (cuda/csrc/encode.cu:183-185) without abs or negation:
//   r >= 0: q = floor((r + h) / d)
//   r <  0: q = ceil((r - h) / d) = floor((r - h + d - 1) / d)
// Adding K*d (K = ceil(max/d)) makes the numerator n non-negative, then
// q = floor(n / d) - K.  uint16 residual batches (bit depth <= 12): n < 2^(bd+2)
// and m = ceil(2^s / d) with s = bd + 2 + ceil(log2 d) keep n*m < 2^32, so one
// 24-bit multiply and a shift divide exactly.  uint32 batches (16-bit frames):
// n < 2^18, mul_hi by ceil(2^32 / d).  Both are checked over every residual of
// every reachable step in tests/test_capi.py.
struct QuantMagic {
  uint32_t m;
  uint32_t s;
  int32_t K;
  int32_t offp;
  int32_t offn;
};

template <typename FT>
__device__ __forceinline__ QuantMagic quant_params(int step, int bit_depth, int max_sample) {
  QuantMagic q;
  const int h = step >> 1;
  q.K = (max_sample + step - 1) / step;
  q.offp = h + q.K * step;
  q.offn = step - 1 - h + q.K * step;
  if (sizeof(FT) == 2) {
    const uint32_t c = step > 1 ? 32 - __builtin_clz((uint32_t)step - 1) : 0;
    q.s = (uint32_t)bit_depth + 2 + c;
    q.m = (uint32_t)(((1ull << q.s) + (uint64_t)step - 1) / (uint64_t)step);
  } else {
    q.s = 32;
    q.m = step > 1 ? quant_magic(step) : 0u;
  }
  return q;
}

template <typename FT>
__device__ __forceinline__ int quantize(int r, const QuantMagic& q) {
  const uint32_t n = (uint32_t)(r + (r < 0 ? q.offn : q.offp));
  uint32_t f;
  if (sizeof(FT) == 2) f = __umul24(n, q.m) >> q.s;
  else f = (uint32_t)(((uint64_t)n * q.m) >> 32);
  return (int)f - q.K;
}

// v_med3_i32: the compiler folds med3(0, hi, v) into a min/max pair.
__device__ __forceinline__ int med3_asm(int a, int b, int c) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// kEncode: IN = uint16 source samples (row pitch), OUT = residuals FT (tile rows)
// decode:  IN = residuals FT (tile rows),          OUT = uint16 samples (row pitch)
template <bool kEncode, bool kMed, typename FT>
__global__ __launch_bounds__(256) void k_band(const FrameDesc* __restrict__ frames,
                                              FrameState* __restrict__ states,
                                              const TileDesc* __restrict__ tiles, int tile_total,
                                              FT* __restrict__ folded, char* gbase) {
  using InT = typename std::conditional<kEncode, uint16_t, FT>::type;
  using OutT = typename std::conditional<kEncode, FT, uint16_t>::type;
  __shared__ uint4 ring_all[16][kBandRingMax / 8];
  const int lane = lane_id();
  const int grp = lane >> 4;
  const int i = lane & 15;
  const int wave = threadIdx.x >> 6;
  const int t_idx = (blockIdx.x * 4 + wave) * 4 + grp;
  uint4* ring = ring_all[wave * 4 + grp];

  // ---- per row-group tile setup
  bool has = t_idx < tile_total;
  TileDesc td = {};
  int step = 1, max_sample = 0, bit_depth = 8;
  FrameState* st = nullptr;
  if (has) {
    td = tiles[t_idx];
    const FrameDesc& f = frames[td.frame];
    st = &states[td.frame];
    has = f.band != 0;
    if (kEncode) {
      has = has && st->active && st->step > 1 && (st->use_med != 0) == kMed;
      step = st->step;
    } else {
      has = has && st->status == 0 && (f.use_med != 0) == kMed;
      step = f.step;
    }
    max_sample = f.max_sample;
    bit_depth = f.bit_depth;
  }
  const int W = has ? td.width : 0;
  const int H = has ? td.height : 0;
  const int nb = (H + kBandRows - 1) / kBandRows;
  const int P = max((W + 7) & ~7, 128);
  const int my_steps = has ? (nb - 1) * P + W + kBandSkew : 0;
  const int total = (int)wave_max((uint32_t)my_steps);
  if (total == 0) return;

  const QuantMagic qm = quant_params<FT>(step, bit_depth, max_sample);
  const InT* in_base = nullptr;
  OutT* out_base = nullptr;
  int64_t in_pitch = 0, out_pitch = 0;
  if (has) {
    const FrameDesc& f = frames[td.frame];
    if (kEncode) {
      in_base = as_global(gbase, (const InT*)(f.plane[td.plane] + td.src_offset));
      in_pitch = f.pitch[td.plane];
      out_base = (OutT*)(folded + td.folded_base);
      out_pitch = W;
    } else {
      in_base = (const InT*)(folded + td.folded_base);
      in_pitch = W;
      out_base = as_global(gbase, (OutT*)(f.out_plane[td.plane] + td.dst_offset));
      out_pitch = f.plane_width[td.plane];
    }
  }
  auto row_ok = [&](int band) { return band < nb && band * kBandRows + i < H; };

  // Lane state.  x is this lane's column for the current 8-step group (lanes
  // start staggered by 8 steps: x < 0 means not started).  Input groups are
  // prefetched four groups ahead into a ring of four register buffers.
  int x = -kBandLag * i;
  int b = 0;
  int L = 0, UL = 0;
  int r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0;
  int bad = 0;
  Vec8<InT> buf0, buf1, buf2, buf3;
  buf0.zero();
  buf1.zero();
  buf2.zero();
  buf3.zero();
  // position of the group `ahead` groups after (x, b)
  auto load_ahead = [&](Vec8<InT>& dst, int ahead) {
    int nx = x + ahead * kBandLag, nbd = b;
    if (nx >= P) {
      nx -= P;
      ++nbd;
    }
    if (nx >= 0 && nx < W && row_ok(nbd)) {
      dst.load(in_base + (int64_t)(nbd * kBandRows + i) * in_pitch + nx);
    }
  };
  load_ahead(buf0, 0);
  load_ahead(buf1, 1);
  load_ahead(buf2, 2);
  load_ahead(buf3, 3);
  uint4 rv = make_uint4(0, 0, 0, 0);
  if (i == 0 && x >= 0 && x < W && b > 0 && row_ok(b)) rv = ring[x >> 3];

  auto group = [&](Vec8<InT>& cur) {
    const bool act = x >= 0 && x < W && row_ok(b);
    if (x == 0) {
      L = 0;
      UL = 0;
    }
    Vec8<OutT> out;
    out.zero();
    int rr[8] = {r0, r1, r2, r3, r4, r5, r6, r7};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int from_prev = dpp_row_shr1(rr[j]);
      const uint32_t rw = j < 2 ? rv.x : j < 4 ? rv.y : j < 6 ? rv.z : rv.w;
      const int ring_v = (int)((j & 1) ? (rw >> 16) : (rw & 0xffffu));
      const int above = i == 0 ? ring_v : from_prev;
      const int grad = L + above - UL;
      const int pred = kMed ? median3(L, above, grad) : med3_asm(0, max_sample, grad);
      int rec;
      if (kEncode) {
        const int sample = (int)cur.get(j);
        bad |= act & (sample > max_sample);
        const int q = quantize<FT>(sample - pred, qm);
        rec = med3_asm(0, max_sample, pred + __mul24(q, step));
        out.set(j, ((uint32_t)q << 1) ^ (uint32_t)(q >> 31));
      } else {
        const uint32_t fv = cur.get(j);
        const int q = (int)(fv >> 1) ^ -(int)(fv & 1u);
        rec = med3_asm(0, max_sample, pred + __mul24(q, step));
        out.set(j, (uint32_t)rec);
      }
      rr[j] = rec;
      L = rec;
      UL = above;
    }
    r0 = rr[0]; r1 = rr[1]; r2 = rr[2]; r3 = rr[3];
    r4 = rr[4]; r5 = rr[5]; r6 = rr[6]; r7 = rr[7];
    if (act) {
      out.store(out_base + (int64_t)(b * kBandRows + i) * out_pitch + x);
      if (i == kBandRows - 1 && b + 1 < nb) {
        ring[x >> 3] = make_uint4((uint32_t)r0 | ((uint32_t)r1 << 16),
                                  (uint32_t)r2 | ((uint32_t)r3 << 16),
                                  (uint32_t)r4 | ((uint32_t)r5 << 16),
                                  (uint32_t)r6 | ((uint32_t)r7 << 16));
      }
    }
    // refill this buffer with the group four ahead, then advance
    load_ahead(cur, 4);
    x += kBandLag;
    if (x >= P) {
      x -= P;
      ++b;
    }
    // ring words for the next group (written by lane 15 at least one group earlier)
    rv = make_uint4(0, 0, 0, 0);
    if (i == 0 && x >= 0 && x < W && b > 0 && row_ok(b)) rv = ring[x >> 3];
  };

  for (int t = 0; t < total; t += 4 * kBandLag) {
    group(buf0);
    if (t + kBandLag >= total) break;
    group(buf1);
    if (t + 2 * kBandLag >= total) break;
    group(buf2);
    if (t + 3 * kBandLag >= total) break;
    group(buf3);
  }
  if (kEncode) {
    const uint64_t any = ballot(bad != 0);
    if (i == 0 && st != nullptr && ((any >> (grp * 16)) & 0xffffull) != 0) {
      set_status(st, kStatusSampleRange);
    }
  }
}

template <bool kEncode, typename FT>
void launch_band(const FrameDesc* frames, FrameState* states, const TileDesc* tiles, int tile_total,
                 FT* folded, bool med, bool plain, hipStream_t stream) {
  const dim3 grid((tile_total + 15) / 16);
  if (med) {
    FVH_TIMED(kEncode ? "k_band_enc" : "k_band_dec", stream);
    hipLaunchKernelGGL((k_band<kEncode, true, FT>), grid, dim3(256), 0, stream, frames, states,
                       tiles, tile_total, folded, (char*)folded);
  }
  if (plain) {
    FVH_TIMED(kEncode ? "k_band_enc_grad" : "k_band_dec_grad", stream);
    hipLaunchKernelGGL((k_band<kEncode, false, FT>), grid, dim3(256), 0, stream, frames, states,
                       tiles, tile_total, folded, (char*)folded);
  }
}

template void launch_band<true, uint16_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                          uint16_t*, bool, bool, hipStream_t);
template void launch_band<true, uint32_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                          uint32_t*, bool, bool, hipStream_t);
template void launch_band<false, uint16_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                           uint16_t*, bool, bool, hipStream_t);
template void launch_band<false, uint32_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                           uint32_t*, bool, bool, hipStream_t);

}  // namespace fvh
