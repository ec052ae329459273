// fvh_encode.hip -- MI355X encoder kernels.
//
// Pipeline (all device-side, one launch per stage for the whole batch, no host
// synchronisation inside a call):
//   (k_plan, fvh_plan.hip: tile geometry, pass states and work counters)
//   k_band         band wavefront predictor + quantiser (fvh_band.hip)
//   k_wavefront    closed-loop predictor + quantiser, one wave64 per tile
//   k_lossless     step-1 predictor (reconstruction == source: fully parallel)
//   k_generic      wide-tile fallback (antidiagonal, recon in global memory)
//   k_analyze      per-shard order-0 histogram/normalisation/sparse table and
//                  exact zero-run / Rice / block-pack costs (wave per shard)
//   k_rans         order-0 rANS lanes, 16 shards per wave, final mode choice
//   k_tile_scan    record offsets inside each tile, tile payload lengths
//   k_frame_finish tile offsets, frame payload, header + directory
//   k_decide       gray10 MED / gray16 refinement decisions between passes
//   k_emit         zero-run / Rice / block-pack records assembled in LDS, copied
//                  to the stream; k_emit_o0 copies order-0 records from scratch
#include "fvh_device.h"
#include "fvh_kernels.h"
#include "fvh_timing.h"

namespace fvh {

// ------------------------------------------------------------ frame upload

constexpr int kUploadChunk = 8;
struct FrameChunk {
  FrameDesc f[kUploadChunk];
  int n;
  int first;
};

__global__ void k_upload(FrameChunk chunk, FrameDesc* __restrict__ dst) {
  const int i = threadIdx.x;
  if (i < chunk.n) dst[chunk.first + i] = chunk.f[i];
}

void launch_upload_frames(const FrameDesc* host_frames, int frame_count, FrameDesc* device_frames,
                          hipStream_t stream) {
  for (int first = 0; first < frame_count; first += kUploadChunk) {
    FrameChunk chunk;
    chunk.n = min(kUploadChunk, frame_count - first);
    chunk.first = first;
    for (int i = 0; i < chunk.n; ++i) chunk.f[i] = host_frames[first + i];
    hipLaunchKernelGGL(k_upload, dim3(1), dim3(64), 0, stream, chunk, device_frames);
  }
}

// ----------------------------------------------------------- pass control

// Between passes: the multi-pass heuristics of encode.cu:812-847.
//   stage 1 (after the first pass): gray10 lossy re-runs with MED when the
//     payload exceeds 0.3 B/sample; gray16 lossy takes the refined step when
//     the payload is < 0.1 or > 0.5 B/sample (direct) or in [0.1, 0.2) (probe).
//   stage 2 (after the refined pass): a probed gray16 frame keeps the refined
//     step only below 0.1 B/sample, otherwise re-runs the baseline step.
//   stage 3: every frame active again for assembly and emission.
__global__ void k_decide(const FrameDesc* __restrict__ frames, FrameState* __restrict__ states,
                         int frame_count, int stage) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= frame_count) return;
  const FrameDesc& f = frames[i];
  FrameState& s = states[i];
  const int64_t samples = f.total_samples;
  if (stage == 1) {
    const int was_active = s.active;
    s.active = 0;
    if (!was_active || s.status != 0) return;
    s.baseline_payload = s.payload;
    if (f.heuristic == 1) {
      s.gray_med = s.payload * 10 > samples * 3;
      if (s.gray_med) {
        s.use_med = 1;
        s.active = 1;
      }
    } else if (f.heuristic == 2) {
      const int64_t base = s.payload;
      s.direct = base < samples / 10 || base > samples / 2;
      const int probe = !s.direct && base < samples / 5;
      if (s.direct || probe) {
        s.baseline_step = s.step;
        s.step = f.step_refined;
        s.use_med = 0;
        s.active = 1;
        s.stage = 2;
      }
    }
  } else if (stage == 2) {
    const int was_active = s.active;
    s.active = 0;
    if (!was_active || s.status != 0 || f.heuristic != 2 || s.stage != 2) return;
    s.refined = s.direct || s.payload < samples / 10;
    if (!s.refined) {
      s.step = s.baseline_step;
      s.active = 1;
    }
  } else {
    s.active = 1;
  }
}

// ---------------------------------------------------- predictor + quantiser

// One wave64 per tile.  Lane l owns rows l, 64+l, 128+l, ... of the tile and
// runs one step behind lane l-1, so at every step lane l needs the sample
// "above" that lane l-1 reconstructed at the previous step: a single wave
// shuffle, no barrier.  Consecutive 64-row bands are pipelined back to back;
// lane 0 takes the last row of the previous band from an LDS row written by
// lane 63 (W-63 steps earlier).  Total steps per tile: bands*max(W,64)+W+63
// instead of the reference's W+H-1 barrier-separated antidiagonals
// (encode.cu:150-192, decode.cu:572-605), with no reconstruction buffer in
// global memory.
//
// kEncode: source -> residual (zigzag, tile order) with closed-loop
//          reconstruction (encode.cu:159-189).
// decode:  residual -> reconstruction written to the output plane
//          (decode.cu:581-602).
template <bool kEncode, typename FT>
__global__ __launch_bounds__(256) void k_wavefront(const FrameDesc* __restrict__ frames,
                                                   FrameState* __restrict__ states,
                                                   const TileDesc* __restrict__ tiles,
                                                   int tile_total, FT* __restrict__ folded,
                                                   char* gbase) {
  __shared__ uint16_t ring_all[4][kRingMax];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int t = blockIdx.x * 4 + wave;
  if (t >= tile_total) return;
  const TileDesc td = tiles[t];
  FrameState* st = &states[td.frame];
  const FrameDesc& f = frames[td.frame];
  int step, med;
  if (kEncode) {
    if (!st->active) return;
    step = st->step;
    med = st->use_med;
    if (step == 1) return;  // lossless: k_lossless
    if (f.band) return;     // k_band
  } else {
    if (f.band) return;     // k_band
    if (st->status != 0) return;
    step = f.step;
    med = f.use_med;
  }
  const int W = td.width, H = td.height;
  if (W > kRingMax) return;  // k_generic
  const int max_sample = f.max_sample;
  const uint32_t half = (uint32_t)(step >> 1);
  const uint32_t magic = step > 1 ? quant_magic(step) : 0u;
  uint16_t* ring = ring_all[wave];
  const int nb = (H + 63) >> 6;
  const int P = W > 64 ? W : 64;
  const int total = (nb - 1) * P + W + 63;
  FT* fold = folded + td.folded_base;
  const uint16_t* src = nullptr;
  int64_t src_pitch = 0;
  uint16_t* dst = nullptr;
  int64_t dst_pitch = 0;
  if (kEncode) {
    src = as_global(gbase, f.plane[td.plane] + td.src_offset);
    src_pitch = f.pitch[td.plane];
  } else {
    dst = as_global(gbase, f.out_plane[td.plane] + td.dst_offset);
    dst_pitch = f.plane_width[td.plane];
  }
  int b = 0, x = -lane;
  int L = 0, UL = 0, prev = 0, bad = 0;
  for (int s = 0; s < total; ++s) {
    const int nbr = shfl_up(prev, 1);
    const bool in_tile = x >= 0 && x < W && b < nb;
    const int y = (b << 6) + lane;
    int ringv = 0;
    if (lane == 0 && b > 0 && in_tile) ringv = ring[x];
    if (in_tile && y < H) {
      const int above = y == 0 ? 0 : (lane == 0 ? ringv : nbr);
      const int left = x > 0 ? L : 0;
      const int ul = (x > 0 && y > 0) ? UL : 0;
      const int grad = left + above - ul;
      const int pred = med ? median3(left, above, grad) : clampi(grad, 0, max_sample);
      int rec;
      if (kEncode) {
        const int sample = src[(int64_t)y * src_pitch + x];
        bad |= sample > max_sample;
        const int r = sample - pred;
        const uint32_t mag = __umulhi((uint32_t)abs(r) + half, magic);
        const int q = r < 0 ? -(int)mag : (int)mag;
        rec = clampi(pred + q * step, 0, max_sample);
        fold[(int64_t)y * W + x] = (FT)(((uint32_t)q << 1) ^ (uint32_t)(q >> 31));
      } else {
        const uint32_t fv = fold[(int64_t)y * W + x];
        const int q = (fv & 1u) ? -(int)(fv >> 1) - 1 : (int)(fv >> 1);
        rec = clampi(pred + q * step, 0, max_sample);
        dst[(int64_t)y * dst_pitch + x] = (uint16_t)rec;
      }
      if (lane == 63 && b + 1 < nb) ring[x] = (uint16_t)rec;
      L = rec;
      UL = above;
      prev = rec;
    }
    if (++x == P) {
      x = 0;
      ++b;
    }
  }
  if (kEncode && ballot(bad != 0) != 0 && lane == 0) set_status(st, kStatusSampleRange);
}

// Step 1: the reconstruction equals the source for in-range samples, so each
// residual depends only on source neighbours and every sample is independent.
template <typename FT>
__global__ __launch_bounds__(256) void k_lossless(const FrameDesc* __restrict__ frames,
                                                  FrameState* __restrict__ states,
                                                  const TileDesc* __restrict__ tiles,
                                                  FT* __restrict__ folded, char* gbase) {
  const int t = blockIdx.x;
  const TileDesc td = tiles[t];
  FrameState* st = &states[td.frame];
  if (!st->active || st->step != 1) return;
  const FrameDesc& f = frames[td.frame];
  const int med = st->use_med;
  const int max_sample = f.max_sample;
  const uint16_t* src = as_global(gbase, f.plane[td.plane] + td.src_offset);
  const int64_t pitch = f.pitch[td.plane];
  FT* fold = folded + td.folded_base;
  const int W = td.width, H = td.height;
  int bad = 0;
  for (int y = threadIdx.x >> 6; y < H; y += 4) {
    const uint16_t* row = src + (int64_t)y * pitch;
    for (int x = lane_id(); x < W; x += 64) {
      const int sample = row[x];
      const int left = x > 0 ? row[x - 1] : 0;
      const int above = y > 0 ? row[x - pitch] : 0;
      const int ul = (x > 0 && y > 0) ? row[x - pitch - 1] : 0;
      const int grad = left + above - ul;
      const int pred = med ? median3(left, above, grad) : clampi(grad, 0, max_sample);
      const int q = sample - pred;
      bad |= sample > max_sample;
      fold[(int64_t)y * W + x] = (FT)(((uint32_t)q << 1) ^ (uint32_t)(q >> 31));
    }
  }
  if (ballot(bad != 0) != 0 && lane_id() == 0) set_status(st, kStatusSampleRange);
}

// Tiles wider than kRingMax: the reference's antidiagonal schedule with the
// reconstruction kept in global memory (tile order for the encoder, the
// output plane for the decoder).
template <bool kEncode, typename FT>
__global__ __launch_bounds__(256) void k_generic(const FrameDesc* __restrict__ frames,
                                                 FrameState* __restrict__ states,
                                                 const TileDesc* __restrict__ tiles,
                                                 FT* __restrict__ folded,
                                                 uint16_t* __restrict__ recon) {
  const int t = blockIdx.x;
  const TileDesc td = tiles[t];
  FrameState* st = &states[td.frame];
  const FrameDesc& f = frames[td.frame];
  const int64_t W = td.width, H = td.height;
  if (W <= kRingMax) return;
  int step, med;
  if (kEncode) {
    if (!st->active) return;
    step = st->step;
    med = st->use_med;
  } else {
    if (st->status != 0) return;
    step = f.step;
    med = f.use_med;
  }
  const int max_sample = f.max_sample;
  const uint32_t half = (uint32_t)(step >> 1);
  const uint32_t magic = step > 1 ? quant_magic(step) : 0u;
  FT* fold = folded + td.folded_base;
  uint16_t* rc;
  int64_t rc_pitch;
  const uint16_t* src = nullptr;
  int64_t src_pitch = 0;
  if (kEncode) {
    rc = recon + td.folded_base;
    rc_pitch = W;
    src = f.plane[td.plane] + td.src_offset;
    src_pitch = f.pitch[td.plane];
  } else {
    rc = f.out_plane[td.plane] + td.dst_offset;
    rc_pitch = f.plane_width[td.plane];
  }
  int bad = 0;
  for (int64_t d = 0; d < W + H - 1; ++d) {
    const int64_t y0 = d - W + 1 > 0 ? d - W + 1 : 0;
    const int64_t y1 = d < H - 1 ? d : H - 1;
    for (int64_t item = threadIdx.x; item <= y1 - y0; item += blockDim.x) {
      const int64_t y = y0 + item;
      const int64_t x = d - y;
      uint16_t* p = rc + y * rc_pitch + x;
      const int left = x > 0 ? p[-1] : 0;
      const int above = y > 0 ? p[-rc_pitch] : 0;
      const int ul = (x > 0 && y > 0) ? p[-rc_pitch - 1] : 0;
      const int grad = left + above - ul;
      const int pred = med ? median3(left, above, grad) : clampi(grad, 0, max_sample);
      if (kEncode) {
        const int sample = src[y * src_pitch + x];
        bad |= sample > max_sample;
        const int r = sample - pred;
        int q;
        if (step == 1) {
          q = r;
        } else {
          const uint32_t mag = __umulhi((uint32_t)abs(r) + half, magic);
          q = r < 0 ? -(int)mag : (int)mag;
        }
        *p = (uint16_t)clampi(pred + q * step, 0, max_sample);
        fold[y * W + x] = (FT)(((uint32_t)q << 1) ^ (uint32_t)(q >> 31));
      } else {
        const uint32_t fv = fold[y * W + x];
        const int q = (fv & 1u) ? -(int)(fv >> 1) - 1 : (int)(fv >> 1);
        *p = (uint16_t)clampi(pred + q * step, 0, max_sample);
      }
    }
    __syncthreads();
  }
  if (kEncode && bad) set_status(st, kStatusSampleRange);
}

// ------------------------------------------------------------- analysis

// Clamp function a -> min(max(a + d, lo), hi).  The order-0 normalisation
// end_i = min(max(E_i, end_{i-1} + 1), 4096 - remaining_i) (encode.cu:370-390)
// is a chain of such functions; clamp functions compose into clamp
// functions, so the chain is evaluated as a wave-wide prefix scan.
struct ClampFn {
  int d, lo, hi;
};

__device__ __forceinline__ ClampFn clamp_identity() { return ClampFn{0, -(1 << 30), 1 << 30}; }

// first g1, then g2
__device__ __forceinline__ ClampFn clamp_compose(ClampFn g1, ClampFn g2) {
  ClampFn r;
  r.d = g1.d + g2.d;
  r.lo = clampi(g1.lo + g2.d, g2.lo, g2.hi);
  r.hi = clampi(g1.hi + g2.d, g2.lo, g2.hi);
  return r;
}

__device__ __forceinline__ int clamp_apply(ClampFn g, int a) { return min(max(a + g.d, g.lo), g.hi); }

__device__ __forceinline__ ClampFn norm_fn(uint32_t prefix, int count, int remaining) {
  if (remaining == 0) return ClampFn{0, 4096, 4096};
  const int e = count == kShardSymbols ? (int)prefix
                                        : (int)(((uint64_t)prefix << kRansTableLog) / (uint64_t)count);
  const int u = 4096 - remaining;
  if (e > u) return ClampFn{0, u, u};
  return ClampFn{1, e, u};
}

__device__ __forceinline__ int write_varint_bytes(uint8_t* out, uint32_t value) {
  int n = 0;
  do {
    uint8_t byte = (uint8_t)(value & 0x7f);
    value >>= 7;
    if (value != 0) byte |= 0x80;
    out[n++] = byte;
  } while (value != 0);
  return n;
}

// Residuals of one lane: eight consecutive symbols per 16-byte quad (u16)
// or two quads (u32), as a native vector so they stay in registers.
using q4u = uint32_t __attribute__((ext_vector_type(4)));
using p16 = uint16_t __attribute__((ext_vector_type(2)));

// Histogram bins: every value from kRansAlphabet on shares the last bin (a
// shard holding one has no order-0 table).  Symbol b of a chunk has class
// b % 4; classes 2p and 2p + 1 count in the low and high halves of the
// words of pair p's rows.  Zeros, the commonest value by far, are counted
// from the zero masks instead: their atomics go to a junk slot past the
// bins, one of eight by lane, so they no longer all meet on bin 0.
constexpr uint32_t kHistLast = kRansAlphabet;  // 511

__device__ __forceinline__ void hist_pair(uint32_t* row, uint32_t lo, uint32_t hi) {
  atomicAdd(&row[lo], 1u);
  atomicAdd(&row[hi], 0x10000u);
}

// The same pair into a lane's private rows (bins 64 words apart; values
// below the small histogram's bins, zeros included: bin 0 is replaced by the
// zero masks' counts)
__device__ __forceinline__ void hist_pair_private(uint32_t* row, uint32_t lo, uint32_t hi) {
  atomicAdd(&row[64 * lo], 1u);
  atomicAdd(&row[64 * hi], 0x10000u);
}

template <typename FT>
struct Chunk8;

template <>
struct Chunk8<uint16_t> {
  q4u v;
  // the order-0 path reads residuals here for the last time: non-temporal
  // loads keep the rank bytes written next in cache (analysis 384 -> 352 us
  // on 16 4K RGB frames)
  __device__ __forceinline__ void load(const uint16_t* p) {
    v = __builtin_nontemporal_load((const q4u*)p);
  }
  // zero the elements from n on (0 <= n <= 8)
  __device__ __forceinline__ void keep(int n) {
    const uint32_t m0 = n >= 2 ? ~0u : n == 1 ? 0xffffu : 0u;
    const uint32_t m1 = n >= 4 ? ~0u : n == 3 ? 0xffffu : 0u;
    const uint32_t m2 = n >= 6 ? ~0u : n == 5 ? 0xffffu : 0u;
    const uint32_t m3 = n >= 8 ? ~0u : n == 7 ? 0xffffu : 0u;
    v.x &= m0;
    v.y &= m1;
    v.z &= m2;
    v.w &= m3;
  }
  // a fresh copy for the optimiser: values unpacked before this point are
  // not reused after it, which keeps them from staying live across phases
  __device__ __forceinline__ void launder() { asm volatile("" : "+v"(v)); }
  __device__ __forceinline__ uint32_t get(int b) const {
    const uint32_t w = b < 2 ? v.x : b < 4 ? v.y : b < 6 ? v.z : v.w;
    return (b & 1) ? (w >> 16) : (w & 0xffffu);
  }
  // low bytes of symbols r + 4 and r as bytes 0 and 1 (r < 4)
  __device__ __forceinline__ uint32_t low_pair(int r) const {
    return __builtin_amdgcn_perm(r < 2 ? v.x : v.y, r < 2 ? v.z : v.w,
                                 (r & 1) ? 0x0c0c0602u : 0x0c0c0400u);
  }
  static __device__ __forceinline__ p16 pk(uint32_t w) { return __builtin_bit_cast(p16, w); }
  static __device__ __forceinline__ uint32_t un(p16 w) { return __builtin_bit_cast(uint32_t, w); }
  // histogram atomics (zeros to the junk slot) and the zero bits: bit b
  // set where symbol b is zero
  __device__ __forceinline__ uint32_t hist_add(uint32_t* row0, uint32_t* row1, uint32_t junk) const {
    const p16 last = {(uint16_t)kHistLast, (uint16_t)kHistLast};
    const p16 one = {1, 1};
    const p16 jk = pk(junk);
    const p16 n0 = __builtin_elementwise_min(pk(v.x), one);  // 1 where nonzero
    const p16 n1 = __builtin_elementwise_min(pk(v.y), one);
    const p16 n2 = __builtin_elementwise_min(pk(v.z), one);
    const p16 n3 = __builtin_elementwise_min(pk(v.w), one);
    const uint32_t w0 = un((one - n0) * jk + __builtin_elementwise_min(pk(v.x), last));
    const uint32_t w1 = un((one - n1) * jk + __builtin_elementwise_min(pk(v.y), last));
    const uint32_t w2 = un((one - n2) * jk + __builtin_elementwise_min(pk(v.z), last));
    const uint32_t w3 = un((one - n3) * jk + __builtin_elementwise_min(pk(v.w), last));
    hist_pair(row0, w0 & 0xffffu, w0 >> 16);
    hist_pair(row1, w1 & 0xffffu, w1 >> 16);
    hist_pair(row0, w2 & 0xffffu, w2 >> 16);
    hist_pair(row1, w3 & 0xffffu, w3 >> 16);
    // nonzero flags of symbols 0, 2, 4, 6 at bits 0, 2, 4, 6 and of 1, 3,
    // 5, 7 at bits 16, 18, 20, 22
    const uint32_t nz = un(n0) | (un(n1) << 2) | (un(n2) << 4) | (un(n3) << 6);
    return ~(nz | (nz >> 15)) & 0xffu;
  }
  // the private histogram's atomics (every value below kSmallBins) and the
  // zero bits
  __device__ __forceinline__ uint32_t hist_add_private(uint32_t* row0, uint32_t* row1) const {
    hist_pair_private(row0, v.x & 0xffffu, v.x >> 16);
    hist_pair_private(row1, v.y & 0xffffu, v.y >> 16);
    hist_pair_private(row0, v.z & 0xffffu, v.z >> 16);
    hist_pair_private(row1, v.w & 0xffffu, v.w >> 16);
    const p16 one = {1, 1};
    const uint32_t nz = un(__builtin_elementwise_min(pk(v.x), one)) |
                        (un(__builtin_elementwise_min(pk(v.y), one)) << 2) |
                        (un(__builtin_elementwise_min(pk(v.z), one)) << 4) |
                        (un(__builtin_elementwise_min(pk(v.w), one)) << 6);
    return ~(nz | (nz >> 15)) & 0xffu;
  }
  __device__ __forceinline__ uint32_t max8() const {
    const uint32_t m = un(__builtin_elementwise_max(__builtin_elementwise_max(pk(v.x), pk(v.y)),
                                                    __builtin_elementwise_max(pk(v.z), pk(v.w))));
    return max(m & 0xffffu, m >> 16);
  }
  __device__ __forceinline__ uint32_t or8() const {
    const uint32_t o = v.x | v.y | v.z | v.w;
    return (o | (o >> 16)) & 0xffffu;
  }
};

// Eight 8-bit residuals, widened on load to the 16-bit layout (four byte
// permutes), after which every step is the 16-bit one
template <>
struct Chunk8<uint8_t> : Chunk8<uint16_t> {
  __device__ __forceinline__ void load(const uint8_t* p) {
    using u32x2 = uint32_t __attribute__((ext_vector_type(2)));
    const u32x2 b = __builtin_nontemporal_load((const u32x2*)p);
    v.x = __builtin_amdgcn_perm(0u, b.x, 0x0c010c00u);
    v.y = __builtin_amdgcn_perm(0u, b.x, 0x0c030c02u);
    v.z = __builtin_amdgcn_perm(0u, b.y, 0x0c010c00u);
    v.w = __builtin_amdgcn_perm(0u, b.y, 0x0c030c02u);
  }
};

template <>
struct Chunk8<uint32_t> {
  q4u a, c;
  __device__ __forceinline__ void load(const uint32_t* p) {
    a = __builtin_nontemporal_load((const q4u*)p);
    c = __builtin_nontemporal_load((const q4u*)(p + 4));
  }
  __device__ __forceinline__ void keep(int n) {
    a.x = n > 0 ? a.x : 0u;
    a.y = n > 1 ? a.y : 0u;
    a.z = n > 2 ? a.z : 0u;
    a.w = n > 3 ? a.w : 0u;
    c.x = n > 4 ? c.x : 0u;
    c.y = n > 5 ? c.y : 0u;
    c.z = n > 6 ? c.z : 0u;
    c.w = n > 7 ? c.w : 0u;
  }
  __device__ __forceinline__ void launder() { asm volatile("" : "+v"(a), "+v"(c)); }
  __device__ __forceinline__ uint32_t get(int b) const {
    switch (b) {
      case 0: return a.x; case 1: return a.y; case 2: return a.z; case 3: return a.w;
      case 4: return c.x; case 5: return c.y; case 6: return c.z; default: return c.w;
    }
  }
  // low bytes of symbols r + 4 and r as bytes 0 and 1 (r < 4)
  __device__ __forceinline__ uint32_t low_pair(int r) const {
    return __builtin_amdgcn_perm(get(r), get(r + 4), 0x0c0c0400u);
  }
  __device__ __forceinline__ uint32_t hist_add(uint32_t* row0, uint32_t* row1, uint32_t junk) const {
    const uint32_t j = junk & 0xffffu;
    auto bin = [&](uint32_t x) { return x == 0 ? j : min(x, kHistLast); };
    hist_pair(row0, bin(a.x), bin(a.y));
    hist_pair(row1, bin(a.z), bin(a.w));
    hist_pair(row0, bin(c.x), bin(c.y));
    hist_pair(row1, bin(c.z), bin(c.w));
    uint32_t z = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) z |= (uint32_t)(get(b) == 0) << b;
    return z;
  }
  __device__ __forceinline__ uint32_t hist_add_private(uint32_t* row0, uint32_t* row1) const {
    hist_pair_private(row0, a.x, a.y);
    hist_pair_private(row1, a.z, a.w);
    hist_pair_private(row0, c.x, c.y);
    hist_pair_private(row1, c.z, c.w);
    uint32_t z = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) z |= (uint32_t)(get(b) == 0) << b;
    return z;
  }
  __device__ __forceinline__ uint32_t max8() const {
    return max(max(max(a.x, a.y), max(a.z, a.w)), max(max(c.x, c.y), max(c.z, c.w)));
  }
  __device__ __forceinline__ uint32_t or8() const { return a.x | a.y | a.z | a.w | c.x | c.y | c.z | c.w; }
};

// Inclusive composition scan of clamp functions over the 64 lanes (lane 0
// first): row shifts then row broadcasts, lanes without a source composing
// with the identity.
__device__ __forceinline__ ClampFn clamp_scan(ClampFn g) {
  const ClampFn id = clamp_identity();
  auto step = [&](auto mov) __attribute__((always_inline)) {
    ClampFn o;
    o.d = (int)mov((uint32_t)id.d, (uint32_t)g.d);
    o.lo = (int)mov((uint32_t)id.lo, (uint32_t)g.lo);
    o.hi = (int)mov((uint32_t)id.hi, (uint32_t)g.hi);
    g = clamp_compose(o, g);
  };
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppRowShr + 1, 0xf, false>(o, v); });
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppRowShr + 2, 0xf, false>(o, v); });
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppRowShr + 4, 0xf, false>(o, v); });
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppRowShr + 8, 0xf, false>(o, v); });
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppBcast15, 0xa, false>(o, v); });
  step([](uint32_t o, uint32_t v) { return dpp_mov<kDppBcast31, 0xc, false>(o, v); });
  return g;
}

// Per-wave LDS of k_analyze.  Histogram rows are kHistRow words apart, so
// the copies of a value fall on different banks; lane l counts into copy
// l % kHistCopies, which halves the atomics that meet on one address (the
// residuals of a shard crowd into a few small values).  Words 512..519 of a
// row are the zeros' junk slots.  The list of present values by rank reuses
// the histogram once its bins are read.  The kernel is bound by the latency
// of each wave's phases, so the copies trade atomic collisions against waves
// per compute unit: on 16 4K RGB frames (most shards on the private
// histogram below) 4 copies and 9 waves took 319 us, 3 and 11 waves 278 us,
// 2 and 16 waves 250 us; sharing the rank map's bytes with the zero words
// (18 waves) 233 us.
constexpr int kHistCopies = 2;
constexpr int kHistRow = 512 + 8;
// Shards whose values are all below kSmallBins (most of them: the residuals
// of 8-bit content crowd into a few small values) count into a private
// histogram per lane instead, word [pair][value][lane]: every lane's
// counters sit in its own bank and no two lanes ever add to one address, so
// the atomics neither collide nor conflict.
constexpr int kSmallBins = 16;

// The rank bytes' staging image (four rows of 1024 bytes) in the histogram
// words past the list of present values (at most 255 entries of 16 bytes)
constexpr int kRankStageOffset = 4096;
constexpr int kRankStageRow = 1024 + 16;

struct AnalyzeLds {
  union {
    uint32_t hist[2][kHistCopies][kHistRow];  // [class pair][copy][bin]
    uint32_t priv[2][kSmallBins][64];         // [class pair][value][lane]
    uint4 list[512];  // by rank: {value, class pair 0 counts, class pair 1 counts, count}
  };
  uint32_t small_tot[2 * kSmallBins];  // [class pair][value]: private counts summed
  // the rank map is dead once the rank bytes are staged, before the zero
  // words are written (512 bytes less: 7 LDS granules, 18 waves per compute
  // unit instead of 16)
  union {
    uint8_t rank_of[512];  // value -> rank
    uint64_t zw[64];       // zero bits of symbols [64 j, 64 j + 64)
  };
};
static_assert(kRankStageOffset >= 255 * 16 &&
                  kRankStageOffset + 4 * kRankStageRow <= (int)sizeof(uint32_t) * 2 * kHistCopies * kHistRow,
              "rank staging inside the histogram, past the list");

// One wave64 per shard (one wave per block).  Lane l holds symbols
// 512 c + 8 l + b (chunk c < 8, b < 8), loaded with eight 16-byte loads
// issued together (every tile's residuals start 16-byte aligned and are
// padded to whole quads), so symbol i has Rice lane i % 4 = b % 4 and a
// 128-symbol block is one 16-lane row of one chunk.  Everything the
// reference computes with shared atomics, a tree reduction and one serial
// thread (encode.cu:195-431) becomes per-lane counts, class-split histogram
// atomics, ballots, popcounts and DPP scans here.  The order-0 table works
// on the compacted list of present values (rank i on lane i % 64), and the
// Rice and zero-run costs come from the same list, so only the histogram,
// the zero masks, the block ORs and the rank bytes touch every symbol.  A
// shard with a value past the histogram (no order-0 table) or more than 511
// distinct values falls back to per-symbol Rice sums.
template <typename FT>
__global__ __launch_bounds__(64) void k_analyze(const FrameDesc* __restrict__ frames,
                                                FrameState* __restrict__ states,
                                                const TileDesc* __restrict__ tiles,
                                                const ShardDesc* __restrict__ shards,
                                                int shard_total, const FT* __restrict__ folded,
                                                ShardRec* __restrict__ recs,
                                                uint8_t* __restrict__ scratch,
                                                uint8_t* __restrict__ rank_area) {
  __shared__ AnalyzeLds lds;
  const int lane = lane_id();
  const int s = blockIdx.x;
  if (s >= shard_total) return;
  const ShardDesc sd = shards[s];
  const TileDesc td = tiles[sd.tile];
  FrameState* st = &states[td.frame];
  if (!st->active) return;
  const int force = frames[td.frame].force_order0;
  const int64_t local0 = (int64_t)sd.local * kShardSymbols;
  const int count = (int)min((int64_t)kShardSymbols, (int64_t)td.width * td.height - local0);
  const FT* v = folded + td.folded_base + local0;
  // chunk c of this lane starts at symbol 512 c + 8 lane; quads past the
  // shard's last symbol are not read
  Chunk8<FT> ch[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int base = 512 * c + 8 * lane;
    ch[c].load(v + (base < count ? base : 0));
  }
  {
    uint4* h = (uint4*)&lds.hist[0][0][0];
    constexpr int kQuads = 2 * kHistCopies * kHistRow / 4;
    for (int i = lane; i < kQuads; i += 64) h[i] = make_uint4(0, 0, 0, 0);
  }
  // symbols past the shard's end read as zeros; they are taken out of the
  // zero masks and of the histogram's zero bins below
  const bool full = count == kShardSymbols;
  if (!full) {
#pragma unroll
    for (int c = 0; c < 8; ++c) ch[c].keep(clampi(count - (512 * c + 8 * lane), 0, 8));
  }
  __syncthreads();

  uint32_t* const row0 = lds.hist[0][lane % kHistCopies];
  uint32_t* const row1 = lds.hist[1][lane % kHistCopies];
  const uint32_t junk = (512u + ((uint32_t)lane >> 2 & 7u)) * 0x10001u;
  uint32_t lmax = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) lmax = max(lmax, ch[c].max8());
  const uint32_t maxv = wave_max(lmax);
  // (wave-uniform) every value of the shard has a bin of the private
  // histogram
  const bool small = maxv < (uint32_t)kSmallBins;
  int zeros = 0, starts = 0;
  uint32_t zmask[8];  // zero bits of each chunk's eight symbols
  uint32_t orv[8];    // OR of each chunk's eight symbols
  uint32_t prev_last = 0;  // symbol before lane 0's chunk is zero
  auto count_chunks = [&](auto add) __attribute__((always_inline)) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      uint32_t zm = add(ch[c]);
      orv[c] = ch[c].or8();
      if (!full) zm &= (1u << clampi(count - (512 * c + 8 * lane), 0, 8)) - 1u;
      zmask[c] = zm;
      zeros += __popc(zm);
      // the symbol before this lane's chunk: lane - 1's last, or for lane 0
      // the previous chunk's last symbol (lane 63)
      const uint32_t up = from_prev_lane(zm >> 7);
      const uint32_t before = lane == 0 ? prev_last : up;
      starts += __popc(zm & ~((zm << 1) | (before & 1u)));
      prev_last = (uint32_t)__builtin_amdgcn_readlane((int)(zm >> 7), 63);
    }
  };
  if (small) {
    uint32_t* const p0 = &lds.priv[0][0][lane];
    uint32_t* const p1 = &lds.priv[1][0][lane];
    count_chunks([&](const Chunk8<FT>& k) { return k.hist_add_private(p0, p1); });
  } else {
    count_chunks([&](const Chunk8<FT>& k) { return k.hist_add(row0, row1, junk); });
  }
  __syncthreads();
  zeros = wave_sum(zeros);
  starts = wave_sum(starts);
  const bool exact = maxv < kHistLast;  // every value in a bin of its own

  // record fields, kept as scalars until the one store at the end
  uint32_t r_mode = kModeOrder0, r_k = 0, r_best3_mode = 0, r_sparse_len = 0;
  uint32_t r_body = 0, r_best3_body = 0xffffffffu;
  uint32_t r_lane[4] = {0, 0, 0, 0};
  uint8_t* sbase = scratch + (int64_t)s * kScratchBytes;

  // ---------------- present values by rank
  // bins 8 lane + k, copies summed; bin 0 from the zero masks (padding
  // excluded), class r's zeros at bits r and r + 4 of each chunk's mask
  int distinct = 0;
  if (exact) {
    const uint32_t z03 = zmask[0] | (zmask[1] << 8) | (zmask[2] << 16) | (zmask[3] << 24);
    const uint32_t z47 = zmask[4] | (zmask[5] << 8) | (zmask[6] << 16) | (zmask[7] << 24);
    uint32_t zc[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t m = 0x11111111u << r;
      zc[r] = (uint32_t)(__popc(z03 & m) + __popc(z47 & m));
    }
    // class pairs as 16-bit halves (at most 1024 zeros a class)
    const uint32_t zero01 = wave_sum(zc[0] | (zc[1] << 16));
    const uint32_t zero23 = wave_sum(zc[2] | (zc[3] << 16));
    uint32_t c01[8], c23[8];
    if (small) {
      // row 16 p + v of the private histogram summed over the 64 lanes:
      // lane l adds half l & 1 of row l >> 1, then the pair's halves meet;
      // the totals of bins 8 l .. 8 l + 7 then go to lanes 0 and 1
      const int row = lane >> 1;
      const uint4* const src = (const uint4*)&lds.priv[row >> 4][row & 15][32 * (lane & 1)];
      uint32_t t = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint4 q = src[i];
        t += (q.x + q.y) + (q.z + q.w);
      }
      t += dpp_mov<kDppXor1>(0u, t);
      if ((lane & 1) == 0) lds.small_tot[row] = t;
      __syncthreads();
      const int at = lane < 2 ? 8 * lane : 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        c01[k] = lane < 2 ? lds.small_tot[at + k] : 0u;
        c23[k] = lane < 2 ? lds.small_tot[kSmallBins + at + k] : 0u;
      }
    } else {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        uint32_t* acc = p == 0 ? c01 : c23;
#pragma unroll
        for (int cp = 0; cp < kHistCopies; ++cp) {
          const uint4 q0 = *(const uint4*)&lds.hist[p][cp][8 * lane];
          const uint4 q1 = *(const uint4*)&lds.hist[p][cp][8 * lane + 4];
          const uint32_t w[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] = cp == 0 ? w[k] : acc[k] + w[k];
        }
      }
    }
    if (lane == 0) {
      c01[0] = zero01;
      c23[0] = zero23;
    }
    uint32_t cnt[8];
    uint32_t pm = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      cnt[k] = (c01[k] & 0xffffu) + (c01[k] >> 16) + (c23[k] & 0xffffu) + (c23[k] >> 16);
      pm |= (uint32_t)(cnt[k] != 0) << k;
    }
    const int npres = __popc(pm);
    const int incl = wave_incl_scan(npres);
    distinct = __builtin_amdgcn_readlane(incl, 63);
    const int rank0 = incl - npres;
    __syncthreads();  // the list overwrites the bins just read (one wave per block)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if ((pm >> k) & 1u) {
        lds.list[rank0 + __popc(pm & ((1u << k) - 1u))] =
            make_uint4((uint32_t)(8 * lane + k), c01[k], c23[k], cnt[k]);
      }
    }
    __syncthreads();
  }
  const bool supported = exact && distinct <= 255;
  const int blocks = (distinct + 63) >> 6;  // 64-rank blocks of the list

  // ---------------- order-0 table (encode.cu:337-392): frequencies by
  // the clamp-function scan of the normalisation, encoder entries by rank,
  // value -> rank map, sparse table (varint value gap, varint frequency
  // except for the last symbol)
  if (supported) {
    uint2* tab = (uint2*)(sbase + kScratchTable);
    uint8_t* sparse = sbase + kScratchSparse;
    const int hdr = 1 + varint_len((uint32_t)distinct);
    if (lane == 0) {
      sparse[0] = (uint8_t)kRansTableLog;
      write_varint_bytes(sparse + 1, (uint32_t)distinct);
    }
    int carry_end = 0, carry_pos = hdr;
    uint32_t carry_p = 0, carry_v = 0;
    for (int t = 0; t < blocks; ++t) {
      const int i = 64 * t + lane;
      const bool valid = i < distinct;
      const uint4 e = valid ? lds.list[i] : make_uint4(0, 0, 0, 0);
      const uint32_t p = carry_p + wave_incl_scan(e.w);
      const int rem = distinct - i - 1;
      const ClampFn g = clamp_scan(valid ? norm_fn(p, count, rem) : clamp_identity());
      const int end = clamp_apply(g, carry_end);
      const uint32_t prev_end = from_prev_lane((uint32_t)end);
      const int cum = lane == 0 ? carry_end : (int)prev_end;
      const uint32_t freq = (uint32_t)(end - cum);
      const uint32_t prev_v = from_prev_lane(e.x);
      const uint32_t gap = e.x - (lane == 0 ? carry_v : prev_v);
      const int len = valid ? varint_len(gap) + (rem ? varint_len(freq) : 0) : 0;
      const int len_incl = wave_incl_scan(len);
      if (valid) {
        tab[i] = make_uint2(freq | ((uint32_t)cum << 13), 0u);
        lds.rank_of[e.x] = (uint8_t)i;
        const int pos = carry_pos + len_incl - len;
        const int n = write_varint_bytes(sparse + pos, gap);
        if (rem) write_varint_bytes(sparse + pos + n, freq);
      }
      carry_end = __builtin_amdgcn_readlane(end, 63);
      carry_p = (uint32_t)__builtin_amdgcn_readlane((int)p, 63);
      carry_v = (uint32_t)__builtin_amdgcn_readlane((int)e.x, 63);
      carry_pos += __builtin_amdgcn_readlane(len_incl, 63);
    }
    r_sparse_len = (uint32_t)carry_pos;
    // rank bytes of rANS lane r (symbols 4k + r): the rank of symbol 4k + r
    // at byte p = 1023 - k, in the group-interleaved chunks of the rank
    // area (the rANS lanes visit symbols from the last one back, so each
    // reads its bytes forwards).  Symbols b = r and b = r + 4 of a chunk
    // are k and k + 1 of lane r: one 16-bit store at byte 1022 - k, which
    // is even, so both bytes fall in one 16-byte chunk.
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 8; ++c) ch[c].launder();
    uint8_t* ranks = rank_area + (int64_t)(s / kRansShards) * kRankGroupBytes +
                     16 * 4 * (s % kRansShards);
    // The pairs go to a staging image of the shard's four lanes in LDS (the
    // histogram's words past the list, dead by now; rows kRankStageRow
    // bytes apart, so the four rows' chunks fall on different banks), which
    // then leaves as whole 16-byte chunks: four stores per lane instead of
    // thirty-two 2-byte ones scattered over 256 partial lines.
    uint8_t* const stage_rk = (uint8_t*)&lds.hist[0][0][0] + kRankStageOffset;
    auto put = [&](int c, int r, uint32_t pair) __attribute__((always_inline)) {
      const int p = 1022 - (128 * c + 2 * lane);  // byte of symbol b = r + 4 (lane r)
      *(uint16_t*)(stage_rk + kRankStageRow * r + p) = (uint16_t)pair;
    };
    if (distinct == (int)maxv + 1) {
      // every value up to the largest is present: a value's rank is the
      // value, and a pair of rank bytes is one byte permute of the symbols
      // (the symbols past a partial shard's end are zeros either way)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) put(c, r, ch[c].low_pair(r));
      }
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t rk0 = lds.rank_of[ch[c].get(r)];
          const uint32_t rk1 = lds.rank_of[ch[c].get(r + 4)];
          put(c, r, rk1 | (rk0 << 8));
        }
      }
    }
    __syncthreads();
    // chunk q of lane r (bytes 16 q .. 16 q + 15) at 1024 q + 16 r; four
    // consecutive lanes write one shard's 64 contiguous bytes of a chunk row
    // (non-temporal: the rank area, a byte per symbol, is read once, by
    // k_rans, which on 16 4K RGB frames then takes 5 % less time)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = 64 * j + lane;
      const int r = i & 3, q = i >> 2;
      const q4u v = *(const q4u*)(stage_rk + kRankStageRow * r + 16 * q);
      __builtin_nontemporal_store(v, (q4u*)(ranks + 1024 * q + 16 * r));
    }
  }

  // ---------------- best of zero-run / Rice / block-pack (encode.cu:226-307)
  if (!force || !supported) {
    // long zero runs (> 64 symbols cost a 2-byte token): runs whose start p
    // has zeros at p..p+64, found from the 64-bit zero word of symbols
    // [64 lane, 64 lane + 64) and the next word
    uint8_t* zb = (uint8_t*)lds.zw;
#pragma unroll
    for (int c = 0; c < 8; ++c) zb[64 * c + lane] = (uint8_t)zmask[c];
    __syncthreads();
    const uint64_t myZ = lds.zw[lane];
    const uint64_t zp = from_prev_lane(myZ);
    const uint64_t zn = from_next_lane(myZ);
    const uint64_t st_bits = myZ & ~((myZ << 1) | (zp >> 63));
    const uint64_t nz = ~myZ, nzn = ~zn;
    const int lo = nz == 0 ? 0 : 64 - __clzll((long long)nz);
    const int hi = nzn == 0 ? 64 : __ffsll((long long)nzn) - 1;
    uint64_t m = 0;
    if (lo < hi) {
      const uint64_t upper = hi == 64 ? ~0ull : ((1ull << hi) - 1);
      m = upper & ~((1ull << lo) - 1);
    }
    const int longs = wave_sum(popc64(st_bits & m));
    // values past 64 and past 8192 take longer tokens: from the list, or
    // symbol by symbol when values fall past the histogram
    int over64 = 0, over8192 = 0;
    if (exact) {
      for (int t = 0; t < blocks; ++t) {
        const int i = 64 * t + lane;
        if (i < distinct) {
          const uint4 e = lds.list[i];
          over64 += e.x > 64 ? (int)e.w : 0;
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const uint32_t x = ch[c].get(b);
          over64 += x > 64;
          over8192 += x > 8192;
        }
      }
      over8192 = wave_sum(over8192);
    }
    over64 = wave_sum(over64);
    const int64_t zero_bytes = (int64_t)(count - zeros) + over64 + over8192 + starts + longs;

    // block b = chunk b / 4, row b % 4 of the wave: OR over the row
    int block_part = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t o = row_or(orv[c]);
      const int blk = 4 * c + (lane >> 4);
      const int n = clampi(count - blk * kBlockSymbols, 0, kBlockSymbols);
      if ((lane & 15) == 0 && n > 0) block_part += 1 + (n * bit_length(o) + 7) / 8;
    }
    const int64_t block_bytes = wave_sum(block_part);

    // Rice: class r (symbol index mod 4) sums of x >> k, from the list's
    // class counts, or symbol by symbol
    const int lane_count = count < 4 ? count : 4;
    const int bl = bit_length(maxv);
    int best_k = 0;
    int64_t best_bytes = INT64_MAX;
    uint32_t best_lane[4] = {0, 0, 0, 0};
    // from k = bl on every x >> k is 0, so a lane's cost n (1 + k) bits only
    // grows with k, and the strict comparison keeps the first of equal
    // costs: no k past bl can be chosen
#pragma unroll 1
    for (int k = 0; k <= min(bl, kMaxRice); ++k) {
      uint32_t cs[4] = {0, 0, 0, 0};
      if (k < bl) {
        if (exact) {
          for (int t = 0; t < blocks; ++t) {
            const int i = 64 * t + lane;
            if (i < distinct) {
              const uint4 e = lds.list[i];
              const uint32_t q = e.x >> k;
              cs[0] += __umul24(e.y & 0xffffu, q);
              cs[1] += __umul24(e.y >> 16, q);
              cs[2] += __umul24(e.z & 0xffffu, q);
              cs[3] += __umul24(e.z >> 16, q);
            }
          }
        } else {
#pragma unroll
          for (int c = 0; c < 8; ++c) {
#pragma unroll
            for (int b = 0; b < 8; ++b) cs[b & 3] += ch[c].get(b) >> k;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) cs[r] = wave_sum(cs[r]);
      }
      int64_t bytes = 0;
      uint32_t lb[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t cls = cs[r];
        const uint32_t n_r = r < lane_count ? (uint32_t)((count - r + 3) / 4) : 0u;
        const uint64_t bits = (uint64_t)cls + (uint64_t)n_r * (uint64_t)(1 + k);
        lb[r] = r < lane_count ? (uint32_t)((bits + 7) / 8) : 0u;
        bytes += lb[r];
      }
      if (bytes < best_bytes) {
        best_bytes = bytes;
        best_k = k;
#pragma unroll
        for (int r = 0; r < 4; ++r) best_lane[r] = lb[r];
      }
    }
    const int64_t rice_bytes = (int64_t)(lane_count - 1) * 4 + best_bytes;
    int mode;
    int64_t body;
    if (block_bytes < zero_bytes && block_bytes < rice_bytes) {
      mode = kModeBlockPack;
      body = block_bytes;
    } else if (zero_bytes <= rice_bytes) {
      mode = kModeZeroRun;
      body = zero_bytes;
    } else {
      mode = kModeRiceBase + best_k;
      body = rice_bytes;
    }
    if (body > 65535 && lane == 0) set_status(st, kStatusBodyTooLarge);
    r_best3_mode = (uint32_t)mode;
    r_mode = (uint32_t)mode;
    r_k = (uint32_t)best_k;
    r_best3_body = (uint32_t)min(body, (int64_t)0xffffffff);
    r_body = r_best3_body;
#pragma unroll
    for (int r = 0; r < 4; ++r) r_lane[r] = min(best_lane[r], 65535u);
  }
  if (lane == 0) {
    ShardRec rec;
    rec.mode = (uint8_t)r_mode;
    rec.k = (uint8_t)r_k;
    rec.supported = (uint8_t)supported;
    rec.best3_mode = (uint8_t)r_best3_mode;
    rec.sparse_len = (uint16_t)r_sparse_len;
    rec.distinct = (uint16_t)(supported ? distinct : 0);
    rec.body = r_body;
    rec.best3_body = r_best3_body;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      rec.lane[r] = (uint16_t)r_lane[r];
      rec.best3_lane[r] = (uint16_t)r_lane[r];
      rec.states[r] = 0;
    }
    rec.offset = 0;
    recs[s] = rec;
  }
}

// Order-0 rANS encode (encode.cu:398-418): one thread per rANS lane, four
// consecutive threads per shard, 16 shards per wave (one wave per block).
// The shard's rank-indexed entries live in LDS as {freq << 19, cum |
// (4096 - freq) << 12, 1/freq rounded up to a double}: the renormalisation
// test is two compares, and floor(x / freq) is trunc(double(x) * 1/freq),
// exact for every state the encoder divides (x < freq * 2^19 <= 2^31).  A
// lane reads its rank bytes forwards, 16 symbols per block from a two-chunk
// window whose next chunk loads while the block runs; the group-interleaved
// rank layout makes each such load one contiguous kilobyte for the wave.  A step's
// renormalisation bytes go into the round's 64-bit word, a round (four
// steps) merges them into a queue of fewer than eight bytes, and when eight
// are queued it stores the oldest eight to the next slot of the lane's
// column in the shard's interleaved lane area, from the top down, which thus
// ends with the lane's bytes in decode order;
// rounds with nothing to store issue a store the hardware drops, so control
// flow and memory waits are the same every round.  Rounds run unpredicated
// while every active lane has 16 symbols left.  The last step decides the
// shard's mode: order-0 when forced or strictly smaller than the best of
// three (encode.cu:423-429).
//
// Table capacity: a wave takes groups of 16 consecutive shards, and one
// instance of the kernel per capacity (16, 64, 256 entries) takes the groups
// whose largest alphabet falls in its range, so the LDS held per wave
// follows the alphabets actually present: 4 KB for small alphabets, which
// lets a compute unit hold every wave its registers allow.  The 16-entry
// instance visits every group and notes each one's tier in a byte; the
// wider instances read those bytes and visit their own groups only.

// Eight bytes starting at byte `pos` (0..15) of the 32-byte window a0:a1:b0:b1,
// the first byte least significant.
__device__ __forceinline__ uint64_t window8_le(uint64_t a0, uint64_t a1, uint64_t b0, int pos) {
  const bool low = pos < 8;
  const uint64_t lo = low ? a0 : a1;
  const uint64_t hi = low ? a1 : b0;
  const uint32_t sh = 8u * (uint32_t)(pos & 7);
  return sh ? (lo >> sh) | (hi << (64u - sh)) : lo;
}

template <typename FT>
__device__ __forceinline__ void rans_group(const FrameDesc* __restrict__ frames,
                                           FrameState* __restrict__ states,
                                           const TileDesc* __restrict__ tiles,
                                           const ShardDesc* __restrict__ shards, int shard_total,
                                           ShardRec* __restrict__ recs,
                                           uint8_t* __restrict__ scratch,
                                           const uint8_t* __restrict__ rank_area, uint4* ent,
                                           int grp, int cap, int cap_below,
                                           uint8_t* __restrict__ tier_of) {
  using q2 = uint64_t __attribute__((ext_vector_type(2)));
  const int lane = lane_id();
  const int r = lane & 3;
  const int s = grp * kRansShards + (lane >> 2);
  bool act = false;
  int count = 0, distinct = 0;
  TileDesc td = {};
  FrameState* st = nullptr;
  if (s < shard_total) {
    const ShardDesc sd = shards[s];
    td = tiles[sd.tile];
    st = &states[td.frame];
    act = st->active && recs[s].supported;
    const int64_t local0 = (int64_t)sd.local * kShardSymbols;
    count = (int)min((int64_t)kShardSymbols, (int64_t)td.width * td.height - local0);
    distinct = recs[s].distinct;
  }
  // the group belongs to the instance whose capacity range holds its largest
  // alphabet
  const int largest = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)(act ? distinct : 0)));
  // the first instance (every group) notes each group's capacity tier for
  // the others, which then visit their own groups only
  if (cap_below < 0 && lane == 0) tier_of[grp] = (uint8_t)(largest > 64 ? 2 : largest > 16 ? 1 : 0);
  if (largest > cap || largest <= cap_below) return;
  uint8_t* sbase = scratch + (int64_t)(act ? s : 0) * kScratchBytes;
  if (act) {
    // LDS entry of a rank: {freq << 19, cum | (4096 - freq) << 16, 1/freq
    // rounded up to a double}; trunc(double(x) * that) == floor(x / freq) for
    // every renormalised state (x < freq * 2^19 <= 2^31), freq 1 included
    const uint2* tab = (const uint2*)(sbase + kScratchTable);
    for (int k = r; k < distinct; k += 4) {
      const uint32_t w = tab[k].x;
      const uint32_t freq = w & 0x1fffu;
      const uint32_t cum = (w >> 13) & 0xfffu;
      double rc = 1.0 / (double)freq;
      if (__fma_rn(rc, (double)freq, -1.0) < 0.0) {
        rc = __longlong_as_double(__double_as_longlong(rc) + 1);
      }
      const uint64_t rb = (uint64_t)__double_as_longlong(rc);
      ent[k] = make_uint4(freq << 19, cum | ((4096u - freq) << 16), (uint32_t)rb,
                          (uint32_t)(rb >> 32));
    }
  }
  __syncthreads();
  int nbytes = 0;
  uint32_t x = kRansL;
  const int nsym = act && count > r ? (count - r + 3) >> 2 : 0;  // symbols of this lane
  // ranks: the lane's bytes start at byte 1024 - nsym of its 1024; aligned
  // chunks ci, ci+1 in a, b (chunk i at cp[64 i]: the group's lanes'
  // chunks interleave); a block of 16 symbols takes the 16 bytes at offset
  // `head` of a:b, while chunk ci+2 loads.  A lane without symbols reads
  // its first chunk and never uses it.
  const int p0 = nsym > 0 ? 1024 - nsym : 0;
  const int head = p0 & 15;
  const q2* cp = (const q2*)(rank_area + (int64_t)grp * kRankGroupBytes) + 64 * (p0 >> 4) + lane;
  const int last = nsym > 0 ? (head + nsym - 1) >> 4 : 0;
  const q2 ca = __builtin_nontemporal_load(cp);  // rank bytes are read once
  const q2 cb = __builtin_nontemporal_load(cp + 64 * min(1, last));
  uint64_t a0 = ca.x, a1 = ca.y, b0 = cb.x, b1 = cb.y;
  int ci = 0;
  // byte queue: pbits / 8 (< 8) bytes, the most recent in the low byte of
  // qq; `stored` bytes (a multiple of 8) already went to scratch.  (The
  // queue's length is kept in bits: as a byte count p, the shift 64 - 8 p
  // was compiled to a quarter-rate multiply by 56 once a round.)
  uint64_t qq = 0;
  uint32_t pbits = 0, stored = 0;
  // byte blocks go through a buffer resource over the group's 16 scratch
  // areas; rounds with nothing to store use an offset past it, which the
  // hardware drops (a shared sink line would take every wave's stores).
  // Word w of the lane goes to slot kLaneSlots - 1 - w of the shard's
  // interleaved lane area (lane r at byte 8 r of each 32-byte slot).
  const uint32_t grp_bytes =
      (uint32_t)(min(shard_total - grp * kRansShards, kRansShards) * kScratchBytes);
  const __amdgpu_buffer_rsrc_t lanes_r = __builtin_amdgcn_make_buffer_rsrc(
      scratch + (int64_t)grp * kRansShards * kScratchBytes, 0, (int)grp_bytes, 0x00020000);
  const uint32_t lane_top_off =
      (uint32_t)((lane >> 2) * kScratchBytes + kScratchLanes + kLaneSlots * 32 + 8 * r);
  // rounds run unpredicated while every active lane of the wave has 16 more
  // symbols; inactive lanes compute nothing of use and store nothing
  // (a group without active lanes runs no rounds at all)
  const int wave_j = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)nsym));
  const int full_j = min(wave_j, __builtin_amdgcn_readfirstlane(
                                     (int)wave_min((uint32_t)(act ? nsym : 0x7fffffff)))) & ~15;
  __builtin_amdgcn_s_waitcnt(0);

  // one block of 16 symbols: four rounds of four steps.  A step emits its
  // renormalisation bytes into the round's word `wd` (most recent lowest);
  // the round then merges them into the queue and stores the oldest eight
  // queued bytes when there are eight.
  auto block = [&](int j0, bool predicated) __attribute__((always_inline)) {
    const uint64_t r_lo = window8_le(a0, a1, b0, head);
    const uint64_t r_hi = window8_le(a1, b0, b1, head);
    const q2 cn = __builtin_nontemporal_load(cp + 64 * min(ci + 2, last));
#pragma unroll
    for (int rd = 0; rd < 4; ++rd) {
      const uint32_t rk4 = (uint32_t)((rd < 2 ? r_lo : r_hi) >> (32 * (rd & 1)));
      uint4 e[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) e[t] = ent[(rk4 >> (8 * t)) & 0xffu];
      uint64_t wd = 0;
      uint32_t wbits = 0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bool live = !predicated || j0 + 4 * rd + t < nsym;
        const uint32_t lim = e[t].x;
        uint32_t sh = (x >> 8) >= lim ? 16u : x >= lim ? 8u : 0u;
        if (!live) sh = 0;
        // the bytes in emission order, low then high, at the top of a dword
        // (x's byte 0 in byte 3, byte 1 in byte 2): the word's new low half
        // is then the high half of {low half, that} << sh, and its new high
        // half the high half of the word << sh
        const uint32_t pv = __builtin_amdgcn_perm(0u, x, 0x00010c0cu);
        const uint32_t lo = (uint32_t)(((((uint64_t)(uint32_t)wd) << 32) | pv) << sh >> 32);
        const uint32_t hi = (uint32_t)((wd << sh) >> 32);
        wd = ((uint64_t)hi << 32) | lo;
        wbits += sh;
        const uint32_t xs = x >> sh;
        const double rc = __longlong_as_double((int64_t)(((uint64_t)e[t].w << 32) | e[t].z));
        const uint32_t q = (uint32_t)((double)xs * rc);
        // (16-bit halves: the add and the multiply read them in place)
        const uint32_t xn = (xs + (e[t].y & 0xffffu)) + __umul24(q, e[t].y >> 16);
        x = live ? xn : x;
      }
      const uint32_t tb = pbits + wbits;  // < 128
      const bool full = tb >= 64u;
      const uint32_t d = (tb - 64u) & 63u;  // the round's bits left after the word
      // the oldest eight bytes: the queue's bytes above the round's
      const uint64_t word = (qq << ((64u - pbits) & 63u)) | (wd >> d);
      const uint64_t keep = wd & ((1ull << d) - 1ull);
      const uint64_t merged = (qq << (wbits & 63u)) | wd;
      qq = full ? keep : merged;
      pbits = full ? d : tb;
      const uint32_t next = stored + 8u;
      const bool fits = next <= (uint32_t)kLaneBytes;
      const uint32_t off = full && fits && act ? lane_top_off - 4u * next : 0x80000000u;
      using u32x2 = uint32_t __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(u32x2{(uint32_t)word, (uint32_t)(word >> 32)}, lanes_r,
                                            off, 0, 0);
      stored = full ? next : stored;
    }
    a0 = b0;
    a1 = b1;
    b0 = cn.x;
    b1 = cn.y;
    ++ci;
  };
  int j0 = 0;
  for (; j0 < full_j; j0 += 16) block(j0, false);
  for (; j0 < wave_j; j0 += 16) block(j0, true);
  if (act) {
    // the last pend (< 8) bytes, most recent lowest in the queue, go to the
    // top of the next slot: they precede the stored words in the lane
    const int pend = (int)(pbits >> 3);
    nbytes = (int)stored + pend;
    if (pend > 0 && (int)stored < kLaneBytes) {
      *(uint64_t*)(sbase + kScratchLanes + 32 * (kLaneSlots - 1 - (int)(stored >> 3)) + 8 * r) =
          qq << ((64u - pbits) & 63u);
    }
    if (nbytes > kLaneBytes) set_status(st, kStatusLaneOverflow);
    recs[s].states[r] = x;
  }
  const int base = lane & ~3;
  const int n0 = shfl(nbytes, base);
  const int n1 = shfl(nbytes, base + 1);
  const int n2 = shfl(nbytes, base + 2);
  const int n3 = shfl(nbytes, base + 3);
  if (act) {
    // the record's head for k_emit_o0: lane lengths 0..2 and final states
    // over spent entries, [mode][body u16] before the sparse table
    const uint32_t body0 =
        (uint32_t)recs[s].sparse_len + 28u + (uint32_t)(n0 + n1 + n2 + n3);
    uint32_t* const head = (uint32_t*)(sbase + kScratchLensStates);
    if (r < 3) head[r] = (uint32_t)nbytes;
    head[3 + r] = x;
    if (r == 0) {
      sbase[kScratchHead] = (uint8_t)kModeOrder0;
      sbase[kScratchHead + 1] = (uint8_t)body0;
      sbase[kScratchHead + 2] = (uint8_t)(body0 >> 8);
      ShardRec& rec = recs[s];
      if (frames[td.frame].force_order0 || body0 < rec.best3_body) {
        rec.mode = kModeOrder0;
        rec.body = body0;
        rec.lane[0] = (uint16_t)n0;
        rec.lane[1] = (uint16_t)n1;
        rec.lane[2] = (uint16_t)n2;
        rec.lane[3] = (uint16_t)n3;
      }
    }
  }
}

template <typename FT, int kCap>
__global__ __launch_bounds__(64) void k_rans(const FrameDesc* __restrict__ frames,
                                             FrameState* __restrict__ states,
                                             const TileDesc* __restrict__ tiles,
                                             const ShardDesc* __restrict__ shards,
                                             int shard_total, const FT* __restrict__ folded,
                                             ShardRec* __restrict__ recs,
                                             uint8_t* __restrict__ scratch,
                                             const uint8_t* __restrict__ rank_area,
                                             int cap_below, uint8_t* __restrict__ tier_of) {
  (void)folded;
  // two entries of padding per row: the shards' rows start eight banks
  // apart, so the four shards a 16-lane pass of a 16-byte LDS read covers
  // take their two commonest ranks from distinct banks
  __shared__ uint4 ent_all[kRansShards][kCap + 2];
  const int groups = (shard_total + kRansShards - 1) / kRansShards;
  if (cap_below < 0) {
    for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
      rans_group<FT>(frames, states, tiles, shards, shard_total, recs, scratch, rank_area,
                     ent_all[lane_id() >> 2], grp, kCap, cap_below, tier_of);
    }
    return;
  }
  // the wider instances: a wave's next 64 groups' tiers in one load, then
  // only the groups of its own tier (none, for small alphabets)
  const uint8_t mine = kCap > 64 ? 2 : 1;
  const int stride = (int)gridDim.x;
  for (int base = blockIdx.x; base < groups; base += 64 * stride) {
    const int g = base + lane_id() * stride;
    uint64_t m = ballot(g < groups && tier_of[g] == mine);
    while (m != 0) {
      const int l = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      rans_group<FT>(frames, states, tiles, shards, shard_total, recs, scratch, rank_area,
                     ent_all[lane_id() >> 2], base + l * stride, kCap, cap_below, tier_of);
    }
  }
}

// ---------------------------------------------------------------- assembly

// Record offsets inside each tile (records are [mode u8][body u16][body]).
// The final pass (all) also lists the shards whose records are not order-0,
// for k_emit (one atomic per wave and 64 shards).  A record is at most
// 3 + 65535 bytes, so 64 of them sum in 32 bits.
__global__ __launch_bounds__(256) void k_tile_scan(FrameState* __restrict__ states,
                                                   const TileDesc* __restrict__ tiles,
                                                   int tile_total, ShardRec* __restrict__ recs,
                                                   int64_t* __restrict__ tile_len, int all,
                                                   int32_t* __restrict__ emit_list,
                                                   int32_t* __restrict__ emit_count) {
  const int lane = lane_id();
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tile_total) return;
  const TileDesc td = tiles[t];
  if (!all && !states[td.frame].active) return;
  int64_t carry = 0;
  for (int c0 = 0; c0 < td.shard_count; c0 += 64) {
    const int i = c0 + lane;
    const bool in = i < td.shard_count;
    const ShardRec* rec = &recs[td.first_shard + (in ? i : 0)];
    const uint32_t rb = in ? rec->body + 3u : 0u;
    const uint32_t incl = wave_incl_scan(rb);
    if (in) recs[td.first_shard + i].offset = (uint32_t)(carry + incl - rb);
    carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (all) {
      const bool other = in && rec->mode != kModeOrder0;
      const uint64_t m = ballot(other);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(emit_count, popc64(m));
        base = shfl(base, leader);
        if (other) emit_list[base + popc64(m & ((1ull << lane) - 1))] = td.first_shard + i;
      }
    }
  }
  if (lane == 0) tile_len[t] = carry;
}

__device__ __forceinline__ void put_words(uint8_t* p, const uint32_t* w, int n) {
  uint32_t* q = (uint32_t*)p;
  for (int i = 0; i < n; ++i) q[i] = w[i];
}

// One workgroup per frame: exclusive scan of tile payload lengths into tile
// offsets, the frame payload (the heuristics' input), and on the final pass
// the 32-byte header and 32-byte directory entries (encode.cu:858-904).
__global__ __launch_bounds__(256) void k_frame_finish(const FrameDesc* __restrict__ frames,
                                                      FrameState* __restrict__ states,
                                                      const TileDesc* __restrict__ tiles,
                                                      const int64_t* __restrict__ tile_len,
                                                      int64_t* __restrict__ tile_off,
                                                      Results* __restrict__ results, int final,
                                                      char* gbase) {
  __shared__ int64_t part[256];
  __shared__ int too_large;
  const FrameDesc& f = frames[blockIdx.x];
  FrameState* st = &states[blockIdx.x];
  if (!final && !st->active) return;
  const int tid = threadIdx.x;
  if (tid == 0) too_large = 0;
  const int T = f.tile_count;
  const int64_t payload_start = kHeaderBytes + (int64_t)T * kDirectoryEntryBytes;
  int64_t running = 0;
  for (int c0 = 0; c0 < T; c0 += 256) {
    const int i = c0 + tid;
    const int64_t len = i < T ? tile_len[f.tile_begin + i] : 0;
    __syncthreads();
    part[tid] = len;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int64_t o = tid >= d ? part[tid - d] : 0;
      __syncthreads();
      part[tid] += o;
      __syncthreads();
    }
    const int64_t off = payload_start + running + part[tid] - len;
    if (i < T) {
      tile_off[f.tile_begin + i] = off;
      if (final) {
        if (len > 0xffffffffll) too_large = 1;
        const TileDesc td = tiles[f.tile_begin + i];
        uint32_t w[8];
        w[0] = (uint32_t)td.plane | ((uint32_t)kEntropyParallelShards << 8) |
               ((uint32_t)kPredictorId << 16);
        w[1] = (uint32_t)td.x;
        w[2] = (uint32_t)td.y;
        w[3] = (uint32_t)td.width;
        w[4] = (uint32_t)td.height;
        w[5] = (uint32_t)(uint64_t)off;
        w[6] = (uint32_t)((uint64_t)off >> 32);
        w[7] = (uint32_t)len;
        if (st->status == 0) {
          put_words(as_global(gbase, f.stream + kHeaderBytes + (int64_t)i * kDirectoryEntryBytes), w, 8);
        }
      }
    }
    running += part[255];
  }
  __syncthreads();
  if (tid == 0) {
    st->payload = running;
    if (final) {
      const int64_t length = payload_start + running;
      if (too_large) set_status(st, kStatusTileTooLarge);
      if (length > f.stream_capacity) set_status(st, kStatusCapacity);
      const int flags = (st->gray_med ? 0x40 : 0) | (st->refined ? 0x80 : 0);
      if (st->status == 0) {
        uint32_t w[8];
        w[0] = 'F' | ('V' << 8) | ('I' << 16) | ((uint32_t)'D' << 24);
        w[1] = 1u | ((uint32_t)f.layout << 8) | ((uint32_t)f.quality << 16) |
               ((uint32_t)((f.bit_depth - 8) | flags) << 24);
        w[2] = (uint32_t)f.width;
        w[3] = (uint32_t)f.height;
        w[4] = (uint32_t)f.tile_w | ((uint32_t)f.tile_h << 16);
        w[5] = f.fps_num;
        w[6] = f.fps_den;
        w[7] = (uint32_t)T;
        put_words(as_global(gbase, f.stream), w, 8);
      }
      st->stream_length = length;
      Results res;
      res.stream_length = length;
      res.status = st->status;
      res.flags = (st->gray_med ? 1 : 0) | (st->refined ? 2 : 0);
      results[blockIdx.x] = res;
    }
  }
}

// ------------------------------------------------------------- emission

// a record, shifted by up to 15 bytes to its 16-byte phase, plus the word
// or_bits may touch past it, in whole 16-byte chunks
constexpr int kBufWords = ((kMaxRecordBytes + 15 + 3) / 4 + 1 + 3) & ~3;

__device__ __forceinline__ void or_bits(uint32_t* buf, uint32_t pos, uint32_t pattern) {
  const uint64_t pp = (uint64_t)pattern << (pos & 31);
  const uint32_t w = pos >> 5;
  if ((uint32_t)pp) atomicOr(&buf[w], (uint32_t)pp);
  if ((uint32_t)(pp >> 32)) atomicOr(&buf[w + 1], (uint32_t)(pp >> 32));
}

// Zero-run, Rice and block-pack records: one wave64 per shard builds the
// record [mode][body u16][body] in LDS and copies it to its offset; the
// reference instead sets bits with global atomics into a zero-filled stream
// (encode.cu:458-474, 906).  The record sits in LDS at its destination's
// offset within a 16-byte chunk, so the copy is one 16-byte store per chunk
// it fills, and byte stores only in the first and last chunks, which it
// shares with the neighbouring records.  Order-0 records are k_emit_o0's;
// k_emit's waves walk the list of the other shards (final k_tile_scan), so
// a batch without such records costs one read of the list's length.
template <typename FT>
__device__ __forceinline__ void emit_record(int s, const FrameDesc* __restrict__ frames,
                                            const FrameState* __restrict__ states,
                                            const TileDesc* __restrict__ tiles,
                                            const ShardDesc* __restrict__ shards,
                                            const FT* __restrict__ folded,
                                            const ShardRec* __restrict__ recs,
                                            const int64_t* __restrict__ tile_off, char* gbase,
                                            uint32_t* buf, uint64_t* zm) {
  const int lane = lane_id();
  const ShardRec rec = recs[s];
  if (rec.mode == kModeOrder0) return;
  const ShardDesc sd = shards[s];
  const TileDesc td = tiles[sd.tile];
  if (states[td.frame].status != 0) return;
  const FrameDesc& f = frames[td.frame];
  const int64_t local0 = (int64_t)sd.local * kShardSymbols;
  const int count = (int)min((int64_t)kShardSymbols, (int64_t)td.width * td.height - local0);
  const FT* v = folded + td.folded_base + local0;
  uint8_t* out = as_global(gbase, f.stream + tile_off[sd.tile] + rec.offset);
  const int total = 3 + (int)rec.body;
  const int sh = (int)((uintptr_t)out & 15);
  const uint32_t hdr = 3u + (uint32_t)sh;  // body's first byte in buf
  uint8_t* bytes = (uint8_t*)buf + sh;
  uint8_t* body = bytes + 3;
  const int words = (sh + total + 3) >> 2;
  for (int i = lane; i < words + 1; i += 64) buf[i] = 0;
  __builtin_amdgcn_wave_barrier();
  const int mode = rec.mode;

  if (mode == kModeZeroRun) {
    // zero-run tokens (encode.cu:514-536): a nonzero v is varint(2v-1), a
    // maximal zero run of R symbols is varint(2(R-1)) at the run's start.
    for (int j = 0; j < 64; ++j) {
      const int i = (j << 6) + lane;
      const uint64_t Z = ballot(i < count && v[i] == 0);
      if (lane == 0) zm[j] = Z;
    }
    if (lane == 0) zm[64] = 0;
    __builtin_amdgcn_wave_barrier();
    int carry = 0;
    uint64_t prevZ = 0;
    for (int j = 0; j < 64 && (j << 6) < count; ++j) {
      const int i = (j << 6) + lane;
      const bool valid = i < count;
      const uint32_t x = valid ? (uint32_t)v[i] : 0u;
      const uint64_t Z = zm[j];
      const uint64_t S = Z & ~((Z << 1) | (prevZ >> 63));
      prevZ = Z;
      uint32_t tok = 0;
      int len = 0;
      if (valid) {
        if (x != 0) {
          tok = 2 * x - 1;
          len = varint_len(tok);
        } else if ((S >> lane) & 1) {
          const uint64_t mask = lane == 63 ? 0ull : (~0ull << (lane + 1));
          uint64_t w = ~Z & mask;
          int end;
          if (w) {
            end = (j << 6) + __ffsll((long long)w) - 1;
          } else {
            int jj = j + 1;
            while (jj < 64 && zm[jj] == ~0ull) ++jj;
            end = jj < 64 ? (jj << 6) + __ffsll((long long)~zm[jj]) - 1 : kShardSymbols;
          }
          end = min(end, count);
          tok = 2u * (uint32_t)(end - i - 1);
          len = varint_len(tok);
        }
      }
      const int incl = wave_incl_scan(len);
      if (len) write_varint_bytes(body + carry + incl - len, tok);
      carry += shfl(incl, 63);
    }
  } else if (mode >= kModeRiceBase && mode <= kModeRiceBase + kMaxRice) {
    // Rice (encode.cu:538-585): lane streams of symbols i%4, each code is q
    // zero bits, a one, then k remainder bits LSB-first.
    const int k = mode - kModeRiceBase;
    const int lane_count = count < 4 ? count : 4;
    uint32_t lane_off[4];
    uint32_t acc = (uint32_t)(lane_count - 1) * 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      lane_off[r] = acc;
      acc += rec.lane[r];
    }
    if (lane < lane_count - 1) {
      const uint32_t len = rec.lane[lane];
      body[lane * 4 + 0] = (uint8_t)len;
      body[lane * 4 + 1] = (uint8_t)(len >> 8);
      body[lane * 4 + 2] = (uint8_t)(len >> 16);
      body[lane * 4 + 3] = (uint8_t)(len >> 24);
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t rmask = k ? ((1u << k) - 1) : 0u;
    const int cls = lane & 3;
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (int j = 0; j < 64 && (j << 6) < count; ++j) {
      const int i = (j << 6) + lane;
      const bool valid = i < count;
      const uint32_t x = valid ? (uint32_t)v[i] : 0u;
      const uint32_t qv = x >> k;
      const uint32_t len = valid ? qv + 1 + (uint32_t)k : 0u;
      uint32_t incl = len;
#pragma unroll
      for (int d = 4; d < 64; d <<= 1) {
        const uint32_t o = shfl_up(incl, d);
        if (lane >= d) incl += o;
      }
      const uint32_t carry = cls == 0 ? c0 : cls == 1 ? c1 : cls == 2 ? c2 : c3;
      if (valid) {
        const uint32_t start = carry + incl - len;
        const uint32_t pos = (hdr + lane_off[cls]) * 8u + start + qv;
        or_bits(buf, pos, 1u | ((x & rmask) << 1));
      }
      c0 += shfl(incl, 60);
      c1 += shfl(incl, 61);
      c2 += shfl(incl, 62);
      c3 += shfl(incl, 63);
    }
  } else if (mode == kModeBlockPack) {
    // block pack (encode.cu:614-682): per 128 symbols a width byte then
    // ceil(n*w/8) bytes, values LSB-first at i*w.
    int myw = 0, mycost = 0;
    for (int b = 0; b < 32; ++b) {
      const int i0 = (2 * b) << 6;
      const uint32_t a = i0 + lane < count ? (uint32_t)v[i0 + lane] : 0u;
      const uint32_t c = i0 + 64 + lane < count ? (uint32_t)v[i0 + 64 + lane] : 0u;
      const uint32_t o = wave_or(a | c);
      const int n = clampi(count - b * kBlockSymbols, 0, kBlockSymbols);
      const int w = bit_length(o);
      if (lane == b) {
        myw = w;
        mycost = n > 0 ? 1 + (n * w + 7) / 8 : 0;
      }
    }
    const int incl = wave_incl_scan(mycost);
    const int off = incl - mycost;
    const int nblocks = (count + kBlockSymbols - 1) / kBlockSymbols;
    if (lane < nblocks) body[off] = (uint8_t)myw;
    __builtin_amdgcn_wave_barrier();
    for (int j = 0; j < 64 && (j << 6) < count; ++j) {
      const int i = (j << 6) + lane;
      const int b = j >> 1;
      const int w = shfl(myw, b);
      const int boff = shfl(off, b);
      if (i < count && w > 0) {
        const uint32_t li = (uint32_t)(((j & 1) << 6) + lane);
        const uint32_t pos = (hdr + (uint32_t)boff + 1u) * 8u + li * (uint32_t)w;
        or_bits(buf, pos, (uint32_t)v[i]);
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    bytes[0] = (uint8_t)mode;
    bytes[1] = (uint8_t)rec.body;
    bytes[2] = (uint8_t)(rec.body >> 8);
  }
  __builtin_amdgcn_wave_barrier();
  uint8_t* const out16 = out - sh;
  const int end = sh + total;
  for (int c = lane; 16 * c < end; c += 64) {
    if (16 * c >= sh && 16 * c + 16 <= end) {
      ((uint4*)out16)[c] = ((const uint4*)buf)[c];
    } else {
      for (int j = max(sh - 16 * c, 0); j < 16 && 16 * c + j < end; ++j) {
        out16[16 * c + j] = ((const uint8_t*)buf)[16 * c + j];
      }
    }
  }
}

template <typename FT>
__global__ __launch_bounds__(256) void k_emit(const FrameDesc* __restrict__ frames,
                                              const FrameState* __restrict__ states,
                                              const TileDesc* __restrict__ tiles,
                                              const ShardDesc* __restrict__ shards,
                                              const int32_t* __restrict__ emit_list,
                                              const int32_t* __restrict__ emit_count,
                                              const FT* __restrict__ folded,
                                              const ShardRec* __restrict__ recs,
                                              const int64_t* __restrict__ tile_off, char* gbase) {
  __shared__ __align__(16) uint32_t buf_all[4][kBufWords];
  __shared__ uint64_t zm_all[4][65];
  const int wave = threadIdx.x >> 6;
  const int n = *emit_count;
  for (int i = blockIdx.x * 4 + wave; i < n; i += 4 * (int)gridDim.x) {
    // the wave's LDS image is rebuilt for each record: keep the previous
    // record's reads ahead of the next one's writes
    asm volatile("" ::: "memory");
    emit_record<FT>(emit_list[i], frames, states, tiles, shards, folded, recs, tile_off, gbase,
                    buf_all[wave], zm_all[wave]);
  }
}

// Order-0 records (encode.cu:587-611), straight from the shard's scratch
// with no LDS staging: the record is [mode][body u16][sparse table] at
// kScratchHead, then the lane lengths 0..2 and final states at
// kScratchLensStates (both written by k_rans), then the four lanes' bytes
// from the interleaved lane area.  kEmitO0Lanes (32) lanes per record, two
// records per wave (16 lanes took 3 % longer on 4K RGB records of about a
// kilobyte, 64 lanes 13 %); lane j of a record writes destination 16-byte
// chunks j, j + kEmitO0Lanes, ....  A chunk takes, from each piece it
// overlaps (usually one), the sixteen bytes at its position: three 8-byte
// words of the piece (consecutive for the head and lengths, 32 bytes apart
// for a lane), byte-aligned and merged under a byte mask; only the chunks at
// the record's two ends, shared with the neighbouring records, store bytes.
// The head and lengths reach at most 15 bytes before and 23 after a piece,
// inside the shard's scratch; a lane's slots are clamped to its area.
__device__ __forceinline__ uint32_t byte_span_mask(int lo, int hi, int d) {
  // bytes [lo, hi) of a 16-byte chunk, restricted to its dword d
  const int l = clampi(lo - 4 * d, 0, 4);
  const int h = clampi(hi - 4 * d, 0, 4);
  const uint32_t below_h = h >= 4 ? ~0u : (1u << (8 * h)) - 1u;
  const uint32_t below_l = l >= 4 ? ~0u : (1u << (8 * l)) - 1u;
  return below_h & ~below_l;
}

constexpr int kEmitO0Lanes = 32;  // lanes per record

template <typename FT>
__global__ __launch_bounds__(256) void k_emit_o0(const FrameDesc* __restrict__ frames,
                                                 const FrameState* __restrict__ states,
                                                 const ShardDesc* __restrict__ shards,
                                                 int shard_total, const ShardRec* __restrict__ recs,
                                                 const uint8_t* __restrict__ scratch,
                                                 const int64_t* __restrict__ tile_off, char* gbase) {
  const int j = (int)threadIdx.x % kEmitO0Lanes;
  const int s = (int)(blockIdx.x * (256 / kEmitO0Lanes) + threadIdx.x / kEmitO0Lanes);
  if (s >= shard_total) return;
  const ShardRec rec = recs[s];
  const ShardDesc sd = shards[s];
  if (rec.mode != kModeOrder0) return;
  if (states[sd.frame].status != 0) return;
  uint8_t* const out = as_global(gbase, frames[sd.frame].stream + tile_off[sd.tile] + rec.offset);
  const uint8_t* const sb = scratch + (int64_t)s * kScratchBytes;
  // piece p covers record bytes [start(p), end(p)); its byte o is at
  // scratch offset off(p) + o
  const int e0 = 3 + rec.sparse_len;
  const int e1 = e0 + 28;
  const int e2 = e1 + rec.lane[0];
  const int e3 = e2 + rec.lane[1];
  const int e4 = e3 + rec.lane[2];
  const int total = e4 + rec.lane[3];  // 3 + rec.body
  auto sel = [](int p, int a0, int a1, int a2, int a3, int a4, int a5) __attribute__((always_inline)) {
    return p == 0 ? a0 : p == 1 ? a1 : p == 2 ? a2 : p == 3 ? a3 : p == 4 ? a4 : a5;
  };
  auto piece = [&](int o) __attribute__((always_inline)) {
    return (o >= e0) + (o >= e1) + (o >= e2) + (o >= e3) + (o >= e4);
  };
  const int sh = (int)((uintptr_t)out & 15);
  uint8_t* const out16 = out - sh;
  const int end = sh + total;
  for (int c = j; 16 * c < end; c += kEmitO0Lanes) {
    const int o0 = 16 * c - sh;  // record byte of the chunk's first byte
    const int tlo = max(-o0, 0);
    const int thi = min(total - o0, 16);
    const int p_last = piece(o0 + thi - 1);
    uint4 v = make_uint4(0, 0, 0, 0);
    for (int p = piece(o0 + tlo); p <= p_last; ++p) {
      const int lo = max(tlo, sel(p, 0, e0, e1, e2, e3, e4) - o0);
      const int hi = min(thi, sel(p, e0, e1, e2, e3, e4, total) - o0);
      if (lo >= hi) continue;  // an empty piece (a lane without bytes)
      // piece byte o0 + t is byte vr + t of the piece's run: scratch byte
      // vr + t of the head and lengths (slots of 8), virtual lane byte of a
      // lane (slots of 32)
      const bool lane_piece = p >= 2;
      const int vr = o0 + sel(p, kScratchHead, kScratchLensStates - e0, kLaneBytes - e2,
                             kLaneBytes - e3, kLaneBytes - e4, kLaneBytes - total);
      const uint8_t* const run = lane_piece ? sb + kScratchLanes + 8 * (p - 2) : sb;
      const int stride = lane_piece ? 32 : 8;
      const int slot = vr >> 3;
      auto word = [&](int t) __attribute__((always_inline)) {
        const int sl = lane_piece ? clampi(slot + t, 0, kLaneSlots - 1) : slot + t;
        return *(const uint2*)(run + stride * sl);
      };
      const uint2 w0 = word(0), w1 = word(1), w2 = word(2);
      const uint32_t k = (uint32_t)(vr & 3);
      const bool up = (vr & 4) != 0;
      const uint32_t d0 = up ? w0.y : w0.x, d1 = up ? w1.x : w0.y, d2 = up ? w1.y : w1.x;
      const uint32_t d3 = up ? w2.x : w1.y, d4 = up ? w2.y : w2.x;
      const uint32_t x0 = __builtin_amdgcn_alignbyte(d1, d0, k);
      const uint32_t x1 = __builtin_amdgcn_alignbyte(d2, d1, k);
      const uint32_t x2 = __builtin_amdgcn_alignbyte(d3, d2, k);
      const uint32_t x3 = __builtin_amdgcn_alignbyte(d4, d3, k);
      const uint32_t m0 = byte_span_mask(lo, hi, 0), m1 = byte_span_mask(lo, hi, 1);
      const uint32_t m2 = byte_span_mask(lo, hi, 2), m3 = byte_span_mask(lo, hi, 3);
      v.x = (x0 & m0) | (v.x & ~m0);
      v.y = (x1 & m1) | (v.y & ~m1);
      v.z = (x2 & m2) | (v.z & ~m2);
      v.w = (x3 & m3) | (v.w & ~m3);
    }
    if (tlo == 0 && thi == 16) {
      ((uint4*)out16)[c] = v;
    } else {
      const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        if (t >= tlo && t < thi) out16[16 * c + t] = (uint8_t)(vw[t >> 2] >> (8 * (t & 3)));
      }
    }
  }
}

// ------------------------------------------------------------ orchestration

template <typename FT>
static void encode_pass(const EncodeBuffers& b, int pass, hipStream_t stream) {
  FT* folded = (FT*)b.folded;
  if (b.any_lossy && b.any_band) {
    launch_band<true, FT>(b.frames, b.states, b.tiles, b.group_first, b.group_total, b.band_ring,
                          folded, b.queues + 2 * pass, b.any_med, b.any_grad, stream);
  }
  if (b.any_lossy && b.any_nonband) {
    {
      FVH_TIMED("k_wavefront_enc", stream);
      hipLaunchKernelGGL((k_wavefront<true, FT>), dim3((b.tile_total + 3) / 4), dim3(256), 0, stream,
                         b.frames, b.states, b.tiles, b.tile_total, folded, (char*)folded);
    }
  }
  if (b.any_lossless) {
    {
      FVH_TIMED("k_lossless", stream);
      hipLaunchKernelGGL((k_lossless<FT>), dim3(b.tile_total), dim3(256), 0, stream, b.frames,
                         b.states, b.tiles, folded, (char*)folded);
    }
  }
  if (b.any_wide) {
    {
      FVH_TIMED("k_generic", stream);
      hipLaunchKernelGGL((k_generic<true, FT>), dim3(b.tile_total), dim3(256), 0, stream, b.frames,
                         b.states, b.tiles, folded, b.recon);
    }
  }
  {
    FVH_TIMED("k_analyze", stream);
    hipLaunchKernelGGL((k_analyze<FT>), dim3(b.shard_total), dim3(64), 0, stream,
                       b.frames, b.states, b.tiles, b.shards, b.shard_total, folded, b.recs,
                       b.scratch, b.ranks);
  }
  {
    FVH_TIMED("k_rans", stream);
    // one persistent instance per table capacity; each takes its groups
    const int groups = (b.shard_total + kRansShards - 1) / kRansShards;
    // each group's capacity tier, one byte a group, in the emit list's
    // bytes (the list is written only by the final tile scan, after every
    // rANS pass)
    uint8_t* const tier_of = (uint8_t*)b.emit_list;
    const dim3 grid(min(groups, 32 * device_cus()));
    hipLaunchKernelGGL((k_rans<FT, 16>), grid, dim3(64), 0, stream, b.frames, b.states, b.tiles,
                       b.shards, b.shard_total, folded, b.recs, b.scratch, b.ranks, -1, tier_of);
    // the wider tables hold fewer waves per compute unit; their instances
    // stride over every group, so a grid of resident waves loses nothing
    static const int o64 = blocks_per_cu((const void*)k_rans<FT, 64>, 0);
    static const int o256 = blocks_per_cu((const void*)k_rans<FT, 256>, 0);
    hipLaunchKernelGGL((k_rans<FT, 64>), dim3(min(groups, o64 * device_cus())), dim3(64), 0,
                       stream, b.frames, b.states, b.tiles, b.shards, b.shard_total, folded,
                       b.recs, b.scratch, b.ranks, 16, tier_of);
    hipLaunchKernelGGL((k_rans<FT, 256>), dim3(min(groups, o256 * device_cus())), dim3(64), 0,
                       stream, b.frames, b.states, b.tiles, b.shards, b.shard_total, folded,
                       b.recs, b.scratch, b.ranks, 64, tier_of);
  }
}

template <typename FT>
static void encode_all(const EncodeBuffers& b, hipStream_t stream) {
  const dim3 fgrid((b.frame_count + 63) / 64);
  const dim3 tgrid((b.tile_total + 3) / 4);
  encode_pass<FT>(b, 0, stream);
  if (b.any_heuristic) {
    {
      FVH_TIMED("k_tile_scan", stream);
      hipLaunchKernelGGL(k_tile_scan, tgrid, dim3(256), 0, stream, b.states, b.tiles, b.tile_total,
                         b.recs, b.tile_len, 0, (int32_t*)nullptr, (int32_t*)nullptr);
    }
    {
      FVH_TIMED("k_frame_finish", stream);
      hipLaunchKernelGGL(k_frame_finish, dim3(b.frame_count), dim3(256), 0, stream, b.frames,
                         b.states, b.tiles, b.tile_len, b.tile_off, b.results, 0, (char*)b.folded);
    }
    {
      FVH_TIMED("k_decide", stream);
      hipLaunchKernelGGL(k_decide, fgrid, dim3(64), 0, stream, b.frames, b.states, b.frame_count, 1);
    }
    encode_pass<FT>(b, 1, stream);
    if (b.any_gray16) {
      {
        FVH_TIMED("k_tile_scan", stream);
        hipLaunchKernelGGL(k_tile_scan, tgrid, dim3(256), 0, stream, b.states, b.tiles,
                           b.tile_total, b.recs, b.tile_len, 0, (int32_t*)nullptr,
                           (int32_t*)nullptr);
      }
      {
        FVH_TIMED("k_frame_finish", stream);
        hipLaunchKernelGGL(k_frame_finish, dim3(b.frame_count), dim3(256), 0, stream, b.frames,
                           b.states, b.tiles, b.tile_len, b.tile_off, b.results, 0, (char*)b.folded);
      }
      {
        FVH_TIMED("k_decide", stream);
        hipLaunchKernelGGL(k_decide, fgrid, dim3(64), 0, stream, b.frames, b.states, b.frame_count,
                           2);
      }
      encode_pass<FT>(b, 2, stream);
    }
    {
      FVH_TIMED("k_decide", stream);
      hipLaunchKernelGGL(k_decide, fgrid, dim3(64), 0, stream, b.frames, b.states, b.frame_count, 3);
    }
  }
  {
    FVH_TIMED("k_tile_scan", stream);
    hipLaunchKernelGGL(k_tile_scan, tgrid, dim3(256), 0, stream, b.states, b.tiles, b.tile_total,
                       b.recs, b.tile_len, 1, b.emit_list, b.queues + kEmitQueue);
  }
  {
    FVH_TIMED("k_frame_finish", stream);
    hipLaunchKernelGGL(k_frame_finish, dim3(b.frame_count), dim3(256), 0, stream, b.frames, b.states,
                       b.tiles, b.tile_len, b.tile_off, b.results, 1, (char*)b.folded);
  }
  {
    FVH_TIMED("k_emit", stream);
    hipLaunchKernelGGL((k_emit<FT>), dim3(min((b.shard_total + 3) / 4, 4 * device_cus())),
                       dim3(256), 0, stream, b.frames, b.states, b.tiles, b.shards, b.emit_list,
                       b.queues + kEmitQueue, (const FT*)b.folded, b.recs, b.tile_off,
                       (char*)b.folded);
  }
  {
    FVH_TIMED("k_emit_o0", stream);
    constexpr int kPerBlock = 256 / kEmitO0Lanes;
    hipLaunchKernelGGL((k_emit_o0<FT>), dim3((b.shard_total + kPerBlock - 1) / kPerBlock), dim3(256), 0, stream,
                       b.frames, b.states, b.shards, b.shard_total, b.recs, b.scratch, b.tile_off,
                       (char*)b.folded);
  }
}

void launch_encode(const EncodeBuffers& b, hipStream_t stream) {
  // the tile geometry, and the pass states and work counters every kernel
  // after it reads
  PlanInit init = {};
  init.kind = 1;
  init.states = b.states;
  init.queues = b.queues;
  init.queue_count = kQueueCount;
  launch_plan(b.frames, b.frame_count, b.tiles, b.shards, b.group_first, b.tile_total, stream,
              init);
  if (b.wide_folded) encode_all<uint32_t>(b, stream);
  else if (b.narrow_folded) encode_all<uint8_t>(b, stream);
  else encode_all<uint16_t>(b, stream);
}

// Decoder reconstruction shares the wavefront kernels.
template <typename FT>
void launch_reconstruct(const DecodeBuffers& b, hipStream_t stream) {
  FT* folded = (FT*)b.folded;
  if (b.any_band) {
    launch_band<false, FT>(b.frames, b.states, b.tiles, b.group_first, b.group_total,
                           b.band_ring, folded, b.queues,
                           b.any_med, b.any_grad, stream);
  }
  if (b.any_nonband) {
    FVH_TIMED("k_wavefront_dec", stream);
    hipLaunchKernelGGL((k_wavefront<false, FT>), dim3((b.tile_total + 3) / 4), dim3(256), 0, stream,
                       b.frames, b.states, b.tiles, b.tile_total, folded, (char*)folded);
  }
  if (b.any_wide) {
    {
      FVH_TIMED("k_generic_dec", stream);
      hipLaunchKernelGGL((k_generic<false, FT>), dim3(b.tile_total), dim3(256), 0, stream, b.frames,
                         b.states, b.tiles, folded, (uint16_t*)nullptr);
    }
  }
}

template void launch_reconstruct<uint16_t>(const DecodeBuffers&, hipStream_t);
template void launch_reconstruct<uint32_t>(const DecodeBuffers&, hipStream_t);

}  // namespace fvh
