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

// Native 4 x u32 vector: plain loads and stores (HIP's uint4 wrapper is
// copied with memcpy, which in the 32-bit case kept chunks in private memory).
using v4u = uint32_t __attribute__((ext_vector_type(4)));

template <typename FT>
struct Vec8;  // eight residuals / samples of one lane

template <>
struct Vec8<uint16_t> {
  v4u v;
  static constexpr int kQuads = 1;  // 16-byte quads per chunk
  __device__ __forceinline__ uint32_t get(int j) const {
    const uint32_t w = j < 2 ? v.x : j < 4 ? v.y : j < 6 ? v.z : v.w;
    return (j & 1) ? (w >> 16) : (w & 0xffffu);
  }
  __device__ __forceinline__ void load(const uint16_t* p) { v = *(const v4u*)p; }
  __device__ __forceinline__ void store(uint16_t* p) const { *(v4u*)p = v; }
  __device__ __forceinline__ void get_q(const v4u* q) { v = q[0]; }
  __device__ __forceinline__ void put_q(v4u* q) const { q[0] = v; }
  __device__ __forceinline__ void zero() { v = v4u{0, 0, 0, 0}; }
  __device__ __forceinline__ void set(int j, uint32_t x) {
    const int k = j >> 1;
    uint32_t w = k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
    if (j & 1) w = (w & 0xffffu) | (x << 16);
    else w = x & 0xffffu;
    if (k == 0) v.x = w;
    else if (k == 1) v.y = w;
    else if (k == 2) v.z = w;
    else v.w = w;
  }
};

template <>
struct Vec8<uint32_t> {
  v4u a, b;
  static constexpr int kQuads = 2;
  __device__ __forceinline__ uint32_t get(int j) const {
    switch (j) {
      case 0: return a.x; case 1: return a.y; case 2: return a.z; case 3: return a.w;
      case 4: return b.x; case 5: return b.y; case 6: return b.z; default: return b.w;
    }
  }
  __device__ __forceinline__ void load(const uint32_t* p) {
    a = *(const v4u*)p;
    b = *(const v4u*)(p + 4);
  }
  __device__ __forceinline__ void store(uint32_t* p) const {
    *(v4u*)p = a;
    *(v4u*)(p + 4) = b;
  }
  __device__ __forceinline__ void get_q(const v4u* q) {
    a = q[0];
    b = q[1];
  }
  __device__ __forceinline__ void put_q(v4u* q) const {
    q[0] = a;
    q[1] = b;
  }
  __device__ __forceinline__ void zero() {
    a = v4u{0, 0, 0, 0};
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
//
// Memory traffic goes through per-row line buffers in LDS.  A "line" is 64
// samples of a tile row (8 chunks of 8 samples, one chunk per group), and a
// lane crosses into a new line every 8 groups; because lane i runs i groups
// behind lane 0, the rows that start (or finish) a line in a given group are
// exactly two, i and i + 8 for i = group mod 8 (start) or (group + 1) mod 8
// (finish).  So in every group the 16 lanes of a tile load one line for each
// of the two starting rows (lane l: row i + 8 (l / 8), chunk l mod 8) and
// store the two finished lines, and each lane exchanges only its own chunk
// with LDS.  Each memory instruction then covers whole 64-sample row
// segments instead of 16 separate rows.  Loads are issued one line period (8
// groups) ahead into eight register slots, one per phase of the period, and
// written to the line buffer in the group the line starts.  The band period
// P is rounded up to whole lines so line starts stay on that schedule.
constexpr int kRowChunks = 9;  // line buffer row: 8 chunks + 1 of padding against bank conflicts

template <typename InT, typename OutT>
struct BandLds {
  v4u lin[4][kBandRows][kRowChunks * Vec8<InT>::kQuads];
  v4u lout[4][kBandRows][kRowChunks * Vec8<OutT>::kQuads];
  uint4 ring[4][kBandRingMax / 8];
};

// One tile group (four tiles, one per 16-lane row of the wave).
template <bool kEncode, bool kMed, typename FT>
__device__ __forceinline__ void band_group(
    const FrameDesc* __restrict__ frames, FrameState* __restrict__ states,
    const TileDesc* __restrict__ tiles, int tile_total, FT* __restrict__ folded, char* gbase,
    uint8_t* __restrict__ sink, int tile_group,
    BandLds<typename std::conditional<kEncode, uint16_t, FT>::type,
            typename std::conditional<kEncode, FT, uint16_t>::type>& lds) {
  using InT = typename std::conditional<kEncode, uint16_t, FT>::type;
  using OutT = typename std::conditional<kEncode, FT, uint16_t>::type;
  constexpr int kQi = Vec8<InT>::kQuads;
  constexpr int kQo = Vec8<OutT>::kQuads;
  auto& lin_all = lds.lin;
  auto& lout_all = lds.lout;
  auto& ring_all = lds.ring;
  const int lane = lane_id();
  const int grp = lane >> 4;
  const int i = lane & 15;
  const int t_idx = tile_group * 4 + grp;
  uint4* ring = ring_all[grp];
  v4u (*lin)[kRowChunks * kQi] = lin_all[grp];
  v4u (*lout)[kRowChunks * kQo] = lout_all[grp];

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
  const int GW = ((W + 63) & ~63) >> 3;          // column groups of a row, whole lines
  // band period in groups: whole lines, and at least the 16 groups after
  // which lane 15 has written the column lane 0 of the next band reads
  const int GP = max(GW, kBandRows);
  const int WG = W >> 3;                          // column groups holding samples
  const int my_groups = has ? (nb - 1) * GP + GW + kBandRows : 0;
  const int total = (int)wave_max((uint32_t)my_groups);
  if (total == 0) return;
  const uint32_t gp_magic = quant_magic(GP);      // floor(t / GP) = mulhi(t, magic)

  const QuantMagic qm = quant_params<FT>(step, bit_depth, max_sample);
  // Loads without a line to fetch read the start of a valid buffer, stores
  // without one go to the lane's sink slot: every group issues the same
  // memory operations on every path.
  const InT* in_base = (const InT*)folded;
  OutT* out_base = (OutT*)sink;
  OutT* const my_sink = (OutT*)(sink + lane * kBandSinkBytes);
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
  // row and first column group of the line that tile row r has reached at
  // local time t (groups since the row group started; lines start at t = 0
  // mod 8): returns false where there is no such line
  auto line_at = [&](int r, int t, int* row, int* cg) __attribute__((always_inline)) {
    if (t < 0) return false;
    const int bb = (int)__umulhi((uint32_t)t, gp_magic);
    *cg = t - bb * GP;
    *row = bb * kBandRows + r;
    return bb < nb && *row < H;
  };
  // cooperative role of this lane: chunk c of a line of row `base + 8 * half`
  const int c = i & 7;
  const int half = i >> 3;
  auto issue_load = [&](Vec8<InT>& slot, int G, int k) __attribute__((always_inline)) {
    // line of row k + 8 half starting at group G + 8
    const int r = k + 8 * half;
    int row = 0, cg = 0;
    const bool ok = line_at(r, G + 8 - r, &row, &cg) && cg + c < WG;
    slot.load(in_base + (ok ? (int64_t)row * in_pitch + 8 * (cg + c) : 0));
  };

  // own lane state: cg = column group (negative before the row group starts),
  // bo = band
  int cg = -i;
  int bo = 0;
  int L = 0, UL = 0;
  int r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0;
  int bad = 0;
  uint4 rv = make_uint4(0, 0, 0, 0);
  auto row_ok = [&](int band) __attribute__((always_inline)) {
    return band < nb && band * kBandRows + i < H;
  };

  Vec8<InT> s0, s1, s2, s3, s4, s5, s6, s7;
  issue_load(s0, -8, 0);
  issue_load(s1, -7, 1);
  issue_load(s2, -6, 2);
  issue_load(s3, -5, 3);
  issue_load(s4, -4, 4);
  issue_load(s5, -3, 5);
  issue_load(s6, -2, 6);
  issue_load(s7, -1, 7);

  auto group = [&](Vec8<InT>& slot, int G, int k) __attribute__((always_inline)) {
    // (1) the line starting now for rows k and k + 8 goes to the line buffer
    slot.put_q(&lin[k + 8 * half][c * kQi]);
    // (2) this lane's chunk
    const bool act = cg >= 0 && cg < WG && row_ok(bo);
    const int x = cg * kBandLag;
    if (cg == 0) {
      L = 0;
      UL = 0;
    }
    Vec8<InT> cur;
    cur.get_q(&lin[i][(cg & 7) * kQi]);
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
    out.put_q(&lout[i][(cg & 7) * kQo]);
    if (act && i == kBandRows - 1 && bo + 1 < nb) {
      ring[x >> 3] = make_uint4((uint32_t)r0 | ((uint32_t)r1 << 16),
                                (uint32_t)r2 | ((uint32_t)r3 << 16),
                                (uint32_t)r4 | ((uint32_t)r5 << 16),
                                (uint32_t)r6 | ((uint32_t)r7 << 16));
    }
    // (3) store the lines finished now, rows k + 1 and k + 9 (mod 16)
    {
      const int r = ((k + 1) & 7) + 8 * half;
      int row = 0, lcg = 0;
      // lcg is the line's last column group here (7 mod 8)
      const bool ok = line_at(r, G - r, &row, &lcg) && lcg - 7 + c < WG;
      Vec8<OutT> done;
      done.get_q(&lout[r][c * kQo]);
      int ok_store = ok;
      asm volatile("" : "+v"(ok_store));
      done.store(ok_store ? out_base + (int64_t)row * out_pitch + 8 * (lcg - 7 + c) : my_sink);
    }
    if (kEncode) asm volatile("" : "+v"(bad));
    // (4) the line starting at G + 8 for rows k and k + 8, into the slot just
    // emptied (after its last use, so the load can target the slot's registers)
    __builtin_amdgcn_sched_barrier(0);
    issue_load(slot, G, k);
    // advance
    ++cg;
    if (cg >= GP) {
      cg = 0;
      ++bo;
    }
    // ring words for the next group (written by lane 15 at least one group earlier)
    rv = make_uint4(0, 0, 0, 0);
    if (i == 0 && cg >= 0 && cg < WG && bo > 0 && row_ok(bo)) rv = ring[cg];
  };

  for (int G = 0; G < total; G += 8) {
    group(s0, G + 0, 0);
    group(s1, G + 1, 1);
    group(s2, G + 2, 2);
    group(s3, G + 3, 3);
    group(s4, G + 4, 4);
    group(s5, G + 5, 5);
    group(s6, G + 6, 6);
    group(s7, G + 7, 7);
  }
  if (kEncode) {
    const uint64_t any = ballot(bad != 0);
    if (i == 0 && st != nullptr && ((any >> (grp * 16)) & 0xffffull) != 0) {
      set_status(st, kStatusSampleRange);
    }
  }
}

// Persistent form: each wave takes tile groups from a device counter until
// none are left, so frames of uneven size in one batch keep every wave busy
// to the end.  The counter is zeroed by the call's state-initialisation
// kernel.
template <bool kEncode, bool kMed, typename FT>
__global__ __launch_bounds__(64) void k_band(const FrameDesc* __restrict__ frames,
                                             FrameState* __restrict__ states,
                                             const TileDesc* __restrict__ tiles, int tile_total,
                                             FT* __restrict__ folded, char* gbase,
                                             uint8_t* __restrict__ sink,
                                             int32_t* __restrict__ queue) {
  using InT = typename std::conditional<kEncode, uint16_t, FT>::type;
  using OutT = typename std::conditional<kEncode, FT, uint16_t>::type;
  __shared__ BandLds<InT, OutT> lds;
  const int groups = (tile_total + 3) / 4;
  for (;;) {
    int g = 0;
    if (lane_id() == 0) g = atomicAdd(queue, 1);
    g = __builtin_amdgcn_readfirstlane(g);
    if (g >= groups) break;
    band_group<kEncode, kMed, FT>(frames, states, tiles, tile_total, folded, gbase, sink, g, lds);
  }
}

template <bool kEncode, typename FT>
void launch_band(const FrameDesc* frames, FrameState* states, const TileDesc* tiles, int tile_total,
                 FT* folded, uint8_t* sink, int32_t* queues, bool med, bool plain,
                 hipStream_t stream) {
  // enough waves to fill the device: eight per compute unit
  const int groups = (tile_total + 3) / 4;
  const dim3 grid(min(groups, 8 * device_cus()));
  if (med) {
    FVH_TIMED(kEncode ? "k_band_enc" : "k_band_dec", stream);
    hipLaunchKernelGGL((k_band<kEncode, true, FT>), grid, dim3(64), 0, stream, frames, states,
                       tiles, tile_total, folded, (char*)folded, sink, queues);
  }
  if (plain) {
    FVH_TIMED(kEncode ? "k_band_enc_grad" : "k_band_dec_grad", stream);
    hipLaunchKernelGGL((k_band<kEncode, false, FT>), grid, dim3(64), 0, stream, frames, states,
                       tiles, tile_total, folded, (char*)folded, sink, queues + 1);
  }
}

template void launch_band<true, uint16_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                          uint16_t*, uint8_t*, int32_t*, bool, bool,
                                          hipStream_t);
template void launch_band<true, uint32_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                          uint32_t*, uint8_t*, int32_t*, bool, bool,
                                          hipStream_t);
template void launch_band<false, uint16_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                           uint16_t*, uint8_t*, int32_t*, bool, bool,
                                           hipStream_t);
template void launch_band<false, uint32_t>(const FrameDesc*, FrameState*, const TileDesc*, int,
                                           uint32_t*, uint8_t*, int32_t*, bool, bool,
                                           hipStream_t);

}  // namespace fvh
