// fvh_band.hip -- throughput wavefront for the closed-loop predictor
// (encoder) and the reconstruction (decoder) on gfx950.
//
// Mapping: a wave64 is four DPP rows of 16 lanes; each row owns one tile, and
// the four tiles of a wave ("tile group") are consecutive tiles of one plane,
// so every plane address is a 32-bit offset from one wave-uniform buffer
// resource.  Inside a tile, rows are taken in bands of 16: lane i of the row
// owns tile row 16*b + i of band b and runs 8 steps behind lane i-1.  Because
// the lag is 8, every lane's column satisfies x == step (mod 8), so each lane
// moves 8 samples per 8 steps with one 16-byte LDS read and write, and inside
// an 8-step group every sample sits at a compile-time register position.  The
// "above" neighbour of lane i is lane i-1's reconstruction from 8 steps
// earlier: one DPP row_shr:1 move per sample, no barrier.  Lane 0 of a row
// takes the previous band's last row from a per-tile LDS ring instead: the
// DPP move keeps its "old" operand on lane 0, and that operand is the ring
// word.  A tile of W x H costs about ceil(H/16) * max(W,128) + 128 steps for
// W*H/16 samples per lane (>= 94% lane utilisation on 256x128 tiles), against
// the reference's W+H-1 barrier-separated antidiagonals with neighbours
// re-read from global memory (cuda/csrc/encode.cu:150-192,
// cuda/csrc/decode.cu:572-605).
//
// Arithmetic per sample (reference: encode.cu:168-190, decode.cu:584-600):
// reconstructions are kept in a biased domain v + beta (beta = -(h + K*d) in
// the encoder, 0 in the decoder).  MED and the gradient are invariant under a
// common shift, so the prediction comes out biased too and the quantiser's
// non-negative numerator is simply n = x - p_beta.  The quotient is exact
// (see quantize) and the reconstruction is med3(beta, max + beta,
// p_beta + q*d).  Residual folding and the decoder's unfolding run two samples
// per instruction in packed 16-bit arithmetic.
//
// Eligible frames (host flag FrameDesc::band): 16-byte aligned planes, every
// tile width, tile origin, pitch and residual base a multiple of 8 samples,
// tile width <= kBandRingMax, planes and pitches small enough for 32-bit
// buffer offsets.  Other frames use k_wavefront.
#include <atomic>
#include <type_traits>

#include "fvh_device.h"
#include "fvh_kernels.h"
#include "fvh_timing.h"

namespace fvh {

constexpr int kBandRows = 16;
constexpr int kRowChunks = 8;              // chunks of 8 samples in one line buffer row
// Line buffer of a tile: rows 0-7, one empty row, rows 8-15 (row r at row
// slot r + r / 8).  Lane i reads and writes chunk (k - i) mod 8 of its own
// row r = i in group k, which puts the 16 lanes of a tile on 16 different
// 16-byte bank slots of a 64-bank LDS (and each half on 8 of a 32-bank one);
// the line loads and stores (lane: row 8 * half + k', chunk c) stay on
// consecutive slots.  A padded row of 9 chunks put every lane of a half on
// the same slot: 8-way conflicts on 2 of the 4 LDS operations per group.
constexpr int kLineRowSlots = kBandRows + 1;
constexpr uint32_t kOob = 0x80000000u;     // buffer offset past every resource: access dropped
constexpr int kRsrcWord3 = 0x00020000;     // raw buffer, 32-bit dwords (gfx9 family)
constexpr int kAuxNt = 2;                  // buffer instruction cache policy: non-temporal

using v4u = uint32_t __attribute__((ext_vector_type(4)));
using v2u = uint32_t __attribute__((ext_vector_type(2)));
using i16x2 = short __attribute__((ext_vector_type(2)));
using u16x2 = unsigned short __attribute__((ext_vector_type(2)));
using rsrc_t = __amdgpu_buffer_rsrc_t;

// row_shr:1 (0x111) within each 16-lane row; lane 0 of a row has no source
// and keeps `old` (bound_ctrl off).
__device__ __forceinline__ int above_dpp(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xf, 0xf, false);
}

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kRsrcWord3);
}

// v_med3_i32: the compiler folds med3(lo, hi, v) into a min/max pair otherwise.
__device__ __forceinline__ int med3_asm(int a, int b, int c) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

template <typename FT>
struct Vec8;  // eight residuals / samples of one lane

template <>
struct Vec8<uint16_t> {
  v4u v;
  static constexpr int kQuads = 1;  // 16-byte quads per chunk
  __device__ __forceinline__ uint32_t word(int k) const {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
  }
  __device__ __forceinline__ uint32_t get(int j) const {
    const uint32_t w = word(j >> 1);
    return (j & 1) ? (w >> 16) : (w & 0xffffu);
  }
  __device__ __forceinline__ uint32_t orall() const { return v.x | v.y | v.z | v.w; }
  // input lines are read once: non-temporal (aux bit 1, "nt")
  __device__ __forceinline__ void load(rsrc_t r, uint32_t off) {
    v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kAuxNt);
  }
  // aux: 0, or kAuxNt for output nothing here reads back
  template <int kAux>
  __device__ __forceinline__ void store(rsrc_t r, uint32_t off) const {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, kAux);
  }
  __device__ __forceinline__ void get_q(const v4u* q) { v = q[0]; }
  __device__ __forceinline__ void put_q(v4u* q) const { q[0] = v; }
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
  __device__ __forceinline__ void set(int j, uint32_t x) {
    switch (j) {
      case 0: a.x = x; break; case 1: a.y = x; break; case 2: a.z = x; break; case 3: a.w = x; break;
      case 4: b.x = x; break; case 5: b.y = x; break; case 6: b.z = x; break; default: b.w = x; break;
    }
  }
  __device__ __forceinline__ void load(rsrc_t r, uint32_t off) {
    a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kAuxNt);
    b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, kAuxNt);
  }
  template <int kAux>
  __device__ __forceinline__ void store(rsrc_t r, uint32_t off) const {
    __builtin_amdgcn_raw_buffer_store_b128(a, r, off, 0, kAux);
    __builtin_amdgcn_raw_buffer_store_b128(b, r, off + 16, 0, kAux);
  }
  __device__ __forceinline__ void get_q(const v4u* q) {
    a = q[0];
    b = q[1];
  }
  __device__ __forceinline__ void put_q(v4u* q) const {
    q[0] = a;
    q[1] = b;
  }
};

// Eight 8-bit residuals (the encoder's output when every residual of the
// batch fits a byte).  In LDS a chunk keeps the whole quad of the 16-bit
// input chunk it replaces, so the line buffers stay in place; in memory it
// is 8 bytes.
template <>
struct Vec8<uint8_t> {
  v2u v;
  static constexpr int kQuads = 1;
  template <int kAux>
  __device__ __forceinline__ void store(rsrc_t r, uint32_t off) const {
    __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, kAux);
  }
  __device__ __forceinline__ void get_q(const v4u* q) { v = *(const v2u*)q; }
  __device__ __forceinline__ void put_q(v4u* q) const { *(v2u*)q = v; }
};

// Quantiser of the reference, q = sign(r) * floor((|r| + h) / d) with
// h = d / 2 (encode.cu:183-185), from the non-negative numerator
//   n = r + h + K*d - [r < 0 and d even],   K = ceil(max / d),
// for which q = floor(n / d) - K (r < 0: floor((r + d - 1 - h) / d) =
// -floor((h - r) / d); for odd d, d - 1 - h == h).  Bit depth <= 12
// (n < 2^(bd+2)): with s = bd + 2 + ceil(log2 d) and m = ceil(2^s / d),
// floor(n / d) = (n * m) >> s, and since n * m < 2^31 and K * 2^s <= 2^31,
// q = (n * m - K * 2^s) >> s as an arithmetic shift of one 24-bit
// multiply-add.  16-bit samples (n < 2^18): floor(n / d) = mul_hi(n, ceil(2^32 / d)).
// Both are checked over every residual of every reachable step in
// tests/test_capi.py.
struct Quant {
  int32_t d;
  int32_t lo, hi;      // reconstruction clamp in the biased domain
  int32_t offp;        // h + K*d
  uint32_t m;
  uint32_t s;
  uint32_t negks;      // -(K << s) mod 2^32 (bit depth <= 12)
  int32_t K;
};

template <typename FT>
__device__ __forceinline__ Quant quant_setup(int step, int max_sample, int bit_depth) {
  Quant q;
  q.d = step;
  const int h = step >> 1;
  q.K = (max_sample + step - 1) / step;
  q.offp = h + q.K * step;
  q.lo = -q.offp;
  q.hi = max_sample - q.offp;
  if (sizeof(FT) <= 2) {
    const uint32_t c = step > 1 ? 32 - __builtin_clz((uint32_t)step - 1) : 0;
    q.s = (uint32_t)bit_depth + 2 + c;
    q.m = (uint32_t)(((1ull << q.s) + (uint64_t)step - 1) / (uint64_t)step);
    q.negks = 0u - ((uint32_t)q.K << q.s);
  } else {
    q.s = 32;
    q.m = quant_magic(step);
    q.negks = 0;
  }
  return q;
}

// v_mad_u32_u24 written out: given mul24(a, b) + (0 - c) the compiler emits a
// multiply and a subtract.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

__device__ __forceinline__ int mad_i24(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// a + b as one instruction the compiler cannot reassociate: the predictor's
// gradient L + (A - UL) keeps A - UL off the reconstruction chain.
__device__ __forceinline__ int add_asm(int a, int b) {
  int r;
  asm("v_add_u32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

template <typename FT, bool kOdd>
__device__ __forceinline__ int quantize(int n, const Quant& q) {
  if (!kOdd) n -= (uint32_t)n < (uint32_t)q.offp ? 1 : 0;
  if (sizeof(FT) <= 2) return (int)mad_u24((uint32_t)n, q.m, q.negks) >> q.s;
  return (int)__umulhi((uint32_t)n, q.m) - q.K;
}

// x * m + c for the low or high 16-bit sample of a word (m <= 2^15 fits
// the 16-bit operand): v_mad_u32_u16 selects the half itself.
__device__ __forceinline__ uint32_t mad_lo16(uint32_t w, uint32_t m, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(w), "v"(m), "v"(c));
  return r;
}
__device__ __forceinline__ uint32_t mad_hi16(uint32_t w, uint32_t m, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(w), "v"(m), "v"(c));
  return r;
}

// quantize(x - p) for sample `hi` of the packed input word w, with the work
// that does not depend on the prediction p done first.  Bit depth <= 12:
// n * m + negks (mod 2^32) for n = x - p equals (x * m + negks) - p * m, so
// after p only one signed 24-bit multiply-add and the shift remain
// (|p| < 2^23, m <= 2^15).  Even d: n < h + K*d exactly when
// p > x - (h + K*d), and n - 1 takes m off the product.
template <typename FT, bool kOdd>
__device__ __forceinline__ int quantize_pred(uint32_t w, bool hi, int p, const Quant& q) {
  const int x = (int)(hi ? w >> 16 : w & 0xffffu);
  if constexpr (sizeof(FT) <= 2) {
    const uint32_t cx = hi ? mad_hi16(w, q.m, q.negks) : mad_lo16(w, q.m, q.negks);
    int base = (int)cx;
    if (!kOdd) base = p > x - q.offp ? (int)(cx - q.m) : (int)cx;
    return mad_i24(p, -(int)q.m, base) >> q.s;
  } else {
    return quantize<FT, kOdd>(x - p, q);
  }
}

template <bool kEncode, typename FT>
struct BandTypes {
  using In = typename std::conditional<kEncode, uint16_t, FT>::type;
  using Out = typename std::conditional<kEncode, FT, uint16_t>::type;
  static constexpr int kQi = Vec8<In>::kQuads;
  static constexpr int kQo = Vec8<Out>::kQuads;
  // equal widths: a lane's output chunk overwrites the input chunk it has
  // just consumed, and one line buffer serves both directions
  static constexpr bool kInPlace = kQi == kQo;
  static constexpr int kLineQuads = 4 * kLineRowSlots * kRowChunks * kQi;
  static constexpr int kOutQuads = kInPlace ? 0 : 4 * kLineRowSlots * kRowChunks * kQo;
};

// Dynamic LDS of one wave: line buffers, then the four rings.
template <bool kEncode, typename FT>
__host__ __device__ constexpr int band_lds_bytes(int ring_groups) {
  return (BandTypes<kEncode, FT>::kLineQuads + BandTypes<kEncode, FT>::kOutQuads) * 16 +
         4 * (2 * ring_groups + 1) * 16;
}

// Memory traffic goes through per-row line buffers in LDS.  A "line" is 64
// samples of a tile row (8 chunks of 8 samples, one chunk per group), and a
// lane crosses into a new line every 8 groups; because lane i runs i groups
// behind lane 0, the rows that start (or finish) a line in group k of an
// 8-group iteration are exactly two, k and k + 8 (start) or k + 1 and k + 9
// mod 16 (finish).  So in every group the 16 lanes of a tile load one line for
// each of the two starting rows (lane l: row k + 8 (l / 8), chunk l mod 8) and
// store the two finished lines, and each lane exchanges only its own chunk
// with LDS.  Loads are issued one iteration ahead into eight register slots,
// one per group of the iteration.  The band period GP (in groups) is rounded
// up to whole lines so line starts stay on that schedule, and all lines
// loaded in one iteration start at the same local time 8 (it + 1 - half),
// all lines stored in groups 0..6 at 8 (it - 1 - half): line positions are
// tracked once per iteration, not per group.
//
// The ring: in every group every lane writes its 8 reconstructions to ring
// slot max(cg, 0) of its tile, and lane 0 reads slot cg + 1 at the end of the
// group.  Lane 0 of band b+1 reads slot x one group before it works on column
// group x, (b+1)*GP + x - 1 groups into the tile; lane 15 of band b wrote it at
// b*GP + 15 + x <= that time (GP >= 16), every other lane of band b earlier,
// and every lane of band b+1 later (lane i at (b+1)*GP + i + x), so the slot
// holds the previous band's last row exactly when it is read.  Slots are
// reset to beta when a tile starts (band 0 reads them); writes clamped to slot
// 0 happen only before lane 15 first writes it.
template <bool kEncode, bool kMed, bool kOdd, typename FT>
__device__ __forceinline__ void band_group(
    const TileDesc* __restrict__ tiles, int t0, int ntiles, rsrc_t in_r, rsrc_t out_r,
    uint32_t in_pitch_u, uint32_t out_pitch_u, int64_t fb0, int ring_groups, const Quant& qz,
    int max_sample, FrameState* st, v4u* smem) {
  using T = BandTypes<kEncode, FT>;
  using InT = typename T::In;
  using OutT = typename T::Out;
  constexpr int kQi = T::kQi;
  constexpr int kQo = T::kQo;
  constexpr uint32_t kCi = 8 * (uint32_t)sizeof(InT);  // bytes of one chunk in memory
  constexpr uint32_t kCo = 8 * (uint32_t)sizeof(OutT);
  // the decoder's samples are the call's output: non-temporal stores keep
  // them from evicting what the next kernels read (decode 16 4K RGB frames:
  // reconstruction -17 us, order-0 decode -27 us); the encoder's residuals
  // are read next by the analysis, which is faster when they stay cached
  constexpr int kStoreAux = kEncode ? 0 : kAuxNt;
  const int lane = lane_id();
  const int grp = lane >> 4;
  const int i = lane & 15;
  const int half = i >> 3;
  const int c = i & 7;

  // ---- this row's tile
  const bool has = grp < ntiles;
  TileDesc td = {};
  if (has) td = tiles[t0 + grp];
  const int W = has ? td.width : 0;
  const int H = has ? td.height : 0;
  const int nb = (H + kBandRows - 1) / kBandRows;
  const int GW = ((W + 63) & ~63) >> 3;   // column groups of a row, whole lines
  // band period in groups: whole lines, and at least the 16 groups after
  // which lane 15 has written the ring slots lane 0 of the next band reads
  const int GP = max(GW, kBandRows);
  const int WG = W >> 3;                  // column groups holding samples
  const int my_groups = has ? (nb - 1) * GP + GW + kBandRows : 0;
  // wave-uniform trip count: a scalar loop, so the memory waits can be exact
  const int total = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)my_groups));
  if (total == 0) return;
  const int iters = (total + 7) >> 3;

  uint32_t in_off, out_off, in_pitch, out_pitch;
  if (kEncode) {
    in_off = (uint32_t)td.src_offset * 2u;
    in_pitch = in_pitch_u;
    out_off = (uint32_t)(td.folded_base - fb0) * (uint32_t)sizeof(FT);
    out_pitch = (uint32_t)W * (uint32_t)sizeof(FT);
  } else {
    in_off = (uint32_t)(td.folded_base - fb0) * (uint32_t)sizeof(FT);
    in_pitch = (uint32_t)W * (uint32_t)sizeof(FT);
    out_off = (uint32_t)td.dst_offset * 2u;
    out_pitch = out_pitch_u;
  }

  // ---- LDS
  v4u* const lin = smem + grp * (kLineRowSlots * kRowChunks * kQi);
  v4u* const lout =
      T::kInPlace ? lin : smem + T::kLineQuads + grp * (kLineRowSlots * kRowChunks * kQo);
  // Ring of a tile: quads [0, ring_groups) hold samples 0-3 of each column
  // group's slot, [ring_groups, 2 ring_groups) samples 4-7, so the lanes'
  // slot writes (lane i: slot cg - i) fall on consecutive quads.  (Rings one
  // quad apart: the four tiles' slots fall on different banks.)
  v4u* const ring = smem + T::kLineQuads + T::kOutQuads + grp * (ring_groups * 2 + 1);
  // row 8 * half + k sits at row slot 9 * half + k
  v4u* const lin_w = lin + ((9 * half) * kRowChunks + c) * kQi;          // + k rows
  const v4u* const lout_r = lout + ((9 * half) * kRowChunks + c) * kQo;  // + ((k+1)&7) rows
  const v4u* const lin_row = lin + (i + half) * kRowChunks * kQi;
  v4u* const lout_row = lout + (i + half) * kRowChunks * kQo;

  const int beta = kEncode ? qz.lo : 0;
  const int lo = kEncode ? qz.lo : 0;
  const int hi = kEncode ? qz.hi : max_sample;

  for (int idx = i; idx < ring_groups * 2; idx += kBandRows) {
    ring[idx] = v4u{(uint32_t)beta, (uint32_t)beta, (uint32_t)beta, (uint32_t)beta};
  }

  // ---- line trackers: local time of the line as band bb and column group cg
  int ld_cg = -8 * half, ld_bb = 0;        // lines loaded in the iteration before `it`
  int st_cg = -8 * (1 + half), st_bb = 0;  // lines finishing in groups 0..6 of `it`
  // lines that do not exist (before the row starts, past the last band or
  // the tile's width) load from an offset the hardware drops: no traffic
  int ld_left = 0;
  auto ld_base = [&]() __attribute__((always_inline)) {
    const bool ok = ld_cg >= 0 && ld_bb < nb && ld_cg + c < WG;
    ld_left = H - 16 * ld_bb - 8 * half;
    return ok ? in_off + __umul24((uint32_t)(16 * ld_bb + 8 * half), in_pitch) +
                    (uint32_t)(ld_cg + c) * kCi
              : kOob;
  };
  auto ld_off = [&](uint32_t b, int k) __attribute__((always_inline)) {
    return k < ld_left ? b + __umul24((uint32_t)k, in_pitch) : kOob;
  };
  uint32_t st_base = kOob;
  int st_left = 0;
  auto st_update = [&]() __attribute__((always_inline)) {
    const bool ok = st_cg >= 0 && st_bb < nb && st_cg + c < WG;
    st_left = H - 16 * st_bb - 8 * half;
    st_base = ok ? out_off + __umul24((uint32_t)(16 * st_bb + 8 * half), out_pitch) +
                       (uint32_t)(st_cg + c) * kCo
                 : kOob;
  };
  st_update();

  // Slot loads stay in slot order, here and in the loop, and every prologue
  // load is preceded by a store the hardware drops (offset past the
  // resource), as every loop load is preceded by a line store.  The memory
  // counter then reads the same on entry as on the back edge, and the wait
  // before a slot's first use is exact (otherwise the loop head waits for
  // every load in flight).
  Vec8<InT> s0, s1, s2, s3, s4, s5, s6, s7;
  {
    const uint32_t b = ld_base();
    Vec8<OutT> none;
    none.get_q(lout_r);
    // (distinct dropped offsets, or they are merged as one store)
    auto prime = [&](Vec8<InT>& slot, uint32_t off, uint32_t k) __attribute__((always_inline)) {
      none.template store<kStoreAux>(out_r, kOob + 64u * k);
      slot.load(in_r, off);
      __builtin_amdgcn_sched_barrier(0);
    };
    prime(s0, ld_off(b, 0), 0);
    prime(s1, ld_off(b, 1), 1);
    prime(s2, ld_off(b, 2), 2);
    prime(s3, ld_off(b, 3), 3);
    prime(s4, ld_off(b, 4), 4);
    prime(s5, ld_off(b, 5), 5);
    prime(s6, ld_off(b, 6), 6);
    prime(s7, ld_off(b, 7), 7);
  }

  // ---- own lane state: cg = column group (negative before the row starts),
  // bo = band
  int cg = -i;
  int bo = 0;
  int L = beta, UL = beta;
  int r0 = beta, r1 = beta, r2 = beta, r3 = beta, r4 = beta, r5 = beta, r6 = beta, r7 = beta;
  v4u rv0 = ring[0], rv1 = ring[ring_groups];  // ring slot of the next group (used by lane 0)
  uint32_t acc = 0;                  // encoder: OR of every active sample

  auto group = [&](Vec8<InT>& slot, uint32_t ldb, int k) __attribute__((always_inline)) {
    // (1) the line starting now for rows k and k + 8 goes to the line buffer
    slot.put_q(lin_w + k * kRowChunks * kQi);
    // (2) this lane's chunk: column group cg, position cg mod 8 of its line
    const int pos = (k - i) & 7;
    Vec8<InT> cur;
    cur.get_q(lin_row + pos * kQi);
    if (cg == 0) {
      L = beta;
      UL = beta;
    }
    const int ringv[8] = {(int)rv0.x, (int)rv0.y, (int)rv0.z, (int)rv0.w,
                          (int)rv1.x, (int)rv1.y, (int)rv1.z, (int)rv1.w};
    int rr[8] = {r0, r1, r2, r3, r4, r5, r6, r7};
    Vec8<OutT> out;
    if constexpr (kEncode) {
      const bool act = (uint32_t)cg < (uint32_t)WG && 16 * bo + i < H;
      acc |= (cur.v.x | cur.v.y | cur.v.z | cur.v.w) & (act ? ~0u : 0u);
      int q[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int A = above_dpp(ringv[j], rr[j]);
        const int g = add_asm(L, A - UL);
        const int p = kMed ? median3(L, A, g) : med3_asm(lo, hi, g);
        q[j] = quantize_pred<FT, kOdd>(cur.word(j >> 1), (j & 1) != 0, p, qz);
        const int rec = med3_asm(lo, hi, p + __mul24(q[j], qz.d));
        rr[j] = rec;
        L = rec;
        UL = A;
      }
      if constexpr (sizeof(FT) <= 2) {
        // zigzag of two residuals per instruction: (q << 1) ^ (q >> 15) on 16-bit halves
        uint32_t w[4];
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
          const i16x2 P = i16x2{(short)q[2 * k2], (short)q[2 * k2 + 1]};
          const i16x2 F = (P << (short)1) ^ (P >> (short)15);
          w[k2] = __builtin_bit_cast(uint32_t, F);
        }
        if constexpr (sizeof(FT) == 2) {
          out.v = v4u{w[0], w[1], w[2], w[3]};
        } else {
          // (every zigzag < 256 here) the halves' low bytes, in order
          out.v = v2u{__builtin_amdgcn_perm(w[1], w[0], 0x06040200u),
                      __builtin_amdgcn_perm(w[3], w[2], 0x06040200u)};
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) out.set(j, ((uint32_t)q[j] << 1) ^ (uint32_t)(q[j] >> 31));
      }
    } else {
      int qd[8];
      if constexpr (sizeof(FT) == 2) {
        // unzigzag and scale two residuals per instruction (|q*d| < 2^15)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
          const u16x2 V = __builtin_bit_cast(u16x2, cur.word(k2));
          const i16x2 Q = (i16x2)(V >> (unsigned short)1) ^ -(i16x2)(V & (unsigned short)1);
          const i16x2 QD = Q * (short)qz.d;
          qd[2 * k2] = QD.x;
          qd[2 * k2 + 1] = QD.y;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t fv = cur.get(j);
          qd[j] = __mul24((int)(fv >> 1) ^ -(int)(fv & 1u), qz.d);
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int A = above_dpp(ringv[j], rr[j]);
        const int g = add_asm(L, A - UL);
        const int p = kMed ? median3(L, A, g) : med3_asm(lo, hi, g);
        const int rec = med3_asm(lo, hi, p + qd[j]);
        rr[j] = rec;
        L = rec;
        UL = A;
      }
      out.v = v4u{(uint32_t)rr[0] | ((uint32_t)rr[1] << 16), (uint32_t)rr[2] | ((uint32_t)rr[3] << 16),
                   (uint32_t)rr[4] | ((uint32_t)rr[5] << 16), (uint32_t)rr[6] | ((uint32_t)rr[7] << 16)};
    }
    r0 = rr[0]; r1 = rr[1]; r2 = rr[2]; r3 = rr[3];
    r4 = rr[4]; r5 = rr[5]; r6 = rr[6]; r7 = rr[7];
    out.put_q(lout_row + pos * kQo);
    // (3) ring slot of this column group
    {
      v4u* rp = ring + max(cg, 0);
      rp[0] = v4u{(uint32_t)r0, (uint32_t)r1, (uint32_t)r2, (uint32_t)r3};
      rp[ring_groups] = v4u{(uint32_t)r4, (uint32_t)r5, (uint32_t)r6, (uint32_t)r7};
    }
    // (4) store the lines finished now, rows k + 1 and k + 9 (mod 16)
    {
      uint32_t off;
      if (k < 7) {
        const int kk = k + 1;
        off = kk < st_left ? st_base + __umul24((uint32_t)kk, out_pitch) : kOob;
      } else {
        st_cg += 8;
        if (st_cg >= GP) {
          st_cg -= GP;
          ++st_bb;
        }
        st_update();
        off = st_left > 0 ? st_base : kOob;
      }
      Vec8<OutT> done;
      done.get_q(lout_r + ((k + 1) & 7) * kRowChunks * kQo);
      done.template store<kStoreAux>(out_r, off);
    }
    // (5) the line starting in the next iteration for rows k and k + 8, into
    // the slot just emptied (pinned in slot order, see above)
    __builtin_amdgcn_sched_barrier(0);
    slot.load(in_r, ld_off(ldb, k));
    __builtin_amdgcn_sched_barrier(0);
    // advance
    ++cg;
    if (cg >= GP) {
      cg = 0;
      ++bo;
    }
    {
      // only lane 0 of a tile uses the slot: every lane reads lane 0's, one
      // address per tile (lane i runs i groups behind lane 0)
      const int cg0 = cg + i >= GP ? cg + i - GP : cg + i;
      const v4u* rq = ring + cg0;
      rv0 = rq[0];
      rv1 = rq[ring_groups];
    }
  };

  // do-while: at least one iteration (total > 0), so no guard the prologue
  // loads could be sunk past
  int it = 0;
  do {
    ld_cg += 8;
    if (ld_cg >= GP) {
      ld_cg -= GP;
      ++ld_bb;
    }
    const uint32_t ldb = ld_base();
    group(s0, ldb, 0);
    group(s1, ldb, 1);
    group(s2, ldb, 2);
    group(s3, ldb, 3);
    group(s4, ldb, 4);
    group(s5, ldb, 5);
    group(s6, ldb, 6);
    group(s7, ldb, 7);
  } while (++it < iters);
  if constexpr (kEncode) {
    const uint32_t mask = ~((uint32_t)max_sample | ((uint32_t)max_sample << 16));
    const uint64_t any = ballot((acc & mask) != 0);
    if (any != 0 && lane == 0) set_status(st, kStatusSampleRange);
  }
}

// Persistent form: each wave takes tile groups from a device counter until
// none are left, so frames of uneven size in one batch keep every wave busy
// to the end.  The counter is zeroed by the call's state-initialisation
// kernel.  A tile group's frame, plane and quantiser are wave-uniform.
template <bool kEncode, bool kMed, typename FT>
__global__ __launch_bounds__(64) void k_band(const FrameDesc* __restrict__ frames,
                                             FrameState* __restrict__ states,
                                             const TileDesc* __restrict__ tiles,
                                             const int32_t* __restrict__ group_first,
                                             int group_total, FT* __restrict__ folded,
                                             int ring_groups, int32_t* __restrict__ queue) {
  extern __shared__ v4u band_smem[];
  for (;;) {
    int g = 0;
    if (lane_id() == 0) g = atomicAdd(queue, 1);
    g = __builtin_amdgcn_readfirstlane(g);
    if (g >= group_total) break;
    const int t0 = group_first[g];
    const int fi = tiles[t0].frame;
    const int p = tiles[t0].plane;
    const FrameDesc& f = frames[fi];
    if (!f.band) continue;
    FrameState* st = states + fi;
    int step;
    if (kEncode) {
      if (!st->active || st->step <= 1 || (st->use_med != 0) != kMed) continue;
      step = st->step;
    } else {
      if (st->status != 0 || (f.use_med != 0) != kMed) continue;
      step = f.step;
    }
    const int ntiles = min(4, f.tile_begin + f.plane_tile_begin[p + 1] - t0);
    const int tl = t0 + ntiles - 1;
    const int64_t fb0 = tiles[t0].folded_base;
    const int64_t fb_end =
        tiles[tl].folded_base + (((int64_t)tiles[tl].width * tiles[tl].height + 7) & ~int64_t{7});
    const uint32_t res_bytes = (uint32_t)((fb_end - fb0) * (int64_t)sizeof(FT));
    const Quant qz = quant_setup<FT>(step, f.max_sample, f.bit_depth);
    rsrc_t in_r, out_r;
    uint32_t in_pitch = 0, out_pitch = 0;
    if (kEncode) {
      const uint32_t plane_bytes =
          (uint32_t)(((f.height - 1) * f.pitch[p] + f.plane_width[p]) * 2);
      in_r = make_rsrc(f.plane[p], plane_bytes);
      out_r = make_rsrc(folded + fb0, res_bytes);
      in_pitch = (uint32_t)f.pitch[p] * 2u;
    } else {
      const uint32_t plane_bytes = (uint32_t)(f.height * f.plane_width[p] * 2);
      in_r = make_rsrc(folded + fb0, res_bytes);
      out_r = make_rsrc(f.out_plane[p], plane_bytes);
      out_pitch = (uint32_t)f.plane_width[p] * 2u;
    }
    if (kEncode && (step & 1) == 0) {
      band_group<kEncode, kMed, false, FT>(tiles, t0, ntiles, in_r, out_r, in_pitch, out_pitch,
                                           fb0, ring_groups, qz, f.max_sample, st, band_smem);
    } else {
      band_group<kEncode, kMed, true, FT>(tiles, t0, ntiles, in_r, out_r, in_pitch, out_pitch,
                                          fb0, ring_groups, qz, f.max_sample, st, band_smem);
    }
  }
}

static std::atomic<int> g_band_waves_per_cu{kBandWavesPerCuDefault};
static std::atomic<int> g_band_decode_waves_per_cu{kBandDecodeWavesPerCuDefault};

int band_waves_per_cu(int set) {
  if (set > 0) g_band_waves_per_cu.store(set);
  return g_band_waves_per_cu.load();
}

int band_decode_waves_per_cu(int set) {
  if (set > 0) g_band_decode_waves_per_cu.store(set);
  return g_band_decode_waves_per_cu.load();
}

template <bool kEncode, typename FT>
void launch_band(const FrameDesc* frames, FrameState* states, const TileDesc* tiles,
                 const int32_t* group_first, int group_total, int ring_groups, FT* folded,
                 int32_t* queues, bool med, bool plain, hipStream_t stream) {
  if (group_total <= 0) return;
  const int lds = band_lds_bytes<kEncode, FT>(ring_groups);
  // persistent waves (a tunable count per compute unit): each takes tile
  // groups of four tiles from the queue until none are left
  const int waves = kEncode ? band_waves_per_cu(0) : band_decode_waves_per_cu(0);
  const dim3 grid(min(group_total, waves * device_cus()));
  if (med) {
    FVH_TIMED(kEncode ? "k_band_enc" : "k_band_dec", stream);
    hipLaunchKernelGGL((k_band<kEncode, true, FT>), grid, dim3(64), lds, stream, frames, states,
                       tiles, group_first, group_total, folded, ring_groups, queues);
  }
  if (plain) {
    FVH_TIMED(kEncode ? "k_band_enc_grad" : "k_band_dec_grad", stream);
    hipLaunchKernelGGL((k_band<kEncode, false, FT>), grid, dim3(64), lds, stream, frames, states,
                       tiles, group_first, group_total, folded, ring_groups, queues + 1);
  }
}

#define FVH_BAND_INST(E, FT)                                                                  \
  template void launch_band<E, FT>(const FrameDesc*, FrameState*, const TileDesc*,            \
                                   const int32_t*, int, int, FT*, int32_t*, bool, bool,        \
                                   hipStream_t);
FVH_BAND_INST(true, uint8_t)
FVH_BAND_INST(true, uint16_t)
FVH_BAND_INST(true, uint32_t)
FVH_BAND_INST(false, uint16_t)
FVH_BAND_INST(false, uint32_t)
#undef FVH_BAND_INST

}  // namespace fvh
