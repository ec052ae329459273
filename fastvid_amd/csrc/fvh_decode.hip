// fvh_decode.hip -- MI355X decoder kernels.
//
//   (k_plan, fvh_plan.hip: tile geometry, frame states and counters)
//   k_dheader      stream headers read on the device (device-fed sizes)
//   k_parse        directory + shard-header validation, one wave per tile,
//                  shards filed by entropy mode (decode.cu:131-220)
//   k_o0_parse     order-0 tables, a thread per shard (k_o0_parse_wave +
//                  k_o0_file: a wave per shard), filed by alphabet tier
//   k_dec_o0       order-0 rANS, 16 shards per wave, one instance per tier
//                  (k_dec_o0_direct: the latency form) (decode.cu:441-549)
//   k_dec_rice     Rice shards, 16 per wave
//   k_dec_other    zero-run and block-pack shards, and Rice shards the
//                  reference rejects inside a long code (decode.cu:222-440)
//   reconstruction: the band wavefront (fvh_band.hip), the shared wavefront
//                  kernels (fvh_encode.hip) or the one-thread-per-tile
//                  diagnostic k_serial (decode.cu:608-656)
//   k_dresults     per-frame results
#include "fvh_device.h"
#include "fvh_kernels.h"
#include "fvh_timing.h"

namespace fvh {

__device__ __forceinline__ uint32_t rd_u16(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8);
}
__device__ __forceinline__ uint32_t rd_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint64_t rd_u64(const uint8_t* p) {
  return (uint64_t)rd_u32(p) | ((uint64_t)rd_u32(p + 4) << 32);
}

// Order-0 alphabet tiers of the decoder (symbols per shard, at most): each
// tier has its own shard list and decoder instance.
constexpr int kO0Tiers = 3;
__host__ __device__ constexpr int o0_tier_cap(int t) { return t == 0 ? 16 : t == 1 ? 64 : 256; }
// decoder shard lists (and their counters): 0 order-0, 1 zero-run and
// block-pack (and Rice shards handed back), 2.. order-0 tiers, then Rice
constexpr int kRiceList = 2 + kO0Tiers;
constexpr int kListCounts = kRiceList + 1;
constexpr int kRiceQueue = 7;

// Device-fed decode: the stream sizes are in device memory and each header is
// read here, one thread per frame.  The header must describe the geometry the
// batch was planned for (dimensions, tiles, layout, bit depth; checks of
// fastvid_cuda.cpp:141-183 on what varies per stream); the quality and the
// gray MED / refined-step flags then give the frame's step and predictor.
__global__ void k_dheader(FrameDesc* __restrict__ frames, FrameState* __restrict__ states,
                          int frame_count, const int64_t* __restrict__ sizes, char* gbase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= frame_count) return;
  FrameDesc& f = frames[i];
  FrameState* st = &states[i];
  const int64_t size = sizes[i];
  if (size < kHeaderBytes) {
    set_status(st, kStatusHeader);
    return;
  }
  const uint8_t* b = as_global(gbase, (const uint8_t*)f.stream);
  uint32_t w[8];
  for (int k = 0; k < 8; ++k) {
    w[k] = (uint32_t)b[4 * k] | ((uint32_t)b[4 * k + 1] << 8) | ((uint32_t)b[4 * k + 2] << 16) |
           ((uint32_t)b[4 * k + 3] << 24);
  }
  const int version = (int)(w[1] & 0xff);
  const int layout = (int)((w[1] >> 8) & 0xff);
  const int quality = (int)((w[1] >> 16) & 0xff);
  const int depth_byte = (int)(w[1] >> 24);
  const int bit_depth = (depth_byte & 0x3f) + 8;
  const bool refined = (depth_byte & 0x80) != 0;
  const bool gray_med = (depth_byte & 0x40) != 0;
  const bool gray = layout == 0;
  bool ok = w[0] == ('F' | ('V' << 8) | ('I' << 16) | ((uint32_t)'D' << 24)) && version == 1 &&
            layout == f.layout && bit_depth == f.bit_depth && quality >= 1 && quality <= 100 &&
            (!refined || (gray && bit_depth == 16 && quality < 100)) &&
            (!gray_med || (gray && bit_depth == 10 && quality < 100)) &&
            (int64_t)w[2] == f.width && (int64_t)w[3] == f.height &&
            (int32_t)(w[4] & 0xffff) == f.tile_w && (int32_t)(w[4] >> 16) == f.tile_h &&
            w[5] != 0 && w[6] != 0 && (int32_t)w[7] == f.tile_count &&
            kHeaderBytes + (int64_t)f.tile_count * kDirectoryEntryBytes <= size;
  if (!ok) {
    set_status(st, kStatusHeader);
    return;
  }
  f.quality = quality;
  f.step = refined ? refined_gray16_step(quality) : quantization_step(layout, bit_depth, quality);
  f.use_med = gray_med || !gray;
  f.header_flags = depth_byte & 0xc0;
  f.stream_capacity = size;
}

// Directory entry and shard-header walk of one tile (decode.cu:139-219).
// One thread per tile; the tile's shards are filed into the per-decoder
// lists with one atomic per wave and list (a wave-wide scan of the lanes'
// counts), not one per thread: twelve thousand tiles' same-address atomics
// serialise at the L2.
__device__ __forceinline__ bool parse_tile(int t, const FrameDesc* __restrict__ frames,
                                           FrameState* __restrict__ states,
                                           const TileDesc* __restrict__ tiles,
                                           DShard* __restrict__ dshards, char* gbase,
                                           int* counts) {
  const TileDesc td = tiles[t];
  const FrameDesc& f = frames[td.frame];
  FrameState* st = &states[td.frame];
  if (st->status != 0) return false;  // e.g. a header rejected by k_dheader
  const uint8_t* enc = as_global(gbase, (const uint8_t*)f.stream);
  const int64_t size = f.stream_capacity;
  const int local = t - f.tile_begin;
  const int64_t T = f.tile_count;
  const int64_t directory_end = kHeaderBytes + T * kDirectoryEntryBytes;
  const uint8_t* entry = enc + kHeaderBytes + (int64_t)local * kDirectoryEntryBytes;
  if (entry[0] != td.plane || entry[1] != kEntropyParallelShards || entry[2] != kPredictorId ||
      entry[3] != 0 || rd_u32(entry + 4) != (uint32_t)td.x || rd_u32(entry + 8) != (uint32_t)td.y ||
      rd_u32(entry + 12) != (uint32_t)td.width || rd_u32(entry + 16) != (uint32_t)td.height) {
    set_status(st, kStatusDirectory);
    return false;
  }
  const uint64_t off_u64 = rd_u64(entry + 20);
  if (off_u64 > (uint64_t)size) {
    set_status(st, kStatusDirectory);
    return false;
  }
  const int64_t payload_offset = (int64_t)off_u64;
  const int64_t payload_length = rd_u32(entry + 28);
  if (payload_length > size - payload_offset) {
    set_status(st, kStatusDirectory);
    return false;
  }
  if (local == 0) {
    if (payload_offset != directory_end) {
      set_status(st, kStatusDirectory);
      return false;
    }
  } else {
    const uint8_t* prev = entry - kDirectoryEntryBytes;
    const uint64_t prev_off = rd_u64(prev + 20);
    const uint64_t prev_len = rd_u32(prev + 28);
    if (prev_off > (uint64_t)size || prev_len > (uint64_t)size - prev_off ||
        off_u64 != prev_off + prev_len) {
      set_status(st, kStatusDirectory);
      return false;
    }
  }
  if (local + 1 == T && payload_offset + payload_length != size) {
    set_status(st, kStatusDirectory);
    return false;
  }
  const int64_t samples = (int64_t)td.width * td.height;
  int64_t cursor = payload_offset;
  int64_t decoded = 0;
  const int64_t tile_end = payload_offset + payload_length;
  for (int ls = 0; ls < td.shard_count; ++ls) {
    if (cursor > tile_end - 3) {
      set_status(st, kStatusShardHeader);
      return false;
    }
    const int mode = enc[cursor];
    const int64_t body_length = rd_u16(enc + cursor + 1);
    const int64_t body_offset = cursor + 3;
    if (mode > 19 || body_length > tile_end - body_offset) {
      set_status(st, kStatusShardHeader);
      return false;
    }
    const int64_t n = samples - decoded < kShardSymbols ? samples - decoded : kShardSymbols;
    DShard d;
    d.body_offset = body_offset;
    d.folded_offset = td.folded_base + decoded;
    d.frame = td.frame;
    d.mode = mode;
    d.body_length = (int32_t)body_length;
    d.count = (int32_t)n;
    dshards[td.first_shard + ls] = d;
    // order-0, other (zero-run, block-pack), Rice
    const bool rice = mode >= 1 && mode <= kModeRiceBase + kMaxRice;
    counts[0] += mode == kModeOrder0;
    counts[2] += rice;
    counts[1] += mode != kModeOrder0 && !rice;
    decoded += n;
    cursor = body_offset + body_length;
  }
  if (decoded != samples || cursor != tile_end) {
    set_status(st, kStatusShardHeader);
    return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void k_parse(const FrameDesc* __restrict__ frames,
                                               FrameState* __restrict__ states,
                                               const TileDesc* __restrict__ tiles, int tile_total,
                                               DShard* __restrict__ dshards,
                                               int32_t* __restrict__ lists,
                                               int32_t* __restrict__ list_count, int shard_total,
                                               char* gbase) {
  // every lane of a wave reaches the scans below (lanes past the last tile,
  // or whose tile failed, file nothing)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int counts[3] = {0, 0, 0};
  bool ok = false;
  if (t < tile_total) {
    ok = parse_tile(t, frames, states, tiles, dshards, gbase, counts);
    if (!ok) counts[0] = counts[1] = counts[2] = 0;
  }
  if (__builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)(t < tile_total))) == 0) return;
  // first list slot of this lane's shards in each list
  auto claim = [&](int n, int32_t* counter) __attribute__((always_inline)) {
    const int incl = (int)wave_incl_scan((uint32_t)n);
    const int total = __builtin_amdgcn_readlane(incl, 63);
    int base = 0;
    if (lane_id() == 0 && total > 0) base = atomicAdd(counter, total);
    return __builtin_amdgcn_readfirstlane(base) + incl - n;
  };
  int at0 = claim(counts[0], &list_count[0]);
  int at1 = claim(counts[1], &list_count[1]);
  int at5 = claim(counts[2], &list_count[kRiceList]);
  if (!ok) return;
  const TileDesc td = tiles[t];
  for (int ls = 0; ls < td.shard_count; ++ls) {
    const int shard = td.first_shard + ls;
    const int mode = dshards[shard].mode;
    if (mode == kModeOrder0) lists[at0++] = shard;
    else if (mode >= 1 && mode <= kModeRiceBase + kMaxRice) lists[kRiceList * shard_total + at5++] = shard;
    else lists[shard_total + at1++] = shard;
  }
}

// Canonical LEB128 with at most five bytes (decode.cu:244-272, 407-426).
__device__ __forceinline__ bool read_varint(const uint8_t* body, int64_t length, int64_t* cursor,
                                            uint32_t* value) {
  uint32_t result = 0;
  int shift = 0;
  for (int used = 0; used < 5; ++used) {
    if (*cursor >= length) return false;
    const uint8_t byte = body[(*cursor)++];
    if (used == 4 && byte > 0x0f) return false;
    result |= (uint32_t)(byte & 0x7f) << shift;
    if ((byte & 0x80) == 0) {
      if (varint_len(result) != used + 1) return false;
      *value = result;
      return true;
    }
    shift += 7;
  }
  return false;
}

// LSB-first bit reader over one Rice lane (the semantics of RiceBitReader,
// decode.cu:63-125): a code is q zeros, a one, then k remainder bits; reading
// past the lane end fails; q and the value must not exceed max_folded.
struct BitReader {
  const uint8_t* bytes;
  int64_t count;
  int64_t next;      // next byte to load
  uint64_t buffer;
  int bits;          // valid bits in buffer
  int64_t consumed;  // bits consumed

  __device__ void refill() {
    while (bits <= 56 && next < count) {
      buffer |= (uint64_t)bytes[next++] << bits;
      bits += 8;
    }
  }

  __device__ bool read(int k, uint32_t max_folded, uint32_t* value) {
    uint32_t q = 0;
    for (;;) {
      if (bits == 0) {
        refill();
        if (bits == 0) return false;
      }
      if (buffer == 0) {
        q += bits;
        consumed += bits;
        bits = 0;
        if (q > max_folded) return false;
        continue;
      }
      const int z = __ffsll((long long)buffer) - 1;
      q += z;
      const int used = z + 1;
      buffer = used == 64 ? 0 : buffer >> used;
      bits -= used;
      consumed += used;
      if (q > max_folded) return false;
      break;
    }
    if (bits < k) refill();
    if (bits < k) return false;
    const uint32_t rem = k ? (uint32_t)buffer & ((1u << k) - 1) : 0u;
    buffer >>= k;
    bits -= k;
    consumed += k;
    const uint64_t result = ((uint64_t)q << k) + rem;
    if (result > max_folded) return false;
    *value = (uint32_t)result;
    return true;
  }
};

// One zero-run, Rice or block-pack body (decode.cu:222-399) decoded by one
// wave; `body` is either the LDS copy or, for bodies too long to stage, the
// stream itself (the function is instantiated for each so LDS reads stay LDS
// reads).
template <typename FT>
__device__ __forceinline__ void dec_other_body(const uint8_t* body, int64_t body_length,
                                               int64_t n, FT* out, int mode, uint32_t max_folded,
                                               FrameState* st, int lane) {
  if (mode == 0) {  // decode.cu:237-295
    if (lane == 0) {
      int64_t cursor = 0, produced = 0;
      int bad = 0;
      while (produced < n && !bad) {
        uint32_t token = 0;
        int shift = 0, used = 0;
        for (;;) {
          if (cursor >= body_length || used == 5) {
            bad = 1;
            break;
          }
          const uint8_t byte = body[cursor++];
          if (used == 4 && byte > 0x0f) {
            bad = 1;
            break;
          }
          token |= (uint32_t)(byte & 0x7f) << shift;
          ++used;
          if ((byte & 0x80) == 0) break;
          shift += 7;
        }
        if (bad) break;
        if (used != varint_len(token)) {
          bad = 1;
          break;
        }
        if ((token & 1) == 0) {
          const int64_t run = token / 2 + 1;
          if (run > n - produced) {
            bad = 1;
            break;
          }
          for (int64_t i = 0; i < run; ++i) out[produced++] = 0;
        } else {
          const uint32_t value = (token + 1) / 2;
          if (value == 0 || value > max_folded) {
            bad = 1;
            break;
          }
          out[produced++] = (FT)value;
        }
      }
      if (bad || cursor != body_length) set_status(st, 1);
    }
  } else if (mode >= 1 && mode <= 17) {  // decode.cu:297-340
    const int lane_count = (int)(n < 4 ? n : 4);
    const int64_t table_bytes = (int64_t)(lane_count - 1) * 4;
    if (lane < lane_count) {
      int ok = table_bytes <= body_length;
      int64_t lane_offset = table_bytes;
      for (int i = 0; ok && i < lane; ++i) lane_offset += rd_u32(body + i * 4);
      const int64_t lane_length = !ok ? 0
          : lane + 1 < lane_count ? (int64_t)rd_u32(body + lane * 4) : body_length - lane_offset;
      if (ok && (lane_offset > body_length || lane_length < 0 ||
                 lane_length > body_length - lane_offset)) {
        ok = 0;
      }
      if (ok) {
        BitReader reader{body + lane_offset, lane_length, 0, 0, 0, 0};
        const int k = mode - 1;
        for (int64_t i = lane; i < n; i += lane_count) {
          uint32_t value = 0;
          if (!reader.read(k, max_folded, &value)) {
            ok = 0;
            break;
          }
          out[i] = (FT)value;
        }
        if (ok) {
          const int64_t used_bytes = (reader.consumed + 7) / 8;
          if (used_bytes != lane_length) ok = 0;
          else if ((reader.consumed & 7) != 0) {
            const uint8_t mask = (uint8_t)~((1u << (reader.consumed & 7)) - 1);
            if ((body[lane_offset + used_bytes - 1] & mask) != 0) ok = 0;
          }
        }
      }
      if (!ok) set_status(st, 2);
    }
  } else if (mode == 18) {  // decode.cu:342-396
    const int64_t block_count = (n + kBlockSymbols - 1) / kBlockSymbols;
    const int64_t block = lane;
    if (block < block_count) {
      int ok = 1;
      int64_t cursor = 0;
      for (int64_t prior = 0; ok && prior < block; ++prior) {
        if (cursor >= body_length) {
          ok = 0;
          break;
        }
        const int w = body[cursor++];
        const int64_t pr = n - prior * kBlockSymbols;
        const int64_t pc = pr < kBlockSymbols ? pr : kBlockSymbols;
        cursor += (pc * w + 7) / 8;
      }
      if (ok && cursor >= body_length) ok = 0;
      const int w = ok ? body[cursor++] : 0;
      if (ok && w > bit_length(max_folded)) ok = 0;
      const int64_t rem = n - block * kBlockSymbols;
      const int64_t cnt = rem < kBlockSymbols ? rem : kBlockSymbols;
      const int64_t payload = (cnt * w + 7) / 8;
      if (ok && payload > body_length - cursor) ok = 0;
      if (ok) {
        const uint8_t* packed = body + cursor;
        int64_t bit = 0;
        for (int64_t i = 0; i < cnt; ++i) {
          uint32_t value = 0;
          for (int bb = 0; bb < w; ++bb, ++bit) {
            value |= (uint32_t)((packed[bit >> 3] >> (bit & 7)) & 1) << bb;
          }
          if (value > max_folded) {
            ok = 0;
            break;
          }
          out[block * kBlockSymbols + i] = (FT)value;
        }
        if (ok && block + 1 == block_count && cursor + payload != body_length) ok = 0;
        if (ok && (bit & 7) != 0) {
          const uint8_t mask = (uint8_t)~((1u << (bit & 7)) - 1);
          if ((packed[payload - 1] & mask) != 0) ok = 0;
        }
      }
      if (!ok) set_status(st, 3);
    }
  } else {
    if (lane == 0) set_status(st, 4);
  }
}

// Zero-run and block-pack bodies staged in LDS (`bl`, length L <= 8960),
// decoded by the whole wave; `stage` is an n-symbol LDS image of the shard.
//
// Zero-run (decode.cu:237-295): a byte with its high bit clear ends a token,
// so every token is found at once; its lane reads back the (at most four)
// bytes before it and checks the varint as the reference does (at most five
// bytes, the fifth <= 0x0f, canonical).  Tokens place their symbols by a
// wave scan of their counts (a run, or one value) in a zero-filled image.
// The reference's sequential walk succeeds exactly when every token is
// valid, the counts sum to n and the body ends on a token's end, which is
// what is checked.
// Block-pack (decode.cu:342-396): lane b finds block b's offset by walking
// the widths before it and checks the block; then every lane extracts
// values from aligned dword pairs.
template <typename FT>
__device__ __forceinline__ void dec_other_wave(const uint8_t* bl, int head, int L, int n,
                                               FT* __restrict__ out, int mode,
                                               uint32_t max_folded, FrameState* st, int lane,
                                               FT* stage, uint2* blocks) {
  constexpr int kPer = 16 / (int)sizeof(FT);  // symbols per 16-byte chunk
  if (mode == 0) {
    const int chunks = (n + kPer - 1) / kPer;
    for (int i = lane; i < chunks * kPer; i += 64) stage[i] = 0;
    __syncthreads();
    bool bad = L == 0 || (bl[L - 1] & 0x80) != 0;  // no tokens, or the last one unfinished
    int carry = 0;  // symbols placed by earlier passes (capped past n)
    for (int base = 0; base < L; base += 64) {
      const int i = base + lane;
      const uint32_t byte = i < L ? bl[i] : 0x80u;
      const bool end = (byte & 0x80u) == 0;
      uint32_t token = 0;
      int count = 0;
      if (end) {
        // the token's first byte: back over at most four continuation bytes
        int j = i;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (j > 0 && (bl[j - 1] & 0x80u)) --j;
        }
        const int used = i - j + 1;
        if (j > 0 && (bl[j - 1] & 0x80u)) bad = true;  // a sixth byte
        for (int q = 0; q < used; ++q) token |= (uint32_t)(bl[j + q] & 0x7fu) << (7 * q);
        if (used == 5 && bl[j + 4] > 0x0fu) bad = true;
        if (used != varint_len(token)) bad = true;
        count = (token & 1u) ? 1 : (int)min(token / 2u + 1u, (uint32_t)n + 1u);
      }
      const int incl = wave_incl_scan(count);
      const int pos = carry + incl - count;
      if (end && (token & 1u)) {
        const uint32_t value = (token + 1u) / 2u;
        if (value > max_folded) bad = true;
        if (pos < n) stage[pos] = (FT)value;
      }
      carry = min(carry + __builtin_amdgcn_readlane(incl, 63), n + 1);
    }
    if (carry != n) bad = true;
    if (ballot(bad) != 0) {
      if (lane == 0) set_status(st, 1);
      return;
    }
    __syncthreads();
    // whole 16-byte chunks of the image (a tile's residual area is padded to
    // a multiple of 8 symbols, so the last chunk stays inside it)
    for (int c = lane; c < chunks; c += 64) ((uint4*)out)[c] = ((const uint4*)stage)[c];
  } else {  // block-pack
    const int block_count = (n + kBlockSymbols - 1) / kBlockSymbols;
    bool bad = false;
    if (lane < block_count) {
      int cursor = 0;
      for (int prior = 0; !bad && prior < lane; ++prior) {
        if (cursor >= L) {
          bad = true;
          break;
        }
        const int w = bl[cursor++];
        const int pr = n - prior * kBlockSymbols;
        cursor += ((pr < kBlockSymbols ? pr : kBlockSymbols) * w + 7) / 8;
      }
      if (!bad && cursor >= L) bad = true;
      const int w = bad ? 0 : bl[cursor++];
      if (w > bit_length(max_folded)) bad = true;
      const int rem = n - lane * kBlockSymbols;
      const int cnt = rem < kBlockSymbols ? rem : kBlockSymbols;
      const int payload = (cnt * w + 7) / 8;
      if (!bad && payload > L - cursor) bad = true;
      if (!bad && lane + 1 == block_count && cursor + payload != L) bad = true;
      if (!bad && ((cnt * w) & 7) != 0) {
        const uint32_t mask = ~((1u << ((cnt * w) & 7)) - 1u) & 0xffu;
        if ((bl[cursor + payload - 1] & mask) != 0) bad = true;
      }
      blocks[lane] = make_uint2((uint32_t)cursor, bad ? 0u : (uint32_t)w);
    }
    if (ballot(bad) != 0) {
      if (lane == 0) set_status(st, 3);
      return;
    }
    __syncthreads();
    const uint32_t* words = (const uint32_t*)(bl - head);  // the staged, aligned chunks
    for (int i = lane; i < n; i += 64) {
      const uint2 blk = blocks[i / kBlockSymbols];
      const uint32_t bit = (uint32_t)(head + (int)blk.x) * 8u + (uint32_t)(i % kBlockSymbols) * blk.y;
      const uint32_t wi = bit >> 5;
      const uint64_t pair = ((uint64_t)words[wi + 1] << 32) | words[wi];
      const uint32_t value = blk.y ? (uint32_t)(pair >> (bit & 31u)) & ((1u << blk.y) - 1u) : 0u;
      if (value > max_folded) bad = true;
      out[i] = (FT)value;
    }
    if (ballot(bad) != 0 && lane == 0) set_status(st, 3);
  }
}

// Zero-run and block-pack shards, and Rice shards k_dec_rice handed back
// (decode.cu:222-399): one wave per shard, shards taken from the list by a
// grid-stride loop.  The body is first copied to LDS; zero-run and
// block-pack bodies are then decoded by the whole wave (dec_other_wave), Rice
// bodies by one lane per Rice lane, and bodies too long to stage (malformed)
// symbol by symbol from global memory.
constexpr int kBodyLds = 65536 / 8 + 768 + 16;  // staged body bytes (+ head)

template <typename FT>
__global__ __launch_bounds__(64) void k_dec_other(const FrameDesc* __restrict__ frames,
                                                  FrameState* __restrict__ states,
                                                  const DShard* __restrict__ dshards,
                                                  const int32_t* __restrict__ lists,
                                                  const int32_t* __restrict__ list_count,
                                                  int shard_total, FT* __restrict__ folded,
                                                  char* gbase) {
  // (16 bytes of slack: block-pack reads the dword after a value's)
  __shared__ __attribute__((aligned(16))) uint8_t body_lds[kBodyLds + 16];
  __shared__ __attribute__((aligned(16))) FT stage[kShardSymbols];
  __shared__ uint2 blocks[kShardSymbols / kBlockSymbols];
  const int lane = lane_id();
  const int total = list_count[1];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = lists[shard_total + item];
    const DShard d = dshards[s];
    FrameState* st = &states[d.frame];
    if (st->status != 0) continue;  // uniform
    const FrameDesc& f = frames[d.frame];
    const uint32_t max_folded = (uint32_t)f.max_sample * 2u;
    const uint8_t* gbody = as_global(gbase, (const uint8_t*)f.stream + d.body_offset);
    const int64_t body_length = d.body_length;
    const int64_t n = d.count;
    FT* out = folded + d.folded_offset;
    const int mode = d.mode;
    // bodies chosen by the encoder are <= 8736 bytes; longer (malformed but
    // parseable) bodies are read from global memory directly.  Staging copies
    // the aligned 16-byte chunks that hold the body: such a chunk never
    // crosses a page, so the bytes before and after the body it covers are
    // readable whenever the body is.
    const int head = (int)((uintptr_t)gbody & 15);
    const bool staged = body_length + head <= (int64_t)kBodyLds;
    if (staged) {
      const uint4* chunks = (const uint4*)(gbody - head);
      const int nchunks = body_length > 0 ? (int)((head + body_length + 15) >> 4) : 0;
      for (int k = lane; k < nchunks; k += 64) ((uint4*)body_lds)[k] = chunks[k];
    }
    __syncthreads();
    if (staged && (mode == 0 || mode == kModeBlockPack)) {
      dec_other_wave(body_lds + head, head, (int)body_length, (int)n, out, mode, max_folded, st,
                     lane, stage, blocks);
    } else if (staged) {
      dec_other_body(body_lds + head, body_length, n, out, mode, max_folded, st, lane);
    } else {
      dec_other_body(gbody, body_length, n, out, mode, max_folded, st, lane);
    }
    __syncthreads();
  }
}

// Forward byte queue over a byte range of global memory, read as the aligned
// 16-byte chunks that hold it (a chunk never crosses a page, so its bytes
// outside the range are readable whenever the range is).  q0:q1 hold the next
// `qn` bytes least significant first; `nxt` is the following chunk, requested
// as soon as the current one is taken, so a byte read waits on memory only
// when a whole chunk was consumed before its successor arrived.  All state is
// shifted, never indexed, so it stays in registers.
struct ByteQueue {
  const uint4* cp;
  uint64_t q0, q1;
  uint4 nxt;
  int qn;    // bytes in q0:q1
  int left;  // bytes of the range not yet in q0:q1
  int c;     // chunk index of q0:q1
  int last;  // index of the last chunk holding bytes of the range

  __device__ __forceinline__ void init(const uint8_t* p, int length) {
    const int head = (int)((uintptr_t)p & 15);
    cp = (const uint4*)(p - head);
    c = 0;
    last = (head + length - 1) >> 4;
    q0 = q1 = 0;
    nxt = make_uint4(0, 0, 0, 0);
    qn = 0;
    left = length;
    if (length > 0) {
      const uint4 first = cp[0];
      nxt = cp[min(1, last)];
      q0 = ((uint64_t)first.y << 32) | first.x;
      q1 = ((uint64_t)first.w << 32) | first.z;
      if (head >= 8) {
        q0 = q1 >> (8 * (head - 8));
        q1 = 0;
      } else if (head > 0) {
        q0 = (q0 >> (8 * head)) | (q1 << (64 - 8 * head));
        q1 >>= 8 * head;
      }
      qn = min(16 - head, length);
      left = length - qn;
    }
  }

  // Next byte of the range; false past its end.
  __device__ __forceinline__ bool take(uint32_t* byte) {
    if (qn == 0) {
      if (left == 0) return false;
      ++c;
      q0 = ((uint64_t)nxt.y << 32) | nxt.x;
      q1 = ((uint64_t)nxt.w << 32) | nxt.z;
      qn = min(16, left);
      left -= qn;
      // the chunk after this one, or this one again at the end of the range;
      // requested only after the old one was copied, so the load can target
      // the same registers instead of a copy that would wait for it
      __builtin_amdgcn_sched_barrier(0);
      nxt = cp[min(c + 1, last)];
    }
    *byte = (uint32_t)q0 & 0xffu;
    q0 = (q0 >> 8) | (q1 << 56);
    q1 >>= 8;
    --qn;
    return true;
  }

  __device__ __forceinline__ bool drained() const { return qn == 0 && left == 0; }
};

// Files the order-0 list positions of a block's threads for the bucket
// decoder instance sized for each alphabet (`nsym` symbols; 0 files
// nothing): counts by wave ballot, summed per block in LDS, one atomic per
// block and tier.  Atomics on one address serialise at the L2, about 8 ns
// each: one per shard took 1.1 ms on 16 4K RGB frames, one per wave and
// tier 55 us on 24 4K RGB 10-bit frames.  Every thread of the block calls it.
template <int kThreads>
__device__ __forceinline__ void file_o0_block(int item, int nsym, int32_t* __restrict__ list_count,
                                              int32_t* __restrict__ tier_lists,
                                              int tier_stride) {
  constexpr int kWaves = kThreads / 64;
  __shared__ int wave_count[kO0Tiers][kWaves];
  __shared__ int block_base[kO0Tiers];
  const int w = (int)threadIdx.x >> 6;
  const int lane = lane_id();
  const int tier = nsym <= 0 ? -1 : nsym <= o0_tier_cap(0) ? 0 : nsym <= o0_tier_cap(1) ? 1 : 2;
  uint64_t mine = 0;  // the lanes of this wave filed in this lane's tier
#pragma unroll
  for (int t = 0; t < kO0Tiers; ++t) {
    const uint64_t m = ballot(tier == t);
    if (tier == t) mine = m;
    if (lane == 0) wave_count[t][w] = popc64(m);
  }
  __syncthreads();
  if (threadIdx.x < kO0Tiers) {
    const int t = (int)threadIdx.x;
    int total = 0;
    for (int i = 0; i < kWaves; ++i) total += wave_count[t][i];
    block_base[t] = total > 0 ? atomicAdd(&list_count[2 + t], total) : 0;
  }
  __syncthreads();
  if (tier >= 0) {
    int at = block_base[tier] + popc64(mine & ((1ull << lane) - 1));
    for (int i = 0; i < w; ++i) at += wave_count[tier][i];
    tier_lists[tier * tier_stride + at] = item;
  }
}

// Order-0 table parse (decode.cu:441-509) of list position `item` by one
// thread: the sparse table (table log, symbol count, value deltas and
// frequencies as canonical LEB128), three lane lengths and four initial
// states, with the reference's checks; a rejected table sets status 5.
// Returns the symbol count, 0 for a rejected table or a failed frame.
__device__ __forceinline__ int parse_o0_thread(int item, const FrameDesc* __restrict__ frames,
                                               FrameState* __restrict__ states,
                                               const DShard* __restrict__ dshards,
                                               const int32_t* __restrict__ lists,
                                               O0Tab* __restrict__ tabs, char* gbase) {
  const DShard d = dshards[lists[item]];
  O0Tab* tab = tabs + item;
  FrameState* st = &states[d.frame];
  if (st->status != 0) {
    tab->nsym = 0;
    return 0;
  }
  const FrameDesc& f = frames[d.frame];
  const uint32_t max_folded = (uint32_t)f.max_sample * 2u;
  const int body_length = d.body_length;
  ByteQueue q;
  q.init(as_global(gbase, (const uint8_t*)f.stream + d.body_offset), body_length);
  int cursor = 0;
  auto varint = [&](uint32_t* value) -> bool {
    uint32_t result = 0;
    for (int used = 0; used < 5; ++used) {
      uint32_t byte;
      if (!q.take(&byte)) return false;
      ++cursor;
      if (used == 4 && byte > 0x0f) return false;
      result |= (byte & 0x7f) << (7 * used);
      if ((byte & 0x80) == 0) {
        if (varint_len(result) != used + 1) return false;
        *value = result;
        return true;
      }
    }
    return false;
  };
  auto u32le = [&](uint32_t* value) -> bool {
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      uint32_t byte;
      if (!q.take(&byte)) return false;
      v |= byte << (8 * k);
    }
    cursor += 4;
    *value = v;
    return true;
  };
  int valid = 1;
  uint32_t log_byte = 0;
  if (!q.take(&log_byte) || log_byte != kRansTableLog) valid = 0;
  cursor = 1;
  uint32_t nsym = 0;
  if (valid && !varint(&nsym)) valid = 0;
  if (nsym == 0 || nsym > 255) valid = 0;
  uint32_t previous = 0, assigned = 0;
  for (uint32_t k = 0; valid && k < nsym; ++k) {
    uint32_t delta = 0;
    if (!varint(&delta) || (k != 0 && delta == 0) || previous + delta > max_folded) {
      valid = 0;
      break;
    }
    const uint32_t value = previous + delta;
    uint32_t fr = (1u << kRansTableLog) - assigned;
    if (k + 1 != nsym && !varint(&fr)) {
      valid = 0;
      break;
    }
    // a u32 wrap of assigned + fr would make the reference fill its table
    // out of bounds; no defined output exists, the stream is rejected
    if ((uint64_t)assigned + fr > (1u << kRansTableLog)) {
      valid = 0;
      break;
    }
    if (fr == 0 || assigned + fr > (1u << kRansTableLog) ||
        (k + 1 != nsym && assigned + fr == (1u << kRansTableLog))) {
      valid = 0;
      break;
    }
    tab->ent[k] = assigned | (value << 13);
    assigned += fr;
    previous = value;
  }
  if (assigned != (1u << kRansTableLog)) valid = 0;
  int32_t len[4] = {0, 0, 0, 0};
  uint32_t total_payload = 0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    uint32_t v = 0;
    if (valid && !u32le(&v)) valid = 0;
    len[l] = (int32_t)v;
    total_payload += v;
  }
  uint32_t state[4] = {0, 0, 0, 0};
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    if (valid && !u32le(&state[l])) valid = 0;
    if (valid && state[l] < kRansL) valid = 0;
  }
  if (valid) {
    len[3] = (int32_t)((uint32_t)(body_length - cursor) - total_payload);
    int64_t at = cursor;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      if (len[l] < 0) valid = 0;  // the reference fails such a lane later
      tab->start[l] = (int32_t)at;
      tab->length[l] = len[l];
      tab->state[l] = state[l];
      at += len[l];
    }
    if (at != body_length) valid = 0;
  }
  if (!valid) {
    set_status(st, 5);
    tab->nsym = 0;
    return 0;
  }
  tab->ent[nsym] = 1u << kRansTableLog;
  tab->nsym = (int32_t)nsym;
  return (int)nsym;
}

// The table parse a thread per order-0 shard, each valid shard filed for the
// decoder instance sized for its alphabet, one atomic per block and tier
// (16 4K RGB frames: 50 us with one per wave and tier, 45 per block of 256,
// 54 per block of 1024, whose fewer blocks leave compute units idle)
constexpr int kParseThreads = 256;

__global__ __launch_bounds__(kParseThreads) void k_o0_parse(const FrameDesc* __restrict__ frames,
                                                            FrameState* __restrict__ states,
                                                            const DShard* __restrict__ dshards,
                                                            const int32_t* __restrict__ lists,
                                                            int32_t* __restrict__ list_count,
                                                            int32_t* __restrict__ tier_lists,
                                                            int tier_stride,
                                                            O0Tab* __restrict__ tabs, char* gbase) {
  const int item = blockIdx.x * kParseThreads + (int)threadIdx.x;
  const int nsym =
      item < list_count[0] ? parse_o0_thread(item, frames, states, dshards, lists, tabs, gbase) : 0;
  file_o0_block<kParseThreads>(item, nsym, list_count, tier_lists, tier_stride);
}

// Latency form of the table parse, one wave per order-0 shard: the
// sparse table (table log, symbol count, value deltas and frequencies as
// canonical LEB128), three lane lengths and four initial states, with the
// reference's checks; a rejected table sets status 5.  The varints are
// found 64 bytes at a time: a byte with its high bit clear ends one, so the
// ballot of such bytes splits the window, and the lane holding a varint's
// last byte assembles its value from the (at most four) bytes before it.
// Symbol values and cumulative frequencies are then wave prefix sums over
// the symbols, 64 at a time, and each symbol's checks run on its own lane.
// The serial parse stops at its first failing check; every failure gives
// status 5, so checking everything and rejecting on any failure decides the
// same.  (The direct-table decoder that follows takes the whole order-0
// list, so the shards are not filed by alphabet size here.)
constexpr int kParseVarints = 512;  // table varints of a shard: 2 * 255 at most

// Parses one shard's table into `tab` (nsym 0 when rejected); returns
// whether it is valid.  Wave-uniform.
__device__ __forceinline__ bool parse_o0_table(const uint8_t* body, int L, uint32_t max_folded,
                                              O0Tab* tab, uint32_t* V, uint16_t* E) {
  const int lane = lane_id();
  bool bad = L < 1 || body[0] != kRansTableLog;
  int base = 1, nvar = 0, need = kParseVarints, nsym = 0;
  uint32_t first_bad = 0xffffffffu;  // lowest index of a malformed varint
  while (!bad && nvar < need) {
    const int p = base + lane;
    const bool in = p < L;
    const uint32_t b = in ? (uint32_t)body[p] : 0x80u;
    const uint64_t T = ballot(in && (b & 0x80u) == 0);
    // no varint ends in the window: one longer than five bytes, or the body
    // ends inside one
    if (T == 0) {
      bad = true;
      break;
    }
    const uint64_t below = T & ((1ull << lane) - 1ull);
    const int start = below ? 64 - __clzll((long long)below) : 0;
    const int seg = lane - start + 1;
    uint32_t v = b & 0x7fu, prev = b;
#pragma unroll
    for (int dd = 1; dd < 5; ++dd) {
      prev = from_prev_lane(prev);
      if (dd < seg) v = (v << 7) | (prev & 0x7fu);
    }
    // (assembled most significant byte first: the varint's last byte is the
    // most significant 7 bits)
    const bool term = (T >> lane) & 1ull;
    const int idx = nvar + popc64(below);
    const bool ok = seg <= 5 && (seg < 5 || b <= 0x0fu) && varint_len(v) == seg;
    if (term && idx < kParseVarints) {
      V[idx] = v;
      E[idx] = (uint16_t)(p + 1);
    }
    first_bad = min(first_bad, wave_min(term && !ok ? (uint32_t)idx : 0xffffffffu));
    nvar += popc64(T);
    base += 64 - __clzll((long long)T);  // past the window's last terminator
    if (need == kParseVarints) {
      __builtin_amdgcn_wave_barrier();
      nsym = (int)V[0];
      if (first_bad == 0 || nsym == 0 || nsym > 255) bad = true;
      need = 2 * nsym;
    }
    if (first_bad < (uint32_t)need) bad = true;
  }
  if (!bad && first_bad < (uint32_t)need) bad = true;
  __builtin_amdgcn_wave_barrier();
  int cursor = 0;
  uint32_t carry_value = 0, carry_assigned = 0;
  if (!bad) {
    cursor = E[need - 1];
    for (int k0 = 0; k0 < nsym; k0 += 64) {
      const int k = k0 + lane;
      const bool valid = k < nsym;
      const uint32_t delta = valid ? V[1 + 2 * k] : 0u;
      const bool has_freq = valid && k + 1 < nsym;
      // frequencies past 4096 fail below either way: clamped, the sums
      // stay in 32 bits
      const uint32_t fr_in = has_freq ? min(V[2 + 2 * k], 4097u) : 0u;
      const uint32_t value = carry_value + wave_incl_scan(delta);
      const uint32_t incl = wave_incl_scan(fr_in);
      const uint32_t assigned = carry_assigned + incl - fr_in;  // before symbol k
      const uint32_t fr = has_freq ? fr_in : (1u << kRansTableLog) - assigned;
      const bool fail = valid && ((k != 0 && delta == 0) || value > max_folded || fr == 0 ||
                                  assigned + fr > (1u << kRansTableLog) ||
                                  (has_freq && assigned + fr == (1u << kRansTableLog)));
      if (valid) tab->ent[k] = assigned | (value << 13);
      if (wave_or(fail ? 1u : 0u)) bad = true;
      carry_value = (uint32_t)__builtin_amdgcn_readlane((int)value, 63);
      carry_assigned += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
  }
  // three lane lengths and four states, little-endian u32s
  if (!bad && cursor + 28 > L) bad = true;
  if (!bad) {
    uint32_t w = 0;
    if (lane < 7) {
      const uint8_t* q = body + cursor + 4 * lane;
      w = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
    }
    const uint32_t w0 = shfl(w, 0), w1 = shfl(w, 1), w2 = shfl(w, 2);
    const int at0 = cursor + 28;
    const int len0 = (int)w0, len1 = (int)w1, len2 = (int)w2;
    const int len3 = (int)((uint32_t)(L - at0) - (w0 + w1 + w2));
    const int64_t at = (int64_t)at0 + len0 + len1 + len2 + len3;
    const bool state_ok = lane < 3 || lane >= 7 || w >= kRansL;
    const uint32_t state = shfl(w, 3 + (lane & 3));  // (every lane takes part in the shuffle)
    if (len0 < 0 || len1 < 0 || len2 < 0 || len3 < 0 || at != L || !wave_min(state_ok ? 1u : 0u)) {
      bad = true;
    } else if (lane < 4) {
      const int len = lane == 0 ? len0 : lane == 1 ? len1 : lane == 2 ? len2 : len3;
      const int start = at0 + (lane > 0 ? len0 : 0) + (lane > 1 ? len1 : 0) + (lane > 2 ? len2 : 0);
      tab->start[lane] = start;
      tab->length[lane] = len;
      tab->state[lane] = state;
    }
  }
  if (bad) {
    if (lane == 0) tab->nsym = 0;
    return false;
  }
  if (lane == 0) {
    tab->ent[nsym] = 1u << kRansTableLog;
    tab->nsym = nsym;
  }
  return true;
}

__global__ __launch_bounds__(64) void k_o0_parse_wave(const FrameDesc* __restrict__ frames,
                                                     FrameState* __restrict__ states,
                                                     const DShard* __restrict__ dshards,
                                                     const int32_t* __restrict__ lists,
                                                     const int32_t* __restrict__ list_count,
                                                     O0Tab* __restrict__ tabs, char* gbase) {
  __shared__ uint32_t value[kParseVarints];
  __shared__ uint16_t end[kParseVarints];
  const int item = (int)blockIdx.x;
  if (item >= list_count[0]) return;
  const DShard d = dshards[lists[item]];
  O0Tab* tab = tabs + item;
  FrameState* st = &states[d.frame];
  if (st->status != 0) {
    if (lane_id() == 0) tab->nsym = 0;
    return;
  }
  const FrameDesc& f = frames[d.frame];
  const uint8_t* body = as_global(gbase, (const uint8_t*)f.stream + d.body_offset);
  if (!parse_o0_table(body, d.body_length, (uint32_t)f.max_sample * 2u, tab, value, end) &&
      lane_id() == 0) {
    set_status(st, 5);
  }
}

// Files the order-0 shards the wave parse accepted (nsym > 0) for the bucket
// decoder instances, a thread per shard (one atomic per block and tier).
constexpr int kFileThreads = 1024;

__global__ __launch_bounds__(kFileThreads) void k_o0_file(int32_t* __restrict__ list_count,
                                                          int32_t* __restrict__ tier_lists,
                                                          int tier_stride,
                                                          const O0Tab* __restrict__ tabs) {
  const int item = blockIdx.x * kFileThreads + (int)threadIdx.x;
  const int nsym = item < list_count[0] ? tabs[item].nsym : 0;
  file_o0_block<kFileThreads>(item, nsym, list_count, tier_lists, tier_stride);
}

// Order-0 rANS lanes (decode.cu:511-549): sixteen shards per wave, one quad
// of lanes per shard, lane r of the quad running rANS lane r.  Waves are
// persistent and take groups of 16 shards from one of three tier lists, by
// alphabet size (at most 16, 64 or 255 symbols; k_o0_parse files each
// shard), so the per-shard tables in LDS are sized by the alphabet and small
// alphabets leave room for more waves on a compute unit.
//
// Slot -> rank: the 4096 slots form buckets of S = 32 - P slots, P the bits
// of the largest rank of the tier (4, 6, 8 for 16, 64, 255 symbols); the
// bucket's word has bit j set where a symbol's slot range starts at slot
// S b + j (the first symbol excluded) and holds in its top P bits the count
// of starts before the bucket, so the rank of slot s is one LDS read and a
// popcount: (w >> S) + popcount(w & ((2 << j) - 1)), b = floor(s / S) by a
// 24-bit multiply.  Rank -> symbol: one 8-byte LDS read of the pair
// {cum_k | value_k << 16, freq_k - 4096} (16-bit samples) or {cum_k |
// (freq_k - 4096 mod 2^13) << 13, value_k} (32-bit).  The state update
// freq*(x >> 12) + (x & 4095) - cum is computed as (freq - 4096)*(x >> 12) +
// (x - cum), one signed 24-bit multiply-add whose low 32 bits equal the
// reference's (exact) u32 result.  Tables: 592 bytes (16 symbols) to 688
// (255) + 8 per symbol, against the reference's 4096-byte slot table
// (decode.cu:519-527).
//
// The symbol loop runs in rounds of four steps with uniform control flow: a
// step renormalises with at most two bytes (a decoded state is >= 2^11), so
// an iteration of two rounds needs at most sixteen, taken big-endian from
// the lane's two current chunks (see `iteration`).  While every lane of the wave has
// four symbols left the rounds run unpredicated; a tail loop finishes lanes of
// uneven length.  Decoded values go to a stage per shard in LDS (O0Block)
// and leave every two rounds as one 16-byte store per lane (16-bit output;
// two for 32-bit).  Errors (decode.cu:536-548): a lane that needs a byte past
// its end consumes more bytes than it has, so the final check "bytes consumed
// == lane length and state == 2^23" covers every failure of the reference.
constexpr int kO0Shards = 16;

// Slot buckets of a tier whose ranks are below kMax <= 255: kPrefixBits for
// the count of symbol starts before a bucket, kSlots slot bits.
template <int kMax>
struct O0Buckets {
  static constexpr int kPrefixBits = kMax <= 16 ? 4 : kMax <= 64 ? 6 : 8;
  static constexpr int kSlots = 32 - kPrefixBits;              // 28, 26, 24
  static constexpr int kCount = (4096 + kSlots - 1) / kSlots;  // 147, 158, 171
  static constexpr int kWords = (kCount + 3) & ~3;             // whole quads
  static constexpr int kPerLane = kWords / 4;                  // prefix pass
  // floor(s / kSlots) == (s * kMagic) >> 16 for every slot s < 4096:
  // kMagic = ceil(2^16 / kSlots) exceeds 2^16 / kSlots by e / kSlots with
  // e = kMagic * kSlots - 2^16, and s * e < 2^16 keeps the error below the
  // gap to the next multiple
  static constexpr uint32_t kMagic = (65536u + kSlots - 1) / kSlots;
  static_assert((kMagic * kSlots - 65536u) * 4095u < 65536u, "bucket division");
  static_assert(kMax - 1 < (1 << kPrefixBits), "prefix count width");
  __device__ __forceinline__ static uint32_t of(uint32_t s) { return __umul24(s, kMagic) >> 16; }
};

// A 16-byte chunk as two little-endian quadwords.  (Chunks are kept as
// scalars: a conditional between two vector-typed objects would be compiled
// as a copy through private memory.)
struct Chunk {
  uint64_t lo, hi;
};

__device__ __forceinline__ Chunk load_chunk(const uint4* p) {
  const uint4 v = *p;
  return Chunk{((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z};
}

// Sixteen bytes starting at byte `pos` (0..15) of the 32-byte window A:B,
// the first byte most significant: bytes pos..pos+7 returned, bytes
// pos+8..pos+15 in `next`, each a funnel shift of two quadwords.
__device__ __forceinline__ uint64_t window16_be(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                                                int pos, uint64_t& next) {
  const bool low = pos < 8;
  const uint64_t q0 = low ? a0 : a1;
  const uint64_t q1 = low ? a1 : b0;
  const uint64_t q2 = low ? b0 : b1;
  const uint32_t sh = 8u * (uint32_t)(pos & 7);
  // (q << 1) << (63 - sh): q << (64 - sh), and nothing for sh = 0
  const uint64_t le0 = (q0 >> sh) | ((q1 << 1) << (63u - sh));
  const uint64_t le1 = (q1 >> sh) | ((q2 << 1) << (63u - sh));
  next = __builtin_bswap64(le1);
  return __builtin_bswap64(le0);
}

// popcount(a) + b in one instruction (the compiler otherwise distributes the
// later scaling of the sum over both terms)
__device__ __forceinline__ uint32_t bcnt_add(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Entry rows: kMax {entry, value} pairs, rounded to a multiple of 16 pairs,
// plus one: the sixteen shards of a wave mostly read their first few ranks,
// and with a row stride of 2 mod 32 dwords their pairs cover the 32 banks
// once.  The stage stride is odd in dwords for the same reason.
template <int kMax>
struct O0Row {
  static constexpr int kPairs = ((kMax + 15) & ~15) + 1;
};

template <typename FT>
struct O0Stage {
  static constexpr int kWords = (32 * (int)sizeof(FT) / 4) | 1;
};

// The order-0 decoders' stage holds a block of each shard's output, flushed
// as four lanes' non-temporal 16-byte stores: whole 128-byte lines
// (`kLines`: 64 symbols of 16 bits, two iterations; 32 of 32 bits, one) or
// one iteration's 32 symbols.  The bucket decoder takes the iteration
// blocks: 64 fewer bytes of LDS per shard for 16-bit output hold it at 12
// waves per compute unit instead of 11, and with non-temporal stores the
// half lines cost no extra traffic.  The direct decoder, bound by one
// wave's chain, keeps whole lines (1080p RGB 10-bit: 147 us, 322 with
// iteration blocks).
template <typename FT, bool kLines>
struct O0Block {
  static constexpr int kPerLane = kLines ? 32 / (int)sizeof(FT) : 8;  // symbols a lane stores
  static constexpr int kSteps = kPerLane;                 // steps of a block (4 lanes)
  static constexpr int kIters = kSteps / 8;               // 8-step iterations
  static constexpr int kWords = kPerLane * (int)sizeof(FT) + 1;  // the block + 1 (odd stride)
};

template <typename FT, int kMax>
struct O0Lds {
  uint32_t bucket[kO0Shards][O0Buckets<kMax>::kWords];
  uint2 row[kO0Shards][O0Row<kMax>::kPairs];
  uint32_t stage[kO0Shards][O0Block<FT, false>::kWords];
};

// Slot -> symbol of the bucket tables (k_dec_o0): slot -> rank by the word
// of the slot's bucket (start bits and the count before it), rank -> the
// {cum | (freq - 4096) << 13, value} pair; returns the decoded state
// freq * (x >> 12) + (x & 4095) - cum before renormalisation, as one signed
// 24-bit multiply-add (its low 32 bits equal the reference's u32 result).
template <typename FT, int kMax>
struct BucketLookup {
  const uint32_t* bucket;
  const uint2* ent;
  __device__ __forceinline__ uint32_t update(uint32_t x, uint32_t& value) const {
    using B = O0Buckets<kMax>;
    const uint32_t s = x & 4095u;
    const uint32_t b = B::of(s);
    const uint32_t w = bucket[b];
    const uint32_t j = s - (uint32_t)B::kSlots * b;
    const uint32_t rank = bcnt_add(w & ((2u << j) - 1u), w >> B::kSlots);
    const uint2 ev = ent[rank];
    if constexpr (sizeof(FT) == 2) {
      // {cum | value << 16, freq - 4096}: the subtraction reads cum in
      // place, and the stage takes the value's half as it is
      value = ev.x >> 16;
      return (uint32_t)__mul24((int)ev.y, (int)(x >> kRansTableLog)) + (x - (ev.x & 0xffffu));
    } else {
      value = ev.y;
      const uint32_t c = ev.x & 0x1fffu;
      const int fneg = ((int)(ev.x << 6)) >> 19;  // freq - 4096, sign-extended
      return (uint32_t)__mul24(fneg, (int)(x >> kRansTableLog)) + (x - c);
    }
  }
};

// The rANS lanes of one shard, one quad of lanes per shard, lane r running
// rANS lane r (decode.cu:530-549), with the slot lookup `look`.  Decoded
// values pass through the shard's `stage` row and leave as whole 64-byte
// (16-bit) or 128-byte (32-bit) blocks; lanes without a shard (item < 0 or a
// rejected table) decode nothing of use and store to their sink slot.
template <typename FT, bool kLines, typename Lookup>
__device__ __forceinline__ void o0_lanes(const Lookup& look, int item, int nsym,
                                         const O0Tab* tab, int r, int lane,
                                         const FrameDesc* __restrict__ frames,
                                         const DShard* __restrict__ dshards,
                                         const int32_t* __restrict__ lists,
                                         FrameState* __restrict__ states,
                                         FT* __restrict__ folded, FT* const stage, char* gbase,
                                         uint8_t* __restrict__ sink,
                                         bool fair = false) {
  DShard d = {};
  if (item >= 0) d = dshards[lists[item]];
  const FrameDesc& f = frames[d.frame];
  const uint8_t* body = as_global(gbase, (const uint8_t*)f.stream + d.body_offset);
  const int length = nsym > 0 ? tab->length[r] : 0;
  uint32_t x = nsym > 0 ? tab->state[r] : kRansL;
  // aligned chunks holding the lane's bytes (a lane without bytes reads the
  // body's first chunk, and never uses it)
  const uint8_t* first = length > 0 ? body + tab->start[r] : body;
  const int head = (int)((uintptr_t)first & 15);
  const uint4* cp = (const uint4*)(first - head);
  const int last = length > 0 ? (head + length - 1) >> 4 : 0;
  const Chunk A0 = load_chunk(cp);
  const Chunk B0 = load_chunk(cp + min(1, last));
  const Chunk C0 = load_chunk(cp + min(2, last));
  uint64_t a0 = A0.lo, a1 = A0.hi, b0 = B0.lo, b1 = B0.hi;
  uint64_t c0 = C0.lo, c1 = C0.hi;  // chunk ci + 2 (its load may be in flight)
  bool pend = false;   // the window moved: chunk ci + 2 is to be loaded into C
  int ci = 0;          // chunk index of A
  int pos = head;      // next byte within A:B, < 16 when an iteration starts
  int consumed = 0;    // lane bits consumed
  const int n = nsym > 0 ? d.count : 0;
  FT* const out = folded + d.folded_offset;
  const int nsteps = n > r ? (n - r + 3) >> 2 : 0;
  // shards without work (past the list's end, rejected tables) do not hold
  // back the unpredicated rounds; their lanes decode nothing of use and
  // store to a sink slot
  const int wave_steps = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)nsteps));
  const int full_steps = min(wave_steps, __builtin_amdgcn_readfirstlane((int)wave_min(
                                             (uint32_t)(nsym > 0 ? nsteps : 0x7fffffff))));
  FT* const flush_base = nsym > 0 ? out : (FT*)(sink + lane * kLaneSinkBytes);
  const int flush_step = nsym > 0 ? 1 : 0;
  // settle the window loads here: otherwise the wait for them stays at the
  // top of the loop, where it would also wait for every earlier store
  __builtin_amdgcn_s_waitcnt(0);

  // one step: slot -> rank -> state update and renormalisation
  auto step = [&](uint64_t& bb, int& used, int slot_pos, bool live) __attribute__((always_inline)) {
    uint32_t value;
    const uint32_t xd = look.update(x, value);
    stage[4 * slot_pos + r] = (FT)value;  // slot_pos: step within the block
    // one byte below 2^23, two below 2^15 (never more: xd >= 2^11): the
    // leading zeros less one, rounded down to a multiple of 8 (none from
    // 2^31, where a state read from a malformed stream may start).  (Two
    // compares and two selects over both renormalised candidates make a
    // shorter chain on paper but run 9 % slower: a select waits longer for
    // a compare's lane mask than for a VGPR.)
    const uint32_t lz = (uint32_t)__builtin_clz(xd);
    int sh = (int)(__builtin_elementwise_sub_sat(lz, 1u) & 0x18u);
    if (!live) sh = 0;
    const uint64_t wide = (((uint64_t)xd << 32) | (uint32_t)(bb >> 32)) << sh;
    x = live ? (uint32_t)(wide >> 32) : x;
    bb <<= sh;
    used += sh;
  };
  // the block's symbols (from step s0): lane r stores kPer of them from
  // symbol kPer * r, 32 bytes
  using Blk = O0Block<FT, kLines>;
  constexpr int kPer = Blk::kPerLane;
  // (non-temporal: the residuals are read once, by the reconstruction, and
  // streaming them past the L2 keeps the lanes' body lines there between
  // their chunk loads)
  auto flush_full = [&](int s0) __attribute__((always_inline)) {
    using q4 = uint32_t __attribute__((ext_vector_type(4)));
    FT* const dst = flush_base + (4 * s0 + kPer * r) * flush_step;
#pragma unroll
    for (int i = 0; i < kPer * (int)sizeof(FT) / 16; ++i)
      __builtin_nontemporal_store(((const q4*)(stage + kPer * r))[i], (q4*)dst + i);
  };
  auto flush = [&](int s0) __attribute__((always_inline)) {
    const int p0 = 4 * s0 + kPer * r;
    if (p0 + kPer <= n) {
      flush_full(s0);
    } else {
      for (int j = 0; p0 + j < n; ++j) out[p0 + j] = stage[kPer * r + j];
    }
  };
  // Eight steps take at most 16 bytes: an iteration reads the 16 bytes at
  // pos of A:B, in two rounds of four steps (the bit window of a round is
  // 64 bits; the second round's continues with the bytes the first did not
  // reach).  After them the window moves by one chunk at most (B <- C); a
  // lane that moved loads the chunk after B into C as the iteration's last
  // memory operation (`refill`), which the next move takes, so each load
  // has a whole iteration to land.
  // (h8: the iteration's first step within its block)
  auto iteration = [&](int s0, int h8, bool predicated) __attribute__((always_inline)) {
    uint64_t rest;
    uint64_t bb = window16_be(a0, a1, b0, b1, pos, rest);
    int used = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) step(bb, used, h8 + t, !predicated || s0 + t < nsteps);
    {
      // bb is the first 8 bytes << used (0..64 bits): append the next
      // `used` bits, rest >> (64 - used), as two shifts of 32 - used / 2
      const uint32_t half = 32u - ((uint32_t)used >> 1);
      bb |= (rest >> half) >> half;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) step(bb, used, h8 + 4 + t, !predicated || s0 + 4 + t < nsteps);
    consumed += used;
    const int p = pos + (used >> 3);  // < 32
    const bool adv = p >= 16;
    // (kept after the rounds: the wait for the last refill goes here)
    __builtin_amdgcn_sched_barrier(0);
    a0 = adv ? b0 : a0;
    a1 = adv ? b1 : a1;
    b0 = adv ? c0 : b0;
    b1 = adv ? c1 : b1;
    // (the moves happen here, before the refill's branch: moved past it,
    // they would keep C's old value alive under the load into C)
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(b0), "+v"(b1));
    pos = p & 15;
    ci += p >> 4;
    pend = adv;
  };
  // (only lanes that moved load, straight into C: the others keep C, which
  // has landed, and lanes that do not load touch no memory)
  auto refill = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_sched_barrier(0);
    if (pend) {
      const Chunk Nn = load_chunk(cp + min(ci + 2, last));
      c0 = Nn.lo;
      c1 = Nn.hi;
    }
    __builtin_amdgcn_sched_barrier(0);
  };

  constexpr int kBlock = Blk::kSteps;
  int s0 = 0;
  for (; s0 + kBlock <= full_steps; s0 += kBlock) {
    // Fair progress (`fair`, a uniform argument): a wave's issue priority
    // falls from 3 to 0 over each quarter of its 1024-step chains, so the
    // waves sharing a SIMD, which the hardware otherwise serves oldest
    // first, finish their groups closer together
    if (fair && (s0 & 255) == 0) {
      const int quarter = s0 >> 8;
      if (quarter == 0) __builtin_amdgcn_s_setprio(3);
      else if (quarter == 1) __builtin_amdgcn_s_setprio(2);
      else if (quarter == 2) __builtin_amdgcn_s_setprio(1);
      else __builtin_amdgcn_s_setprio(0);
    }
#pragma unroll
    for (int h = 0; h < Blk::kIters; ++h) {
      iteration(s0 + 8 * h, 8 * h, false);
      if (h + 1 < Blk::kIters) refill();
    }
    // every lane of the wave has the whole block here (full_steps >=
    // s0 + kBlock): unconditional stores, so the memory waits stay exact
    flush_full(s0);
    refill();
  }
  // (a block past a lane's last step decodes nothing: its steps are not live)
  for (; s0 < wave_steps; s0 += kBlock) {
#pragma unroll
    for (int h = 0; h < Blk::kIters; ++h) {
      iteration(s0 + 8 * h, 8 * h, true);
      if (h + 1 < Blk::kIters) refill();
    }
    flush(s0);
    refill();
  }
  if (nsym > 0 && (consumed != 8 * length || x != kRansL)) set_status(&states[d.frame], 5);
}

template <typename FT, int kMax>
__global__ __launch_bounds__(64) void k_dec_o0(const FrameDesc* __restrict__ frames,
                                               const DShard* __restrict__ dshards,
                                               const int32_t* __restrict__ lists,
                                               const int32_t* __restrict__ tier_list,
                                               const int32_t* __restrict__ tier_count,
                                               O0Tab* __restrict__ tabs,
                                               FrameState* __restrict__ states,
                                               FT* __restrict__ folded, char* gbase,
                                               uint8_t* __restrict__ sink,
                                               int32_t* __restrict__ queue, int fair) {
  __shared__ O0Lds<FT, kMax> lds;
  const int lane = lane_id();
  const int g = lane >> 2;
  const int r = lane & 3;
  const int count = *tier_count;
  using B = O0Buckets<kMax>;
  uint32_t* const bucket = lds.bucket[g];
  uint2* const ent = lds.row[g];
  FT* const stage = (FT*)lds.stage[g];
  // a wave's first group is its block index, later ones come from the
  // queue: blocks past the list's end (all of them, for an empty tier) leave
  // without touching the queue counter every wave shares
  for (int grp = blockIdx.x;;) {
    if (grp * kO0Shards >= count) break;
    const int at = grp * kO0Shards + g;
    const int item = at < count ? tier_list[at] : -1;
    const O0Tab* tab = tabs + (item >= 0 ? item : 0);
    const int nsym = item >= 0 ? tab->nsym : 0;
    // ---- tables: entries and values, bucket start bits, counts before
    // each bucket (a parsed table has 1 <= cum_k < 4096 for k >= 1)
    for (int q = r; q < B::kWords / 4; q += 4) ((uint4*)bucket)[q] = make_uint4(0, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();
    for (int k = r; k < nsym; k += 4) {
      const uint32_t e0 = tab->ent[k];
      const uint32_t c = e0 & 0x1fffu;
      const uint32_t fr = (tab->ent[k + 1] & 0x1fffu) - c;
      if constexpr (sizeof(FT) == 2) {
        ent[k] = make_uint2(c | ((e0 >> 13) << 16), fr - 4096u);
      } else {
        ent[k] = make_uint2(c | (((fr - 4096u) & 0x1fffu) << 13), e0 >> 13);
      }
      if (k > 0) {
        const uint32_t b = B::of(c);
        atomicOr(&bucket[b], 1u << (c - (uint32_t)B::kSlots * b));
      }
    }
    __builtin_amdgcn_wave_barrier();
    {
      // lane r: buckets kPerLane r .. kPerLane (r + 1) - 1
      uint32_t* const mine = bucket + B::kPerLane * r;
      uint32_t total = 0;
      for (int i = 0; i < B::kPerLane; ++i) total += __popc(mine[i]);
      // exclusive scan of the four lane totals of the quad
      const uint32_t t1 = __shfl_up(total, 1, 4);
      const uint32_t s1 = total + (r >= 1 ? t1 : 0u);
      const uint32_t t2 = __shfl_up(s1, 2, 4);
      const uint32_t incl = s1 + (r >= 2 ? t2 : 0u);
      uint32_t run = incl - total;
      for (int i = 0; i < B::kPerLane; ++i) {
        const uint32_t w = mine[i];
        mine[i] = w | (run << B::kSlots);
        run += __popc(w);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // ---- the shard's rANS lanes
    BucketLookup<FT, kMax> look{bucket, ent};
    o0_lanes<FT, false>(look, item, nsym, tab, r, lane, frames, dshards, lists, states, folded,
                        stage, gbase, sink, fair != 0);
    int next = 0;
    if (lane == 0) next = atomicAdd(queue, 1);
    grp = (int)gridDim.x + __builtin_amdgcn_readfirstlane(next);
  }
}

// Latency form of the order-0 decoder, for batches whose order-0 shards fit
// the device at once (a frame or a few): four shards per wave, each with a
// direct 4096-slot table in LDS, entry {cum : 12, freq - 1 : 12, rank : 8},
// so a step has one dependent LDS read instead of the bucket tables' two:
// freq * (x >> 12) + (x & 4095) - cum = (freq - 1) * (x >> 12) + ((x >> 12)
// + (x & 4095) - cum), one unsigned 24-bit multiply-add of operands that
// depend only on x and the entry.  The symbol's value comes from the rank
// by a second read off the state chain.  Lanes 16..63 of the wave have no
// shard.  Shards come from the whole order-0 list (any alphabet of at most
// 255 symbols), block by block.
constexpr int kO0DirectShards = 4;

template <typename FT>
struct O0DirectLds {
  uint32_t slot[kO0DirectShards][4096];
  uint32_t value[kO0DirectShards][256];
  uint32_t cum[kO0DirectShards][257];
  uint32_t stage[kO0Shards][O0Block<FT, true>::kWords];
};

struct DirectLookup {
  const uint32_t* slot;
  const uint32_t* value;
  __device__ __forceinline__ uint32_t update(uint32_t x, uint32_t& v) const {
    const uint32_t e = slot[x & 4095u];
    const uint32_t q = x >> kRansTableLog;
    const uint32_t t = q + (x & 4095u);
    v = value[e >> 24];
    return __umul24((e >> 12) & 4095u, q) + (t - (e & 4095u));
  }
};

template <typename FT>
__global__ __launch_bounds__(64) void k_dec_o0_direct(const FrameDesc* __restrict__ frames,
                                                      const DShard* __restrict__ dshards,
                                                      const int32_t* __restrict__ lists,
                                                      const int32_t* __restrict__ o0_count,
                                                      O0Tab* __restrict__ tabs,
                                                      FrameState* __restrict__ states,
                                                      FT* __restrict__ folded, char* gbase,
                                                      uint8_t* __restrict__ sink) {
  __shared__ O0DirectLds<FT> lds;
  const int lane = lane_id();
  const int g = lane >> 2;
  const int r = lane & 3;
  const int gs = g < kO0DirectShards ? g : 0;  // lanes without a shard read shard 0's table
  const int count = *o0_count;
  for (int blk = blockIdx.x; blk * kO0DirectShards < count; blk += gridDim.x) {
    const int at = blk * kO0DirectShards + g;
    const int item = g < kO0DirectShards && at < count ? at : -1;
    const O0Tab* tab = tabs + (item >= 0 ? item : 0);
    const int nsym = item >= 0 ? tab->nsym : 0;
    uint32_t* const slot = lds.slot[gs];
    uint32_t* const val = lds.value[gs];
    uint32_t* const cum = lds.cum[gs];
    // values and cumulative counts by rank (cum[nsym] = 4096)
    for (int k = r; k <= nsym; k += 4) {
      const uint32_t e = tab->ent[k];
      cum[k] = e & 0x1fffu;
      if (k < nsym) val[k] = e >> 13;
    }
    __builtin_amdgcn_wave_barrier();
    // slots: lane r fills [1024 r, 1024 r + 1024), symbol by symbol
    const int lo = 1024 * r, hi = lo + 1024;
    const int wave_nsym = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)nsym));
    for (int k = 0; k < wave_nsym; ++k) {
      if (k >= nsym) continue;
      const int c0 = (int)cum[k], c1 = (int)cum[k + 1];
      const int a = max(c0, lo), b = min(c1, hi);
      if (a >= b) continue;
      const uint32_t e = (uint32_t)c0 | ((uint32_t)(c1 - c0 - 1) << 12) | ((uint32_t)k << 24);
      int sl = a;
      for (; sl < b && (sl & 3); ++sl) slot[sl] = e;
      for (; sl + 4 <= b; sl += 4) *(uint4*)&slot[sl] = make_uint4(e, e, e, e);
      for (; sl < b; ++sl) slot[sl] = e;
    }
    __builtin_amdgcn_wave_barrier();
    DirectLookup look{slot, val};
    o0_lanes<FT, true>(look, item, nsym, tab, r, lane, frames, dshards, lists, states, folded,
                       (FT*)lds.stage[g], gbase, sink);
    __builtin_amdgcn_wave_barrier();
  }
}

// Rice shards (decode.cu:297-340): sixteen shards per wave, a quad of lanes
// per shard, lane r decoding Rice lane r (symbols r, r + 4, ...), waves
// persistent over the Rice list like k_dec_o0.  A lane keeps a 64-bit bit
// buffer (next bit lowest) that one 4-byte refill per step holds at >= 32
// bits, so a code of at most 32 bits (its zero run found by one ctz, then
// the one, then k remainder bits) decodes in one step.  Refills take the
// bytes at `pos` of the lane's two-chunk window A:B: a round of four steps
// reads at most 16 bytes from pos < 16, so the window moves by one chunk at
// most per round, through a chunk pipeline (C, D behind it, the next chunk
// loaded as the round's last memory operation by lanes that moved).  Decoded values are staged per shard in LDS and leave as one
// 16-byte store per lane every two rounds.  A longer code (a zero run past
// the buffer) is read from the lane's chunks in memory and the window is
// reloaded where it ends; after that the round's remaining refills still
// read inside A:B (pos <= 19 + 8).  A lane whose stream the reference
// rejects inside such a code stops, and its shard goes back on the zero-run
// / block-pack list, where k_dec_other decodes it bit by bit and reports it;
// bytes read past a lane's end fail the final "consumed bits" check exactly
// where the reference runs out of bits.
constexpr int kRiceShards = 16;

// Four bytes starting at byte `pos` (0..27) of the 32-byte window
// a0:a1:b0:b1, the first byte least significant.  (Taken by value: a select
// among references would address the window through memory.)
__device__ __forceinline__ uint32_t window4_le(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                                               int pos) {
  const int idx = pos >> 3;
  const uint64_t lo = idx == 0 ? a0 : idx == 1 ? a1 : idx == 2 ? b0 : b1;
  const uint64_t hi = idx == 0 ? a1 : idx == 1 ? b0 : b1;
  const uint32_t sh = 8u * (uint32_t)(pos & 7);
  return sh ? (uint32_t)((lo >> sh) | (hi << (64u - sh))) : (uint32_t)lo;
}

// Eight bytes starting at byte `o` (0..15) of two consecutive 16-byte chunks,
// the first byte least significant.
__device__ __forceinline__ uint64_t bytes8_le(uint64_t x0, uint64_t x1, uint64_t y0, int o) {
  const uint64_t lo = o < 8 ? x0 : x1;
  const uint64_t hi = o < 8 ? x1 : y0;
  const uint32_t sh = 8u * (uint32_t)(o & 7);
  return sh ? (lo >> sh) | (hi << (64u - sh)) : lo;
}

template <typename FT>
__global__ __launch_bounds__(64) void k_dec_rice(const FrameDesc* __restrict__ frames,
                                                 FrameState* __restrict__ states,
                                                 const DShard* __restrict__ dshards,
                                                 const int32_t* __restrict__ rice_list,
                                                 const int32_t* __restrict__ rice_count,
                                                 int32_t* __restrict__ other_list,
                                                 int32_t* __restrict__ other_count,
                                                 FT* __restrict__ folded, char* gbase,
                                                 uint8_t* __restrict__ sink,
                                                 int32_t* __restrict__ queue) {
  __shared__ uint32_t stage_all[kRiceShards][O0Stage<FT>::kWords];
  const int lane = lane_id();
  const int g = lane >> 2;
  const int r = lane & 3;
  const int count = *rice_count;
  FT* const stage = (FT*)stage_all[g];
  for (int grp = blockIdx.x;;) {
    if (grp * kRiceShards >= count) break;
    const int at = grp * kRiceShards + g;
    const int item = at < count ? rice_list[at] : -1;
    DShard d = {};
    if (item >= 0) d = dshards[item];
    const FrameDesc& f = frames[d.frame];
    FrameState* const st = &states[d.frame];
    const uint32_t max_folded = (uint32_t)f.max_sample * 2u;
    const uint8_t* body = as_global(gbase, (const uint8_t*)f.stream + d.body_offset);
    const int n = item >= 0 ? d.count : 0;
    const int k = item >= 0 ? d.mode - 1 : 0;
    const int64_t body_length = d.body_length;
    // lane table (decode.cu:301-313): lane lengths 0..2, the last lane takes
    // the rest of the body
    const int lane_count = n < 4 ? n : 4;
    const int64_t table_bytes = (int64_t)(lane_count - 1) * 4;
    const bool mine = r < lane_count;
    bool ok = mine && table_bytes <= body_length;
    int64_t lane_offset = table_bytes;
    int64_t lane_length = 0;
    if (ok) {
      for (int i = 0; i < r; ++i) lane_offset += rd_u32(body + 4 * i);
      lane_length = r + 1 < lane_count ? (int64_t)rd_u32(body + 4 * r) : body_length - lane_offset;
      if (lane_offset > body_length || lane_length < 0 || lane_length > body_length - lane_offset) {
        ok = false;
      }
    }
    const int length = ok ? (int)lane_length : 0;
    // aligned chunks holding the lane's bytes (a lane without bytes reads the
    // body's first chunk and never uses it)
    const uint8_t* first = length > 0 ? body + lane_offset : body;
    const int head = (int)((uintptr_t)first & 15);
    const uint4* cp = (const uint4*)(first - head);
    const int last = length > 0 ? (head + length - 1) >> 4 : 0;
    const Chunk A0 = load_chunk(cp);
    const Chunk B0 = load_chunk(cp + min(1, last));
    const Chunk C0 = load_chunk(cp + min(2, last));
    const Chunk D0 = load_chunk(cp + min(3, last));
    uint64_t a0 = A0.lo, a1 = A0.hi, b0 = B0.lo, b1 = B0.hi, c0 = C0.lo, c1 = C0.hi;
    uint64_t d0 = D0.lo, d1 = D0.hi;  // chunk ci + 3, or in flight into n0:n1
    uint64_t n0 = 0, n1 = 0;
    bool pend = false;  // n0:n1 holds chunk ci + 3 once its load lands
    int ci = 0;         // chunk index of A
    int pos = head;     // next byte within A:B, < 16 when a round starts
    uint64_t buf = 0;   // bit buffer, next bit lowest
    int nb = 0;         // bits in buf
    int consumed = 0;   // lane bits consumed
    bool slow = false;  // met a code longer than 32 bits
    const uint32_t kmask = (1u << k) - 1u;
    FT* const out = folded + d.folded_offset;
    const int nsteps = ok ? (n - r + 3) >> 2 : 0;
    const int wave_steps = __builtin_amdgcn_readfirstlane((int)wave_max((uint32_t)nsteps));
    const int full_steps = min(wave_steps, __builtin_amdgcn_readfirstlane((int)wave_min(
                                               (uint32_t)(item >= 0 && mine ? nsteps : 0x7fffffff))));
    // lanes without symbols store to their sink slot
    FT* const flush_base = nsteps > 0 ? out : (FT*)(sink + lane * kLaneSinkBytes);
    const int flush_step = nsteps > 0 ? 1 : 0;
    __builtin_amdgcn_s_waitcnt(0);

    auto step = [&](int slot_pos, bool live) __attribute__((always_inline)) {
      // refill: the four bytes at pos of A:B (pos < 28 within a round)
      const bool fill = nb < 32;
      const uint32_t word = window4_le(a0, a1, b0, b1, pos);
      buf = fill ? buf | ((uint64_t)word << nb) : buf;
      pos += fill ? 4 : 0;
      nb += fill ? 32 : 0;
      // one code: z zero bits, a one, k remainder bits
      const uint32_t low = (uint32_t)buf;
      const int z = low ? (int)__builtin_ctz(low) : 32;
      const int used = z + 1 + k;
      const bool fits = used <= 32;
      uint32_t value = ((uint32_t)z << k) | ((low >> ((z + 1) & 31)) & kmask);
      const bool take = live && fits && !slow;
      if (live && !fits && !slow) {
        // a longer code: read it from the lane's aligned chunks in memory
        // (a chunk never crosses a page, so bytes past the lane's end are
        // readable and only fail the final check), then reload the window
        // where it ends
        auto bits_at = [&](int bit) __attribute__((always_inline)) {
          const int a = head + (bit >> 3);
          const Chunk x = load_chunk(cp + min(a >> 4, last));
          const Chunk y = load_chunk(cp + min((a >> 4) + 1, last));
          return bytes8_le(x.lo, x.hi, y.lo, a & 15) >> (bit & 7);
        };
        int bit = consumed;
        uint64_t q = 0;
        bool good = true;
        for (;;) {
          if ((bit >> 3) >= length || q > max_folded) {
            good = false;
            break;
          }
          const uint64_t w = bits_at(bit);
          const int avail = 64 - (bit & 7);
          if (w == 0) {
            q += (uint64_t)avail;
            bit += avail;
            continue;
          }
          const int zz = (int)__builtin_ctzll(w);
          q += (uint64_t)zz;
          bit += zz + 1;
          break;
        }
        uint64_t v = q << k;
        if (good && k > 0) {
          v |= bits_at(bit) & kmask;
          bit += k;
        }
        if (!good || q > max_folded || v > max_folded) {
          slow = true;  // the reference fails here: k_dec_other reports it
        } else {
          value = (uint32_t)v;
          consumed = bit;
          const int a = head + (bit >> 3);
          ci = a >> 4;
          pos = a & 15;
          const Chunk A1 = load_chunk(cp + min(ci, last));
          const Chunk B1 = load_chunk(cp + min(ci + 1, last));
          const Chunk C1 = load_chunk(cp + min(ci + 2, last));
          const Chunk D1 = load_chunk(cp + min(ci + 3, last));
          a0 = A1.lo, a1 = A1.hi, b0 = B1.lo, b1 = B1.hi;
          c0 = C1.lo, c1 = C1.hi, d0 = D1.lo, d1 = D1.hi;
          pend = false;  // the chunk in flight belongs to the old window
          buf = (uint64_t)(window4_le(a0, a1, b0, b1, pos) >> (bit & 7));
          nb = 32 - (bit & 7);
          pos += 4;
        }
        // nothing of this branch stays in flight past it
        __builtin_amdgcn_s_waitcnt(0);
      }
      stage[4 * slot_pos + r] = (FT)value;
      ok = ok && !(take && value > max_folded);
      const int u = take ? used : 0;
      buf >>= u;
      nb -= u;
      consumed += u;
    };
    // the 32 symbols of two rounds: lane r stores symbols 8 r .. 8 r + 7
    auto flush_full = [&](int s0) __attribute__((always_inline)) {
      FT* const dst = flush_base + (4 * s0 + 8 * r) * flush_step;
      if constexpr (sizeof(FT) == 2) {
        *(uint4*)dst = *(const uint4*)(stage + 8 * r);
      } else {
        ((uint4*)dst)[0] = ((const uint4*)(stage + 8 * r))[0];
        ((uint4*)dst)[1] = ((const uint4*)(stage + 8 * r))[1];
      }
    };
    auto flush = [&](int s0) __attribute__((always_inline)) {
      const int p0 = 4 * s0 + 8 * r;
      if (p0 + 8 <= n) {
        flush_full(s0);
      } else {
        for (int j = 0; p0 + j < n; ++j) out[p0 + j] = stage[8 * r + j];
      }
    };
    // one round: four steps, then the window moves by one chunk at most
    auto round = [&](int s0, int half, bool predicated) __attribute__((always_inline)) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        step(4 * half + t, !predicated || s0 + 4 * half + t < nsteps);
      }
      const bool adv = pos >= 16;
      // (kept after the steps: the wait for last round's load goes here)
      __builtin_amdgcn_sched_barrier(0);
      d0 = pend ? n0 : d0;
      d1 = pend ? n1 : d1;
      a0 = adv ? b0 : a0;
      a1 = adv ? b1 : a1;
      b0 = adv ? c0 : b0;
      b1 = adv ? c1 : b1;
      c0 = adv ? d0 : c0;
      c1 = adv ? d1 : c1;
      pos -= adv ? 16 : 0;
      ci += adv ? 1 : 0;
      pend = adv;
    };
    // lanes that did not move read one line every such lane of the wave
    // shares, so the load stays unconditional and n0:n1 needs no merge
    auto refill = [&]() __attribute__((always_inline)) {
      __builtin_amdgcn_sched_barrier(0);
      const Chunk Nn = load_chunk(pend ? cp + min(ci + 3, last) : (const uint4*)dshards);
      n0 = Nn.lo;
      n1 = Nn.hi;
      __builtin_amdgcn_sched_barrier(0);
    };

    int s0 = 0;
    for (; s0 + 8 <= full_steps; s0 += 8) {
      round(s0, 0, false);
      refill();
      round(s0, 1, false);
      flush_full(s0);
      refill();
    }
    for (; s0 < wave_steps; s0 += 8) {
      round(s0, 0, true);
      refill();
      round(s0, 1, true);
      flush(s0);
      refill();
    }
    // end of the lane (decode.cu:328-338): every byte used, zero padding bits
    if (ok && !slow) {
      const int used_bytes = (consumed + 7) >> 3;
      if (used_bytes != length) {
        ok = false;
      } else if ((consumed & 7) != 0) {
        const uint32_t pad = 8u - (uint32_t)(consumed & 7);
        if ((buf & ((1u << pad) - 1u)) != 0) ok = false;
      }
    }
    // a shard with a lane that stopped goes to k_dec_other whole
    const uint64_t redo_lanes = ballot(item >= 0 && slow);
    const uint64_t redo = redo_lanes & (0xfull << (4 * g));
    if (item >= 0 && mine && !ok && redo == 0) set_status(st, 2);
    if (r == 0 && redo != 0) other_list[atomicAdd(other_count, 1)] = item;
    int next = 0;
    if (lane == 0) next = atomicAdd(queue, 1);
    grp = (int)gridDim.x + __builtin_amdgcn_readfirstlane(next);
  }
}

// predictor="serial": one thread per tile in raster order (decode.cu:608-656).
template <typename FT>
__global__ __launch_bounds__(256) void k_serial(const FrameDesc* __restrict__ frames,
                                                const FrameState* __restrict__ states,
                                                const TileDesc* __restrict__ tiles, int tile_total,
                                                const FT* __restrict__ folded) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tile_total) return;
  const TileDesc td = tiles[t];
  if (states[td.frame].status != 0) return;
  const FrameDesc& f = frames[td.frame];
  const int step = f.step, max_sample = f.max_sample, med = f.use_med;
  const int64_t pitch = f.plane_width[td.plane];
  uint16_t* dst = f.out_plane[td.plane] + td.dst_offset;
  const FT* fold = folded + td.folded_base;
  for (int64_t y = 0; y < td.height; ++y) {
    int left = 0, ul = 0;
    for (int64_t x = 0; x < td.width; ++x) {
      const int above = y > 0 ? dst[(y - 1) * pitch + x] : 0;
      const int grad = left + above - ul;
      const int pred = med ? median3(left, above, grad) : clampi(grad, 0, max_sample);
      const uint32_t fv = fold[y * td.width + x];
      const int q = (fv & 1u) ? -(int)(fv >> 1) - 1 : (int)(fv >> 1);
      const int rec = clampi(pred + q * step, 0, max_sample);
      dst[y * pitch + x] = (uint16_t)rec;
      ul = above;
      left = rec;
    }
  }
}

__global__ void k_dresults(const FrameState* __restrict__ states, Results* __restrict__ results,
                           int frame_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= frame_count) return;
  Results r;
  r.stream_length = 0;
  r.status = states[i].status;
  r.flags = 0;
  results[i] = r;
}

template <typename FT>
static void decode_all(const DecodeBuffers& b, hipStream_t stream) {
  const dim3 fgrid((b.frame_count + 63) / 64);
  {
    // the tile geometry, and the frame states, list counts and work counters
    // every kernel after it reads
    PlanInit init = {};
    init.kind = 2;
    init.states = b.states;
    init.queues = b.queues;
    init.queue_count = kQueueCount;
    init.list_count = b.list_count;
    init.list_counts = kListCounts;
    launch_plan(b.frames, b.frame_count, b.tiles, b.shards, b.group_first, b.tile_total, stream,
                init);
  }
  if (b.fed_sizes != nullptr) {
    FVH_TIMED("k_dheader", stream);
    hipLaunchKernelGGL(k_dheader, fgrid, dim3(64), 0, stream, b.frames, b.states, b.frame_count,
                       b.fed_sizes, (char*)b.folded);
  }
  {
    FVH_TIMED("k_parse", stream);
    // one wave per block: a batch has only thousands of tiles, and blocks of
    // four waves left most compute units idle (25 -> under 20 us on 16 4K
    // RGB frames)
    hipLaunchKernelGGL(k_parse, dim3((b.tile_total + 63) / 64), dim3(64), 0, stream, b.frames,
                       b.states, b.tiles, b.tile_total, b.dshards, b.lists, b.list_count,
                       b.shard_total, (char*)b.folded);
  }
  // a batch whose shards all fit the device at once takes the latency forms
  // of the table parse (a wave per shard) and of the order-0 decoder (four
  // shards per wave, direct slot tables): a frame's decode time is then one
  // shard's chain of steps, each one LDS read long
  static const int direct_per_cu = blocks_per_cu((const void*)k_dec_o0_direct<FT>, 0);
  const bool latency_form = b.shard_total <= kO0DirectShards * direct_per_cu * device_cus();
  int32_t* const tier_lists = b.lists + 2 * b.shard_total;
  // Batches holding a frame of more than 8 bits (whose shards' tables run to
  // hundreds of symbols) parse with the wave form too, then file the shards
  // for the bucket decoder; a thread's serial walk of a long table took 147
  // us on four 4K gray 10-bit lossless frames, the wave form 10.  Short
  // tables, as in 8-bit frames, parse faster a thread each (16 4K RGB
  // frames: 47 us, the wave form 51).
  if (latency_form || b.any_deep) {
    {
      FVH_TIMED("k_o0_parse_wave", stream);
      hipLaunchKernelGGL(k_o0_parse_wave, dim3(max(b.shard_total, 1)), dim3(64), 0, stream,
                         b.frames, b.states, b.dshards, b.lists, b.list_count, b.o0tabs,
                         (char*)b.folded);
    }
    if (!latency_form) {
      FVH_TIMED("k_o0_file", stream);
      hipLaunchKernelGGL(k_o0_file, dim3((b.shard_total + kFileThreads - 1) / kFileThreads),
                         dim3(kFileThreads), 0, stream, b.list_count, tier_lists, b.shard_total,
                         b.o0tabs);
    }
  } else {
    FVH_TIMED("k_o0_parse", stream);
    hipLaunchKernelGGL(k_o0_parse, dim3((b.shard_total + kParseThreads - 1) / kParseThreads),
                       dim3(kParseThreads), 0, stream,
                       b.frames, b.states, b.dshards, b.lists, b.list_count, tier_lists,
                       b.shard_total, b.o0tabs, (char*)b.folded);
  }
  if (latency_form) {
    FVH_TIMED("k_dec_o0_direct", stream);
    const int blocks = (b.shard_total + kO0DirectShards - 1) / kO0DirectShards;
    hipLaunchKernelGGL((k_dec_o0_direct<FT>), dim3(max(blocks, 1)), dim3(64), 0, stream, b.frames,
                       b.dshards, b.lists, b.list_count, b.o0tabs, b.states, (FT*)b.folded,
                       (char*)b.folded, b.sink);
  } else {
    FVH_TIMED("k_dec_o0", stream);
    // One persistent instance per alphabet tier; each drains its list.  The
    // grid is the waves the device holds at once: a wave past them starts
    // only when a resident one leaves, at the end of the list, and then
    // decodes the group its block index names alone (16 4K RGB 8-bit
    // frames, 6075 groups at 11 waves per compute unit: 451 -> 427 us).
    // Every group is 1024 steps long and each SIMD runs its waves VALU-bound,
    // oldest first, so the groups of a round end spread out; that
    // pipelines a last round that is only partly filled, while a last round
    // that is nearly full (or a single round) ends sooner with the waves kept
    // level (`fair`): 6075 groups at 12 waves (1.98 rounds) 441 -> 412 us,
    // but at 11 waves (2.16 rounds) 427 -> 450.  The count of groups is the
    // order-0 shards of every tier, which for 8-bit frames are this tier's.
    const int groups = (b.shard_total + kO0Shards - 1) / kO0Shards;
    static const int o0 = blocks_per_cu((const void*)k_dec_o0<FT, o0_tier_cap(0)>, 0);
    const int r0 = o0 * device_cus();
    const dim3 grid(min(groups, r0));
    const int part = groups % r0;
    const int fair = groups <= r0 || part == 0 || 4 * part >= 3 * r0;
    int32_t* tiers = b.lists + 2 * b.shard_total;
    hipLaunchKernelGGL((k_dec_o0<FT, o0_tier_cap(0)>), grid, dim3(64), 0, stream, b.frames,
                       b.dshards, b.lists, tiers, b.list_count + 2, b.o0tabs, b.states,
                       (FT*)b.folded, (char*)b.folded, b.sink, b.queues + 4, fair);
    // wider tables: fewer resident waves, and waves beyond them only queue
    static const int o1 = blocks_per_cu((const void*)k_dec_o0<FT, o0_tier_cap(1)>, 0);
    static const int o2 = blocks_per_cu((const void*)k_dec_o0<FT, o0_tier_cap(2) - 1>, 0);
    const int r1 = o1 * device_cus(), r2 = o2 * device_cus();
    hipLaunchKernelGGL((k_dec_o0<FT, o0_tier_cap(1)>), dim3(min((int)grid.x, r1)), dim3(64), 0,
                       stream, b.frames, b.dshards, b.lists, tiers + b.shard_total,
                       b.list_count + 3, b.o0tabs, b.states, (FT*)b.folded, (char*)b.folded,
                       b.sink, b.queues + 5, 0);
    hipLaunchKernelGGL((k_dec_o0<FT, o0_tier_cap(2) - 1>), dim3(min((int)grid.x, r2)), dim3(64),
                       0, stream, b.frames, b.dshards, b.lists, tiers + 2 * b.shard_total,
                       b.list_count + 4, b.o0tabs, b.states, (FT*)b.folded, (char*)b.folded,
                       b.sink, b.queues + 6, 0);
  }
  {
    FVH_TIMED("k_dec_rice", stream);
    const dim3 grid(min((b.shard_total + kRiceShards - 1) / kRiceShards, 16 * device_cus()));
    hipLaunchKernelGGL((k_dec_rice<FT>), grid, dim3(64), 0, stream, b.frames, b.states, b.dshards,
                       b.lists + kRiceList * b.shard_total, b.list_count + kRiceList,
                       b.lists + b.shard_total, b.list_count + 1, (FT*)b.folded, (char*)b.folded,
                       b.sink, b.queues + kRiceQueue);
  }
  {
    FVH_TIMED("k_dec_other", stream);
    static const int per_cu = blocks_per_cu((const void*)k_dec_other<FT>, 0);
    const int grid = min(b.shard_total, per_cu * device_cus());
    hipLaunchKernelGGL((k_dec_other<FT>), dim3(grid), dim3(64), 0, stream, b.frames, b.states,
                       b.dshards, b.lists, b.list_count, b.shard_total, (FT*)b.folded,
                       (char*)b.folded);
  }
  if (b.wavefront) {
    launch_reconstruct<FT>(b, stream);
  } else {
    FVH_TIMED("k_serial", stream);
    hipLaunchKernelGGL((k_serial<FT>), dim3((b.tile_total + 255) / 256), dim3(256), 0, stream,
                       b.frames, b.states, b.tiles, b.tile_total, (const FT*)b.folded);
  }
  {
    FVH_TIMED("k_dresults", stream);
    hipLaunchKernelGGL(k_dresults, fgrid, dim3(64), 0, stream, b.states, b.results, b.frame_count);
  }
}

void launch_decode(const DecodeBuffers& b, hipStream_t stream) {
  if (b.wide_folded) decode_all<uint32_t>(b, stream);
  else decode_all<uint16_t>(b, stream);
}

}  // namespace fvh
