// fvh_decode.hip -- MI355X decoder kernels.
//
//   k_dstate_init  per-frame status
//   k_parse        directory + shard-header validation, one thread per tile
//                  (checks of cuda/csrc/decode.cu:131-220)
//   k_entropy      entropy decode of every mode, one wave64 per shard
//                  (decode.cu:222-550)
//   reconstruction: the shared wavefront kernel (fvh_encode.hip) or the
//                  one-thread-per-tile diagnostic k_serial (decode.cu:608-656)
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

__global__ void k_dstate_init(FrameState* __restrict__ states, int frame_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= frame_count) return;
  FrameState s = {};
  s.active = 1;
  states[i] = s;
}

// Directory entry and shard-header walk of one tile (decode.cu:139-219).
__global__ __launch_bounds__(256) void k_parse(const FrameDesc* __restrict__ frames,
                                               FrameState* __restrict__ states,
                                               const TileDesc* __restrict__ tiles, int tile_total,
                                               DShard* __restrict__ dshards) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tile_total) return;
  const TileDesc td = tiles[t];
  const FrameDesc& f = frames[td.frame];
  FrameState* st = &states[td.frame];
  const uint8_t* enc = f.stream;
  const int64_t size = f.stream_capacity;
  const int local = t - f.tile_begin;
  const int64_t T = f.tile_count;
  const int64_t directory_end = kHeaderBytes + T * kDirectoryEntryBytes;
  const uint8_t* entry = enc + kHeaderBytes + (int64_t)local * kDirectoryEntryBytes;
  if (entry[0] != td.plane || entry[1] != kEntropyParallelShards || entry[2] != kPredictorId ||
      entry[3] != 0 || rd_u32(entry + 4) != (uint32_t)td.x || rd_u32(entry + 8) != (uint32_t)td.y ||
      rd_u32(entry + 12) != (uint32_t)td.width || rd_u32(entry + 16) != (uint32_t)td.height) {
    set_status(st, kStatusDirectory);
    return;
  }
  const uint64_t off_u64 = rd_u64(entry + 20);
  if (off_u64 > (uint64_t)size) {
    set_status(st, kStatusDirectory);
    return;
  }
  const int64_t payload_offset = (int64_t)off_u64;
  const int64_t payload_length = rd_u32(entry + 28);
  if (payload_length > size - payload_offset) {
    set_status(st, kStatusDirectory);
    return;
  }
  if (local == 0) {
    if (payload_offset != directory_end) {
      set_status(st, kStatusDirectory);
      return;
    }
  } else {
    const uint8_t* prev = entry - kDirectoryEntryBytes;
    const uint64_t prev_off = rd_u64(prev + 20);
    const uint64_t prev_len = rd_u32(prev + 28);
    if (prev_off > (uint64_t)size || prev_len > (uint64_t)size - prev_off ||
        off_u64 != prev_off + prev_len) {
      set_status(st, kStatusDirectory);
      return;
    }
  }
  if (local + 1 == T && payload_offset + payload_length != size) {
    set_status(st, kStatusDirectory);
    return;
  }
  const int64_t samples = (int64_t)td.width * td.height;
  int64_t cursor = payload_offset;
  int64_t decoded = 0;
  const int64_t tile_end = payload_offset + payload_length;
  for (int ls = 0; ls < td.shard_count; ++ls) {
    if (cursor > tile_end - 3) {
      set_status(st, kStatusShardHeader);
      return;
    }
    const int mode = enc[cursor];
    const int64_t body_length = rd_u16(enc + cursor + 1);
    const int64_t body_offset = cursor + 3;
    if (mode > 19 || body_length > tile_end - body_offset) {
      set_status(st, kStatusShardHeader);
      return;
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
    decoded += n;
    cursor = body_offset + body_length;
  }
  if (decoded != samples || cursor != tile_end) set_status(st, kStatusShardHeader);
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

// One wave64 per shard.  Modes: zero-run (one lane), Rice (one lane per Rice
// lane), block-pack (one lane per 128-symbol block), order-0 rANS (table
// parse by one lane, LUT striped over the wave, one lane per rANS lane).
template <typename FT>
__global__ __launch_bounds__(256) void k_entropy(const FrameDesc* __restrict__ frames,
                                                 FrameState* __restrict__ states,
                                                 const TileDesc* __restrict__ tiles,
                                                 const ShardDesc* __restrict__ shards,
                                                 const DShard* __restrict__ dshards,
                                                 int shard_total, FT* __restrict__ folded) {
  __shared__ uint16_t values_all[4][256];
  __shared__ uint16_t freq_all[4][256];
  __shared__ uint16_t cum_all[4][256];
  __shared__ uint8_t lut_all[4][4096];
  __shared__ int meta_all[4][12];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int s = blockIdx.x * 4 + wave;
  if (s >= shard_total) return;
  // shard metadata is only defined once the frame's parse succeeded
  const int frame = tiles[shards[s].tile].frame;
  FrameState* st = &states[frame];
  if (st->status != 0) return;
  const DShard d = dshards[s];
  const FrameDesc& f = frames[d.frame];
  const uint32_t max_folded = (uint32_t)f.max_sample * 2u;
  const uint8_t* body = f.stream + d.body_offset;
  const int64_t body_length = d.body_length;
  const int64_t n = d.count;
  FT* out = folded + d.folded_offset;
  const int mode = d.mode;

  if (mode == 0) {  // decode.cu:237-295
    if (lane != 0) return;
    int64_t cursor = 0, produced = 0;
    while (produced < n) {
      uint32_t token = 0;
      int shift = 0, used = 0;
      for (;;) {
        if (cursor >= body_length || used == 5) {
          set_status(st, 1);
          return;
        }
        const uint8_t byte = body[cursor++];
        if (used == 4 && byte > 0x0f) {
          set_status(st, 1);
          return;
        }
        token |= (uint32_t)(byte & 0x7f) << shift;
        ++used;
        if ((byte & 0x80) == 0) break;
        shift += 7;
      }
      if (used != varint_len(token)) {
        set_status(st, 1);
        return;
      }
      if ((token & 1) == 0) {
        const int64_t run = token / 2 + 1;
        if (run > n - produced) {
          set_status(st, 1);
          return;
        }
        for (int64_t i = 0; i < run; ++i) out[produced++] = 0;
      } else {
        const uint32_t value = (token + 1) / 2;
        if (value == 0 || value > max_folded) {
          set_status(st, 1);
          return;
        }
        out[produced++] = (FT)value;
      }
    }
    if (cursor != body_length) set_status(st, 1);
    return;
  }

  if (mode >= 1 && mode <= 17) {  // decode.cu:297-340
    const int lane_count = (int)(n < 4 ? n : 4);
    if (lane >= lane_count) return;
    const int64_t table_bytes = (int64_t)(lane_count - 1) * 4;
    if (table_bytes > body_length) {
      set_status(st, 2);
      return;
    }
    int64_t lane_offset = table_bytes;
    for (int i = 0; i < lane; ++i) lane_offset += rd_u32(body + i * 4);
    const int64_t lane_length = lane + 1 < lane_count ? (int64_t)rd_u32(body + lane * 4)
                                                      : body_length - lane_offset;
    if (lane_offset > body_length || lane_length < 0 || lane_length > body_length - lane_offset) {
      set_status(st, 2);
      return;
    }
    BitReader reader{body + lane_offset, lane_length, 0, 0, 0, 0};
    const int k = mode - 1;
    for (int64_t i = lane; i < n; i += lane_count) {
      uint32_t value = 0;
      if (!reader.read(k, max_folded, &value)) {
        set_status(st, 2);
        return;
      }
      out[i] = (FT)value;
    }
    const int64_t used_bytes = (reader.consumed + 7) / 8;
    if (used_bytes != lane_length) {
      set_status(st, 2);
      return;
    }
    if ((reader.consumed & 7) != 0) {
      const uint8_t mask = (uint8_t)~((1u << (reader.consumed & 7)) - 1);
      if ((body[lane_offset + used_bytes - 1] & mask) != 0) set_status(st, 2);
    }
    return;
  }

  if (mode == 18) {  // decode.cu:342-396
    const int64_t block_count = (n + kBlockSymbols - 1) / kBlockSymbols;
    const int64_t block = lane;
    if (block >= block_count) return;
    int64_t cursor = 0;
    for (int64_t prior = 0; prior < block; ++prior) {
      if (cursor >= body_length) {
        set_status(st, 3);
        return;
      }
      const int w = body[cursor++];
      const int64_t pr = n - prior * kBlockSymbols;
      const int64_t pc = pr < kBlockSymbols ? pr : kBlockSymbols;
      cursor += (pc * w + 7) / 8;
    }
    if (cursor >= body_length) {
      set_status(st, 3);
      return;
    }
    const int w = body[cursor++];
    if (w > bit_length(max_folded)) {
      set_status(st, 3);
      return;
    }
    const int64_t rem = n - block * kBlockSymbols;
    const int64_t cnt = rem < kBlockSymbols ? rem : kBlockSymbols;
    const int64_t payload = (cnt * w + 7) / 8;
    if (payload > body_length - cursor) {
      set_status(st, 3);
      return;
    }
    const uint8_t* packed = body + cursor;
    int64_t bit = 0;
    for (int64_t i = 0; i < cnt; ++i) {
      uint32_t value = 0;
      for (int b = 0; b < w; ++b, ++bit) {
        value |= (uint32_t)((packed[bit >> 3] >> (bit & 7)) & 1) << b;
      }
      if (value > max_folded) {
        set_status(st, 3);
        return;
      }
      out[block * kBlockSymbols + i] = (FT)value;
    }
    if (block + 1 == block_count && cursor + payload != body_length) set_status(st, 3);
    if ((bit & 7) != 0) {
      const uint8_t mask = (uint8_t)~((1u << (bit & 7)) - 1);
      if ((packed[payload - 1] & mask) != 0) set_status(st, 3);
    }
    return;
  }

  // order-0 rANS (decode.cu:441-549)
  uint16_t* values = values_all[wave];
  uint16_t* freq = freq_all[wave];
  uint16_t* cum = cum_all[wave];
  uint8_t* lut = lut_all[wave];
  int* meta = meta_all[wave];  // [0]=valid [1]=n [2..5]=start [6..9]=len
  __shared__ uint32_t st_all[4][4];
  uint32_t* lane_state = st_all[wave];
  if (lane == 0) {
    int valid = 1;
    int64_t cursor = 0;
    if (body_length < 1 || body[cursor++] != kRansTableLog) valid = 0;
    uint32_t nsym = 0;
    if (valid && !read_varint(body, body_length, &cursor, &nsym)) valid = 0;
    if (nsym == 0 || nsym > 255) valid = 0;
    uint32_t previous = 0, assigned = 0;
    for (uint32_t i = 0; valid && i < nsym; ++i) {
      uint32_t delta = 0;
      if (!read_varint(body, body_length, &cursor, &delta) || (i != 0 && delta == 0) ||
          previous + delta > max_folded) {
        valid = 0;
        break;
      }
      const uint32_t value = previous + delta;
      uint32_t fr = (1u << kRansTableLog) - assigned;
      if (i + 1 != nsym && !read_varint(body, body_length, &cursor, &fr)) {
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
          (i + 1 != nsym && assigned + fr == (1u << kRansTableLog))) {
        valid = 0;
        break;
      }
      values[i] = (uint16_t)value;
      freq[i] = (uint16_t)fr;
      cum[i] = (uint16_t)assigned;
      assigned += fr;
      previous = value;
    }
    if (assigned != (1u << kRansTableLog)) valid = 0;
    int32_t total_payload = 0;
    int32_t len[4] = {0, 0, 0, 0};
    for (int l = 0; valid && l < 3; ++l) {
      if (cursor + 4 > body_length) {
        valid = 0;
      } else {
        len[l] = (int32_t)rd_u32(body + cursor);
        total_payload = (int32_t)((uint32_t)total_payload + (uint32_t)len[l]);
        cursor += 4;
      }
    }
    if (valid && cursor + 16 > body_length) valid = 0;
    for (int l = 0; valid && l < 4; ++l) {
      lane_state[l] = rd_u32(body + cursor + l * 4);
      if (lane_state[l] < kRansL) valid = 0;
    }
    if (valid) {
      cursor += 16;
      len[3] = (int32_t)((uint32_t)(int32_t)(body_length - cursor) - (uint32_t)total_payload);
      if (len[3] < 0) valid = 0;
      for (int l = 0; l < 3; ++l) {
        if (len[l] < 0) valid = 0;  // the reference fails such a lane later
      }
      int64_t c = cursor;
      for (int l = 0; l < 4; ++l) {
        meta[2 + l] = (int)c;
        meta[6 + l] = len[l];
        c += len[l];
      }
      if (c != body_length) valid = 0;
    }
    meta[0] = valid;
    meta[1] = (int)nsym;
    if (!valid) set_status(st, 5);
  }
  __builtin_amdgcn_wave_barrier();
  if (!meta[0]) return;
  const int nsym = meta[1];
  for (int sym = 0; sym < nsym; ++sym) {
    const int b = cum[sym];
    const int e = b + freq[sym];
    for (int slot = b + lane; slot < e; slot += 64) lut[slot] = (uint8_t)sym;
  }
  __builtin_amdgcn_wave_barrier();
  if (lane >= 4) return;
  uint32_t x = lane_state[lane];
  const int start = meta[2 + lane];
  const int length = meta[6 + lane];
  int read = 0;
  for (int64_t i = lane; i < n; i += 4) {
    const uint32_t slot = x & ((1u << kRansTableLog) - 1);
    const int sym = lut[slot];
    out[i] = (FT)values[sym];
    x = (uint32_t)freq[sym] * (x >> kRansTableLog) + slot - cum[sym];
    while (x < kRansL) {
      if (read >= length) {
        set_status(st, 5);
        return;
      }
      x = (x << 8) | body[start + read++];
    }
  }
  if (read != length || x != kRansL) set_status(st, 5);
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
    FVH_TIMED("k_dstate_init", stream);
    hipLaunchKernelGGL(k_dstate_init, fgrid, dim3(64), 0, stream, b.states, b.frame_count);
  }
  launch_plan(b.frames, b.frame_count, b.tiles, b.shards, b.tile_total, stream);
  {
    FVH_TIMED("k_parse", stream);
    hipLaunchKernelGGL(k_parse, dim3((b.tile_total + 255) / 256), dim3(256), 0, stream, b.frames,
                       b.states, b.tiles, b.tile_total, b.dshards);
  }
  {
    FVH_TIMED("k_entropy", stream);
    hipLaunchKernelGGL((k_entropy<FT>), dim3((b.shard_total + 3) / 4), dim3(256), 0, stream,
                       b.frames, b.states, b.tiles, b.shards, b.dshards, b.shard_total,
                       (FT*)b.folded);
  }
  if (b.wavefront) {
    launch_reconstruct<FT>(b, stream);
  } else {
    {
      FVH_TIMED("k_serial", stream);
      hipLaunchKernelGGL((k_serial<FT>), dim3((b.tile_total + 255) / 256), dim3(256), 0, stream,
                         b.frames, b.states, b.tiles, b.tile_total, (const FT*)b.folded);
    }
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
