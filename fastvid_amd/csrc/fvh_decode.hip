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

__global__ void k_dstate_init(FrameState* __restrict__ states, int frame_count,
                              int32_t* __restrict__ list_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    list_count[0] = 0;
    list_count[1] = 0;
  }
  if (i >= frame_count) return;
  FrameState s = {};
  s.active = 1;
  states[i] = s;
}

// Directory entry and shard-header walk of one tile (decode.cu:139-219).
__global__ __launch_bounds__(256) void k_parse(const FrameDesc* __restrict__ frames,
                                               FrameState* __restrict__ states,
                                               const TileDesc* __restrict__ tiles, int tile_total,
                                               DShard* __restrict__ dshards,
                                               int32_t* __restrict__ lists,
                                               int32_t* __restrict__ list_count, int shard_total,
                                               char* gbase) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tile_total) return;
  const TileDesc td = tiles[t];
  const FrameDesc& f = frames[td.frame];
  FrameState* st = &states[td.frame];
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
  int n_order0 = 0;
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
    n_order0 += mode == kModeOrder0;
    decoded += n;
    cursor = body_offset + body_length;
  }
  if (decoded != samples || cursor != tile_end) {
    set_status(st, kStatusShardHeader);
    return;
  }
  // file the tile's shards by decoder kernel
  int at0 = n_order0 ? atomicAdd(&list_count[0], n_order0) : 0;
  int at1 = td.shard_count - n_order0 ? atomicAdd(&list_count[1], td.shard_count - n_order0) : 0;
  for (int ls = 0; ls < td.shard_count; ++ls) {
    const int shard = td.first_shard + ls;
    if (dshards[shard].mode == kModeOrder0) lists[at0++] = shard;
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

// Zero-run, Rice and block-pack shards (decode.cu:222-399): one wave per
// shard, shards taken from the "other" list by a grid-stride loop.  The body
// is first copied to LDS so the serial token walks and block-offset walks read
// LDS instead of issuing dependent global loads.
template <typename FT>
__global__ __launch_bounds__(64) void k_dec_other(const FrameDesc* __restrict__ frames,
                                                  FrameState* __restrict__ states,
                                                  const DShard* __restrict__ dshards,
                                                  const int32_t* __restrict__ lists,
                                                  const int32_t* __restrict__ list_count,
                                                  int shard_total, FT* __restrict__ folded,
                                                  char* gbase) {
  __shared__ __attribute__((aligned(16))) uint8_t body_lds[65536 / 8 + 768 + 16];
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
    const bool staged = body_length + head <= (int64_t)sizeof(body_lds);
    if (staged) {
      const uint4* chunks = (const uint4*)(gbody - head);
      const int nchunks = body_length > 0 ? (int)((head + body_length + 15) >> 4) : 0;
      for (int k = lane; k < nchunks; k += 64) ((uint4*)body_lds)[k] = chunks[k];
    }
    __syncthreads();
    if (staged) {
      dec_other_body(body_lds + head, body_length, n, out, mode, max_folded, st, lane);
    } else {
      dec_other_body(gbody, body_length, n, out, mode, max_folded, st, lane);
    }
    __syncthreads();
  }
}

// Order-0 rANS shards (decode.cu:441-549): four shards per wave, 16 lanes per
// shard.  Per shard: the table bytes are staged in LDS and parsed by one lane
// with the reference's checks; the 16 lanes fill the 4096-slot rank LUT; four
// lanes run the rANS lanes, each with an 8-byte lookahead of its byte stream
// so the state recurrence only waits on two LDS lookups per symbol.
constexpr int kO0Groups = 4;
constexpr int kO0Stage = 1312;  // >= any valid table + lane lengths + states

struct O0Meta {
  int valid;
  int n;
  int start[4];
  int length[4];
  uint32_t state[4];
};

template <typename FT>
__global__ __launch_bounds__(64) void k_dec_o0(const FrameDesc* __restrict__ frames,
                                               FrameState* __restrict__ states,
                                               const DShard* __restrict__ dshards,
                                               const int32_t* __restrict__ lists,
                                               const int32_t* __restrict__ list_count,
                                               FT* __restrict__ folded, char* gbase) {
  __shared__ __attribute__((aligned(16))) uint8_t lut_all[kO0Groups][4096];
  __shared__ uint32_t ent_all[kO0Groups][256];
  __shared__ uint16_t val_all[kO0Groups][256];
  __shared__ O0Meta meta_all[kO0Groups];
  const int lane = lane_id();
  const int g = lane >> 4;
  const int i = lane & 15;
  const int total = list_count[0];
  const int item = blockIdx.x * kO0Groups + g;
  const bool has = item < total;
  uint8_t* lut = lut_all[g];
  uint32_t* ent = ent_all[g];
  uint16_t* val = val_all[g];
  O0Meta& meta = meta_all[g];
  DShard d = {};
  FrameState* st = nullptr;
  const FrameDesc* f = nullptr;
  bool go = false;
  if (has) {
    d = dshards[lists[item]];
    st = &states[d.frame];
    f = &frames[d.frame];
    go = st->status == 0;
  }
  const uint8_t* body =
      as_global(gbase, go ? (const uint8_t*)f->stream + d.body_offset : (const uint8_t*)gbase);
  const int64_t body_length = go ? d.body_length : 0;
  // stage the table bytes (aliasing the LUT) as the aligned 16-byte chunks
  // that hold them; such a chunk never crosses a page
  const int head = (int)((uintptr_t)body & 15);
  const int nstage = (int)min(body_length, (int64_t)kO0Stage);
  const int nchunks = nstage > 0 ? (head + nstage + 15) >> 4 : 0;
  const uint4* chunks = (const uint4*)(body - head);
  for (int k = i; k < nchunks; k += 16) ((uint4*)lut)[k] = chunks[k];
  __syncthreads();
  if (go && i == 0) {
    const uint32_t max_folded = (uint32_t)f->max_sample * 2u;
    const uint8_t* b = lut + head;
    // reads past the stage only happen for tables no valid stream can have
    const int64_t limit = min(body_length, (int64_t)kO0Stage);
    int valid = 1;
    int64_t cursor = 0;
    if (body_length < 1 || b[cursor++] != kRansTableLog) valid = 0;
    uint32_t nsym = 0;
    if (valid && !read_varint(b, limit, &cursor, &nsym)) valid = 0;
    if (nsym == 0 || nsym > 255) valid = 0;
    uint32_t previous = 0, assigned = 0;
    for (uint32_t k = 0; valid && k < nsym; ++k) {
      uint32_t delta = 0;
      if (!read_varint(b, limit, &cursor, &delta) || (k != 0 && delta == 0) ||
          previous + delta > max_folded) {
        valid = 0;
        break;
      }
      const uint32_t value = previous + delta;
      uint32_t fr = (1u << kRansTableLog) - assigned;
      if (k + 1 != nsym && !read_varint(b, limit, &cursor, &fr)) {
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
      val[k] = (uint16_t)value;
      ent[k] = (fr & 0xffffu) | ((uint32_t)assigned << 16);
      assigned += fr;
      previous = value;
    }
    if (assigned != (1u << kRansTableLog)) valid = 0;
    int32_t total_payload = 0;
    int* len = meta.length;  // LDS: indexed by a loop counter
    for (int l = 0; l < 4; ++l) len[l] = 0;
    for (int l = 0; valid && l < 3; ++l) {
      if (cursor + 4 > limit) {
        valid = 0;
      } else {
        len[l] = (int32_t)rd_u32(b + cursor);
        total_payload = (int32_t)((uint32_t)total_payload + (uint32_t)len[l]);
        cursor += 4;
      }
    }
    if (valid && cursor + 16 > limit) valid = 0;
    for (int l = 0; valid && l < 4; ++l) {
      meta.state[l] = rd_u32(b + cursor + l * 4);
      if (meta.state[l] < kRansL) valid = 0;
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
        meta.start[l] = (int)c;
        c += len[l];
      }
      if (c != body_length) valid = 0;
    }
    meta.valid = valid;
    meta.n = (int)nsym;
    if (!valid) set_status(st, 5);
  }
  __syncthreads();
  const bool ok = go && meta.valid;
  // rank LUT: lane i fills slots [256 i, 256 i + 256), four per store
  if (ok) {
    const int nsym = meta.n;
    const int lo_slot = i * 256;
    int rk = 0;
    int next_cum = nsym > 1 ? (int)(ent[1] >> 16) : (1 << kRansTableLog);
    while (next_cum <= lo_slot) {
      ++rk;
      next_cum = rk + 1 < nsym ? (int)(ent[rk + 1] >> 16) : (1 << kRansTableLog);
    }
    uint32_t* lut_w = (uint32_t*)lut + i * 64;
    for (int w = 0; w < 64; ++w) {
      uint32_t word = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int slot = lo_slot + w * 4 + q;
        while (next_cum <= slot) {
          ++rk;
          next_cum = rk + 1 < nsym ? (int)(ent[rk + 1] >> 16) : (1 << kRansTableLog);
        }
        word |= (uint32_t)rk << (8 * q);
      }
      lut_w[w] = word;
    }
  }
  __syncthreads();
  if (!ok || i >= 4) return;
  // rANS lane i.  Its bytes are read as aligned 16-byte chunks: `cur` is the
  // chunk being consumed (next byte at `pos`), `nxt` the following chunk,
  // requested as soon as `cur` becomes current.  Only chunks holding at least
  // one byte of the lane are read.  The window `bb` holds `avail` bytes, most
  // significant first, and is topped up whenever it falls to four bytes, so
  // the (at most two) renormalisation reads of a step never wait on memory.
  const int r = i;
  const uint8_t* src = body + meta.start[r];
  const int length = meta.length[r];
  uint32_t x = meta.state[r];
  FT* out = folded + d.folded_offset;
  const int n = d.count;
  int pos = (int)((uintptr_t)src & 15);
  const uint4* cp = (const uint4*)(src - pos);
  int left = length;  // lane bytes not yet moved into the window
  int c = 0;
  uint4 cur = make_uint4(0, 0, 0, 0);
  uint4 nxt = cur;
  if (left > 0) cur = cp[0];
  if (left > 16 - pos) nxt = cp[1];
  uint64_t bb = 0;
  int avail = 0;
  int bad = 0;
  for (int idx = r; idx < n; idx += 4) {
    while (avail <= 4 && left > 0) {
      // dword pos/4 of the chunk, by shifts
      const uint64_t half = (pos & 8) ? (((uint64_t)cur.w << 32) | cur.z)
                                      : (((uint64_t)cur.y << 32) | cur.x);
      const uint32_t w = (uint32_t)(half >> ((pos & 4) << 3));
      const int skip = pos & 3;
      const int take = min(4 - skip, left);
      // byte `pos` to the top, later bytes below it
      const uint32_t msb_first = __builtin_bswap32(w) << (8 * skip);
      bb |= ((uint64_t)msb_first << 32) >> (8 * avail);
      avail += take;
      left -= take;
      pos += take;
      if (pos == 16) {
        ++c;
        pos = 0;
        cur = nxt;
        if (left > 16) nxt = cp[c + 1];
      }
    }
    const uint32_t slot = x & ((1u << kRansTableLog) - 1);
    const int sym = lut[slot];
    const uint32_t e = ent[sym];
    out[idx] = (FT)val[sym];
    x = __umul24(e & 0xffffu, x >> kRansTableLog) + slot - (e >> 16);
    while (x < kRansL) {
      if (avail == 0) {
        bad = 1;
        break;
      }
      x = (x << 8) | (uint32_t)(bb >> 56);
      bb <<= 8;
      --avail;
    }
    if (bad) break;
  }
  if (bad || left != 0 || avail != 0 || x != kRansL) set_status(st, 5);
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
    hipLaunchKernelGGL(k_dstate_init, fgrid, dim3(64), 0, stream, b.states, b.frame_count,
                       b.list_count);
  }
  launch_plan(b.frames, b.frame_count, b.tiles, b.shards, b.tile_total, stream);
  {
    FVH_TIMED("k_parse", stream);
    hipLaunchKernelGGL(k_parse, dim3((b.tile_total + 255) / 256), dim3(256), 0, stream, b.frames,
                       b.states, b.tiles, b.tile_total, b.dshards, b.lists, b.list_count,
                       b.shard_total, (char*)b.folded);
  }
  {
    FVH_TIMED("k_dec_o0", stream);
    hipLaunchKernelGGL((k_dec_o0<FT>), dim3((b.shard_total + kO0Groups - 1) / kO0Groups), dim3(64),
                       0, stream, b.frames, b.states, b.dshards, b.lists, b.list_count,
                       (FT*)b.folded, (char*)b.folded);
  }
  {
    FVH_TIMED("k_dec_other", stream);
    const int grid = min(b.shard_total, 4096);
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
