// fvh_common.h -- shared host/device definitions of the MI355X fastvid codec.
//
// Format constants follow the version-1 bitstream (cuda/csrc/encode.cu:14-28,
// cuda/csrc/decode.cu:12-21).  Batch descriptors are this library's own
// design: a batch is a flat list of frames; tiles and shards of all frames are
// numbered globally so one launch covers the whole batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvh {

constexpr int kHeaderBytes = 32;
constexpr int kDirectoryEntryBytes = 32;
constexpr int kShardSymbols = 4096;
constexpr int kBlockSymbols = 128;
constexpr int kMaxRice = 16;
constexpr int kModeZeroRun = 0;
constexpr int kModeRiceBase = 1;
constexpr int kModeBlockPack = 18;
constexpr int kModeOrder0 = 19;
constexpr int kEntropyParallelShards = 19;
constexpr int kPredictorId = 6;
constexpr int kRansAlphabet = 511;
constexpr int kRansTableLog = 12;
constexpr uint32_t kRansL = 1u << 23;

// Worst-case record size of one full shard: block-pack bound
// 32 * (1 + 128*17/8) = 8736 body bytes (>= order-0 bound 1024+28+4*1536)
// plus the 3-byte record header.
constexpr int kMaxRecordBytes = 8739;

// Per-shard order-0 scratch layout (bytes).  Once the rANS encoder has read
// the entries, it writes the order-0 record's head around the sparse table:
// [mode][body u16] at kScratchHead, and the lane lengths 0..2 and final
// states (7 x u32) over spent entries at kScratchLensStates (not at 0, so a
// dword-aligned 16-byte read that starts up to 15 bytes before them stays in
// the shard).
constexpr int kScratchTable = 0;        // 255 x {freq | cum << 13, 0} by rank
constexpr int kScratchLensStates = 20;  // 28 bytes of the order-0 record
constexpr int kScratchHead = 2045;      // 3 record header bytes, just before the sparse table
constexpr int kScratchSparse = 2048;    // sparse table bytes (<= 1024)
constexpr int kScratchLanes = 3072;     // 4 lanes x kLaneBytes, interleaved by 8-byte slot
constexpr int kLaneBytes = 1536;        // proven bound for 1024 symbols
// The lane area is kLaneSlots slots of 32 bytes, lane r at byte 8 r of each.
// A lane's n bytes, in stream order, are bytes 1536 - n .. 1535 of its
// virtual run, virtual byte v at slot v / 8, byte v % 8: the rANS encoder
// fills slots from the last one down, eight bytes at a time (the final
// partial word at the top of the next slot), and the four lanes of a shard,
// which advance nearly in step, fill each slot's 32 bytes together.
constexpr int kLaneSlots = kLaneBytes / 8;  // 192
constexpr int kScratchBytes = kScratchLanes + 4 * kLaneBytes;  // 9216

// Order-0 rank bytes, by group of kRansShards consecutive shards.  rANS lane
// r of shard s is lane L = 4 (s % kRansShards) + r of its group; its 1024
// rank bytes (byte p: the rank of symbol 4 (1023 - p) + r, the order the
// encoder visits them) are 64 chunks of 16, chunk q at byte (64 q + L) * 16
// of the group's area, so one chunk index of all 64 lanes is one contiguous
// kilobyte.
constexpr int kRansShards = 16;
constexpr int kRankGroupBytes = kRansShards * kShardSymbols;  // 65536

// Frame-level status codes (first error wins).
constexpr int kStatusSampleRange = 1;     // encoder: sample > max (encode.cu:160-162)
constexpr int kStatusBodyTooLarge = 2;    // encoder: body > u16 (encode.cu:305-307)
constexpr int kStatusTileTooLarge = 100;  // encoder: tile payload > u32 (encode.cu:898-899)
constexpr int kStatusCapacity = 101;      // encoder: internal output bound violated
constexpr int kStatusLaneOverflow = 102;  // encoder: rANS lane exceeded its bound
constexpr int kStatusDirectory = 10;      // decoder: directory (decode.cu:131-186)
constexpr int kStatusShardHeader = 11;    // decoder: shard walk (decode.cu:188-219)
constexpr int kStatusHeader = 12;         // decoder, sizes on the device: a stream header that
                                          // does not match the geometry the batch was planned for

// quantization_step and refined_gray16_step (encode.cu:71-89), for host and
// device: the device-fed decoder derives the step from each stream's header.
__host__ __device__ inline int32_t quantization_step(int64_t layout, int64_t bit_depth,
                                                     int64_t quality) {
  if ((layout == 0 && bit_depth == 8) || (layout == 1 && bit_depth == 8)) return 1;
  int32_t denominator = 6;
  if (layout == 1 && bit_depth == 10) denominator = 20;
  else if (layout == 1 && bit_depth == 16) denominator = 24;
  else if (layout == 2 && bit_depth == 10) denominator = 10;
  const int32_t scale = int32_t{1} << (bit_depth - 8);
  return 1 + (int32_t)(((100 - quality) * scale + denominator - 1) / denominator);
}

__host__ __device__ inline int32_t refined_gray16_step(int64_t quality) {
  return 1 + (int32_t)(((100 - quality) * 256 + 7) / 8);
}

struct FrameDesc {
  const uint16_t* plane[3];    // encode input planes (device)
  uint16_t* out_plane[3];      // decode output planes (device, dense rows)
  int64_t pitch[3];            // row pitch of plane[] in samples
  int64_t plane_width[3];      // samples per row of each plane
  int64_t plane_sample_base[3];// first residual of each plane within the frame's residuals
  int64_t plane_shard_base[3]; // first shard of each plane within the frame
  int64_t row_shards[3];       // shards in one full-height row of tiles
  int64_t row_samples[3];      // residuals in one full-height row of tiles (every tile's
                               // residuals start on a multiple of 8 elements)
  uint8_t* stream;             // encode: output stream; decode: input stream
  int64_t stream_capacity;     // encode: output capacity; decode: stream size
  int64_t folded_base;         // first folded element of the frame in the batch
  int64_t total_samples;
  int64_t width, height;
  int32_t tile_w, tile_h;      // nominal luma tile size
  int32_t ntw[3];              // nominal tile width per plane
  int32_t cols[3];             // tile columns per plane
  int32_t rows;                // tile rows (shared by all planes)
  int32_t plane_count;
  int32_t plane_tile_begin[4]; // local tile index where each plane starts; [3] = tile_count
  int32_t tile_begin, tile_count;   // global tile range
  int32_t shard_begin, shard_count; // global shard range
  int32_t layout, bit_depth, quality, max_sample;
  uint32_t fps_num, fps_den;
  int32_t step0;               // quantization_step (encode.cu:71-85)
  int32_t step_refined;        // refined_gray16_step (encode.cu:87-89)
  int32_t force_order0;        // encode.cu:778
  int32_t heuristic;           // 0 none, 1 gray10 MED candidate, 2 gray16 refinement candidate
  int32_t use_med;             // decode: predictor choice
  int32_t step;                // decode: step from the header
  int32_t header_flags;        // decode: header byte 7 flag bits
  int32_t band;                // tiles eligible for the band wavefront (k_band)
  int32_t group_begin[4];      // global index of each plane's first band tile group (groups
                               // of 4 consecutive tiles of one plane); [3] = end of frame
};

// Per-frame state written on the device during a call.
struct FrameState {
  int32_t status;          // first error code (0 = ok)
  int32_t step;            // active quantiser step of the current pass
  int32_t use_med;         // active predictor of the current pass
  int32_t active;          // current pass runs for this frame
  int64_t payload;         // sum of record bytes of the last pass
  int64_t baseline_payload;
  int64_t stream_length;
  int32_t gray_med;
  int32_t refined;
  int32_t baseline_step;
  int32_t direct;
  int32_t stage;           // heuristic stage machine
  int32_t pad;
};

struct TileDesc {
  int64_t folded_base;     // batch-global folded offset
  int64_t src_offset;      // y*pitch + x within its plane
  int64_t dst_offset;      // y*plane_width + x within its dense plane
  int32_t frame, plane;
  int32_t x, y, width, height;
  int32_t first_shard, shard_count;   // global shard range
};

struct ShardDesc {
  int32_t tile;            // global tile index
  int32_t local;           // shard index within the tile
  int32_t frame;           // batch frame index (the tile's)
};

// Encoder per-shard analysis record.
struct ShardRec {
  uint8_t mode;            // chosen mode
  uint8_t k;               // Rice parameter of the best-of-three analysis
  uint8_t supported;       // order-0 feasible
  uint8_t best3_mode;
  uint16_t sparse_len;     // order-0 sparse table bytes
  uint16_t distinct;
  uint32_t body;           // chosen body bytes
  uint32_t best3_body;     // best of zero-run / Rice / block-pack
  uint16_t lane[4];        // chosen lane bytes (Rice or order-0)
  uint16_t best3_lane[4];
  uint32_t offset;         // record offset within its tile payload
  uint32_t states[4];      // order-0 final states
};

// Decoder per-shard metadata (decode.cu:209-213).
struct DShard {
  int64_t body_offset;     // byte offset of the body within the stream
  int64_t folded_offset;   // batch-global folded offset
  int32_t frame;
  int32_t mode;
  int32_t body_length;
  int32_t count;
};

// Parsed order-0 table of one shard (k_o0_parse -> k_dec_o0), indexed by the
// shard's position in the order-0 list.  ent[k] = cum_k | value_k << 13 for
// the nsym ranks in table order and ent[nsym] = 4096 (the end of the slot
// range), so a rank's frequency is the difference of two adjacent entries.
struct O0Tab {
  int32_t nsym;            // 0: the shard's table was rejected or not decoded
  int32_t pad[3];
  int32_t start[4];        // rANS lane byte offsets within the body
  int32_t length[4];
  uint32_t state[4];
  uint32_t sink[4];        // per-lane target of the decoder's masked-off stores
  uint32_t ent[256];
};
static_assert(sizeof(O0Tab) == 1104, "O0Tab layout");

struct Results {           // per frame, read back by the host
  int64_t stream_length;
  int32_t status;
  int32_t flags;           // bit0 gray_med, bit1 refined
};

}  // namespace fvh
