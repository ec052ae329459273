// fvh_kernels.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include "fvh_common.h"

namespace fvh {

constexpr int kRingMax = 4096;     // widest tile handled by the wave-per-tile wavefront
constexpr int kBandRingMax = 512;  // widest tile handled by the band wavefront
constexpr int kLaneSinkBytes = 32; // per-lane sink slot for masked-off stores (8 x uint32)
constexpr int kSinkBytes = 256 * kLaneSinkBytes;  // workspace sink region
constexpr int kQueueCount = 8;     // work counters per call (two per predictor pass)
constexpr int kEmitQueue = 7;      // encoder: length of the emit list

struct EncodeBuffers {
  FrameDesc* frames;
  FrameState* states;
  TileDesc* tiles;
  ShardDesc* shards;
  ShardRec* recs;
  int64_t* tile_len;
  int64_t* tile_off;
  void* folded;          // uint16 or uint32 residuals, tile order
  uint8_t* scratch;      // kScratchBytes per shard
  uint8_t* ranks;        // kRankGroupBytes per group of kRansShards shards
  uint16_t* recon;       // tile-order reconstruction for tiles wider than kRingMax (or null)
  uint8_t* sink;         // kSinkBytes target of masked-off stores
  int32_t* queues;       // kQueueCount work counters, zeroed by k_plan
  int32_t* group_first;  // first tile of each band tile group (written by k_plan)
  int32_t* emit_list;    // shards whose records are not order-0 (final k_tile_scan),
                         // counted in queues[kEmitQueue]
  int group_total;
  int band_ring;         // ring slots per tile of the band wavefront (widest band tile)
  Results* results;
  int frame_count;
  int tile_total;
  int shard_total;
  int wide_folded;       // residuals stored as uint32 (16-bit frames present)
  int narrow_folded;     // residuals stored as uint8 (every one fits a byte)
  int any_heuristic;     // gray10 / gray16 lossy frames present
  int any_gray16;
  int any_lossless;
  int any_lossy;
  int any_wide;
  int any_band;          // some frame uses the band wavefront
  int any_nonband;       // some frame needs k_wavefront
  int any_med;           // some frame may predict with MED
  int any_grad;          // some frame may predict with the clamped gradient
};

struct DecodeBuffers {
  FrameDesc* frames;
  FrameState* states;
  TileDesc* tiles;
  ShardDesc* shards;
  DShard* dshards;
  int32_t* lists;        // [0, n): order-0 shards, [n, 2n): zero-run / block-pack,
                         // [(2+t) n, (3+t) n): order-0 list positions of alphabet
                         // tier t, [5n, 6n): Rice shards (n = shard_total)
  int32_t* list_count;   // [0] order-0, [1] zero-run / block-pack, [2 + t] tier t, [5] Rice
  O0Tab* o0tabs;         // parsed order-0 tables, by order-0 list position
  uint8_t* sink;         // kSinkBytes target of masked-off stores
  int32_t* queues;       // kQueueCount work counters, zeroed by k_plan
  int32_t* group_first;
  int group_total;
  int band_ring;
  const int64_t* fed_sizes;  // device stream sizes (headers then parsed on the device), or null
  void* folded;
  Results* results;
  int frame_count;
  int tile_total;
  int shard_total;
  int wide_folded;
  int any_wide;
  int wavefront;
  int any_band;
  int any_nonband;
  int any_med;
  int any_grad;
  int any_deep;  // a frame of more than 8 bits
};

// Compute units of the current device (cached per device).
int device_cus();
// 64-thread blocks of `kernel` one compute unit holds at once: a persistent
// kernel's grid beyond this times device_cus() only adds waves that start
// after others have ended, which for an empty work list is pure dispatch.
int blocks_per_cu(const void* kernel, int dyn_lds);
// LDS allocation unit of a gfx950 compute unit (160 KB = 128 of them)
constexpr size_t kLdsGranule = 1280;
constexpr size_t kLdsGranules = 128;

// Bytes of the library-owned encoder workspace of fvh_encode_launch_static
// (one per device, held until the process ends): one 3840x2160 frame of any
// layout and depth at the default tiles, or a batch of smaller frames.
constexpr uint64_t kStaticWorkspaceBytes = 256ull << 20;
// Device address of that workspace on `stream`'s device (looked up once per
// device with hipGetSymbolAddress, then cached).
hipError_t static_workspace(hipStream_t stream, void** out);

// Persistent waves per compute unit of the band wavefront (k_band), the
// predictor's (the launch-geometry knob the autotuner sets beside the tile
// size) and the reconstruction's.  set > 0 changes it for later launches;
// returns the current value.  The reconstruction is faster with fewer
// streams of tile rows in flight (16 4K RGB frames: 325 us at 4 waves, 339
// at 8, 341 at 5, 349 at 6), the predictor with more (268 us at 8).
constexpr int kBandWavesPerCuDefault = 8;
constexpr int kBandDecodeWavesPerCuDefault = 4;
int band_waves_per_cu(int set);
int band_decode_waves_per_cu(int set);

// State a call starts from, set by the planning kernel's threads beside the
// tile geometry (one launch fewer than a kernel of its own): the work
// counters and list counts zeroed, each frame's state reset (the encoder's
// with its first pass's step and predictor).
struct PlanInit {
  int kind;             // 0 nothing, 1 encoder states, 2 decoder states
  FrameState* states;
  int32_t* queues;      // queue_count work counters
  int queue_count;
  int32_t* list_count;  // list_counts counters (decoder)
  int list_counts;
};
void launch_plan(const FrameDesc* frames, int frame_count, TileDesc* tiles, ShardDesc* shards,
                 int32_t* group_first, int tile_total, hipStream_t stream,
                 const PlanInit& init = PlanInit{});
void launch_upload_frames(const FrameDesc* host_frames, int frame_count, FrameDesc* device_frames,
                          hipStream_t stream);
void launch_encode(const EncodeBuffers& b, hipStream_t stream);
void launch_decode(const DecodeBuffers& b, hipStream_t stream);
template <typename FT>
void launch_reconstruct(const DecodeBuffers& b, hipStream_t stream);
template <bool kEncode, typename FT>
void launch_band(const FrameDesc* frames, FrameState* states, const TileDesc* tiles,
                 const int32_t* group_first, int group_total, int ring_groups, FT* folded,
                 int32_t* queues, bool med, bool plain, hipStream_t stream);

}  // namespace fvh
