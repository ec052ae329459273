// fvh_api.hip -- extern "C" entry points of libfastvid_hip.so (host side).
//
// Argument validation mirrors fastvid_encode_cuda (cuda/csrc/encode.cu:695-735)
// and the header/directory checks of decode (cuda/csrc/fastvid_cuda.cpp:141-333),
// with the same messages.  Geometry and buffer sizing are computed here; all
// per-tile and per-shard work happens in the kernels.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/fastvid_hip.h"
#include "fvh_kernels.h"

namespace fvh {
namespace {

void set_msg(char* msg, size_t len, const char* text) {
  if (msg && len) std::snprintf(msg, len, "%s", text);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Worst-case bytes of one shard record of n symbols: the larger of the
// block-pack bound (which also bounds the zero-run and Rice bodies whenever
// they are chosen) and the order-0 bound (table + 28 + four rANS lanes of at
// most 12 bits per symbol plus one byte).
int64_t shard_bound(int64_t n) {
  const int64_t block = ceil_div(n, kBlockSymbols) + ceil_div(17 * n, 8);
  const int64_t lane_symbols = ceil_div(n, 4);
  const int64_t order0 = 31 + 4 * std::min<int64_t>(n, 255) + 4 * (ceil_div(3 * lane_symbols, 2) + 1);
  return 3 + std::max(block, order0);
}

int64_t tile_bound(int64_t w, int64_t h) {
  const int64_t samples = w * h;
  const int64_t full = samples / kShardSymbols;
  const int64_t tail = samples - full * kShardSymbols;
  return full * shard_bound(kShardSymbols) + (tail ? shard_bound(tail) : 0);
}

int64_t align256(int64_t v) { return (v + 255) & ~int64_t{255}; }

// Fills the geometry part of a FrameDesc (expected_tiles, encode.cu:41-69,
// in closed form) and returns tile/shard/sample counts and the stream bound.
struct Geometry {
  int64_t tiles, shards, samples, bound;
  bool wide;
};

Geometry fill_geometry(FrameDesc& f, int64_t width, int64_t height, int64_t tw, int64_t th,
                       int64_t layout) {
  Geometry g = {0, 0, 0, 0, false};
  f.width = width;
  f.height = height;
  f.tile_w = (int32_t)tw;
  f.tile_h = (int32_t)th;
  f.layout = (int32_t)layout;
  f.plane_count = layout == 0 ? 1 : 3;
  f.rows = (int32_t)ceil_div(height, th);
  const int64_t rows = f.rows;
  const int64_t last_h = height - (rows - 1) * th;
  int64_t tile_acc = 0, shard_acc = 0, sample_acc = 0, folded_acc = 0, bound = kHeaderBytes;
  auto pad8 = [](int64_t v) { return (v + 7) & ~int64_t{7}; };
  for (int p = 0; p < 3; ++p) {
    f.plane_tile_begin[p] = (int32_t)tile_acc;
    if (p >= f.plane_count) {
      f.plane_width[p] = 0;
      f.plane_sample_base[p] = folded_acc;
      f.plane_shard_base[p] = shard_acc;
      f.row_shards[p] = 0;
      f.row_samples[p] = 0;
      f.ntw[p] = 1;
      f.cols[p] = 1;
      continue;
    }
    const bool sub = layout == 1 && p != 0;
    const int64_t pw = sub ? (width + 1) / 2 : width;
    const int64_t ntw = sub ? (tw + 1) / 2 : tw;
    const int64_t cols = ceil_div(pw, ntw);
    const int64_t last_w = pw - (cols - 1) * ntw;
    f.plane_width[p] = pw;
    f.ntw[p] = (int32_t)ntw;
    f.cols[p] = (int32_t)cols;
    f.plane_sample_base[p] = folded_acc;
    f.plane_shard_base[p] = shard_acc;
    // residuals: tile-raster order, each tile padded to a multiple of 8
    // elements so every tile and shard starts 16-byte aligned
    const int64_t full_row_res = (cols - 1) * pad8(ntw * th) + pad8(last_w * th);
    const int64_t last_row_res = (cols - 1) * pad8(ntw * last_h) + pad8(last_w * last_h);
    f.row_samples[p] = full_row_res;
    folded_acc += (rows - 1) * full_row_res + last_row_res;
    const int64_t full_row = (cols - 1) * ceil_div(ntw * th, kShardSymbols) +
        ceil_div(last_w * th, kShardSymbols);
    const int64_t last_row = (cols - 1) * ceil_div(ntw * last_h, kShardSymbols) +
        ceil_div(last_w * last_h, kShardSymbols);
    f.row_shards[p] = full_row;
    tile_acc += rows * cols;
    shard_acc += (rows - 1) * full_row + last_row;
    sample_acc += pw * height;
    bound += kDirectoryEntryBytes * rows * cols;
    bound += (rows - 1) * ((cols - 1) * tile_bound(ntw, th) + tile_bound(last_w, th));
    bound += (cols - 1) * tile_bound(ntw, last_h) + tile_bound(last_w, last_h);
    if (std::min(ntw, pw) > kRingMax) g.wide = true;
  }
  f.plane_tile_begin[3] = (int32_t)tile_acc;
  for (int p = f.plane_count; p < 3; ++p) f.plane_tile_begin[p] = (int32_t)tile_acc;
  f.total_samples = sample_acc;
  g.tiles = tile_acc;
  g.shards = shard_acc;
  g.samples = folded_acc;
  g.bound = bound;
  return g;
}

struct WorkspaceLayout {
  int64_t frames, states, tiles, shards, recs, tile_len, tile_off, folded, scratch, ranks, recon,
      dshards, sink, queues, groups, emit_list;
  int64_t total;
};

int64_t tile_count_closed(int64_t width, int64_t height, int64_t tw, int64_t th, int64_t layout) {
  const int64_t rows = ceil_div(height, th);
  const int64_t luma = ceil_div(width, tw);
  if (layout == 0) return rows * luma;
  if (layout == 2) return 3 * rows * luma;
  return rows * (luma + 2 * ceil_div((width + 1) / 2, (tw + 1) / 2));
}

// Frames whose every tile fits the band wavefront's vector accesses: planes
// 16-byte aligned, pitches, plane widths and nominal tile widths multiples of
// 8 samples, tiles no wider than kBandRingMax, and every plane and every
// 16 rows of pitch addressable with 31-bit byte offsets (the kernel drops
// masked stores by pushing their offset past 2^31).
int band_eligible(const FrameDesc& f, const void* const* planes) {
  for (int p = 0; p < f.plane_count; ++p) {
    if (((uintptr_t)planes[p] & 15) != 0) return 0;
    if (f.pitch[p] % 8 != 0 || f.plane_width[p] % 8 != 0 || f.ntw[p] % 8 != 0) return 0;
    if (std::min<int64_t>(f.ntw[p], f.plane_width[p]) > kBandRingMax) return 0;
    const int64_t span = ((f.height - 1) * f.pitch[p] + f.plane_width[p]) * 2;
    if (span >= (int64_t{1} << 31) || 32 * f.pitch[p] >= (int64_t{1} << 31)) return 0;
  }
  return 1;
}

// Band tile groups (4 consecutive tiles of one plane) of one frame, numbered
// from `first`; returns the frame's group count.
int64_t number_groups(FrameDesc& f, int64_t first) {
  int64_t g = first;
  for (int p = 0; p < 3; ++p) {
    f.group_begin[p] = (int32_t)g;
    g += (f.plane_tile_begin[p + 1] - f.plane_tile_begin[p] + 3) / 4;
  }
  f.group_begin[3] = (int32_t)g;
  return g - first;
}

// Ring slots per tile the band wavefront needs for a frame: its band period
// in column groups, max(whole 64-sample lines, 16).
int band_ring_groups(const FrameDesc& f) {
  int64_t r = 0;
  for (int p = 0; p < f.plane_count; ++p) {
    const int64_t w = std::min<int64_t>(f.ntw[p], f.plane_width[p]);
    r = std::max<int64_t>(r, std::max<int64_t>(((w + 63) / 64) * 8, 16));
  }
  return (int)r;
}

struct EncodePlan {
  std::vector<FrameDesc> frames;
  std::vector<int64_t> output_offsets;
  int64_t output_bytes = 0;
  int64_t tiles = 0, shards = 0, samples = 0;
  bool wide_folded = false, any_heuristic = false, any_gray16 = false;
  bool narrow_folded = true;  // until a frame's residuals may exceed a byte
  bool any_lossless = false, any_lossy = false, any_wide = false;
  bool any_band = false, any_nonband = false, any_med = false, any_grad = false;
  int64_t groups = 0;
  int band_ring = 16;
  WorkspaceLayout ws = {};
};

int build_encode_plan(const fvh_encode_frame* in, int32_t count, EncodePlan& plan, char* msg,
                      size_t msg_len) {
  if (count <= 0) {
    set_msg(msg, msg_len, "no frames to encode");
    return FVH_EARG;
  }
  plan.frames.assign(count, FrameDesc{});
  plan.output_offsets.assign(count, 0);
  int64_t out = 0;
  for (int32_t i = 0; i < count; ++i) {
    const fvh_encode_frame& fr = in[i];
    // encode.cu:695-706
    if (!(fr.layout >= 0 && fr.layout <= 2)) {
      set_msg(msg, msg_len, "layout must be gray, yuv422, or rgb444");
      return FVH_EARG;
    }
    if (fr.plane_count != (fr.layout == 0 ? 1 : 3)) {
      set_msg(msg, msg_len, "plane count does not match layout");
      return FVH_EARG;
    }
    if (!(fr.bit_depth == 8 || fr.bit_depth == 10 || fr.bit_depth == 12 || fr.bit_depth == 16)) {
      set_msg(msg, msg_len, "bit_depth must be 8, 10, 12, or 16");
      return FVH_EARG;
    }
    if (!(fr.quality >= 1 && fr.quality <= 100)) {
      set_msg(msg, msg_len, "quality must be in 1..=100");
      return FVH_EARG;
    }
    if (!(fr.fps_numerator > 0 && fr.fps_numerator <= 0xffffffffLL && fr.fps_denominator > 0 &&
          fr.fps_denominator <= 0xffffffffLL)) {
      set_msg(msg, msg_len, "frame-rate components must be nonzero u32 values");
      return FVH_EARG;
    }
    if (!(fr.tile_width > 0 && fr.tile_width <= 65535 && fr.tile_height > 0 &&
          fr.tile_height <= 65535)) {
      set_msg(msg, msg_len, "tile dimensions must be nonzero u16 values");
      return FVH_EARG;
    }
    // encode.cu:711-725
    const int64_t height = fr.plane_heights[0];
    const int64_t width = fr.plane_widths[0];
    if (!(width > 0 && height > 0 && width <= 0xffffffffLL && height <= 0xffffffffLL)) {
      set_msg(msg, msg_len, "frame dimensions must be nonzero u32 values");
      return FVH_EARG;
    }
    const int64_t secondary = fr.layout == 1 ? (width + 1) / 2 : width;
    for (int64_t p = 0; p < fr.plane_count; ++p) {
      const int64_t expected = p == 0 ? width : secondary;
      if (fr.plane_heights[p] != height || fr.plane_widths[p] != expected) {
        set_msg(msg, msg_len, "plane dimensions do not match declared layout");
        return FVH_EARG;
      }
      if (fr.pitches[p] < fr.plane_widths[p] || fr.planes[p] == nullptr) {
        set_msg(msg, msg_len, "plane pointer or pitch is invalid");
        return FVH_EARG;
      }
    }
    if (tile_count_closed(width, height, fr.tile_width, fr.tile_height, fr.layout) > (1 << 20)) {
      set_msg(msg, msg_len, "too many tiles");  // encode.cu:735
      return FVH_EARG;
    }
    FrameDesc& f = plan.frames[i];
    const Geometry g = fill_geometry(f, width, height, fr.tile_width, fr.tile_height, fr.layout);
    for (int p = 0; p < 3; ++p) {
      f.plane[p] = p < fr.plane_count ? (const uint16_t*)fr.planes[p] : nullptr;
      f.pitch[p] = p < fr.plane_count ? fr.pitches[p] : 0;
      f.out_plane[p] = nullptr;
    }
    plan.samples = (plan.samples + 7) & ~int64_t{7};  // 16-byte aligned residual rows
    f.folded_base = plan.samples;
    f.band = band_eligible(f, fr.planes);
    plan.groups += number_groups(f, plan.groups);
    if (f.band) plan.band_ring = std::max(plan.band_ring, band_ring_groups(f));
    f.tile_begin = (int32_t)plan.tiles;
    f.tile_count = (int32_t)g.tiles;
    f.shard_begin = (int32_t)plan.shards;
    f.shard_count = (int32_t)g.shards;
    f.bit_depth = (int32_t)fr.bit_depth;
    f.quality = (int32_t)fr.quality;
    f.max_sample = (1 << fr.bit_depth) - 1;
    f.fps_num = (uint32_t)fr.fps_numerator;
    f.fps_den = (uint32_t)fr.fps_denominator;
    f.step0 = quantization_step(fr.layout, fr.bit_depth, fr.quality);
    f.step_refined = refined_gray16_step(fr.quality);
    f.force_order0 = fr.layout == 2 && (fr.bit_depth != 10 || width < 3840);  // encode.cu:778
    f.heuristic = 0;
    if (fr.layout == 0 && fr.bit_depth == 10 && f.step0 > 1) f.heuristic = 1;
    if (fr.layout == 0 && fr.bit_depth == 16 && f.step0 > 1) f.heuristic = 2;
    f.stream_capacity = g.bound;
    plan.output_offsets[i] = out;
    out += align256(g.bound);
    plan.tiles += g.tiles;
    plan.shards += g.shards;
    plan.samples += g.samples;
    plan.wide_folded |= fr.bit_depth == 16;
    // a lossy frame's quantised residuals lie in [-K, K], K = ceil(max / step)
    // (encode.cu:183-185), so zigzagged they are at most 2K; frames whose
    // step the heuristics may change keep 16 bits
    plan.narrow_folded &= f.heuristic == 0 && f.step0 > 1 &&
                          2 * ((f.max_sample + f.step0 - 1) / f.step0) <= 255;
    plan.any_heuristic |= f.heuristic != 0;
    plan.any_gray16 |= f.heuristic == 2;
    plan.any_lossless |= f.step0 == 1;
    plan.any_lossy |= f.step0 > 1;
    plan.any_wide |= g.wide;
    if (f.step0 > 1) {
      plan.any_band |= f.band != 0;
      plan.any_nonband |= f.band == 0;
    }
    plan.any_med |= fr.layout != 0 || f.heuristic == 1;
    plan.any_grad |= fr.layout == 0;
  }
  if (plan.tiles > INT32_MAX / 2 || plan.shards > INT32_MAX / 4) {
    set_msg(msg, msg_len, "batch is too large");
    return FVH_EARG;
  }
  plan.output_bytes = out;
  WorkspaceLayout& w = plan.ws;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = off;
    off = align256(off + bytes);
    return at;
  };
  w.frames = take((int64_t)count * sizeof(FrameDesc));
  w.states = take((int64_t)count * sizeof(FrameState));
  w.tiles = take(plan.tiles * (int64_t)sizeof(TileDesc));
  w.shards = take(plan.shards * (int64_t)sizeof(ShardDesc));
  w.recs = take(plan.shards * (int64_t)sizeof(ShardRec));
  w.tile_len = take(plan.tiles * 8);
  w.tile_off = take(plan.tiles * 8);
  w.folded = take((plan.samples + 8) * (plan.wide_folded ? 4 : 2));
  // (64 bytes of slack: order-0 record copies read a few bytes past a piece)
  w.scratch = take(plan.shards * (int64_t)kScratchBytes + 64);
  w.ranks = take((plan.shards + kRansShards - 1) / kRansShards * (int64_t)kRankGroupBytes);
  w.recon = plan.any_wide ? take(plan.samples * 2) : -1;
  w.dshards = -1;
  w.sink = take(kSinkBytes);
  w.queues = take(kQueueCount * 4);
  w.groups = take(plan.groups * 4);
  w.emit_list = take(plan.shards * 4);
  w.total = off;
  return FVH_OK;
}

}  // namespace
}  // namespace fvh

using namespace fvh;

extern "C" {

int fvh_encode_plan(const fvh_encode_frame* frames, int32_t frame_count, uint64_t* workspace_bytes,
                    uint64_t* output_bytes, uint64_t* output_offsets, char* msg, size_t msg_len) {
  EncodePlan plan;
  const int rc = build_encode_plan(frames, frame_count, plan, msg, msg_len);
  if (rc) return rc;
  if (workspace_bytes) *workspace_bytes = (uint64_t)plan.ws.total;
  if (output_bytes) *output_bytes = (uint64_t)plan.output_bytes;
  if (output_offsets) {
    for (int32_t i = 0; i < frame_count; ++i) output_offsets[i] = (uint64_t)plan.output_offsets[i];
  }
  return FVH_OK;
}

}  // extern "C"

namespace fvh {
namespace {

// Enqueues the encode of a planned batch with workspace `workspace`.
int encode_launch_planned(EncodePlan& plan, int32_t frame_count, void* workspace,
                          uint64_t workspace_bytes, void* output, uint64_t output_bytes,
                          fvh_result* results, fvh_stream_t stream, char* msg, size_t msg_len) {
  if (workspace_bytes < (uint64_t)plan.ws.total || output_bytes < (uint64_t)plan.output_bytes ||
      workspace == nullptr || output == nullptr || results == nullptr) {
    set_msg(msg, msg_len, "workspace or output buffer is too small");
    return FVH_EARG;
  }
  uint8_t* ws = (uint8_t*)workspace;
  for (int32_t i = 0; i < frame_count; ++i) {
    plan.frames[i].stream = (uint8_t*)output + plan.output_offsets[i];
  }
  EncodeBuffers b = {};
  b.frames = (FrameDesc*)(ws + plan.ws.frames);
  b.states = (FrameState*)(ws + plan.ws.states);
  b.tiles = (TileDesc*)(ws + plan.ws.tiles);
  b.shards = (ShardDesc*)(ws + plan.ws.shards);
  b.recs = (ShardRec*)(ws + plan.ws.recs);
  b.tile_len = (int64_t*)(ws + plan.ws.tile_len);
  b.tile_off = (int64_t*)(ws + plan.ws.tile_off);
  b.folded = ws + plan.ws.folded;
  b.scratch = ws + plan.ws.scratch;
  b.ranks = ws + plan.ws.ranks;
  b.recon = plan.ws.recon >= 0 ? (uint16_t*)(ws + plan.ws.recon) : nullptr;
  b.sink = ws + plan.ws.sink;
  b.queues = (int32_t*)(ws + plan.ws.queues);
  b.group_first = (int32_t*)(ws + plan.ws.groups);
  b.emit_list = (int32_t*)(ws + plan.ws.emit_list);
  b.group_total = (int)plan.groups;
  b.band_ring = plan.band_ring;
  b.results = (Results*)results;
  b.frame_count = frame_count;
  b.tile_total = (int)plan.tiles;
  b.shard_total = (int)plan.shards;
  b.wide_folded = plan.wide_folded;
  b.narrow_folded = plan.narrow_folded && !plan.wide_folded;
  b.any_heuristic = plan.any_heuristic;
  b.any_gray16 = plan.any_gray16;
  b.any_lossless = plan.any_lossless;
  b.any_lossy = plan.any_lossy;
  b.any_wide = plan.any_wide;
  b.any_band = plan.any_band;
  b.any_nonband = plan.any_nonband;
  b.any_med = plan.any_med;
  b.any_grad = plan.any_grad;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  launch_upload_frames(plan.frames.data(), frame_count, b.frames, s);
  launch_encode(b, s);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_msg(msg, msg_len, hipGetErrorString(err));
    return FVH_EHIP;
  }
  return FVH_OK;
}

}  // namespace
}  // namespace fvh

extern "C" {

int fvh_encode_launch(const fvh_encode_frame* frames, int32_t frame_count, void* workspace,
                      uint64_t workspace_bytes, void* output, uint64_t output_bytes,
                      fvh_result* results, fvh_stream_t stream, char* msg, size_t msg_len) {
  EncodePlan plan;
  const int rc = build_encode_plan(frames, frame_count, plan, msg, msg_len);
  if (rc) return rc;
  return encode_launch_planned(plan, frame_count, workspace, workspace_bytes, output,
                               output_bytes, results, stream, msg, msg_len);
}

int fvh_encode_launch_static(const fvh_encode_frame* frames, int32_t frame_count, void* output,
                             uint64_t output_bytes, fvh_result* results, fvh_stream_t stream,
                             char* msg, size_t msg_len) {
  EncodePlan plan;
  const int rc = build_encode_plan(frames, frame_count, plan, msg, msg_len);
  if (rc) return rc;
  if ((uint64_t)plan.ws.total > kStaticWorkspaceBytes) {
    if (msg && msg_len) {
      std::snprintf(msg, msg_len,
                    "batch needs %lld workspace bytes, more than the %llu of the static workspace",
                    (long long)plan.ws.total, (unsigned long long)kStaticWorkspaceBytes);
    }
    return FVH_EARG;
  }
  void* ws = nullptr;
  const hipError_t err = static_workspace((hipStream_t)stream, &ws);
  if (err != hipSuccess) {
    set_msg(msg, msg_len, hipGetErrorString(err));
    return FVH_EHIP;
  }
  return encode_launch_planned(plan, frame_count, ws, kStaticWorkspaceBytes, output,
                               output_bytes, results, stream, msg, msg_len);
}

// fastvid_cuda.cpp:141-183 (messages verbatim).
int fvh_decode_header(const uint8_t* bytes, int64_t size, fvh_stream_info* h, char* msg,
                      size_t msg_len) {
  std::memset(h, 0, sizeof(*h));
  h->stream_size = size;
  if (size < kHeaderBytes) { set_msg(msg, msg_len, "truncated Fastvid header"); return FVH_EARG; }
  if (std::memcmp(bytes, "FVID", 4) != 0) { set_msg(msg, msg_len, "bad Fastvid magic"); return FVH_EARG; }
  if (bytes[4] != 1) { set_msg(msg, msg_len, "unsupported Fastvid bitstream version"); return FVH_EARG; }
  if (bytes[5] > 2) { set_msg(msg, msg_len, "unknown pixel layout"); return FVH_EARG; }
  h->layout = bytes[5];
  const bool gray = h->layout == 0;
  h->quality = bytes[6];
  if (!(h->quality >= 1 && h->quality <= 100)) { set_msg(msg, msg_len, "quality is out of range"); return FVH_EARG; }
  const bool refined = (bytes[7] & 0x80) != 0;
  const bool gray_med = (bytes[7] & 0x40) != 0;
  h->bit_depth = (int64_t)(bytes[7] & 0x3f) + 8;
  if (!(h->bit_depth == 8 || h->bit_depth == 10 || h->bit_depth == 12 || h->bit_depth == 16)) {
    set_msg(msg, msg_len, "unsupported high-bit depth");
    return FVH_EARG;
  }
  if (refined && !(gray && h->bit_depth == 16 && h->quality < 100)) {
    set_msg(msg, msg_len, "invalid gray16 refinement flag");
    return FVH_EARG;
  }
  if (gray_med && !(gray && h->bit_depth == 10 && h->quality < 100)) {
    set_msg(msg, msg_len, "invalid gray MED flag");
    return FVH_EARG;
  }
  auto u32 = [&](int o) {
    return (int64_t)((uint32_t)bytes[o] | ((uint32_t)bytes[o + 1] << 8) |
                     ((uint32_t)bytes[o + 2] << 16) | ((uint32_t)bytes[o + 3] << 24));
  };
  h->width = u32(8);
  h->height = u32(12);
  h->tile_width = (int64_t)bytes[16] | ((int64_t)bytes[17] << 8);
  h->tile_height = (int64_t)bytes[18] | ((int64_t)bytes[19] << 8);
  if (!(h->width > 0 && h->height > 0 && h->tile_width > 0 && h->tile_height > 0)) {
    set_msg(msg, msg_len, "zero dimension");
    return FVH_EARG;
  }
  if (!(u32(20) != 0 && u32(24) != 0)) { set_msg(msg, msg_len, "zero frame-rate component"); return FVH_EARG; }
  h->tile_count = u32(28);
  if (h->tile_count != tile_count_closed(h->width, h->height, h->tile_width, h->tile_height, h->layout)) {
    set_msg(msg, msg_len, "tile count does not match dimensions");
    return FVH_EARG;
  }
  if (h->tile_count > (1 << 20)) { set_msg(msg, msg_len, "too many tiles"); return FVH_EARG; }
  if (kHeaderBytes + h->tile_count * kDirectoryEntryBytes > size) {
    set_msg(msg, msg_len, "truncated tile directory");
    return FVH_EARG;
  }
  const int64_t secondary_width = h->layout == 1 ? (h->width + 1) / 2 : h->width;
  h->y_samples = h->width * h->height;
  h->secondary_samples = gray ? 0 : secondary_width * h->height;
  h->total_samples = h->y_samples + 2 * h->secondary_samples;
  FrameDesc scratch_desc = {};
  const Geometry g = fill_geometry(scratch_desc, h->width, h->height, h->tile_width, h->tile_height, h->layout);
  h->total_shards = g.shards;
  h->step = refined ? refined_gray16_step(h->quality) : quantization_step(h->layout, h->bit_depth, h->quality);
  h->max_sample = (int64_t{1} << h->bit_depth) - 1;
  h->use_med = gray_med || !gray;
  h->flags = bytes[7] & 0xc0;
  return FVH_OK;
}

// DRAM-input directory walk: fastvid_cuda.cpp:255-333 (messages verbatim).
int fvh_decode_check_host(const uint8_t* bytes, int64_t size, const fvh_stream_info* h, char* msg,
                          size_t msg_len) {
  FrameDesc f = {};
  fill_geometry(f, h->width, h->height, h->tile_width, h->tile_height, h->layout);
  const int64_t T = h->tile_count;
  auto u32 = [&](int64_t o) {
    return (uint32_t)bytes[o] | ((uint32_t)bytes[o + 1] << 8) | ((uint32_t)bytes[o + 2] << 16) |
           ((uint32_t)bytes[o + 3] << 24);
  };
  std::vector<int64_t> offsets(T), lengths(T), tw(T), th(T);
  int64_t next = kHeaderBytes + T * kDirectoryEntryBytes;
  for (int64_t i = 0; i < T; ++i) {
    const int plane = i >= f.plane_tile_begin[2] ? 2 : (i >= f.plane_tile_begin[1] ? 1 : 0);
    const int64_t li = i - f.plane_tile_begin[plane];
    const int64_t r = li / f.cols[plane], c = li % f.cols[plane];
    const int64_t x = c * f.ntw[plane], y = r * f.tile_h;
    const int64_t w = std::min<int64_t>(f.ntw[plane], f.plane_width[plane] - x);
    const int64_t hh = std::min<int64_t>(f.tile_h, h->height - y);
    tw[i] = w;
    th[i] = hh;
    const int64_t e = kHeaderBytes + i * kDirectoryEntryBytes;
    if (bytes[e] != plane) { set_msg(msg, msg_len, "non-canonical tile plane"); return FVH_EARG; }
    if (bytes[e + 1] != kEntropyParallelShards) { set_msg(msg, msg_len, "tile requires bounded-shard entropy"); return FVH_EARG; }
    if (bytes[e + 2] != kPredictorId) { set_msg(msg, msg_len, "tile requires full-tile clamp-gradient prediction"); return FVH_EARG; }
    if (bytes[e + 3] != 0) { set_msg(msg, msg_len, "nonzero reserved directory byte"); return FVH_EARG; }
    if (!(u32(e + 4) == (uint32_t)x && u32(e + 8) == (uint32_t)y && u32(e + 12) == (uint32_t)w &&
          u32(e + 16) == (uint32_t)hh)) {
      set_msg(msg, msg_len, "non-canonical tile directory");
      return FVH_EARG;
    }
    const uint64_t off = (uint64_t)u32(e + 20) | ((uint64_t)u32(e + 24) << 32);
    if (off > (uint64_t)INT64_MAX) { set_msg(msg, msg_len, "payload offset is too large"); return FVH_EARG; }
    offsets[i] = (int64_t)off;
    lengths[i] = u32(e + 28);
    if (offsets[i] != next) { set_msg(msg, msg_len, "payloads are not contiguous"); return FVH_EARG; }
    if (lengths[i] > size - next) { set_msg(msg, msg_len, "truncated tile payload"); return FVH_EARG; }
    next += lengths[i];
  }
  if (next != size) { set_msg(msg, msg_len, "trailing stream bytes"); return FVH_EARG; }
  for (int64_t i = 0; i < T; ++i) {
    int64_t cursor = offsets[i], decoded = 0;
    const int64_t samples = tw[i] * th[i];
    const int64_t end = cursor + lengths[i];
    while (decoded < samples) {
      if (!(cursor + 3 <= end)) { set_msg(msg, msg_len, "truncated entropy shard header"); return FVH_EARG; }
      const int64_t mode = bytes[cursor];
      const int64_t body = (int64_t)bytes[cursor + 1] | ((int64_t)bytes[cursor + 2] << 8);
      if (mode > 19) { set_msg(msg, msg_len, "unknown entropy shard mode"); return FVH_EARG; }
      if (body > end - (cursor + 3)) { set_msg(msg, msg_len, "truncated entropy shard body"); return FVH_EARG; }
      decoded += std::min<int64_t>(kShardSymbols, samples - decoded);
      cursor += 3 + body;
    }
    if (cursor != end) { set_msg(msg, msg_len, "trailing bounded-shard tile bytes"); return FVH_EARG; }
  }
  return FVH_OK;
}

struct DecodePlanLayout {
  int64_t frames, states, tiles, shards, dshards, lists, o0tabs, folded, sink, queues, groups, total;
  int64_t tiles_n, shards_n, samples, groups_n;
  bool wide_folded, any_wide;
};

static DecodePlanLayout decode_layout(const fvh_stream_info* infos, int32_t count) {
  DecodePlanLayout L = {};
  for (int32_t i = 0; i < count; ++i) {
    L.tiles_n += infos[i].tile_count;
    L.shards_n += infos[i].total_shards;
    FrameDesc scratch_desc = {};
    const Geometry g = fill_geometry(scratch_desc, infos[i].width, infos[i].height,
                                     infos[i].tile_width, infos[i].tile_height, infos[i].layout);
    L.samples = ((L.samples + 7) & ~int64_t{7}) + g.samples;
    L.groups_n += number_groups(scratch_desc, L.groups_n);
    L.wide_folded |= infos[i].bit_depth == 16;
    const int64_t tw = infos[i].tile_width;
    const int64_t ctw = infos[i].layout == 1 ? (tw + 1) / 2 : tw;
    L.any_wide |= std::min(tw, infos[i].width) > kRingMax ||
                  (infos[i].layout != 0 && std::min(ctw, infos[i].width) > kRingMax);
  }
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = off;
    off = align256(off + bytes);
    return at;
  };
  L.frames = take((int64_t)count * sizeof(FrameDesc));
  L.states = take((int64_t)count * sizeof(FrameState));
  L.tiles = take(L.tiles_n * (int64_t)sizeof(TileDesc));
  L.shards = take(L.shards_n * (int64_t)sizeof(ShardDesc));
  L.dshards = take(L.shards_n * (int64_t)sizeof(DShard));
  // order-0 list, other list, three order-0 tier lists, Rice list, then the
  // counters
  L.lists = take((6 * L.shards_n + 8) * 4);
  L.o0tabs = take(L.shards_n * (int64_t)sizeof(O0Tab));
  L.folded = take((L.samples + 8) * (L.wide_folded ? 4 : 2));
  L.sink = take(kSinkBytes);
  L.queues = take(kQueueCount * 4);
  L.groups = take(L.groups_n * 4);
  L.total = off;
  return L;
}

int fvh_decode_plan(const fvh_stream_info* infos, int32_t count, uint64_t* workspace_bytes,
                    char* msg, size_t msg_len) {
  if (count <= 0) {
    set_msg(msg, msg_len, "no streams to decode");
    return FVH_EARG;
  }
  const DecodePlanLayout L = decode_layout(infos, count);
  if (L.tiles_n > INT32_MAX / 2 || L.shards_n > INT32_MAX / 4) {
    set_msg(msg, msg_len, "batch is too large");
    return FVH_EARG;
  }
  *workspace_bytes = (uint64_t)L.total;
  return FVH_OK;
}

}  // extern "C"

namespace fvh {
namespace {

int decode_launch(const fvh_stream_info* infos, const void* const* streams,
                  const int64_t* fed_sizes, void* const* outputs, int32_t count, void* workspace,
                  uint64_t workspace_bytes, fvh_result* results, int32_t wavefront,
                  fvh_stream_t stream, char* msg, size_t msg_len) {
  if (count <= 0) {
    set_msg(msg, msg_len, "no streams to decode");
    return FVH_EARG;
  }
  const DecodePlanLayout L = decode_layout(infos, count);
  if (workspace_bytes < (uint64_t)L.total || workspace == nullptr || results == nullptr) {
    set_msg(msg, msg_len, "workspace is too small");
    return FVH_EARG;
  }
  std::vector<FrameDesc> frames(count);
  int64_t tiles = 0, shards = 0, samples = 0, groups = 0;
  int band_ring = 16;
  bool any_band = false, any_nonband = false, any_med = false, any_grad = false;
  bool any_deep = false;
  for (int32_t i = 0; i < count; ++i) {
    const fvh_stream_info& h = infos[i];
    FrameDesc& f = frames[i];
    const Geometry g = fill_geometry(f, h.width, h.height, h.tile_width, h.tile_height, h.layout);
    any_deep |= h.bit_depth > 8;
    f.stream = (uint8_t*)streams[i];
    // with device sizes, k_dheader sets the size once the header checks out
    f.stream_capacity = fed_sizes != nullptr ? 0 : h.stream_size;
    uint16_t* base = (uint16_t*)outputs[i];
    f.out_plane[0] = base;
    f.out_plane[1] = h.layout ? base + h.y_samples : nullptr;
    f.out_plane[2] = h.layout ? base + h.y_samples + h.secondary_samples : nullptr;
    for (int p = 0; p < 3; ++p) {
      f.plane[p] = nullptr;
      f.pitch[p] = f.plane_width[p];
    }
    samples = (samples + 7) & ~int64_t{7};
    f.folded_base = samples;
    const void* outs[3] = {f.out_plane[0], f.out_plane[1], f.out_plane[2]};
    f.band = band_eligible(f, outs);
    groups += number_groups(f, groups);
    if (f.band) band_ring = std::max(band_ring, band_ring_groups(f));
    any_band |= f.band != 0;
    any_nonband |= f.band == 0;
    if (fed_sizes != nullptr) {
      // the predictor comes from each stream's header on the device: gray
      // frames may use either
      any_med |= h.layout != 0 || h.bit_depth == 10;
      any_grad |= h.layout == 0;
    } else {
      any_med |= h.use_med != 0;
      any_grad |= h.use_med == 0;
    }
    f.tile_begin = (int32_t)tiles;
    f.tile_count = (int32_t)g.tiles;
    f.shard_begin = (int32_t)shards;
    f.shard_count = (int32_t)g.shards;
    f.bit_depth = (int32_t)h.bit_depth;
    f.quality = (int32_t)h.quality;
    f.max_sample = (int32_t)h.max_sample;
    f.step = (int32_t)h.step;
    f.use_med = (int32_t)h.use_med;
    f.header_flags = (int32_t)h.flags;
    tiles += g.tiles;
    shards += g.shards;
    samples += g.samples;
  }
  uint8_t* ws = (uint8_t*)workspace;
  DecodeBuffers b = {};
  b.frames = (FrameDesc*)(ws + L.frames);
  b.states = (FrameState*)(ws + L.states);
  b.tiles = (TileDesc*)(ws + L.tiles);
  b.shards = (ShardDesc*)(ws + L.shards);
  b.dshards = (DShard*)(ws + L.dshards);
  b.lists = (int32_t*)(ws + L.lists);
  b.list_count = b.lists + 6 * L.shards_n;
  b.o0tabs = (O0Tab*)(ws + L.o0tabs);
  b.sink = ws + L.sink;
  b.queues = (int32_t*)(ws + L.queues);
  b.group_first = (int32_t*)(ws + L.groups);
  b.group_total = (int)groups;
  b.band_ring = band_ring;
  b.folded = ws + L.folded;
  b.results = (Results*)results;
  b.frame_count = count;
  b.tile_total = (int)tiles;
  b.shard_total = (int)shards;
  b.wide_folded = L.wide_folded;
  b.any_wide = L.any_wide;
  b.wavefront = wavefront;
  b.any_band = any_band;
  b.any_nonband = any_nonband;
  b.any_med = any_med;
  b.any_grad = any_grad;
  b.any_deep = any_deep;
  b.fed_sizes = fed_sizes;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  launch_upload_frames(frames.data(), count, b.frames, s);
  launch_decode(b, s);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_msg(msg, msg_len, hipGetErrorString(err));
    return FVH_EHIP;
  }
  return FVH_OK;
}

}  // namespace
}  // namespace fvh

extern "C" {

int fvh_decode_launch(const fvh_stream_info* infos, const void* const* streams,
                      void* const* outputs, int32_t count, void* workspace, uint64_t workspace_bytes,
                      fvh_result* results, int32_t wavefront, fvh_stream_t stream, char* msg,
                      size_t msg_len) {
  return fvh::decode_launch(infos, streams, nullptr, outputs, count, workspace, workspace_bytes,
                            results, wavefront, stream, msg, msg_len);
}

int fvh_decode_launch_sized(const fvh_stream_info* infos, const void* const* streams,
                            const int64_t* stream_sizes, void* const* outputs, int32_t count,
                            void* workspace, uint64_t workspace_bytes, fvh_result* results,
                            int32_t wavefront, fvh_stream_t stream, char* msg, size_t msg_len) {
  if (stream_sizes == nullptr) {
    set_msg(msg, msg_len, "stream_sizes must be a device pointer");
    return FVH_EARG;
  }
  return fvh::decode_launch(infos, streams, stream_sizes, outputs, count, workspace,
                            workspace_bytes, results, wavefront, stream, msg, msg_len);
}

int fvh_results_read(const fvh_result* results, fvh_result* host, int32_t count,
                     fvh_stream_t stream, char* msg, size_t msg_len) {
  if (results == nullptr || host == nullptr || count <= 0) {
    set_msg(msg, msg_len, "results, host and count must be set");
    return FVH_EARG;
  }
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(host, results, sizeof(fvh_result) * (size_t)count,
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_msg(msg, msg_len, hipGetErrorString(e));
    return FVH_EHIP;
  }
  return FVH_OK;
}

int fvh_results_copy(const fvh_result* results, fvh_result* host, int32_t count,
                     fvh_stream_t stream, char* msg, size_t msg_len) {
  if (results == nullptr || host == nullptr || count <= 0) {
    set_msg(msg, msg_len, "results, host and count must be set");
    return FVH_EARG;
  }
  const hipError_t e = hipMemcpyAsync(host, results, sizeof(fvh_result) * (size_t)count,
                                      hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_msg(msg, msg_len, hipGetErrorString(e));
    return FVH_EHIP;
  }
  return FVH_OK;
}

int fvh_set_band_waves(int32_t waves_per_cu) {
  if (waves_per_cu < 0 || waves_per_cu > 64) return -1;
  return fvh::band_waves_per_cu(waves_per_cu);
}

int fvh_set_band_waves_decode(int32_t waves_per_cu) {
  if (waves_per_cu < 0 || waves_per_cu > 64) return -1;
  return fvh::band_decode_waves_per_cu(waves_per_cu);
}

const char* fvh_build_info(void) { return "fastvid_hip 0.1 gfx950"; }

}  // extern "C"
