// fvh_plan.hip -- batch geometry on the device.
//
// Tiles follow the canonical order of expected_tiles (cuda/csrc/encode.cu:41-69):
// plane-major, then tile rows, then columns; yuv422 chroma planes use half
// width and half nominal tile width; edge tiles are clipped.  Instead of the
// reference's host-built metadata vectors and H2D copies
// (cuda/csrc/encode.cu:739-767), each tile's origin, size, residual base and
// first shard come from closed forms over the frame descriptor, one thread per
// tile, so a batch of any number of frames is planned by one launch.
#include <algorithm>

#include "fvh_device.h"
#include "fvh_kernels.h"
#include "fvh_timing.h"

namespace fvh {

__global__ __launch_bounds__(256) void k_plan(const FrameDesc* __restrict__ frames,
                                              int frame_count, TileDesc* __restrict__ tiles,
                                              ShardDesc* __restrict__ shards,
                                              int32_t* __restrict__ group_first, int tile_total,
                                              PlanInit init) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (init.kind != 0) {
    if (t < init.queue_count) init.queues[t] = 0;
    if (t < init.list_counts) init.list_count[t] = 0;
    if (t < frame_count) {
      FrameState s = {};
      if (init.kind == 1) {
        s.step = frames[t].step0;
        s.use_med = frames[t].layout != 0;  // first pass: MED for colour, gradient for gray (encode.cu:812)
      }
      s.active = 1;
      init.states[t] = s;
    }
  }
  if (t >= tile_total) return;
  int lo = 0, hi = frame_count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (frames[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
  }
  const FrameDesc& f = frames[lo];
  const int local = t - f.tile_begin;
  const int plane = local >= f.plane_tile_begin[2] ? 2 : (local >= f.plane_tile_begin[1] ? 1 : 0);
  const int i = local - f.plane_tile_begin[plane];
  const int r = i / f.cols[plane];
  const int c = i - r * f.cols[plane];
  const int64_t ntw = f.ntw[plane];
  const int64_t pw = f.plane_width[plane];
  const int64_t x = (int64_t)c * ntw;
  const int64_t y = (int64_t)r * f.tile_h;
  const int64_t w = min(ntw, pw - x);
  const int64_t h = min((int64_t)f.tile_h, f.height - y);
  TileDesc td;
  td.folded_base = f.folded_base + f.plane_sample_base[plane] + (int64_t)r * f.row_samples[plane] +
      (int64_t)c * ((ntw * h + 7) & ~int64_t{7});
  td.src_offset = y * f.pitch[plane] + x;
  td.dst_offset = y * pw + x;
  td.frame = lo;
  td.plane = plane;
  td.x = (int32_t)x;
  td.y = (int32_t)y;
  td.width = (int32_t)w;
  td.height = (int32_t)h;
  const int64_t full_tile_shards = (ntw * h + kShardSymbols - 1) / kShardSymbols;
  const int64_t first = f.shard_begin + f.plane_shard_base[plane] + (int64_t)r * f.row_shards[plane] +
      (int64_t)c * full_tile_shards;
  const int64_t count = (w * h + kShardSymbols - 1) / kShardSymbols;
  td.first_shard = (int32_t)first;
  td.shard_count = (int32_t)count;
  tiles[t] = td;
  // band tile groups: 4 consecutive tiles of one plane
  if ((i & 3) == 0) group_first[f.group_begin[plane] + (i >> 2)] = t;
  for (int64_t s = 0; s < count; ++s) {
    ShardDesc sd;
    sd.tile = t;
    sd.local = (int32_t)s;
    sd.frame = lo;
    shards[first + s] = sd;
  }
}

int blocks_per_cu(const void* kernel, int dyn_lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, (size_t)dyn_lds) !=
          hipSuccess ||
      per_cu <= 0) {
    per_cu = 16;
  }
  // The occupancy query counts LDS by the byte; gfx950 hands it out in
  // granules of 1280 bytes (128 per compute unit): one-wave blocks of 12800
  // bytes run 12 to a compute unit, of 12801 bytes 11, where the query
  // answers 12 for anything up to 13653.
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, kernel) == hipSuccess) {
    const size_t granules = (attr.sharedSizeBytes + (size_t)dyn_lds + kLdsGranule - 1) / kLdsGranule;
    if (granules > 0) per_cu = std::min(per_cu, (int)(kLdsGranules / granules));
  }
  return per_cu;
}

int device_cus() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
      cus = 256;
    }
    cached[dev] = cus;
  }
  return cached[dev];
}

void launch_plan(const FrameDesc* frames, int frame_count, TileDesc* tiles, ShardDesc* shards,
                 int32_t* group_first, int tile_total, hipStream_t stream, const PlanInit& init) {
  int threads = max(tile_total, 0);
  if (init.kind != 0) {
    threads = max(threads, max(frame_count, max(init.queue_count, init.list_counts)));
  }
  if (threads <= 0) return;
  {
    FVH_TIMED("k_plan", stream);
    hipLaunchKernelGGL(k_plan, dim3((threads + 255) / 256), dim3(256), 0, stream, frames,
                       frame_count, tiles, shards, group_first, tile_total, init);
  }
}

}  // namespace fvh
