/*
 * fastvid_hip.h -- C ABI of libfastvid_hip.so, the MI355X (gfx950) fastvid
 * version-1 intra codec.
 *
 * Plain C: device pointers, sizes, a hipStream_t and int status codes.  No
 * torch or C++ types cross this boundary.  Host code (fastvid_amd/) allocates
 * device buffers (for instance as torch tensors), calls a *_plan function to
 * learn the sizes, then a *_launch function that enqueues every kernel on the
 * given stream without blocking.  Status words and stream lengths land in a
 * caller-provided device "results" array that the caller reads after
 * synchronising.
 *
 * Return codes of the functions themselves: 0 = ok, FVH_EARG = argument
 * rejected before any launch (msg holds the reference's message),
 * FVH_EHIP = a HIP runtime call failed (msg holds hipGetErrorString).
 */
#ifndef FASTVID_HIP_H
#define FASTVID_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* fvh_stream_t; /* same type as hipStream_t */

#define FVH_OK 0
#define FVH_EARG 1
#define FVH_EHIP 2

/* One frame to encode.  Replaces the argument list of the pybind entry
 * `_C.encode(planes, layout, bit_depth, quality, fps_numerator,
 * fps_denominator, tile_width, tile_height)` (cuda/csrc/fastvid_cuda.cpp:366-377,
 * fastvid_encode_cuda at cuda/csrc/encode.cu:686-694). */
typedef struct fvh_encode_frame {
  const void* planes[3];       /* device uint16 planes: (Y) or (Y, Cb, Cr) */
  int64_t pitches[3];          /* row pitch in samples (>= width of the plane) */
  int64_t plane_widths[3];     /* columns of each plane */
  int64_t plane_heights[3];    /* rows of each plane */
  int64_t plane_count;
  int64_t layout;              /* 0 gray, 1 yuv422, 2 rgb444 */
  int64_t bit_depth;           /* 8, 10, 12, 16 */
  int64_t quality;             /* 1..100 */
  int64_t fps_numerator, fps_denominator;
  int64_t tile_width, tile_height;
} fvh_encode_frame;

/* Per-frame outcome written to device memory by the launch. */
typedef struct fvh_result {
  int64_t stream_length;       /* bytes of the encoded stream (encode) */
  int32_t status;              /* 0 ok; reference-compatible code otherwise */
  int32_t flags;               /* bit0: gray MED pass taken, bit1: gray16 refinement */
} fvh_result;

/* Validates the frames (checks and messages of encode.cu:695-735) and sizes
 * the buffers of fvh_encode_launch.  output_offsets[i] receives the byte
 * offset of frame i's stream inside the output buffer.  Host-only. */
int fvh_encode_plan(const fvh_encode_frame* frames, int32_t frame_count,
                    uint64_t* workspace_bytes, uint64_t* output_bytes,
                    uint64_t* output_offsets, char* msg, size_t msg_len);

/* Enqueues the whole encode of a batch on `stream`: predictor + quantiser,
 * entropy analysis, order-0 rANS, the gray10/gray16 multi-pass heuristics,
 * stream assembly and emission (replaces fastvid_encode_cuda,
 * cuda/csrc/encode.cu:686-931).  Frame i's stream is written at
 * output + output_offsets[i]; results[i] (device memory) receives its length
 * and status (1 = sample above 2^bit_depth-1, "CUDA encoder analysis failed
 * (status 1)" in the reference).  Does not block. */
int fvh_encode_launch(const fvh_encode_frame* frames, int32_t frame_count,
                      void* workspace, uint64_t workspace_bytes, void* output,
                      uint64_t output_bytes, fvh_result* results,
                      fvh_stream_t stream, char* msg, size_t msg_len);

/* As fvh_encode_launch, with the workspace taken from a library-owned
 * device array instead of a caller buffer, so the call allocates no device
 * memory.  No counterpart in the reference (whose encoder allocates its
 * scratch per call, cuda/csrc/encode.cu:768-773): it is for callers that
 * cannot allocate -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The array holds
 * FVH_STATIC_WORKSPACE_BYTES per GPU, from the first call on that GPU until
 * the process ends; a batch whose fvh_encode_plan workspace is larger is
 * rejected with FVH_EARG before any launch.  Every call on one GPU shares
 * that GPU's array, so calls whose work overlaps on one GPU (other streams,
 * other threads) race: serialise them.  The caller still provides the
 * output and results buffers (sizes and offsets from fvh_encode_plan). */
#define FVH_STATIC_WORKSPACE_BYTES (256ull << 20)
int fvh_encode_launch_static(const fvh_encode_frame* frames, int32_t frame_count, void* output,
                             uint64_t output_bytes, fvh_result* results, fvh_stream_t stream,
                             char* msg, size_t msg_len);

/* Header fields of one stream after host-side validation. */
typedef struct fvh_stream_info {
  int64_t layout, quality, bit_depth, width, height, tile_width, tile_height;
  int64_t tile_count, total_samples, total_shards, y_samples, secondary_samples;
  int64_t step, max_sample, use_med, flags;
  int64_t stream_size;
} fvh_stream_info;

/* Validates the 32-byte header (checks and messages of
 * cuda/csrc/fastvid_cuda.cpp:142-183) and derives the frame geometry.
 * `header` is host memory holding min(32, stream_size) bytes.  Host-only. */
int fvh_decode_header(const uint8_t* header, int64_t stream_size,
                      fvh_stream_info* info, char* msg, size_t msg_len);

/* Host-side walk of a host-resident stream's directory and shard headers
 * (checks and messages of the DRAM-input path,
 * cuda/csrc/fastvid_cuda.cpp:255-333).  Host-only. */
int fvh_decode_check_host(const uint8_t* stream, int64_t stream_size,
                          const fvh_stream_info* info, char* msg, size_t msg_len);

/* Workspace bytes for decoding a batch of validated streams. */
int fvh_decode_plan(const fvh_stream_info* infos, int32_t count,
                    uint64_t* workspace_bytes, char* msg, size_t msg_len);

/* Enqueues the decode of a batch: device directory/shard parse, entropy
 * decode of all modes and reconstruction (replaces fastvid_decode_cuda,
 * cuda/csrc/decode.cu:660-730, called from decode at
 * cuda/csrc/fastvid_cuda.cpp:141-357).  streams[i] is a device pointer to
 * infos[i].stream_size bytes; outputs[i] is a device pointer to
 * infos[i].total_samples uint16 values (planes concatenated, dense rows).
 * wavefront=0 selects the one-thread-per-tile diagnostic reconstruction
 * (predictor="serial").  results[i].status receives 0 or the reference
 * status code (1-5 entropy, 10 directory, 11 shard headers). */
int fvh_decode_launch(const fvh_stream_info* infos, const void* const* streams,
                      void* const* outputs, int32_t count, void* workspace,
                      uint64_t workspace_bytes, fvh_result* results, int32_t wavefront,
                      fvh_stream_t stream, char* msg, size_t msg_len);

/* As fvh_decode_launch, with the stream sizes in device memory
 * (stream_sizes[i], int64) and each header read on the device, so the call
 * can be captured once in a graph and replayed for new streams of the same
 * geometry (no counterpart in the reference, whose decode parses the header
 * on the host, cuda/csrc/fastvid_cuda.cpp:141-183).  infos[i] gives the
 * geometry the batch is planned for: width, height, tile size, layout and
 * bit depth; quality, step, predictor and stream_size are taken from the
 * stream.  streams[i] must hold at least stream_sizes[i] bytes.  A header
 * that does not match its info yields status 12 (otherwise the statuses of
 * fvh_decode_launch). */
int fvh_decode_launch_sized(const fvh_stream_info* infos, const void* const* streams,
                            const int64_t* stream_sizes, void* const* outputs, int32_t count,
                            void* workspace, uint64_t workspace_bytes, fvh_result* results,
                            int32_t wavefront, fvh_stream_t stream, char* msg, size_t msg_len);

/* Copies the `count` result records of a job at `results` (device) to
 * `host` (pinned host memory) on the job's stream and waits for that stream:
 * a job's only device-to-host transfer, as the reference's `.cpu()` of its
 * status and analysis tensors (cuda/csrc/encode.cu:806-807, decode.cu:718). */
int fvh_results_read(const fvh_result* results, fvh_result* host, int32_t count,
                     fvh_stream_t stream, char* msg, size_t msg_len);

/* The same copy, only enqueued on `stream` (no wait): a job that enqueues it
 * right behind its kernels can later wait for just that copy (an event
 * recorded after it), so the read-back of one job does not hold back the
 * work queued behind it on the stream (same reference counterpart). */
int fvh_results_copy(const fvh_result* results, fvh_result* host, int32_t count,
                     fvh_stream_t stream, char* msg, size_t msg_len);

/* Per-kernel timing (no counterpart in the reference, which is timed from
 * outside with CUDA events around whole calls, scripts/evaluate.py:447-463).
 * While enabled, every kernel launch of the library is bracketed by two
 * hipEvents on its launch stream.  fvh_timing_read waits for the pending
 * events and fills up to `max` entries (one per kernel name) with launch
 * counts and total milliseconds; reset=1 clears the totals afterwards.
 * Returns the number of entries written. */
typedef struct fvh_kernel_time {
  char name[32];
  int64_t launches;
  double total_ms;
} fvh_kernel_time;

int fvh_timing_enable(int32_t enable);
int fvh_timing_read(fvh_kernel_time* out, int32_t max, int32_t reset);

/* Launch geometry of the predictor wavefront (no counterpart in the
 * reference, whose predict_tiles_kernel runs one block per tile,
 * cuda/csrc/encode.cu:780-783): persistent waves per compute unit, each
 * taking groups of four tiles (16 lanes per tile) from a device queue.
 * 1..64 sets it for launches enqueued afterwards (graphs keep the value they
 * were captured with); 0 only reads it.  Returns the value in force
 * (default 8), or -1 for an argument outside 0..64. */
int fvh_set_band_waves(int32_t waves_per_cu);

/* The same for the decoder's reconstruction wavefront (no counterpart in
 * the reference, whose reconstruct_tiles_kernel runs one block per tile,
 * cuda/csrc/decode.cu:696-697); default 4. */
int fvh_set_band_waves_decode(int32_t waves_per_cu);

/* Seeded synthetic plane on the device (no counterpart in the reference;
 * the serving workload generates each request's frame with it).  Writes
 * height rows of width uint16 samples at `out` (row pitch `pitch` samples):
 * ramp + three sinusoids + Gaussian noise + rectangles + offset, in 8-bit
 * units, scaled by (2^bit_depth - 1)/255, rounded and clipped.  Enqueued on
 * `stream`; FVH_EARG for invalid parameters. */
typedef struct fvh_synth_params {
  int64_t width, height, pitch;
  int32_t bit_depth, rect_count;
  float ramp_cx, ramp_cy, ramp_lo, ramp_scale;  /* ramp: (cy*y/H - lo)*scale + cx*x/W*scale */
  float kx[3], ky[3], phase[3];                 /* 20*sin(kx*x + ky*y + phase) each */
  float offset, noise_sigma;                    /* 8-bit units */
  uint64_t noise_seed;
  int32_t rects_under_noise;                    /* 1: rectangles before offset and noise */
  int32_t rects[20][5];                         /* x, y, w, h, value */
} fvh_synth_params;

int fvh_synth_plane(const fvh_synth_params* params, void* out, fvh_stream_t stream);

/* Library version string and the offload target it was built for. */
const char* fvh_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
