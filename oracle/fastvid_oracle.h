/*
 * fastvid_oracle.h -- CPU restatement of the fastvid version-1 intra codec.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in fastvid_amd/csrc and the CPU baseline timed by bench.py.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path never links or calls it.
 *
 * Every function restates the reference CUDA extension (cuda/csrc/encode.cu,
 * cuda/csrc/decode.cu, cuda/csrc/fastvid_cuda.cpp) in plain scalar C; the
 * reference file:line each one follows is cited beside its definition.
 */
#ifndef FASTVID_ORACLE_H
#define FASTVID_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One tile of the canonical geometry (encode.cu:30-39). */
typedef struct {
  int64_t plane, x, y, width, height, folded_base, first_shard, shard_count;
} fvo_tile;

/* Per-shard analysis record, same fields as encode.cu:308-316. */
typedef struct {
  int64_t mode, param, body, record, lane[4];
} fvo_analysis;

int32_t fvo_quantization_step(int64_t layout, int64_t bit_depth, int64_t quality);
int32_t fvo_refined_gray16_step(int64_t quality);

/* Number of tiles of the canonical geometry; tiles==NULL only counts. */
int64_t fvo_tiles(int64_t width, int64_t height, int64_t tile_width,
                  int64_t tile_height, int64_t layout, fvo_tile* tiles);

/* Predictor + quantizer of one tile, raster order (equivalent to the
 * reference antidiagonal wavefront because both visit every causal
 * neighbour first).  plane points at sample (0,0) of the plane; pitch is the
 * row pitch in samples.  folded receives width*height values in tile-raster
 * order.  Returns 1 if any sample exceeds max_sample, else 0. */
int fvo_predict_tile(const uint16_t* plane, int64_t pitch, const fvo_tile* tile,
                     int32_t step, int32_t max_sample, int use_med,
                     uint32_t* folded);

/* Zero-run / Rice / block-pack costs and the mode choice of one shard
 * (analyze_shards_kernel).  Returns the status code (0 or 2). */
int fvo_analyze_shard(const uint32_t* folded, int count, fvo_analysis* out);

/* Order-0 rANS analysis of one shard (analyze_order0_shards_kernel).
 * Returns 1 if the shard is supported.  On success fills the sparse table
 * (sparse_len bytes), lane byte streams in emission order (lanes[l][0..]),
 * lane byte counts, final states and the body length. */
int fvo_order0_shard(const uint32_t* folded, int count, uint8_t* sparse,
                     int* sparse_len, uint8_t lanes[4][2048], int lane_bytes[4],
                     uint32_t states[4], int64_t* body);

/* Normalised order-0 frequencies of one shard: freq/cum indexed by folded
 * value (511 entries each).  Returns the distinct symbol count, or -1 if the
 * shard is unsupported (value >= 511).  Distinct counts above 255 are
 * returned as-is (the caller rejects them). */
int fvo_order0_frequencies(const uint32_t* folded, int count, uint16_t* freq,
                           uint16_t* cum);

/* Encoder.  planes[p] points at a row-major uint16 plane with row pitch
 * pitches[p] samples; width/height are the luma plane's dimensions.  On success *out receives a malloc'd stream of *out_len
 * bytes and 0 is returned; on failure a nonzero code is returned and msg
 * holds the reference's error text.  threads>1 splits tiles and shards over
 * POSIX threads.  Optional: folded_out (total samples, tile order) and
 * analysis_out (total shards) receive the final pass's intermediates. */
int fvo_encode(const uint16_t* const* planes, const int64_t* pitches,
               int64_t plane_count, int64_t width, int64_t height,
               int64_t layout, int64_t bit_depth, int64_t quality,
               int64_t fps_num, int64_t fps_den, int64_t tile_width,
               int64_t tile_height, int threads, uint8_t** out,
               int64_t* out_len, uint32_t* folded_out,
               fvo_analysis* analysis_out, char* msg, size_t msg_len);

/* Which multi-pass heuristic of encode.cu:812-847 an encode took, with the
 * payload sums (record bytes over all shards) its thresholds compared. */
enum {
  FVO_BRANCH_NONE = 0,              /* one pass */
  FVO_BRANCH_GRAY10_MED = 1,        /* gray10 lossy re-run with MED (flag 0x40) */
  FVO_BRANCH_GRAY16_DIRECT_LOW = 2, /* baseline < samples/10: refined (flag 0x80) */
  FVO_BRANCH_GRAY16_DIRECT_HIGH = 3,/* baseline > samples/2: refined (flag 0x80) */
  FVO_BRANCH_GRAY16_PROBE_KEPT = 4, /* probe pass below samples/10: refined */
  FVO_BRANCH_GRAY16_PROBE_REVERTED = 5 /* probe pass not below samples/10: baseline again */
};

typedef struct {
  int64_t baseline_payload, refined_payload; /* refined_payload -1 if no probe/refinement */
  int64_t branch, passes;
} fvo_trace;

/* fvo_encode plus the heuristic trace (trace may be NULL). */
int fvo_encode_traced(const uint16_t* const* planes, const int64_t* pitches,
                      int64_t plane_count, int64_t width, int64_t height,
                      int64_t layout, int64_t bit_depth, int64_t quality,
                      int64_t fps_num, int64_t fps_den, int64_t tile_width,
                      int64_t tile_height, int threads, uint8_t** out,
                      int64_t* out_len, uint32_t* folded_out,
                      fvo_analysis* analysis_out, fvo_trace* trace, char* msg,
                      size_t msg_len);

/* Header fields after validation (fastvid_cuda.cpp:141-183). */
typedef struct {
  int64_t layout, quality, bit_depth, width, height, tile_width, tile_height;
  int64_t tile_count, total_samples, total_shards, y_samples, secondary_samples;
  int64_t step, max_sample, use_med, refined_gray16, gray_med;
} fvo_header;

int fvo_parse_header(const uint8_t* stream, int64_t size, fvo_header* header,
                     char* msg, size_t msg_len);

/* Decoder.  output receives total_samples uint16 values (planes
 * concatenated, each row-major and dense).  Returns 0 on success, a nonzero
 * code with the reference's message otherwise.  threads>1 splits shards and
 * tiles over POSIX threads. */
int fvo_decode(const uint8_t* stream, int64_t size, uint16_t* output,
               int64_t output_capacity, int threads, char* msg, size_t msg_len);

/* Decode only the folded residuals (tile order) of a valid stream; used to
 * cross-check the entropy decoders in isolation. */
int fvo_decode_folded(const uint8_t* stream, int64_t size, uint32_t* folded,
                      int64_t capacity, char* msg, size_t msg_len);

void fvo_free(void* pointer);

#ifdef __cplusplus
}
#endif
#endif
