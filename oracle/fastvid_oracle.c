/*
 * fastvid_oracle.c -- scalar CPU restatement of the fastvid version-1 codec.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation
 * and the CPU baseline of bench.py.  It is never part of the product path.
 *
 * Each function follows the reference CUDA extension text; the file:line it
 * restates is cited beside it (paths relative to the reference repository).
 * The CUDA kernels run one thread per sample/shard/lane in parallel; here the
 * same arithmetic runs in scalar loops whose visiting order does not change
 * any result (each per-sample, per-shard and per-lane computation only reads
 * values that the reference also has available at that point).
 */
#include "fastvid_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HEADER_BYTES 32
#define DIRECTORY_ENTRY_BYTES 32
#define SHARD_SYMBOLS 4096
#define BLOCK_SYMBOLS 128
#define MAX_RICE 16
#define MODE_ZERO_RUN 0
#define MODE_RICE_BASE 1
#define MODE_BLOCK_PACK 18
#define MODE_ORDER0 19
#define ENTROPY_PARALLEL_SHARDS 19
#define PREDICTOR_ID 6
#define RANS_ALPHABET 511
#define RANS_TABLE_LOG 12
#define RANS_L (1u << 23)
#define SCRATCH_BYTES 10240

static void set_msg(char* msg, size_t len, const char* text) {
  if (msg && len) {
    snprintf(msg, len, "%s", text);
  }
}

void fvo_free(void* pointer) { free(pointer); }

/* ---------------------------------------------------------------- threads */

typedef void (*range_fn)(void* ctx, int64_t index);

typedef struct {
  range_fn fn;
  void* ctx;
  int64_t count;
  int64_t next;
} pool_job;

static void* pool_worker(void* arg) {
  pool_job* job = (pool_job*)arg;
  for (;;) {
    const int64_t index = __atomic_fetch_add(&job->next, 1, __ATOMIC_RELAXED);
    if (index >= job->count) break;
    job->fn(job->ctx, index);
  }
  return NULL;
}

static void parallel_for(int threads, int64_t count, range_fn fn, void* ctx) {
  if (threads <= 1 || count <= 1) {
    for (int64_t i = 0; i < count; ++i) fn(ctx, i);
    return;
  }
  if (threads > 256) threads = 256;
  pool_job job = {fn, ctx, count, 0};
  pthread_t ids[256];
  int started = 0;
  for (int t = 1; t < threads; ++t) {
    if (pthread_create(&ids[started], NULL, pool_worker, &job) == 0) ++started;
  }
  pool_worker(&job);
  for (int t = 0; t < started; ++t) pthread_join(ids[t], NULL);
}

/* --------------------------------------------------------------- geometry */

/* quantization_step: encode.cu:71-85 (duplicated fastvid_cuda.cpp:48-62). */
int32_t fvo_quantization_step(int64_t layout, int64_t bit_depth, int64_t quality) {
  if ((layout == 0 && bit_depth == 8) || (layout == 1 && bit_depth == 8)) {
    return 1;
  }
  int32_t denominator = 6;
  if (layout == 1 && bit_depth == 10) {
    denominator = 20;
  } else if (layout == 1 && bit_depth == 16) {
    denominator = 24;
  } else if (layout == 2 && bit_depth == 10) {
    denominator = 10;
  }
  const int32_t scale = (int32_t)1 << (bit_depth - 8);
  return 1 + (int32_t)(((100 - quality) * scale + denominator - 1) / denominator);
}

/* refined_gray16_step: encode.cu:87-89. */
int32_t fvo_refined_gray16_step(int64_t quality) {
  return 1 + (int32_t)(((100 - quality) * 256 + 7) / 8);
}

/* expected_tiles: encode.cu:41-69 (plane-major, then y, then x; edge tiles
 * clipped; yuv422 chroma planes use half width and half nominal tile width). */
int64_t fvo_tiles(int64_t width, int64_t height, int64_t tile_width,
                  int64_t tile_height, int64_t layout, fvo_tile* tiles) {
  int64_t count = 0;
  int64_t folded_base = 0;
  int64_t first_shard = 0;
  const int64_t plane_count = layout == 0 ? 1 : 3;
  for (int64_t plane = 0; plane < plane_count; ++plane) {
    const int subsampled = layout == 1 && plane != 0;
    const int64_t plane_width = subsampled ? (width + 1) / 2 : width;
    const int64_t nominal_width = subsampled ? (tile_width + 1) / 2 : tile_width;
    for (int64_t y = 0; y < height; y += tile_height) {
      for (int64_t x = 0; x < plane_width; x += nominal_width) {
        const int64_t w = nominal_width < plane_width - x ? nominal_width : plane_width - x;
        const int64_t h = tile_height < height - y ? tile_height : height - y;
        const int64_t samples = w * h;
        const int64_t shards = (samples + SHARD_SYMBOLS - 1) / SHARD_SYMBOLS;
        if (tiles) {
          fvo_tile t = {plane, x, y, w, h, folded_base, first_shard, shards};
          tiles[count] = t;
        }
        ++count;
        folded_base += samples;
        first_shard += shards;
      }
    }
  }
  return count;
}

/* Closed-form tile count (same value as fvo_tiles) so that a malformed
 * header with huge dimensions does not enumerate tiles. */
static int64_t tile_count_closed(int64_t width, int64_t height, int64_t tile_width,
                                 int64_t tile_height, int64_t layout) {
  const int64_t rows = (height + tile_height - 1) / tile_height;
  const int64_t luma = (width + tile_width - 1) / tile_width;
  if (layout == 0) return rows * luma;
  if (layout == 2) return 3 * rows * luma;
  const int64_t cw = (width + 1) / 2;
  const int64_t ctw = (tile_width + 1) / 2;
  return rows * (luma + 2 * ((cw + ctw - 1) / ctw));
}

/* ------------------------------------------------------------- predictor */

static inline uint32_t zigzag(int32_t value) {
  /* encode.cu:118-122 */
  return value >= 0 ? (uint32_t)value * 2 : (uint32_t)(-value) * 2 - 1;
}

static inline int32_t unzigzag(uint32_t value) {
  /* decode.cu:401-405 */
  return (value & 1) == 0 ? (int32_t)(value / 2) : -(int32_t)(value / 2) - 1;
}

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* MED / clamped-gradient prediction: encode.cu:168-182, decode.cu:584-598. */
static inline int32_t predict(int32_t left, int32_t above, int32_t upper_left,
                              int use_med, int32_t max_sample) {
  int32_t prediction;
  if (use_med) {
    if (upper_left >= imax(left, above)) {
      prediction = imin(left, above);
    } else if (upper_left <= imin(left, above)) {
      prediction = imax(left, above);
    } else {
      prediction = left + above - upper_left;
    }
  } else {
    prediction = left + above - upper_left;
    prediction = imax(0, imin(max_sample, prediction));
  }
  return prediction;
}

/* predict_tiles_kernel: encode.cu:128-193, visited in raster order. */
int fvo_predict_tile(const uint16_t* plane, int64_t pitch, const fvo_tile* tile,
                     int32_t step, int32_t max_sample, int use_med,
                     uint32_t* folded) {
  const int64_t w = tile->width, h = tile->height;
  uint16_t* recon = (uint16_t*)malloc((size_t)(w * h) * sizeof(uint16_t));
  int status = 0;
  for (int64_t y = 0; y < h; ++y) {
    const uint16_t* row = plane + (tile->y + y) * pitch + tile->x;
    for (int64_t x = 0; x < w; ++x) {
      const uint16_t sample = row[x];
      if (sample > max_sample) status = 1;
      const int32_t left = x > 0 ? recon[y * w + x - 1] : 0;
      const int32_t above = y > 0 ? recon[(y - 1) * w + x] : 0;
      const int32_t upper_left = x > 0 && y > 0 ? recon[(y - 1) * w + x - 1] : 0;
      const int32_t prediction = predict(left, above, upper_left, use_med, max_sample);
      const int32_t residual = (int32_t)sample - prediction;
      const int32_t magnitude = (abs(residual) + step / 2) / step;
      const int32_t quantized = residual < 0 ? -magnitude : magnitude;
      const int32_t value = imax(0, imin(max_sample, prediction + quantized * step));
      recon[y * w + x] = (uint16_t)value;
      folded[y * w + x] = zigzag(quantized);
    }
  }
  free(recon);
  return status;
}

/* --------------------------------------------------------------- analysis */

static inline int varint_length(uint32_t value) {
  /* encode.cu:108-114 */
  return value < (1u << 7) ? 1
      : value < (1u << 14) ? 2
      : value < (1u << 21) ? 3
      : value < (1u << 28) ? 4
      : 5;
}

static inline int write_varint(uint8_t* output, uint32_t value) {
  /* encode.cu:445-456 (LEB128) */
  int bytes = 0;
  do {
    uint8_t byte = (uint8_t)(value & 0x7f);
    value >>= 7;
    if (value != 0) byte |= 0x80;
    output[bytes++] = byte;
  } while (value != 0);
  return bytes;
}

static inline int bit_length(uint32_t value) {
  return value == 0 ? 0 : 32 - __builtin_clz(value);
}

/* analyze_shards_kernel: encode.cu:195-317. */
int fvo_analyze_shard(const uint32_t* v, int count, fvo_analysis* out) {
  const int lane_count = count < 4 ? count : 4;
  uint64_t lane_bits[MAX_RICE + 1][4];
  memset(lane_bits, 0, sizeof(lane_bits));
  int64_t zero_bytes = 0;
  for (int i = 0; i < count; ++i) {
    const uint32_t value = v[i];
    const int lane = i % lane_count;
    for (int k = 0; k <= MAX_RICE; ++k) {
      lane_bits[k][lane] += (uint64_t)(value >> k) + 1 + (uint64_t)k;
    }
    if (value != 0) {
      zero_bytes += varint_length(value * 2 - 1);
    } else if (i == 0 || v[i - 1] != 0) {
      uint32_t run = 1;
      while (i + (int)run < count && v[i + run] == 0) ++run;
      zero_bytes += varint_length((run - 1) * 2);
    }
  }
  int64_t block_bytes = 0;
  for (int start = 0; start < count; start += BLOCK_SYMBOLS) {
    const int symbols = count - start < BLOCK_SYMBOLS ? count - start : BLOCK_SYMBOLS;
    uint32_t maximum = 0;
    for (int i = 0; i < symbols; ++i) {
      if (v[start + i] > maximum) maximum = v[start + i];
    }
    block_bytes += 1 + (symbols * bit_length(maximum) + 7) / 8;
  }
  int best = 0;
  int64_t best_lane_bytes = INT64_MAX;
  for (int k = 0; k <= MAX_RICE; ++k) {
    int64_t bytes = 0;
    for (int lane = 0; lane < lane_count; ++lane) {
      bytes += (int64_t)((lane_bits[k][lane] + 7) / 8);
    }
    if (bytes < best_lane_bytes) {  /* first minimum: encode.cu:280 */
      best_lane_bytes = bytes;
      best = k;
    }
  }
  const int64_t rice_bytes = (lane_count - 1) * 4 + best_lane_bytes;
  int64_t mode, body;
  /* encode.cu:295-304: block only if strictly smaller than both; zero-run
   * wins ties against Rice. */
  if (block_bytes < zero_bytes && block_bytes < rice_bytes) {
    mode = MODE_BLOCK_PACK;
    body = block_bytes;
  } else if (zero_bytes <= rice_bytes) {
    mode = MODE_ZERO_RUN;
    body = zero_bytes;
  } else {
    mode = MODE_RICE_BASE + best;
    body = rice_bytes;
  }
  out->mode = mode;
  out->param = best;
  out->body = body;
  out->record = body + 3;
  for (int lane = 0; lane < 4; ++lane) {
    out->lane[lane] = lane < lane_count ? (int64_t)((lane_bits[best][lane] + 7) / 8) : 0;
  }
  return body > 65535 ? 2 : 0;
}

/* Histogram + normalisation of analyze_order0_shards_kernel,
 * encode.cu:337-392. */
int fvo_order0_frequencies(const uint32_t* v, int count, uint16_t* frequency,
                           uint16_t* cumulative) {
  uint32_t histogram[RANS_ALPHABET];
  memset(histogram, 0, sizeof(histogram));
  for (int i = 0; i < count; ++i) {
    if (v[i] >= RANS_ALPHABET) return -1;
    ++histogram[v[i]];
  }
  int distinct = 0;
  for (int value = 0; value < RANS_ALPHABET; ++value) distinct += histogram[value] != 0;
  memset(frequency, 0, RANS_ALPHABET * sizeof(uint16_t));
  memset(cumulative, 0, RANS_ALPHABET * sizeof(uint16_t));
  uint32_t assigned = 0, prefix = 0;
  int remaining = distinct;
  for (int value = 0; value < RANS_ALPHABET; ++value) {
    if (histogram[value] == 0) continue;
    --remaining;
    prefix += histogram[value];
    uint32_t end;
    if (remaining == 0) {
      end = 1u << RANS_TABLE_LOG;
    } else {
      end = (uint32_t)(((uint64_t)prefix << RANS_TABLE_LOG) / (uint64_t)count);
      if (end < assigned + 1) end = assigned + 1;
      if (end > (1u << RANS_TABLE_LOG) - (uint32_t)remaining) {
        end = (1u << RANS_TABLE_LOG) - (uint32_t)remaining;
      }
    }
    frequency[value] = (uint16_t)(end - assigned);
    cumulative[value] = (uint16_t)assigned;
    assigned = end;
  }
  return distinct;
}

/* analyze_order0_shards_kernel: encode.cu:319-431 (table bytes :361-391,
 * four byte-wise rANS lanes :398-418, body size :420-422). */
int fvo_order0_shard(const uint32_t* v, int count, uint8_t* sparse, int* sparse_len,
                     uint8_t lanes[4][2048], int lane_bytes[4], uint32_t states[4],
                     int64_t* body) {
  uint16_t frequency[RANS_ALPHABET], cumulative[RANS_ALPHABET];
  const int distinct = fvo_order0_frequencies(v, count, frequency, cumulative);
  if (distinct < 0 || distinct > 255) return 0;
  int cursor = 0;
  sparse[cursor++] = RANS_TABLE_LOG;
  cursor += write_varint(sparse + cursor, (uint32_t)distinct);
  int remaining = distinct;
  int previous = 0;
  for (int value = 0; value < RANS_ALPHABET; ++value) {
    if (frequency[value] == 0) continue;
    --remaining;
    cursor += write_varint(sparse + cursor, (uint32_t)(value - previous));
    if (remaining != 0) cursor += write_varint(sparse + cursor, frequency[value]);
    previous = value;
  }
  *sparse_len = cursor;
  int64_t total = cursor + 12 + 16;
  for (int lane = 0; lane < 4; ++lane) {
    uint32_t state = RANS_L;
    int bytes = 0;
    for (int index = count - 1 - ((count - 1 - lane) & 3); index >= 0; index -= 4) {
      const uint32_t value = v[index];
      const uint32_t freq = frequency[value];
      const uint64_t threshold = ((uint64_t)(RANS_L >> RANS_TABLE_LOG) << 8) * freq;
      while (state >= threshold) {
        lanes[lane][bytes++] = (uint8_t)state;
        state >>= 8;
      }
      state = ((state / freq) << RANS_TABLE_LOG) + state % freq + cumulative[value];
    }
    lane_bytes[lane] = bytes;
    states[lane] = state;
    total += bytes;
  }
  *body = total;
  return 1;
}

/* ---------------------------------------------------------------- encoder */

typedef struct {
  const uint16_t* const* planes;
  const int64_t* pitches;
  const fvo_tile* tiles;
  int64_t tile_count;
  int64_t total_shards;
  int64_t* shard_base;
  int32_t* shard_count;
  int64_t* shard_tile;
  uint32_t* folded;
  int32_t step;
  int32_t max_sample;
  int use_med;
  int force_order0;
  fvo_analysis* analysis;
  uint8_t* scratch;   /* SCRATCH_BYTES per shard: sparse table | lane streams */
  uint32_t* states;   /* 4 per shard */
  int* tile_status;
  int* shard_status;
} encode_ctx;

static void predict_task(void* arg, int64_t index) {
  encode_ctx* c = (encode_ctx*)arg;
  const fvo_tile* t = &c->tiles[index];
  c->tile_status[index] = fvo_predict_tile(
      c->planes[t->plane], c->pitches[t->plane], t, c->step, c->max_sample,
      c->use_med, c->folded + t->folded_base);
}

/* One shard of run_analysis (encode.cu:779-805): with force_order0 the
 * order-0 coder is taken whenever supported, else the best of three; without
 * it order-0 replaces the best of three only when strictly smaller. */
static void analyze_task(void* arg, int64_t shard) {
  encode_ctx* c = (encode_ctx*)arg;
  const uint32_t* v = c->folded + c->shard_base[shard];
  const int count = c->shard_count[shard];
  fvo_analysis* a = &c->analysis[shard];
  uint8_t* scratch = c->scratch + shard * SCRATCH_BYTES;
  uint8_t lanes[4][2048];
  int lane_bytes[4];
  int sparse_len = 0;
  int64_t body0 = 0;
  int status = 0;
  if (!c->force_order0) status = fvo_analyze_shard(v, count, a);
  const int supported = fvo_order0_shard(v, count, scratch, &sparse_len, lanes,
                                         lane_bytes, c->states + shard * 4, &body0);
  if (supported && (c->force_order0 || body0 < a->body)) {
    a->mode = MODE_ORDER0;
    a->param = RANS_TABLE_LOG;
    a->body = body0;
    a->record = body0 + 3;
    for (int lane = 0; lane < 4; ++lane) {
      a->lane[lane] = lane_bytes[lane];
      /* emission copies lane bytes reversed (encode.cu:603-609) */
      uint8_t* dst = scratch + 2048 + lane * 2048;
      for (int i = 0; i < lane_bytes[lane]; ++i) dst[i] = lanes[lane][lane_bytes[lane] - 1 - i];
    }
  } else if (c->force_order0) {
    status = fvo_analyze_shard(v, count, a);
  }
  c->shard_status[shard] = status;
}

static int run_pass(encode_ctx* c, int threads, int32_t step, int use_med) {
  c->step = step;
  c->use_med = use_med;
  parallel_for(threads, c->tile_count, predict_task, c);
  parallel_for(threads, c->total_shards, analyze_task, c);
  /* Status words are first-error; the predictor's code 1 is raised before
   * the analysis kernels run (encode.cu:784-808). */
  for (int64_t i = 0; i < c->tile_count; ++i) {
    if (c->tile_status[i]) return c->tile_status[i];
  }
  for (int64_t i = 0; i < c->total_shards; ++i) {
    if (c->shard_status[i]) return c->shard_status[i];
  }
  return 0;
}

static int64_t payload_bytes(const encode_ctx* c) {
  int64_t total = 0;
  for (int64_t i = 0; i < c->total_shards; ++i) total += c->analysis[i].record;
  return total;
}

static void put_u16(uint8_t* p, uint16_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
}
static void put_u32(uint8_t* p, uint32_t v) {
  for (int s = 0; s < 4; ++s) p[s] = (uint8_t)(v >> (8 * s));
}
static void put_u64(uint8_t* p, uint64_t v) {
  for (int s = 0; s < 8; ++s) p[s] = (uint8_t)(v >> (8 * s));
}

static inline void set_bit(uint8_t* out, int64_t position) {
  out[position >> 3] |= (uint8_t)(1u << (position & 7));
}

/* emit_shards_kernel / emit_block_shards_kernel: encode.cu:488-682.
 * out must be zeroed (encode.cu:906). */
static void emit_shard(const encode_ctx* c, int64_t shard, uint8_t* out) {
  const uint32_t* v = c->folded + c->shard_base[shard];
  const int count = c->shard_count[shard];
  const fvo_analysis* a = &c->analysis[shard];
  const int mode = (int)a->mode;
  const int body_bytes = (int)a->body;
  out[0] = (uint8_t)mode;
  put_u16(out + 1, (uint16_t)body_bytes);
  uint8_t* body = out + 3;
  if (mode == MODE_ZERO_RUN) {  /* encode.cu:514-536 */
    int cursor = 0;
    uint32_t run = 0;
    for (int i = 0; i < count; ++i) {
      if (v[i] == 0) {
        ++run;
      } else {
        if (run != 0) {
          cursor += write_varint(body + cursor, (run - 1) * 2);
          run = 0;
        }
        cursor += write_varint(body + cursor, v[i] * 2 - 1);
      }
    }
    if (run != 0) write_varint(body + cursor, (run - 1) * 2);
    return;
  }
  if (mode >= MODE_RICE_BASE && mode <= MODE_RICE_BASE + MAX_RICE) {  /* :538-585 */
    const int k = (int)a->param;
    const int lane_count = count < 4 ? count : 4;
    int64_t lane_offset = (lane_count - 1) * 4;
    for (int lane = 0; lane < lane_count; ++lane) {
      if (lane + 1 < lane_count) put_u32(body + lane * 4, (uint32_t)a->lane[lane]);
      uint8_t* lane_out = body + lane_offset;
      int64_t bit = 0;
      for (int i = lane; i < count; i += lane_count) {
        const uint32_t value = v[i];
        const uint32_t quotient = value >> k;
        set_bit(lane_out, bit + quotient);
        for (int b = 0; b < k; ++b) {
          if ((value >> b) & 1u) set_bit(lane_out, bit + quotient + 1 + b);
        }
        bit += (int64_t)quotient + 1 + k;
      }
      lane_offset += a->lane[lane];
    }
    return;
  }
  if (mode == MODE_BLOCK_PACK) {  /* encode.cu:614-682 */
    int offset = 0;
    for (int start = 0; start < count; start += BLOCK_SYMBOLS) {
      const int symbols = count - start < BLOCK_SYMBOLS ? count - start : BLOCK_SYMBOLS;
      uint32_t maximum = 0;
      for (int i = 0; i < symbols; ++i) {
        if (v[start + i] > maximum) maximum = v[start + i];
      }
      const int width = bit_length(maximum);
      body[offset] = (uint8_t)width;
      uint8_t* packed = body + offset + 1;
      for (int i = 0; i < symbols; ++i) {
        for (int b = 0; b < width; ++b) {
          if ((v[start + i] >> b) & 1u) set_bit(packed, (int64_t)i * width + b);
        }
      }
      offset += 1 + (symbols * width + 7) / 8;
    }
    return;
  }
  /* order-0: encode.cu:587-611 */
  const uint8_t* scratch = c->scratch + shard * SCRATCH_BYTES;
  const int sparse = body_bytes - (int)(a->lane[0] + a->lane[1] + a->lane[2] + a->lane[3]) - 28;
  memcpy(body, scratch, (size_t)sparse);
  for (int lane = 0; lane < 3; ++lane) put_u32(body + sparse + lane * 4, (uint32_t)a->lane[lane]);
  int start = sparse + 28;
  for (int lane = 0; lane < 4; ++lane) {
    put_u32(body + sparse + 12 + lane * 4, c->states[shard * 4 + lane]);
    memcpy(body + start, scratch + 2048 + lane * 2048, (size_t)a->lane[lane]);
    start += (int)a->lane[lane];
  }
}

typedef struct {
  const encode_ctx* c;
  const int64_t* offsets;
  uint8_t* stream;
} emit_ctx;

static void emit_task(void* arg, int64_t shard) {
  emit_ctx* e = (emit_ctx*)arg;
  emit_shard(e->c, shard, e->stream + e->offsets[shard]);
}

int fvo_encode(const uint16_t* const* planes, const int64_t* pitches,
               int64_t plane_count, int64_t width, int64_t height,
               int64_t layout, int64_t bit_depth, int64_t quality,
               int64_t fps_num, int64_t fps_den, int64_t tile_width,
               int64_t tile_height, int threads, uint8_t** out,
               int64_t* out_len, uint32_t* folded_out,
               fvo_analysis* analysis_out, char* msg, size_t msg_len) {
  return fvo_encode_traced(planes, pitches, plane_count, width, height, layout, bit_depth,
                           quality, fps_num, fps_den, tile_width, tile_height, threads, out,
                           out_len, folded_out, analysis_out, NULL, msg, msg_len);
}

/* fastvid_encode_cuda host orchestration: encode.cu:686-931. */
int fvo_encode_traced(const uint16_t* const* planes, const int64_t* pitches,
                      int64_t plane_count, int64_t width, int64_t height,
                      int64_t layout, int64_t bit_depth, int64_t quality,
                      int64_t fps_num, int64_t fps_den, int64_t tile_width,
                      int64_t tile_height, int threads, uint8_t** out,
                      int64_t* out_len, uint32_t* folded_out,
                      fvo_analysis* analysis_out, fvo_trace* trace, char* msg,
                      size_t msg_len) {
  *out = NULL;
  if (trace) {
    trace->baseline_payload = trace->refined_payload = -1;
    trace->branch = FVO_BRANCH_NONE;
    trace->passes = 0;
  }
  *out_len = 0;
  /* argument checks: encode.cu:695-716 */
  if (!(layout >= 0 && layout <= 2)) {
    set_msg(msg, msg_len, "layout must be gray, yuv422, or rgb444");
    return -1;
  }
  if (plane_count != (layout == 0 ? 1 : 3)) {
    set_msg(msg, msg_len, "plane count does not match layout");
    return -1;
  }
  if (!(bit_depth == 8 || bit_depth == 10 || bit_depth == 12 || bit_depth == 16)) {
    set_msg(msg, msg_len, "bit_depth must be 8, 10, 12, or 16");
    return -1;
  }
  if (!(quality >= 1 && quality <= 100)) {
    set_msg(msg, msg_len, "quality must be in 1..=100");
    return -1;
  }
  if (!(fps_num > 0 && fps_num <= 0xffffffffLL && fps_den > 0 && fps_den <= 0xffffffffLL)) {
    set_msg(msg, msg_len, "frame-rate components must be nonzero u32 values");
    return -1;
  }
  if (!(tile_width > 0 && tile_width <= 65535 && tile_height > 0 && tile_height <= 65535)) {
    set_msg(msg, msg_len, "tile dimensions must be nonzero u16 values");
    return -1;
  }
  if (!(width > 0 && height > 0 && width <= 0xffffffffLL && height <= 0xffffffffLL)) {
    set_msg(msg, msg_len, "frame dimensions must be nonzero u32 values");
    return -1;
  }
  const int64_t tile_count = tile_count_closed(width, height, tile_width, tile_height, layout);
  if (tile_count > (1 << 20)) {
    set_msg(msg, msg_len, "too many tiles");
    return -1;
  }
  fvo_tile* tiles = (fvo_tile*)malloc((size_t)tile_count * sizeof(fvo_tile));
  fvo_tiles(width, height, tile_width, tile_height, layout, tiles);
  const int64_t total_samples = tiles[tile_count - 1].folded_base +
      tiles[tile_count - 1].width * tiles[tile_count - 1].height;
  const int64_t total_shards = tiles[tile_count - 1].first_shard + tiles[tile_count - 1].shard_count;

  encode_ctx c;
  memset(&c, 0, sizeof(c));
  c.planes = planes;
  c.pitches = pitches;
  c.tiles = tiles;
  c.tile_count = tile_count;
  c.total_shards = total_shards;
  c.shard_base = (int64_t*)malloc((size_t)total_shards * sizeof(int64_t));
  c.shard_count = (int32_t*)malloc((size_t)total_shards * sizeof(int32_t));
  c.shard_tile = (int64_t*)malloc((size_t)total_shards * sizeof(int64_t));
  for (int64_t t = 0; t < tile_count; ++t) {
    const int64_t samples = tiles[t].width * tiles[t].height;
    for (int64_t s = 0; s < tiles[t].shard_count; ++s) {
      const int64_t index = tiles[t].first_shard + s;
      const int64_t local = s * SHARD_SYMBOLS;
      c.shard_base[index] = tiles[t].folded_base + local;
      c.shard_count[index] = (int32_t)(samples - local < SHARD_SYMBOLS ? samples - local : SHARD_SYMBOLS);
      c.shard_tile[index] = t;
    }
  }
  c.folded = (uint32_t*)malloc((size_t)total_samples * sizeof(uint32_t));
  c.analysis = (fvo_analysis*)calloc((size_t)total_shards, sizeof(fvo_analysis));
  c.scratch = (uint8_t*)malloc((size_t)total_shards * SCRATCH_BYTES);
  c.states = (uint32_t*)malloc((size_t)total_shards * 4 * sizeof(uint32_t));
  c.tile_status = (int*)calloc((size_t)tile_count, sizeof(int));
  c.shard_status = (int*)calloc((size_t)total_shards, sizeof(int));
  c.max_sample = (int32_t)((1 << bit_depth) - 1);
  /* encode.cu:778 */
  c.force_order0 = layout == 2 && (bit_depth != 10 || width < 3840);

  int rc = 0;
  int32_t step = fvo_quantization_step(layout, bit_depth, quality);
  int passes = 1;
  int branch = FVO_BRANCH_NONE;
  int status = run_pass(&c, threads, step, layout != 0);
  const int64_t baseline_payload = payload_bytes(&c);
  int64_t refined_payload = -1;
  int gray_med = 0, refined = 0;
  /* gray10 MED re-run: encode.cu:818-823 */
  if (!status) {
    gray_med = layout == 0 && bit_depth == 10 && step > 1 &&
        baseline_payload * 10 > total_samples * 3;
    if (gray_med) {
      status = run_pass(&c, threads, step, 1);
      ++passes;
      branch = FVO_BRANCH_GRAY10_MED;
    }
  }
  /* gray16 refinement: encode.cu:824-847 */
  if (!status && layout == 0 && bit_depth == 16 && step > 1) {
    const int direct = baseline_payload < total_samples / 10 ||
        baseline_payload > total_samples / 2;
    const int probe = !direct && baseline_payload < total_samples / 5;
    if (direct || probe) {
      const int32_t baseline_step = step;
      step = fvo_refined_gray16_step(quality);
      status = run_pass(&c, threads, step, 0);
      ++passes;
      if (!status) {
        refined_payload = payload_bytes(&c);
        refined = direct || refined_payload < total_samples / 10;
        branch = direct ? (baseline_payload < total_samples / 10 ? FVO_BRANCH_GRAY16_DIRECT_LOW
                                                                 : FVO_BRANCH_GRAY16_DIRECT_HIGH)
                        : (refined ? FVO_BRANCH_GRAY16_PROBE_KEPT : FVO_BRANCH_GRAY16_PROBE_REVERTED);
        if (!refined) {
          step = baseline_step;
          status = run_pass(&c, threads, step, 0);
          ++passes;
        }
      }
    }
  }
  if (trace) {
    trace->baseline_payload = baseline_payload;
    trace->refined_payload = refined_payload;
    trace->branch = branch;
    trace->passes = passes;
  }
  if (status) {
    char text[96];
    snprintf(text, sizeof(text), "CUDA encoder analysis failed (status %d)", status);
    set_msg(msg, msg_len, text);
    rc = status;
    goto done;
  }
  {
    /* assembly: encode.cu:848-904 */
    int64_t* tile_lengths = (int64_t*)calloc((size_t)tile_count, sizeof(int64_t));
    int64_t* offsets = (int64_t*)malloc((size_t)total_shards * sizeof(int64_t));
    const int64_t payload_start = HEADER_BYTES + tile_count * DIRECTORY_ENTRY_BYTES;
    int64_t length = payload_start;
    for (int64_t s = 0; s < total_shards; ++s) {
      tile_lengths[c.shard_tile[s]] += c.analysis[s].record;
      offsets[s] = length;
      length += c.analysis[s].record;
    }
    uint8_t* stream = (uint8_t*)calloc((size_t)length + 3, 1);
    memcpy(stream, "FVID", 4);
    stream[4] = 1;
    stream[5] = (uint8_t)layout;
    stream[6] = (uint8_t)quality;
    stream[7] = (uint8_t)((bit_depth - 8) | (gray_med ? 0x40 : 0) | (refined ? 0x80 : 0));
    put_u32(stream + 8, (uint32_t)width);
    put_u32(stream + 12, (uint32_t)height);
    put_u16(stream + 16, (uint16_t)tile_width);
    put_u16(stream + 18, (uint16_t)tile_height);
    put_u32(stream + 20, (uint32_t)fps_num);
    put_u32(stream + 24, (uint32_t)fps_den);
    put_u32(stream + 28, (uint32_t)tile_count);
    int64_t tile_offset = payload_start;
    for (int64_t t = 0; t < tile_count; ++t) {
      uint8_t* e = stream + HEADER_BYTES + t * DIRECTORY_ENTRY_BYTES;
      e[0] = (uint8_t)tiles[t].plane;
      e[1] = ENTROPY_PARALLEL_SHARDS;
      e[2] = PREDICTOR_ID;
      e[3] = 0;
      put_u32(e + 4, (uint32_t)tiles[t].x);
      put_u32(e + 8, (uint32_t)tiles[t].y);
      put_u32(e + 12, (uint32_t)tiles[t].width);
      put_u32(e + 16, (uint32_t)tiles[t].height);
      put_u64(e + 20, (uint64_t)tile_offset);
      if (tile_lengths[t] > 0xffffffffLL) {
        set_msg(msg, msg_len, "tile payload is too large");
        free(stream);
        free(tile_lengths);
        free(offsets);
        rc = -1;
        goto done;
      }
      put_u32(e + 28, (uint32_t)tile_lengths[t]);
      tile_offset += tile_lengths[t];
    }
    emit_ctx e = {&c, offsets, stream};
    parallel_for(threads, total_shards, emit_task, &e);
    *out = stream;
    *out_len = length;
    if (folded_out) memcpy(folded_out, c.folded, (size_t)total_samples * sizeof(uint32_t));
    if (analysis_out) memcpy(analysis_out, c.analysis, (size_t)total_shards * sizeof(fvo_analysis));
    free(tile_lengths);
    free(offsets);
  }
done:
  free(tiles);
  free(c.shard_base);
  free(c.shard_count);
  free(c.shard_tile);
  free(c.folded);
  free(c.analysis);
  free(c.scratch);
  free(c.states);
  free(c.tile_status);
  free(c.shard_status);
  return rc;
}

/* ---------------------------------------------------------------- decoder */

static uint16_t get_u16(const uint8_t* b) { return (uint16_t)(b[0] | (b[1] << 8)); }
static uint32_t get_u32(const uint8_t* b) {
  return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}
static uint64_t get_u64(const uint8_t* b) {
  return (uint64_t)get_u32(b) | ((uint64_t)get_u32(b + 4) << 32);
}

/* Header validation: fastvid_cuda.cpp:141-183 (messages verbatim). */
int fvo_parse_header(const uint8_t* bytes, int64_t size, fvo_header* h,
                     char* msg, size_t msg_len) {
  memset(h, 0, sizeof(*h));
  if (size < HEADER_BYTES) { set_msg(msg, msg_len, "truncated Fastvid header"); return -1; }
  if (memcmp(bytes, "FVID", 4) != 0) { set_msg(msg, msg_len, "bad Fastvid magic"); return -1; }
  if (bytes[4] != 1) { set_msg(msg, msg_len, "unsupported Fastvid bitstream version"); return -1; }
  if (bytes[5] > 2) { set_msg(msg, msg_len, "unknown pixel layout"); return -1; }
  h->layout = bytes[5];
  const int gray = h->layout == 0;
  h->quality = bytes[6];
  if (!(h->quality >= 1 && h->quality <= 100)) { set_msg(msg, msg_len, "quality is out of range"); return -1; }
  h->refined_gray16 = (bytes[7] & 0x80) != 0;
  h->gray_med = (bytes[7] & 0x40) != 0;
  h->bit_depth = (int64_t)(bytes[7] & 0x3f) + 8;
  if (!(h->bit_depth == 8 || h->bit_depth == 10 || h->bit_depth == 12 || h->bit_depth == 16)) {
    set_msg(msg, msg_len, "unsupported high-bit depth");
    return -1;
  }
  if (h->refined_gray16 && !(gray && h->bit_depth == 16 && h->quality < 100)) {
    set_msg(msg, msg_len, "invalid gray16 refinement flag");
    return -1;
  }
  if (h->gray_med && !(gray && h->bit_depth == 10 && h->quality < 100)) {
    set_msg(msg, msg_len, "invalid gray MED flag");
    return -1;
  }
  h->width = get_u32(bytes + 8);
  h->height = get_u32(bytes + 12);
  h->tile_width = get_u16(bytes + 16);
  h->tile_height = get_u16(bytes + 18);
  if (!(h->width > 0 && h->height > 0 && h->tile_width > 0 && h->tile_height > 0)) {
    set_msg(msg, msg_len, "zero dimension");
    return -1;
  }
  if (!(get_u32(bytes + 20) != 0 && get_u32(bytes + 24) != 0)) {
    set_msg(msg, msg_len, "zero frame-rate component");
    return -1;
  }
  h->tile_count = get_u32(bytes + 28);
  if (h->tile_count != tile_count_closed(h->width, h->height, h->tile_width, h->tile_height, h->layout)) {
    set_msg(msg, msg_len, "tile count does not match dimensions");
    return -1;
  }
  if (h->tile_count > (1 << 20)) { set_msg(msg, msg_len, "too many tiles"); return -1; }
  if (HEADER_BYTES + h->tile_count * DIRECTORY_ENTRY_BYTES > size) {
    set_msg(msg, msg_len, "truncated tile directory");
    return -1;
  }
  const int64_t secondary_width = h->layout == 1 ? (h->width + 1) / 2 : h->width;
  h->y_samples = h->width * h->height;
  h->secondary_samples = gray ? 0 : secondary_width * h->height;
  h->total_samples = h->y_samples + 2 * h->secondary_samples;
  h->step = h->refined_gray16 ? fvo_refined_gray16_step(h->quality)
                              : fvo_quantization_step(h->layout, h->bit_depth, h->quality);
  h->max_sample = (1LL << h->bit_depth) - 1;
  h->use_med = h->gray_med || !gray;
  return 0;
}

typedef struct {
  int64_t mode, body_offset, body_length, sample_count, output_offset;
} shard_meta;

/* Directory + shard-header walk of the DRAM-input path,
 * fastvid_cuda.cpp:255-333 (messages verbatim).  For any stream both this
 * walk and the device parse_metadata_kernel (decode.cu:131-220) accept
 * exactly the same directories. */
static int walk_directory(const uint8_t* bytes, int64_t size, const fvo_header* h,
                          fvo_tile* tiles, shard_meta* meta, int64_t* shard_total,
                          char* msg, size_t msg_len) {
  const int64_t directory_end = HEADER_BYTES + h->tile_count * DIRECTORY_ENTRY_BYTES;
  int64_t next_payload = directory_end;
  int64_t* offsets = (int64_t*)malloc((size_t)h->tile_count * sizeof(int64_t));
  int64_t* lengths = (int64_t*)malloc((size_t)h->tile_count * sizeof(int64_t));
  int rc = 0;
  for (int64_t i = 0; i < h->tile_count; ++i) {
    const uint8_t* e = bytes + HEADER_BYTES + i * DIRECTORY_ENTRY_BYTES;
    if (e[0] != tiles[i].plane) { set_msg(msg, msg_len, "non-canonical tile plane"); rc = -2; break; }
    if (e[1] != ENTROPY_PARALLEL_SHARDS) { set_msg(msg, msg_len, "tile requires bounded-shard entropy"); rc = -2; break; }
    if (e[2] != PREDICTOR_ID) { set_msg(msg, msg_len, "tile requires full-tile clamp-gradient prediction"); rc = -2; break; }
    if (e[3] != 0) { set_msg(msg, msg_len, "nonzero reserved directory byte"); rc = -2; break; }
    if (!(get_u32(e + 4) == tiles[i].x && get_u32(e + 8) == tiles[i].y &&
          get_u32(e + 12) == tiles[i].width && get_u32(e + 16) == tiles[i].height)) {
      set_msg(msg, msg_len, "non-canonical tile directory");
      rc = -2;
      break;
    }
    const uint64_t offset = get_u64(e + 20);
    if (offset > (uint64_t)INT64_MAX) { set_msg(msg, msg_len, "payload offset is too large"); rc = -2; break; }
    offsets[i] = (int64_t)offset;
    lengths[i] = get_u32(e + 28);
    if (offsets[i] != next_payload) { set_msg(msg, msg_len, "payloads are not contiguous"); rc = -2; break; }
    if (lengths[i] > size - next_payload) { set_msg(msg, msg_len, "truncated tile payload"); rc = -2; break; }
    next_payload += lengths[i];
  }
  if (!rc && next_payload != size) { set_msg(msg, msg_len, "trailing stream bytes"); rc = -2; }
  int64_t shard_index = 0;
  for (int64_t i = 0; !rc && i < h->tile_count; ++i) {
    int64_t cursor = offsets[i];
    int64_t decoded = 0;
    const int64_t samples = tiles[i].width * tiles[i].height;
    const int64_t tile_end = cursor + lengths[i];
    while (decoded < samples) {
      if (!(cursor + 3 <= tile_end)) { set_msg(msg, msg_len, "truncated entropy shard header"); rc = -2; break; }
      const int64_t mode = bytes[cursor];
      const int64_t body_length = get_u16(bytes + cursor + 1);
      const int64_t body_offset = cursor + 3;
      if (mode > 19) { set_msg(msg, msg_len, "unknown entropy shard mode"); rc = -2; break; }
      if (body_length > tile_end - body_offset) { set_msg(msg, msg_len, "truncated entropy shard body"); rc = -2; break; }
      const int64_t count = samples - decoded < SHARD_SYMBOLS ? samples - decoded : SHARD_SYMBOLS;
      shard_meta m = {mode, body_offset, body_length, count, tiles[i].folded_base + decoded};
      meta[shard_index++] = m;
      decoded += count;
      cursor = body_offset + body_length;
    }
    if (!rc && cursor != tile_end) { set_msg(msg, msg_len, "trailing bounded-shard tile bytes"); rc = -2; }
  }
  *shard_total = shard_index;
  free(offsets);
  free(lengths);
  return rc;
}

/* Zero-run / Rice / block shard decode: decode_legacy_shard,
 * decode.cu:222-399.  Returns the reference status code (0 on success). */
static int decode_legacy(const uint8_t* body, int64_t body_length, int64_t mode,
                         int64_t sample_count, uint32_t max_folded, uint32_t* output) {
  if (mode == 0) {  /* decode.cu:237-295 */
    int64_t cursor = 0, produced = 0;
    while (produced < sample_count) {
      uint32_t token = 0;
      int shift = 0, used = 0;
      for (;;) {
        if (cursor >= body_length || used == 5) return 1;
        const uint8_t byte = body[cursor++];
        if (used == 4 && byte > 0x0f) return 1;
        token |= (uint32_t)(byte & 0x7f) << shift;
        ++used;
        if ((byte & 0x80) == 0) break;
        shift += 7;
      }
      if (used != varint_length(token)) return 1;
      if ((token & 1) == 0) {
        const int64_t run = token / 2 + 1;
        if (run > sample_count - produced) return 1;
        for (int64_t i = 0; i < run; ++i) output[produced++] = 0;
      } else {
        const uint32_t value = (token + 1) / 2;
        if (value == 0 || value > max_folded) return 1;
        output[produced++] = value;
      }
    }
    return cursor != body_length ? 1 : 0;
  }
  if (mode >= 1 && mode <= 17) {  /* decode.cu:297-340 with RiceBitReader :63-125 */
    const int k = (int)(mode - 1);
    const int lane_count = (int)(sample_count < 4 ? sample_count : 4);
    const int64_t table_bytes = (int64_t)(lane_count - 1) * 4;
    for (int lane = 0; lane < lane_count; ++lane) {
      if (table_bytes > body_length) return 2;
      int64_t lane_offset = table_bytes;
      for (int i = 0; i < lane; ++i) lane_offset += get_u32(body + i * 4);
      const int64_t lane_length = lane + 1 < lane_count ? (int64_t)get_u32(body + lane * 4)
                                                        : body_length - lane_offset;
      if (lane_offset > body_length || lane_length < 0 || lane_length > body_length - lane_offset) return 2;
      const uint8_t* data = body + lane_offset;
      int64_t bit = 0;
      const int64_t total_bits = lane_length * 8;
      for (int64_t index = lane; index < sample_count; index += lane_count) {
        uint64_t quotient = 0;
        for (;;) {  /* unary part: zeros then a one; fails past the end */
          if (bit >= total_bits) return 2;
          const int b = (data[bit >> 3] >> (bit & 7)) & 1;
          ++bit;
          if (b) break;
          ++quotient;
        }
        if (quotient > max_folded) return 2;
        if (bit + k > total_bits) return 2;
        uint32_t remainder = 0;
        for (int b = 0; b < k; ++b, ++bit) {
          remainder |= (uint32_t)((data[bit >> 3] >> (bit & 7)) & 1) << b;
        }
        const uint64_t result = (quotient << k) + remainder;
        if (result > max_folded) return 2;
        output[index] = (uint32_t)result;
      }
      const int64_t used_bytes = (bit + 7) / 8;
      if (used_bytes != lane_length) return 2;
      if ((bit & 7) != 0) {
        const uint8_t mask = (uint8_t)~((1u << (bit & 7)) - 1);
        if ((data[used_bytes - 1] & mask) != 0) return 2;
      }
    }
    return 0;
  }
  if (mode == 18) {  /* decode.cu:342-396 as one sequential walk */
    const int64_t block_count = (sample_count + BLOCK_SYMBOLS - 1) / BLOCK_SYMBOLS;
    const int max_width = bit_length(max_folded);
    int64_t cursor = 0;
    for (int64_t block = 0; block < block_count; ++block) {
      if (cursor >= body_length) return 3;
      const int width = body[cursor++];
      if (width > max_width) return 3;
      const int64_t remaining = sample_count - block * BLOCK_SYMBOLS;
      const int64_t count = remaining < BLOCK_SYMBOLS ? remaining : BLOCK_SYMBOLS;
      const int64_t payload = (count * width + 7) / 8;
      if (payload > body_length - cursor) return 3;
      const uint8_t* packed = body + cursor;
      int64_t bit = 0;
      for (int64_t i = 0; i < count; ++i) {
        uint32_t value = 0;
        for (int b = 0; b < width; ++b, ++bit) {
          value |= (uint32_t)((packed[bit >> 3] >> (bit & 7)) & 1) << b;
        }
        if (value > max_folded) return 3;
        output[block * BLOCK_SYMBOLS + i] = value;
      }
      if (block + 1 == block_count && cursor + payload != body_length) return 3;
      if ((bit & 7) != 0) {
        const uint8_t mask = (uint8_t)~((1u << (bit & 7)) - 1);
        if ((packed[payload - 1] & mask) != 0) return 3;
      }
      cursor += payload;
    }
    return 0;
  }
  return 4;
}

/* read_varint_order0: decode.cu:407-426 (canonical LEB128, at most 5 bytes). */
static int read_varint_order0(const uint8_t* body, int64_t length, int64_t* cursor, uint32_t* value) {
  uint32_t result = 0;
  int shift = 0;
  for (int used = 0; used < 5; ++used) {
    if (*cursor >= length) return 0;
    const uint8_t byte = body[(*cursor)++];
    if (used == 4 && byte > 0x0f) return 0;
    result |= (uint32_t)(byte & 0x7f) << shift;
    if ((byte & 0x80) == 0) {
      if (varint_length(result) != used + 1) return 0;
      *value = result;
      return 1;
    }
    shift += 7;
  }
  return 0;
}

/* decode_order0_shards_kernel, order-0 branch: decode.cu:441-549.  One
 * deliberate difference: a frequency whose u32 sum with the assigned total
 * wraps past 2^32 is rejected here.  The reference accepts it and then fills
 * its 4096-slot lookup table out of bounds (undefined behaviour), so no
 * defined reference output exists for such a stream. */
static int decode_order0(const uint8_t* body, int64_t body_length, int64_t sample_count,
                         uint32_t max_folded, uint32_t* output) {
  uint16_t values[RANS_ALPHABET], frequency[RANS_ALPHABET], cumulative[RANS_ALPHABET];
  uint8_t table[1 << RANS_TABLE_LOG];
  uint32_t states[4];
  int64_t lane_start[4], lane_length[4];
  int64_t cursor = 0;
  if (body_length < 1 || body[cursor++] != RANS_TABLE_LOG) return 5;
  uint32_t n = 0;
  if (!read_varint_order0(body, body_length, &cursor, &n)) return 5;
  if (n == 0 || n > 255) return 5;
  uint32_t previous = 0, assigned = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t delta = 0;
    if (!read_varint_order0(body, body_length, &cursor, &delta) || (i != 0 && delta == 0) ||
        previous + delta > max_folded) {
      return 5;
    }
    const uint32_t value = previous + delta;
    uint32_t freq = (1u << RANS_TABLE_LOG) - assigned;
    if (i + 1 != n && !read_varint_order0(body, body_length, &cursor, &freq)) return 5;
    if ((uint64_t)assigned + freq > (1u << RANS_TABLE_LOG)) return 5;  /* wrap rejected */
    if (freq == 0 || assigned + freq > (1u << RANS_TABLE_LOG) ||
        (i + 1 != n && assigned + freq == (1u << RANS_TABLE_LOG))) {
      return 5;
    }
    values[i] = (uint16_t)value;
    frequency[i] = (uint16_t)freq;
    cumulative[i] = (uint16_t)assigned;
    assigned += freq;
    previous = value;
  }
  if (assigned != (1u << RANS_TABLE_LOG)) return 5;
  int32_t total_payload = 0;
  for (int lane = 0; lane < 3; ++lane) {
    if (cursor + 4 > body_length) return 5;
    lane_length[lane] = (int32_t)get_u32(body + cursor);  /* static_cast<int> */
    total_payload = (int32_t)((uint32_t)total_payload + (uint32_t)lane_length[lane]);
    cursor += 4;
  }
  if (cursor + 16 > body_length) return 5;
  for (int lane = 0; lane < 4; ++lane) {
    states[lane] = get_u32(body + cursor + lane * 4);
    if (states[lane] < RANS_L) return 5;
  }
  cursor += 16;
  lane_length[3] = (int32_t)((uint32_t)(int32_t)(body_length - cursor) - (uint32_t)total_payload);
  if (lane_length[3] < 0) return 5;
  /* A negative length for lanes 0-2 makes the reference fail that lane's
   * first byte read or its final length check; rejecting it up front gives
   * the same verdict without reading outside the body. */
  for (int lane = 0; lane < 3; ++lane) {
    if (lane_length[lane] < 0) return 5;
  }
  for (int lane = 0; lane < 4; ++lane) {
    lane_start[lane] = (int32_t)cursor;
    cursor += lane_length[lane];
  }
  if (cursor != body_length) return 5;
  for (uint32_t s = 0; s < n; ++s) {
    for (int slot = cumulative[s]; slot < cumulative[s] + frequency[s]; ++slot) table[slot] = (uint8_t)s;
  }
  for (int lane = 0; lane < 4; ++lane) {
    uint32_t state = states[lane];
    int64_t read = 0;
    for (int64_t index = lane; index < sample_count; index += 4) {
      const uint32_t slot = state & ((1u << RANS_TABLE_LOG) - 1);
      const uint8_t symbol = table[slot];
      output[index] = values[symbol];
      state = (uint32_t)frequency[symbol] * (state >> RANS_TABLE_LOG) + slot - cumulative[symbol];
      while (state < RANS_L) {
        if (read >= lane_length[lane]) return 5;
        state = (state << 8) | body[lane_start[lane] + read++];
      }
    }
    if (read != lane_length[lane] || state != RANS_L) return 5;
  }
  return 0;
}

typedef struct {
  const uint8_t* bytes;
  const shard_meta* meta;
  const fvo_tile* tiles;
  uint32_t max_folded;
  uint32_t* folded;
  uint16_t* output;
  const fvo_header* h;
  int64_t plane_base[3], plane_width[3];
  int* status;
} decode_ctx;

static void entropy_task(void* arg, int64_t shard) {
  decode_ctx* d = (decode_ctx*)arg;
  const shard_meta* m = &d->meta[shard];
  const uint8_t* body = d->bytes + m->body_offset;
  uint32_t* out = d->folded + m->output_offset;
  d->status[shard] = m->mode == MODE_ORDER0
      ? decode_order0(body, m->body_length, m->sample_count, d->max_folded, out)
      : decode_legacy(body, m->body_length, m->mode, m->sample_count, d->max_folded, out);
}

/* reconstruct_tiles_serial_kernel: decode.cu:608-656 (the wavefront kernel
 * decode.cu:552-606 produces the same values). */
static void reconstruct_task(void* arg, int64_t index) {
  decode_ctx* d = (decode_ctx*)arg;
  const fvo_tile* t = &d->tiles[index];
  const int64_t pitch = d->plane_width[t->plane];
  uint16_t* plane = d->output + d->plane_base[t->plane];
  const uint32_t* f = d->folded + t->folded_base;
  const int32_t step = (int32_t)d->h->step;
  const int32_t max_sample = (int32_t)d->h->max_sample;
  const int use_med = (int)d->h->use_med;
  for (int64_t y = 0; y < t->height; ++y) {
    int32_t left = 0, upper_left = 0;
    for (int64_t x = 0; x < t->width; ++x) {
      const int64_t dst = (t->y + y) * pitch + t->x + x;
      const int32_t above = y > 0 ? plane[dst - pitch] : 0;
      const int32_t prediction = predict(left, above, upper_left, use_med, max_sample);
      int32_t value = prediction + unzigzag(f[y * t->width + x]) * step;
      value = imax(0, imin(max_sample, value));
      plane[dst] = (uint16_t)value;
      upper_left = above;
      left = value;
    }
  }
}

static int decode_common(const uint8_t* bytes, int64_t size, uint16_t* output,
                         int64_t capacity, uint32_t* folded_only, int threads,
                         char* msg, size_t msg_len) {
  fvo_header h;
  if (fvo_parse_header(bytes, size, &h, msg, msg_len)) return -1;
  if (output && capacity < h.total_samples) { set_msg(msg, msg_len, "output too small"); return -1; }
  if (folded_only && capacity < h.total_samples) { set_msg(msg, msg_len, "output too small"); return -1; }
  fvo_tile* tiles = (fvo_tile*)malloc((size_t)h.tile_count * sizeof(fvo_tile));
  fvo_tiles(h.width, h.height, h.tile_width, h.tile_height, h.layout, tiles);
  const int64_t shards = tiles[h.tile_count - 1].first_shard + tiles[h.tile_count - 1].shard_count;
  shard_meta* meta = (shard_meta*)malloc((size_t)shards * sizeof(shard_meta));
  int64_t walked = 0;
  int rc = walk_directory(bytes, size, &h, tiles, meta, &walked, msg, msg_len);
  if (rc) {
    free(tiles);
    free(meta);
    return rc;
  }
  decode_ctx d;
  memset(&d, 0, sizeof(d));
  d.bytes = bytes;
  d.meta = meta;
  d.tiles = tiles;
  d.max_folded = (uint32_t)(h.max_sample * 2);
  d.folded = folded_only ? folded_only : (uint32_t*)malloc((size_t)h.total_samples * sizeof(uint32_t));
  d.output = output;
  d.h = &h;
  const int64_t secondary_width = h.layout == 1 ? (h.width + 1) / 2 : h.width;
  d.plane_base[0] = 0;
  d.plane_base[1] = h.y_samples;
  d.plane_base[2] = h.y_samples + h.secondary_samples;
  d.plane_width[0] = h.width;
  d.plane_width[1] = secondary_width;
  d.plane_width[2] = secondary_width;
  d.status = (int*)calloc((size_t)shards, sizeof(int));
  parallel_for(threads, shards, entropy_task, &d);
  for (int64_t s = 0; s < shards; ++s) {
    if (d.status[s]) {
      char text[96];
      snprintf(text, sizeof(text), "malformed entropy payload (CUDA status %d)", d.status[s]);
      set_msg(msg, msg_len, text);
      rc = d.status[s];
      break;
    }
  }
  if (!rc && output) parallel_for(threads, h.tile_count, reconstruct_task, &d);
  if (!folded_only) free(d.folded);
  free(d.status);
  free(tiles);
  free(meta);
  return rc;
}

int fvo_decode(const uint8_t* stream, int64_t size, uint16_t* output,
               int64_t output_capacity, int threads, char* msg, size_t msg_len) {
  return decode_common(stream, size, output, output_capacity, NULL, threads, msg, msg_len);
}

int fvo_decode_folded(const uint8_t* stream, int64_t size, uint32_t* folded,
                      int64_t capacity, char* msg, size_t msg_len) {
  return decode_common(stream, size, NULL, capacity, folded, 1, msg, msg_len);
}
