/*
 * isp_oracle.c - TEST INFRASTRUCTURE ONLY: the CPU oracle of the ISP.
 *
 * A plain-C restatement, expression by expression, of the reference's CPU
 * backend: per-pixel arithmetic from modules/isp/src/isp_math.hpp, the frame
 * set-up, border handling and orchestration from modules/isp/src/cpu_pipeline.cpp,
 * and the tone table from modules/isp/src/isp_params.cpp:166-182. Built with
 * -ffp-contract=off and no fast-math, like the reference (CMakeLists.txt:62-70),
 * so every float operation rounds as the reference's does.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
 * as the checker. The product library never links or loads it.
 *
 * Parity pins (tests/test_oracle.py):
 *  - every function against the reference CPU backend compiled from its own
 *    sources (oracle/_ref, built by oracle/Makefile), byte for byte on seeded
 *    frames of every pattern, format, demosaic, white balance, tone curve and
 *    output format, and against the golden hashes in tests/golden/ that the
 *    compiled reference wrote;
 *  - closed forms: mirror() against reflection by brute force, RAW10/RAW12
 *    against the MIPI CSI-2 bit layout, bilinear against an independent double
 *    reference (< 1e-6), Malvar on linear gradients (< 2e-6), colour matrix,
 *    tone formulas, flat-field statistics and AE (error 1 EV, scale 2, 20 ms at
 *    gain 1.5), histogram rebuild from the image, black frame.
 *  - the 1024-bin histogram extension (no reference counterpart) is pinned by
 *    rebuilding it from the written RGB16 / RGB32F image only.
 *  - the colour conversion (oracle_color_convert) against the published colour
 *    bars (ITU-R BT.801, SMPTE RP 219, JFIF) and the module's own CPU converter
 *    compiled from its sources (oracle/_ref), byte for byte.
 *  - the frame checker (oracle_frame_check) by the reference demo's published
 *    run (applications/isp/README.md:89: mean and worst PSNR, mean luma and AE
 *    error of 30 checked 640x480 frames) and by closed forms: a perfect image
 *    scores 0 error and 99 dB, a pixel off by d adds d^2, a changed bin counts
 *    one mismatch, and the rebuilt rows equal the ISP's histograms.
 */
#include "isp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define FIXED_ONE 1048576.0F /* isp_math.hpp:30 */
#define AWB_BINS 1024        /* isp_math.hpp:32 */

/* FrameParams (isp_math.hpp:35-60). */
typedef struct frame_params {
  int width, height;
  const unsigned char* raw;
  size_t raw_pitch;
  int raw_format;
  int cfa[4];
  float black[4];
  float inv_range[4];
  int white;
  float wb[3];
  const float* lsc;
  int lsc_width, lsc_height, lsc_channels;
  float lsc_sx, lsc_sy;
  float ccm[9];
  const float* lut;
  int lut_size;
  float sharpen;
  unsigned char* out;
  size_t out_pitch;
  int out_format;
} frame_params;

/* AwbAccum (isp_math.hpp:64-68). */
typedef struct awb_accum {
  uint64_t sum[3];
  uint64_t quads;
  uint32_t hist[3][AWB_BINS];
} awb_accum;

struct oracle_pipeline {
  isp_params params;
  float* lsc;  /* own copy */
  float* lut;  /* tone_lut() */
  int lut_size;
  float gains[3]; /* AwbState (isp_math.hpp:71-74) */
  int valid;
  float* samples; /* CpuScratch (cpu_pipeline.hpp:28-32) */
  float* rgb;
  size_t samples_cap, rgb_cap;
};

/* kCfa (cpu_pipeline.cpp:15-20). */
static const int kCfa[4][4] = {
    {ISP_CHANNEL_R, ISP_CHANNEL_GR, ISP_CHANNEL_GB, ISP_CHANNEL_B},
    {ISP_CHANNEL_B, ISP_CHANNEL_GB, ISP_CHANNEL_GR, ISP_CHANNEL_R},
    {ISP_CHANNEL_GR, ISP_CHANNEL_R, ISP_CHANNEL_B, ISP_CHANNEL_GB},
    {ISP_CHANNEL_GB, ISP_CHANNEL_B, ISP_CHANNEL_R, ISP_CHANNEL_GR},
};

/* mirror (isp_math.hpp:82-91). */
static int mirror(int v, int n) {
  const int period = 2 * (n - 1);
  v %= period;
  if (v < 0) v += period;
  return v < n ? v : period - v;
}

/* color_of, channel_at (isp_math.hpp:93-99). */
static int color_of(int channel) {
  return channel == ISP_CHANNEL_R ? 0 : channel == ISP_CHANNEL_B ? 2 : 1;
}
static int channel_at(const frame_params* p, int x, int y) { return p->cfa[((y & 1) << 1) | (x & 1)]; }

/* load_raw (isp_math.hpp:101-119). */
static int load_raw(const frame_params* p, int x, int y) {
  const unsigned char* row = p->raw + (size_t)y * p->raw_pitch;
  switch (p->raw_format) {
    case ISP_RAW_U8:
      return row[x];
    case ISP_RAW_U16: {
      uint16_t v;
      memcpy(&v, row + 2 * (size_t)x, 2);
      return v;
    }
    case ISP_RAW_RAW10: {
      const unsigned char* group = row + 5 * (x >> 2);
      const int k = x & 3;
      return ((int)group[k] << 2) | ((group[4] >> (2 * k)) & 3);
    }
    default: {
      const unsigned char* group = row + 3 * (x >> 1);
      const int k = x & 1;
      return ((int)group[k] << 4) | ((group[2] >> (4 * k)) & 15);
    }
  }
}

/* lens_shading_gain (isp_math.hpp:121-140). */
static float lens_shading_gain(const frame_params* p, int x, int y, int channel) {
  if (p->lsc == NULL) return 1.0F;
  const float fx = (float)x * p->lsc_sx;
  const float fy = (float)y * p->lsc_sy;
  int ix = (int)fx;
  int iy = (int)fy;
  ix = ix > p->lsc_width - 2 ? p->lsc_width - 2 : ix;
  iy = iy > p->lsc_height - 2 ? p->lsc_height - 2 : iy;
  const float tx = fx - (float)ix;
  const float ty = fy - (float)iy;
  const int step = p->lsc_channels;
  const int row = p->lsc_width * step;
  const float* g = p->lsc + ((size_t)iy * p->lsc_width + ix) * step + (step == 4 ? channel : 0);
  const float top = g[0] + (g[step] - g[0]) * tx;
  const float bottom = g[row] + (g[row + step] - g[row]) * tx;
  return top + (bottom - top) * ty;
}

/* linearize (isp_math.hpp:144-148). */
static float linearize(const frame_params* p, int raw, int x, int y, int channel) {
  float v = ((float)raw - p->black[channel]) * p->inv_range[channel];
  v = v > 0.0F ? v : 0.0F;
  return v * lens_shading_gain(p, x, y, channel);
}

/* preprocess (isp_math.hpp:151-155). */
static float preprocess(const frame_params* p, const float gains[3], int x, int y) {
  const int channel = channel_at(p, x, y);
  const float v = linearize(p, load_raw(p, x, y), x, y, channel) * gains[color_of(channel)];
  return v < 1.0F ? v : 1.0F;
}

/* clamp01 (isp_math.hpp:157). */
static float clamp01(float v) { return v < 0.0F ? 0.0F : v > 1.0F ? 1.0F : v; }

/* demosaic<kMalvar> (isp_math.hpp:161-215). */
static void demosaic(float n[5][5], int malvar, int channel, float rgb[3]) {
  const float c = n[2][2];
  float at_green, diagonal, across, down;
  if (!malvar) {
    at_green = (n[1][2] + n[3][2] + n[2][1] + n[2][3]) * 0.25F;
    diagonal = (n[1][1] + n[1][3] + n[3][1] + n[3][3]) * 0.25F;
    across = (n[2][1] + n[2][3]) * 0.5F;
    down = (n[1][2] + n[3][2]) * 0.5F;
  } else {
    const float axial1 = n[1][2] + n[3][2] + n[2][1] + n[2][3];
    const float axial2 = n[0][2] + n[4][2] + n[2][0] + n[2][4];
    const float diagonal1 = n[1][1] + n[1][3] + n[3][1] + n[3][3];
    const float horizontal2 = n[2][0] + n[2][4];
    const float vertical2 = n[0][2] + n[4][2];
    at_green = (4.0F * c + 2.0F * axial1 - axial2) * 0.125F;
    diagonal = (6.0F * c + 2.0F * diagonal1 - 1.5F * axial2) * 0.125F;
    across = (5.0F * c + 4.0F * (n[2][1] + n[2][3]) - horizontal2 - diagonal1 + 0.5F * vertical2) *
             0.125F;
    down = (5.0F * c + 4.0F * (n[1][2] + n[3][2]) - vertical2 - diagonal1 + 0.5F * horizontal2) *
           0.125F;
    at_green = clamp01(at_green);
    diagonal = clamp01(diagonal);
    across = clamp01(across);
    down = clamp01(down);
  }
  switch (channel) {
    case ISP_CHANNEL_R:
      rgb[0] = c; rgb[1] = at_green; rgb[2] = diagonal;
      break;
    case ISP_CHANNEL_B:
      rgb[0] = diagonal; rgb[1] = at_green; rgb[2] = c;
      break;
    case ISP_CHANNEL_GR:
      rgb[0] = across; rgb[1] = c; rgb[2] = down;
      break;
    default:
      rgb[0] = down; rgb[1] = c; rgb[2] = across;
      break;
  }
}

/* sharpen (isp_math.hpp:219-227). */
static void sharpen(float m[3][3][3], float amount, float rgb[3]) {
  for (int k = 0; k < 3; ++k) {
    const float c = m[1][1][k];
    const float blur = (4.0F * c + 2.0F * (m[0][1][k] + m[2][1][k] + m[1][0][k] + m[1][2][k]) +
                        (m[0][0][k] + m[0][2][k] + m[2][0][k] + m[2][2][k])) *
                       0.0625F;
    rgb[k] = clamp01(c + amount * (c - blur));
  }
}

/* correct_color (isp_math.hpp:229-233). */
static void correct_color(const frame_params* p, const float in[3], float out[3]) {
  for (int k = 0; k < 3; ++k) {
    out[k] = clamp01(p->ccm[3 * k] * in[0] + p->ccm[3 * k + 1] * in[1] + p->ccm[3 * k + 2] * in[2]);
  }
}

/* tone (isp_math.hpp:235-243). */
static float tone(const frame_params* p, float x) {
  const float f = x * (float)(p->lut_size - 1);
  const int i = (int)f;
  if (i >= p->lut_size - 1) return p->lut[p->lut_size - 1];
  const float t = f - (float)i;
  return p->lut[i] + t * (p->lut[i + 1] - p->lut[i]);
}

/* to_fixed (isp_math.hpp:245-247). */
static unsigned int to_fixed(float v) { return (unsigned int)(v * FIXED_ONE + 0.5F); }

/* store (isp_math.hpp:250-276). */
static void store(const frame_params* p, int x, int y, const float t[3]) {
  unsigned char* row = p->out + (size_t)y * p->out_pitch;
  switch (p->out_format) {
    case ISP_OUT_RGB8:
      for (int k = 0; k < 3; ++k) row[3 * x + k] = (unsigned char)(t[k] * 255.0F + 0.5F);
      break;
    case ISP_OUT_RGBA8:
      for (int k = 0; k < 3; ++k) row[4 * x + k] = (unsigned char)(t[k] * 255.0F + 0.5F);
      row[4 * x + 3] = 255;
      break;
    case ISP_OUT_RGB16:
      for (int k = 0; k < 3; ++k) {
        const uint16_t v = (uint16_t)(t[k] * 65535.0F + 0.5F);
        memcpy(row + 2 * (size_t)(3 * x + k), &v, 2);
      }
      break;
    default:
      for (int k = 0; k < 3; ++k) memcpy(row + 4 * (size_t)(3 * x + k), &t[k], 4);
      break;
  }
}

/* output_codes (isp_math.hpp:279-301); with 1024 bins, the 10-bit extension
 * codes of RGB16 ((v * 1023 + 32767) / 65535) and RGB32F (clamp01 * 1023 + 0.5). */
static void output_codes(const frame_params* p, int x, int y, int bins, int q[3]) {
  const unsigned char* row = p->out + (size_t)y * p->out_pitch;
  const unsigned top = (unsigned)bins - 1U;
  switch (p->out_format) {
    case ISP_OUT_RGB8:
      for (int k = 0; k < 3; ++k) q[k] = row[3 * x + k];
      break;
    case ISP_OUT_RGBA8:
      for (int k = 0; k < 3; ++k) q[k] = row[4 * x + k];
      break;
    case ISP_OUT_RGB16:
      for (int k = 0; k < 3; ++k) {
        uint16_t v;
        memcpy(&v, row + 2 * (size_t)(3 * x + k), 2);
        q[k] = (int)((v * top + 32767U) / 65535U);
      }
      break;
    default:
      for (int k = 0; k < 3; ++k) {
        float v;
        memcpy(&v, row + 4 * (size_t)(3 * x + k), 4);
        q[k] = (int)(clamp01(v) * (float)top + 0.5F);
      }
      break;
  }
}

/* luma_code (isp_math.hpp:303-305). */
static int luma_code(const int q[3]) { return (2126 * q[0] + 7152 * q[1] + 722 * q[2] + 5000) / 10000; }

/* white_balance_quad (isp_math.hpp:311-332). */
static int white_balance_quad(const frame_params* p, int qx, int qy, int color[4],
                              unsigned int fixed[4], int bin[4]) {
  int raw[4];
  for (int k = 0; k < 4; ++k) {
    raw[k] = load_raw(p, 2 * qx + (k & 1), 2 * qy + (k >> 1));
    if (raw[k] >= p->white) return 0;
  }
  for (int k = 0; k < 4; ++k) {
    const int x = 2 * qx + (k & 1);
    const int y = 2 * qy + (k >> 1);
    const int channel = channel_at(p, x, y);
    float v = linearize(p, raw[k], x, y, channel);
    v = v < 4.0F ? v : 4.0F;
    color[k] = color_of(channel);
    fixed[k] = to_fixed(v);
    const int b = (int)(v * (float)AWB_BINS);
    bin[k] = b < AWB_BINS - 1 ? b : AWB_BINS - 1;
  }
  return 1;
}

/* histogram_percentile (isp_math.hpp:336-352). */
float oracle_histogram_percentile(const uint32_t* hist, float percentile) {
  uint64_t total = 0;
  for (int b = 0; b < AWB_BINS; ++b) total += hist[b];
  if (total == 0) return 0.0F;
  const double target = (double)percentile * (double)total;
  uint64_t below = 0;
  for (int b = 0; b < AWB_BINS; ++b) {
    if (hist[b] != 0U && (double)(below + hist[b]) >= target) {
      const double within = (target - (double)below) / (double)hist[b];
      return (float)(((double)b + within) / AWB_BINS);
    }
    below += hist[b];
  }
  return 1.0F;
}

/* finish_white_balance (isp_math.hpp:357-384). */
static void finish_white_balance(const awb_accum* a, int mode, float percentile, float smoothing,
                                 float gains[3], int* valid) {
  if (a->quads == 0) return;
  float level[3];
  if (mode == ISP_WB_GRAY_WORLD) {
    const double quads = (double)a->quads;
    level[0] = (float)((double)a->sum[0] / quads);
    level[1] = (float)((double)a->sum[1] / (2.0 * quads));
    level[2] = (float)((double)a->sum[2] / quads);
  } else {
    for (int k = 0; k < 3; ++k) level[k] = oracle_histogram_percentile(a->hist[k], percentile);
  }
  if (!(level[0] > 0.0F && level[1] > 0.0F && level[2] > 0.0F)) return;
  float target[3] = {level[1] / level[0], 1.0F, level[1] / level[2]};
  for (int k = 0; k < 3; ++k) {
    target[k] = target[k] < 0.0625F ? 0.0625F : target[k] > 16.0F ? 16.0F : target[k];
  }
  for (int k = 0; k < 3; ++k) {
    gains[k] = *valid != 0 && smoothing < 1.0F ? gains[k] + smoothing * (target[k] - gains[k])
                                               : target[k];
  }
  *valid = 1;
}

/* finish_stats (isp_math.hpp:387-425). */
void oracle_finish_stats(const uint64_t sums[4], int64_t pixels, const float gains[3],
                         const isp_auto_exposure* ae, const isp_frame_exposure* exposure,
                         float* out) {
  const double n = (double)pixels * (double)FIXED_ONE;
  const double mean[3] = {(double)sums[0] / n, (double)sums[1] / n, (double)sums[2] / n};
  const double luma = 0.2126 * mean[0] + 0.7152 * mean[1] + 0.0722 * mean[2];
  double error = luma > 0.0 ? log2((double)ae->target / luma) : 8.0;
  error = error < -8.0 ? -8.0 : error > 8.0 ? 8.0 : error;
  const double scale = exp2((double)ae->damping * error);
  double next_exposure = 0.0;
  double next_gain = 0.0;
  if (exposure->exposure_ns > 0) {
    const double gain = exposure->gain > 0.0F ? exposure->gain : ae->min_gain;
    const double wanted = (double)exposure->exposure_ns * gain * scale;
    next_exposure = wanted / ae->min_gain;
    const double low = (double)ae->min_exposure_ns;
    const double high = (double)ae->max_exposure_ns;
    next_exposure = next_exposure < low ? low : next_exposure > high ? high : next_exposure;
    next_gain = wanted / next_exposure;
    next_gain = next_gain < ae->min_gain ? ae->min_gain
                : next_gain > ae->max_gain ? ae->max_gain
                                           : next_gain;
  }
  out[ISP_STAT_MEAN_R] = (float)mean[0];
  out[ISP_STAT_MEAN_G] = (float)mean[1];
  out[ISP_STAT_MEAN_B] = (float)mean[2];
  out[ISP_STAT_MEAN_LUMA] = (float)luma;
  out[ISP_STAT_AE_ERROR_EV] = (float)error;
  out[ISP_STAT_EXPOSURE_SCALE] = (float)scale;
  out[ISP_STAT_SUGGESTED_EXPOSURE_NS] = (float)next_exposure;
  out[ISP_STAT_SUGGESTED_GAIN] = (float)next_gain;
  out[ISP_STAT_WB_GAIN_R] = gains[0];
  out[ISP_STAT_WB_GAIN_G] = gains[1];
  out[ISP_STAT_WB_GAIN_B] = gains[2];
  out[ISP_STAT_SATURATED_FRACTION] = (float)((double)sums[3] / (double)pixels);
}

/* oracle_finish_stats over `count` cases (isp_epilogue_case, include/isp_c.h):
 * out holds ISP_STATS_SIZE floats per case. */
void oracle_finish_stats_cases(const isp_epilogue_case* cases, float* out, int count) {
  for (int i = 0; i < count; ++i) {
    const isp_epilogue_case* c = &cases[i];
    oracle_finish_stats(c->sums, c->pixels, c->gains, &c->ae, &c->exposure, out + (size_t)i * ISP_STATS_SIZE);
  }
}

/* tone_lut (isp_params.cpp:166-182). Returns the size, writes min(size, capacity). */
int oracle_tone_lut(const isp_params* params, float* out, int capacity) {
  if (params->tone_curve == ISP_TONE_LUT) {
    for (int i = 0; i < params->lut_size && i < capacity; ++i) out[i] = params->lut[i];
    return params->lut_size;
  }
  for (int i = 0; i < ISP_TONE_LUT_SIZE && i < capacity; ++i) {
    const double x = (double)i / (ISP_TONE_LUT_SIZE - 1);
    double y = x;
    if (params->tone_curve == ISP_TONE_SRGB) {
      y = x <= 0.0031308 ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
    } else if (params->tone_curve == ISP_TONE_GAMMA) {
      y = pow(x, 1.0 / (double)params->gamma);
    }
    out[i] = (float)(y < 0.0 ? 0.0 : y > 1.0 ? 1.0 : y);
  }
  return ISP_TONE_LUT_SIZE;
}

int oracle_mirror(int v, int n) { return mirror(v, n); }

/* make_frame_params (cpu_pipeline.cpp:50-88). */
static frame_params make_frame_params(const oracle_pipeline* o, const isp_raw_view* raw,
                                      const isp_rgb_view* out) {
  const isp_params* params = &o->params;
  frame_params p;
  memset(&p, 0, sizeof(p));
  p.width = raw->width;
  p.height = raw->height;
  p.raw = (const unsigned char*)raw->data;
  p.raw_pitch = raw->pitch;
  p.raw_format = raw->format;
  for (int k = 0; k < 4; ++k) p.cfa[k] = kCfa[raw->pattern][k];
  const float white = params->white_level > 0.0F
                          ? params->white_level
                          : (float)((1U << (unsigned)raw->bit_depth) - 1U);
  for (int k = 0; k < 4; ++k) {
    p.black[k] = params->has_channel_black_levels ? params->channel_black_levels[k]
                                                  : params->black_level;
    p.inv_range[k] = 1.0F / (white - p.black[k]);
  }
  p.white = (int)ceilf(white);
  for (int k = 0; k < 3; ++k) p.wb[k] = params->wb_gains[k];
  if (params->has_lens_shading) {
    p.lsc = o->lsc;
    p.lsc_width = params->lens_shading.width;
    p.lsc_height = params->lens_shading.height;
    p.lsc_channels = params->lens_shading.channels;
    p.lsc_sx = (float)(params->lens_shading.width - 1) / (float)(raw->width - 1);
    p.lsc_sy = (float)(params->lens_shading.height - 1) / (float)(raw->height - 1);
  }
  for (int k = 0; k < 9; ++k) p.ccm[k] = params->ccm[k];
  p.lut = o->lut;
  p.lut_size = o->lut_size;
  p.sharpen = params->sharpen;
  p.out = (unsigned char*)out->data;
  p.out_pitch = out->pitch;
  p.out_format = out->format;
  return p;
}

int oracle_load_raw(const isp_raw_view* raw, int x, int y) {
  frame_params p;
  memset(&p, 0, sizeof(p));
  p.raw = (const unsigned char*)raw->data;
  p.raw_pitch = raw->pitch;
  p.raw_format = raw->format;
  return load_raw(&p, x, y);
}

int oracle_create(const isp_params* params, oracle_pipeline** out) {
  oracle_pipeline* o = (oracle_pipeline*)calloc(1, sizeof(*o));
  if (o == NULL) return 1;
  o->params = *params;
  if (params->has_lens_shading) {
    const size_t n = (size_t)params->lens_shading.width * params->lens_shading.height *
                     params->lens_shading.channels;
    o->lsc = (float*)malloc(n * sizeof(float));
    memcpy(o->lsc, params->lens_shading.gains, n * sizeof(float));
    o->params.lens_shading.gains = o->lsc;
  }
  const int size = params->tone_curve == ISP_TONE_LUT ? params->lut_size : ISP_TONE_LUT_SIZE;
  o->lut = (float*)malloc((size_t)size * sizeof(float));
  o->lut_size = oracle_tone_lut(params, o->lut, size);
  o->params.lut = NULL;
  oracle_reset(o);
  *out = o;
  return 0;
}

void oracle_destroy(oracle_pipeline* o) {
  if (o == NULL) return;
  free(o->lsc);
  free(o->lut);
  free(o->samples);
  free(o->rgb);
  free(o);
}

/* initial_white_balance (cpu_pipeline.cpp:90-95). */
void oracle_reset(oracle_pipeline* o) {
  for (int k = 0; k < 3; ++k) o->gains[k] = o->params.wb_gains[k];
  o->valid = 0;
}

void oracle_awb_state(const oracle_pipeline* o, float gains[3], int* valid) {
  for (int k = 0; k < 3; ++k) gains[k] = o->gains[k];
  *valid = o->valid;
}

/* ---- row-band passes ---------------------------------------------------- */

typedef struct band_job {
  const frame_params* p;
  const oracle_pipeline* o;
  const float* gains;
  int border, malvar, bins, with_stats, with_hist;
  int y0, y1; /* rows of this band (quad rows for the white balance pass) */
  int pass;
  awb_accum* awb;
  uint64_t sums[4];
  uint32_t* hist;
} band_job;

/* demosaic_frame<kMalvar> (cpu_pipeline.cpp:22-46), rows [y0, y1) of the bordered frame. */
static void demosaic_rows(const band_job* j) {
  const frame_params* p = j->p;
  const int border = j->border;
  const int w = p->width + 2 * border;
  for (int ty = j->y0; ty < j->y1; ++ty) {
    const int vy = ty - border;
    for (int tx = 0; tx < w; ++tx) {
      const int vx = tx - border;
      float n[5][5];
      memset(n, 0, sizeof(n));
      for (int dy = -2; dy <= 2; ++dy) {
        const int y = mirror(vy + dy, p->height);
        for (int dx = -2; dx <= 2; ++dx) {
          n[dy + 2][dx + 2] = j->o->samples[(size_t)y * p->width + mirror(vx + dx, p->width)];
        }
      }
      demosaic(n, j->malvar, channel_at(p, vx, vy), &j->o->rgb[((size_t)ty * w + tx) * 3U]);
    }
  }
}

static void* band_run(void* arg) {
  band_job* j = (band_job*)arg;
  const frame_params* p = j->p;
  switch (j->pass) {
    case 0: /* white balance statistics (cpu_pipeline.cpp:105-119) */
      for (int qy = j->y0; qy < j->y1; ++qy) {
        for (int qx = 0; qx < p->width / 2; ++qx) {
          int color[4];
          unsigned int fixed[4];
          int bin[4];
          if (!white_balance_quad(p, qx, qy, color, fixed, bin)) continue;
          for (int k = 0; k < 4; ++k) {
            j->awb->sum[color[k]] += fixed[k];
            ++j->awb->hist[color[k]][bin[k]];
          }
          ++j->awb->quads;
        }
      }
      break;
    case 1: /* preprocess (cpu_pipeline.cpp:126-132) */
      for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < p->width; ++x) {
          j->o->samples[(size_t)y * p->width + x] = preprocess(p, j->gains, x, y);
        }
      }
      break;
    case 2: /* demosaic (cpu_pipeline.cpp:134-140) */
      demosaic_rows(j);
      break;
    case 3: { /* sharpen, colour, stats sums, tone, store (cpu_pipeline.cpp:143-172) */
      const int border = j->border;
      const int w = p->width + 2 * border;
      const float* rgbs = j->o->rgb;
      for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < p->width; ++x) {
          float rgb[3];
          const float* centre = &rgbs[((size_t)(y + border) * w + x + border) * 3U];
          if (border != 0) {
            float m[3][3][3];
            for (int dy = 0; dy < 3; ++dy) {
              for (int dx = 0; dx < 3; ++dx) {
                const float* v = &rgbs[((size_t)(y + dy) * w + x + dx) * 3U];
                for (int k = 0; k < 3; ++k) m[dy][dx][k] = v[k];
              }
            }
            sharpen(m, p->sharpen, rgb);
          } else {
            for (int k = 0; k < 3; ++k) rgb[k] = centre[k];
          }
          float linear[3];
          correct_color(p, rgb, linear);
          if (j->with_stats) {
            for (int k = 0; k < 3; ++k) j->sums[k] += to_fixed(linear[k]);
            j->sums[3] += load_raw(p, x, y) >= p->white ? 1U : 0U;
          }
          float toned[3];
          for (int k = 0; k < 3; ++k) toned[k] = tone(p, linear[k]);
          store(p, x, y, toned);
        }
      }
      break;
    }
    case 4: /* histograms of the written output (cpu_pipeline.cpp:174-186) */
      for (int y = j->y0; y < j->y1; ++y) {
        for (int x = 0; x < p->width; ++x) {
          int q[3];
          output_codes(p, x, y, j->bins, q);
          ++j->hist[0 * j->bins + luma_code(q)];
          ++j->hist[1 * j->bins + q[0]];
          ++j->hist[2 * j->bins + q[1]];
          ++j->hist[3 * j->bins + q[2]];
        }
      }
      break;
    default:
      break;
  }
  return NULL;
}

/* Runs `pass` over rows [0, rows) split into `threads` bands. */
static void run_bands(band_job* jobs, int threads, int pass, int rows) {
  pthread_t tid[256];
  for (int t = 0; t < threads; ++t) {
    jobs[t].pass = pass;
    jobs[t].y0 = (int)((int64_t)rows * t / threads);
    jobs[t].y1 = (int)((int64_t)rows * (t + 1) / threads);
  }
  if (threads == 1) {
    band_run(&jobs[0]);
    return;
  }
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, band_run, &jobs[t]);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* The white balance statistics of a frame (cpu_pipeline.cpp:101-119): the
 * white-balance pass over row bands of quads, the bands' sums and histograms
 * added into `a`. */
static int measure_white_balance(band_job* jobs, int threads, const frame_params* p, awb_accum* a) {
  awb_accum* parts = (awb_accum*)calloc((size_t)threads, sizeof(awb_accum));
  if (parts == NULL) return 1;
  for (int t = 0; t < threads; ++t) jobs[t].awb = &parts[t];
  run_bands(jobs, threads, 0, p->height / 2);
  *a = parts[0];
  for (int t = 1; t < threads; ++t) {
    for (int k = 0; k < 3; ++k) a->sum[k] += parts[t].sum[k];
    a->quads += parts[t].quads;
    for (int k = 0; k < 3; ++k)
      for (int b = 0; b < AWB_BINS; ++b) a->hist[k][b] += parts[t].hist[k][b];
  }
  for (int t = 0; t < threads; ++t) jobs[t].awb = NULL;
  free(parts);
  return 0;
}

struct oracle_awb_stats {
  awb_accum a;
};

oracle_awb_stats* oracle_awb_measure(const oracle_pipeline* o, const isp_raw_view* raw, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  const isp_rgb_view none = {NULL, 0, 0};
  const frame_params fp = make_frame_params(o, raw, &none);
  band_job* jobs = (band_job*)calloc((size_t)threads, sizeof(band_job));
  oracle_awb_stats* s = (oracle_awb_stats*)malloc(sizeof(oracle_awb_stats));
  if (jobs == NULL || s == NULL) {
    free(jobs);
    free(s);
    return NULL;
  }
  for (int t = 0; t < threads; ++t) {
    jobs[t].p = &fp;
    jobs[t].o = o;
  }
  const int rc = measure_white_balance(jobs, threads, &fp, &s->a);
  free(jobs);
  if (rc != 0) {
    free(s);
    return NULL;
  }
  return s;
}

/* finish_white_balance with the pipeline's settings (cpu_pipeline.cpp:120-122); nothing under
 * manual white balance, as in oracle_process. */
void oracle_awb_apply(oracle_pipeline* o, const oracle_awb_stats* s) {
  if (o->params.white_balance == ISP_WB_MANUAL) return; /* oracle_process keeps the state */
  finish_white_balance(&s->a, o->params.white_balance, o->params.white_patch_percentile,
                       o->params.awb_smoothing, o->gains, &o->valid);
}

void oracle_awb_free(oracle_awb_stats* s) { free(s); }

/* cpu_process (cpu_pipeline.cpp:97-191). */
int oracle_process(oracle_pipeline* o, const isp_raw_view* raw, const isp_rgb_view* out,
                   const isp_stats_out* stats, const isp_frame_exposure* exposure, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  const frame_params fp = make_frame_params(o, raw, out);
  const frame_params* p = &fp;
  const isp_params* params = &o->params;
  const int bins = params->histogram_bins == ISP_HISTOGRAM_BINS_WIDE ? ISP_HISTOGRAM_BINS_WIDE
                                                                     : ISP_HISTOGRAM_BINS;
  const isp_frame_exposure none = {0, 0.0F};
  if (exposure == NULL) exposure = &none;
  band_job* jobs = (band_job*)calloc((size_t)threads, sizeof(band_job));
  if (jobs == NULL) return 1;
  float gains[3] = {p->wb[0], p->wb[1], p->wb[2]};
  for (int t = 0; t < threads; ++t) {
    jobs[t].p = p;
    jobs[t].o = o;
    jobs[t].gains = gains;
    jobs[t].bins = bins;
  }

  /* White balance: this frame's estimate, from its own statistics. */
  if (params->white_balance != ISP_WB_MANUAL) {
    awb_accum* a = (awb_accum*)malloc(sizeof(awb_accum));
    if (a == NULL || measure_white_balance(jobs, threads, p, a) != 0) {
      free(a);
      free(jobs);
      return 1;
    }
    finish_white_balance(a, params->white_balance, params->white_patch_percentile,
                         params->awb_smoothing, o->gains, &o->valid);
    for (int k = 0; k < 3; ++k) gains[k] = o->gains[k];
    free(a);
  }

  /* Black level, lens shading, white balance. */
  const size_t pixels = (size_t)p->width * p->height;
  if (o->samples_cap < pixels) {
    free(o->samples);
    o->samples = (float*)malloc(pixels * sizeof(float));
    o->samples_cap = pixels;
  }
  run_bands(jobs, threads, 1, p->height);

  /* Demosaic, with a border for the sharpening's neighbours. */
  const int border = params->sharpen > 0.0F ? 1 : 0;
  const size_t rgb_n = (size_t)(p->width + 2 * border) * (p->height + 2 * border) * 3U;
  if (o->rgb_cap < rgb_n) {
    free(o->rgb);
    o->rgb = (float*)malloc(rgb_n * sizeof(float));
    o->rgb_cap = rgb_n;
  }
  for (int t = 0; t < threads; ++t) {
    jobs[t].border = border;
    jobs[t].malvar = params->demosaic == ISP_DEMOSAIC_MALVAR;
    jobs[t].with_stats = stats != NULL && stats->stats != NULL;
  }
  run_bands(jobs, threads, 2, p->height + 2 * border);

  /* Sharpening, colour correction, tone curve, and the sums for the statistics. */
  run_bands(jobs, threads, 3, p->height);
  uint64_t sums[4] = {0, 0, 0, 0};
  for (int t = 0; t < threads; ++t)
    for (int k = 0; k < 4; ++k) sums[k] += jobs[t].sums[k];

  if (stats != NULL && stats->histogram != NULL) {
    const size_t hn = (size_t)ISP_HISTOGRAM_CHANNELS * bins;
    uint32_t* parts = (uint32_t*)calloc((size_t)threads * hn, sizeof(uint32_t));
    for (int t = 0; t < threads; ++t) jobs[t].hist = parts + (size_t)t * hn;
    run_bands(jobs, threads, 4, p->height);
    memset(stats->histogram, 0, hn * sizeof(uint32_t));
    for (int t = 0; t < threads; ++t)
      for (size_t i = 0; i < hn; ++i) stats->histogram[i] += parts[(size_t)t * hn + i];
    free(parts);
  }
  if (stats != NULL && stats->stats != NULL) {
    oracle_finish_stats(sums, (int64_t)pixels, gains, &params->auto_exposure, exposure,
                        stats->stats);
  }
  free(jobs);
  return 0;
}

/* ---- the demo's frame checker ------------------------------------------- */

/* tone (applications/isp/src/frame_check.hpp:45-51): the table at v in 0..1,
 * interpolated linearly. */
static float check_tone(const float* lut, int size, float v) {
  const float f = v * (float)(size - 1);
  const int i = (int)f;
  return i >= size - 1 ? lut[size - 1] : lut[i] + (f - (float)i) * (lut[i + 1] - lut[i]);
}

/* check_frame_host (check_op.cpp:44-72), with truth_codes (frame_check.hpp:54-62)
 * and luma_bin (frame_check.hpp:64-67). */
void oracle_frame_check(const float* truth, int width, int height, const float* lut, int lut_size,
                        const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                        const float* stats, isp_frame_check_result* result) {
  uint32_t rebuilt[ISP_HISTOGRAM_CHANNELS * ISP_HISTOGRAM_BINS];
  memset(rebuilt, 0, sizeof(rebuilt));
  memset(result, 0, sizeof(*result));
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      const float* t = truth + ((size_t)y * width + x) * 3;
      const uint8_t* px = rgb + (size_t)y * pitch + (size_t)x * 3;
      for (int k = 0; k < 3; ++k) {
        const int code = (int)round(255.0 * (double)check_tone(lut, lut_size, t[k]));
        const int d = (int)px[k] - code;
        result->squared_error += (uint64_t)(d * d);
      }
      const int q[3] = {px[0], px[1], px[2]};
      ++rebuilt[0 * ISP_HISTOGRAM_BINS + luma_code(q)];
      ++rebuilt[1 * ISP_HISTOGRAM_BINS + q[0]];
      ++rebuilt[2 * ISP_HISTOGRAM_BINS + q[1]];
      ++rebuilt[3 * ISP_HISTOGRAM_BINS + q[2]];
    }
  }
  for (int i = 0; i < ISP_HISTOGRAM_CHANNELS * ISP_HISTOGRAM_BINS; ++i)
    result->histogram_mismatches += rebuilt[i] != histogram[i] ? 1U : 0U;
  for (int k = 0; k < ISP_STATS_SIZE; ++k) result->stats[k] = stats != NULL ? stats[k] : 0.0F;
}

double oracle_frame_psnr(const isp_frame_check_result* result, int width, int height) {
  const double squared = (double)result->squared_error;
  return squared == 0.0 ? 99.0 : 10.0 * log10(255.0 * 255.0 * 3.0 * (double)width * (double)height / squared);
}

/* ---- colour conversion (modules/color-convert) --------------------------- */

/* from_rgb (conversion.cpp:176-198) as floats, as make_plan rounds its map
 * (conversion.cpp:428-436): c_i = m[3i] R + m[3i+1] G + m[3i+2] B + o[i]. */
static void color_affine(int gray, int matrix, int range, float m[9], float o[3]) {
  double kr = 0.299, kb = 0.114;
  if (matrix == ISP_MATRIX_BT709) {
    kr = 0.2126;
    kb = 0.0722;
  } else if (matrix == ISP_MATRIX_BT2020) {
    kr = 0.2627;
    kb = 0.0593;
  }
  const double kg = 1.0 - kr - kb;
  double a[3][4];
  memset(a, 0, sizeof(a));
  if (gray) {
    a[0][0] = kr;
    a[0][1] = kg;
    a[0][2] = kb;
  } else {
    const int limited = range == ISP_RANGE_LIMITED;
    const double ys = limited ? 219.0 : 255.0, yo = limited ? 16.0 : 0.0, cs = limited ? 224.0 : 255.0;
    const double cb_den = 2.0 * (1.0 - kb), cr_den = 2.0 * (1.0 - kr);
    const double s = ys / 255.0, c = cs / 255.0;
    a[0][0] = kr * s;
    a[0][1] = kg * s;
    a[0][2] = kb * s;
    a[0][3] = yo;
    a[1][0] = -kr / cb_den * c;
    a[1][1] = -kg / cb_den * c;
    a[1][2] = (1.0 - kb) / cb_den * c;
    a[1][3] = 128.0;
    a[2][0] = (1.0 - kr) / cr_den * c;
    a[2][1] = -kg / cr_den * c;
    a[2][2] = -kb / cr_den * c;
    a[2][3] = 128.0;
  }
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) m[3 * i + j] = (float)a[i][j];
    o[i] = (float)a[i][3];
  }
}

/* to_u8 (kernel_math.hpp:80-83). */
static uint8_t color_code(float v) {
  const float r = floorf(v + 0.5F);
  return (uint8_t)(r <= 0.0F ? 0.0F : (r >= 255.0F ? 255.0F : r));
}

int oracle_color_convert(const isp_image_view* src, const isp_image_view* dst, int matrix, int range) {
  const int s = src->format, t = dst->format;
  const int bpp = s == ISP_PIX_RGB8 || s == ISP_PIX_BGR8 ? 3 : s == ISP_PIX_RGBA8 || s == ISP_PIX_BGRA8 ? 4 : 0;
  const int bgr = s == ISP_PIX_BGR8 || s == ISP_PIX_BGRA8;
  if (bpp == 0 || (t > ISP_PIX_UYVY && t != ISP_PIX_MONO8)) return -1;
  if (src->width != dst->width || src->height != dst->height || src->width <= 0 || src->height <= 0) return -1;
  const int W = src->width, H = src->height;
  float m[9], o[3];
  color_affine(t == ISP_PIX_MONO8, matrix, range, m, o);
  /* Chroma block of the destination (convert_block, kernel_math.hpp:335-361). */
  const int sx = t == ISP_PIX_MONO8 ? 0 : 1;
  const int sy = t == ISP_PIX_NV12 || t == ISP_PIX_I420 ? 1 : 0;
  const int bw = (W + (1 << sx) - 1) >> sx, bh = (H + (1 << sy) - 1) >> sy;
  for (int by = 0; by < bh; ++by) {
    for (int bx = 0; bx < bw; ++bx) {
      const int x0 = bx << sx, y0 = by << sy;
      const int x1 = x0 + (1 << sx) < W ? x0 + (1 << sx) : W;
      const int y1 = y0 + (1 << sy) < H ? y0 + (1 << sy) : H;
      float cb = 0.0F, cr = 0.0F;
      int n = 0;
      for (int y = y0; y < y1; ++y) {
        const uint8_t* row = src->planes[0].data + (size_t)y * src->planes[0].pitch;
        uint8_t* out = dst->planes[0].data + (size_t)y * dst->planes[0].pitch;
        for (int x = x0; x < x1; ++x) {
          const uint8_t* p = row + (size_t)bpp * x;
          const float c0 = (float)p[bgr ? 2 : 0], c1 = (float)p[1], c2 = (float)p[bgr ? 0 : 2];
          /* apply (kernel_math.hpp:234-242) */
          const float y_ = m[0] * c0 + m[1] * c1 + m[2] * c2 + o[0];
          const float u_ = m[3] * c0 + m[4] * c1 + m[5] * c2 + o[1];
          const float v_ = m[6] * c0 + m[7] * c1 + m[8] * c2 + o[2];
          /* store (kernel_math.hpp:253-300): the luma */
          if (t == ISP_PIX_YUYV)
            out[4 * (x >> 1) + 2 * (x & 1)] = color_code(y_);
          else if (t == ISP_PIX_UYVY)
            out[4 * (x >> 1) + 1 + 2 * (x & 1)] = color_code(y_);
          else
            out[x] = color_code(y_);
          cb += u_;
          cr += v_;
          ++n;
        }
      }
      if (t == ISP_PIX_MONO8) continue;
      const float count = (float)n;
      const uint8_t qb = color_code(cb / count), qr = color_code(cr / count);
      /* store_chroma (kernel_math.hpp:303-331) */
      if (t == ISP_PIX_NV12 || t == ISP_PIX_NV16) {
        uint8_t* q = dst->planes[1].data + (size_t)by * dst->planes[1].pitch + 2 * (size_t)bx;
        q[0] = qb;
        q[1] = qr;
      } else if (t == ISP_PIX_I420) {
        dst->planes[1].data[(size_t)by * dst->planes[1].pitch + bx] = qb;
        dst->planes[2].data[(size_t)by * dst->planes[2].pitch + bx] = qr;
      } else {
        uint8_t* q = dst->planes[0].data + (size_t)by * dst->planes[0].pitch + 4 * (size_t)bx;
        q[t == ISP_PIX_YUYV ? 1 : 0] = qb;
        q[t == ISP_PIX_YUYV ? 3 : 2] = qr;
      }
    }
  }
  return 0;
}
