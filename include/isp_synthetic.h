/*
 * isp_synthetic.h - test and benchmark inputs of the ISP, C ABI.
 *
 * 1. The synthetic raw camera with a closed-form truth: the reference's
 *    holoscan::isp::SyntheticRawConfig / make_synthetic_scene /
 *    synthetic_isp_params / render_synthetic_raw / synthetic_truth
 *    (modules/isp/include/holoscan/isp/synthetic_raw.hpp:40-354,
 *    modules/isp/src/synthetic_raw.cpp:46-154), rendered on the host.
 * 2. The seeded noisy benchmark scene (no reference counterpart): a smooth
 *    colour gradient plus random rectangles, cos^4 vignetting, Poisson-Gaussian
 *    noise, black level and clipping, mosaicked and packed.
 */
#ifndef ISP_SYNTHETIC_H_
#define ISP_SYNTHETIC_H_

#include "isp_c.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ISP_SCENE_CHART = 0, ISP_SCENE_GRAY = 1 }; /* SceneKind (synthetic_raw.hpp:35-38) */

/* SyntheticRawConfig (synthetic_raw.hpp:40-59). */
typedef struct isp_synthetic_config {
  int32_t width;  /* 640, 16..16384 */
  int32_t height; /* 480, 16..16384 */
  uint8_t pattern;
  uint8_t format;
  int32_t bit_depth; /* 12 */
  float black_levels[4];
  float wb_gains[3];
  float ccm[9];
  float vignetting; /* 0..0.9 */
  uint8_t kind;     /* ISP_SCENE_* */
  float exposure;   /* (0, 64] */
  uint32_t frame;
} isp_synthetic_config;

void isp_synthetic_config_init(isp_synthetic_config* config);
/* render_synthetic_raw(make_synthetic_scene(config)) into host memory. */
isp_status isp_synthetic_render(const isp_synthetic_config* config, void* data, size_t pitch);
/* synthetic_isp_params(config, grid): `lsc_gains` receives grid*grid gains when
 * vignetting > 0 (params->lens_shading.gains points at it). */
isp_status isp_synthetic_params(const isp_synthetic_config* config, int32_t grid,
                                isp_params* params, float* lsc_gains);
/* synthetic_truth at (x, y): linear RGB before the tone curve. */
isp_status isp_synthetic_truth(const isp_synthetic_config* config, int32_t x, int32_t y,
                               float linear[3]);
/* The whole truth image, width*height*3 floats. */
isp_status isp_synthetic_truth_image(const isp_synthetic_config* config, float* linear);

/* The seeded noisy scene. */
typedef struct isp_scene_config {
  int32_t width;
  int32_t height;
  uint8_t pattern;
  uint8_t format;          /* ISP_RAW_U16 / RAW10 / RAW12 / U8 */
  int32_t bit_depth;       /* white = 2^bit_depth - 1 */
  float black_level;       /* added after the noise, then clipped to [0, white] */
  uint64_t seed;
  int32_t rectangles;      /* 40 */
  float corner_falloff;    /* cos^4 vignetting value at the corners: 0.6 */
  float noise_gain;        /* DN per electron: 0.5 */
  float read_sigma;        /* DN: 4 */
  float illumination[3];   /* per-colour sensor response (1 = none), for white balance */
} isp_scene_config;

void isp_scene_config_init(isp_scene_config* config);
/* Renders the frame into host memory, rows of isp_raw_row_bytes at `pitch`. */
isp_status isp_scene_render(const isp_scene_config* config, void* data, size_t pitch);
/* The lens shading grid (gw x gh x 1) equal to the inverse of the falloff at the
 * grid points, as a calibration would measure it. */
isp_status isp_scene_lens_shading(const isp_scene_config* config, int32_t gw, int32_t gh,
                                  float* gains);

#ifdef __cplusplus
}
#endif

#endif /* ISP_SYNTHETIC_H_ */
