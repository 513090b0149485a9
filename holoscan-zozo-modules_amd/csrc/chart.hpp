// chart.hpp - the synthetic chart camera's per-pixel arithmetic: the scene,
// lens blur, vignetting, inverse colour matrix, white balance, quantisation and
// truth of the reference's SyntheticRawScene (synthetic_raw.hpp:102-312).
//
// The host renderer (synthetic.cpp) and the device renderer and frame checker
// (isp_source.hip) include these same functions, and both sides are built
// without contraction and with IEEE division, so a sample or truth value is
// the same float expression evaluated in the same order on either side: the
// device frame is byte for byte the host's.
//
// The lens blur is split in two steps, chart_point() and chart_blur(), so that
// a tile of pixels can evaluate each point of the half-pixel lattice once
// (neighbouring pixels share 12 of their 16 points) and still add every
// pixel's 16 points in the reference's order.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/isp_synthetic.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ISP_HD __host__ __device__ inline
#else
#define ISP_HD inline
#endif

namespace isp_amd {

struct Chart {
  int width, height, kind, format;
  int cfa[4];
  float black[4];
  float white;
  float wb[3], inv_wb[3];
  float ccm[9], inv_ccm[9];
  float vignetting, exposure;
  float cx, cy, inv_cx2, inv_cy2;
  int shift;
};

// make_synthetic_scene (synthetic_raw.cpp:46-101): validates the config and
// derives the scene (synthetic.cpp).
Chart make_chart(const isp_synthetic_config& c);

// The 24 colour chart patches, linear sRGB (checker_color's table, synthetic_raw.hpp:102-116).
constexpr float kChartPatch[24][3] = {
    {0.1714F, 0.0844F, 0.0578F}, {0.5395F, 0.3050F, 0.2232F}, {0.1221F, 0.1946F, 0.3372F},
    {0.0953F, 0.1500F, 0.0561F}, {0.2346F, 0.2159F, 0.4397F}, {0.1356F, 0.5089F, 0.4020F},
    {0.6724F, 0.2086F, 0.0252F}, {0.0802F, 0.1046F, 0.3813F}, {0.5333F, 0.1022F, 0.1248F},
    {0.1119F, 0.0452F, 0.1500F}, {0.3372F, 0.5029F, 0.0513F}, {0.7454F, 0.3663F, 0.0273F},
    {0.0395F, 0.0467F, 0.3050F}, {0.0612F, 0.2961F, 0.0666F}, {0.4287F, 0.0369F, 0.0452F},
    {0.7991F, 0.5711F, 0.0137F}, {0.4969F, 0.0931F, 0.3005F}, {0.0024F, 0.2346F, 0.3564F},
    {0.8963F, 0.8963F, 0.8879F}, {0.5776F, 0.5776F, 0.5776F}, {0.3515F, 0.3515F, 0.3515F},
    {0.1946F, 0.1946F, 0.1912F}, {0.0908F, 0.0908F, 0.0908F}, {0.0343F, 0.0343F, 0.0343F},
};

ISP_HD int chart_colour(int channel) { return channel == ISP_CHANNEL_R ? 0 : channel == ISP_CHANNEL_B ? 2 : 1; }

ISP_HD void chart_set3(float* rgb, float r, float g, float b) {
  rgb[0] = r;
  rgb[1] = g;
  rgb[2] = b;
}

// The coordinate of blur point i (0..3) of pixel x: x + 0.5 i - 0.75. Every
// value is a multiple of 0.25 below 2^14, so the result is exact and equals
// chart_coord(x', i') whenever 2x + i = 2x' + i'.
ISP_HD float chart_coord(int x, int i) { return static_cast<float>(x) + 0.5F * static_cast<float>(i) - 0.75F; }

// The scene at point (fx, fy) (synthetic_scene_point, synthetic_raw.hpp:152-236).
ISP_HD void chart_point(const Chart& s, float fx, float fy, float rgb[3]) {
  const int w = s.width;
  const int h = s.height;
  if (s.kind == ISP_SCENE_GRAY) {
    float v;
    if (fx < static_cast<float>(w / 2)) {
      v = 0.03F + 0.8F * fy / static_cast<float>(h - 1);
    } else {
      const float cell = static_cast<float>(h / 4 > 1 ? h / 4 : 1);
      const int col = static_cast<int>(std::floor((fx - static_cast<float>(w / 2 - s.shift)) / cell)) & 3;
      const int row = static_cast<int>(std::floor(fy / cell)) & 3;
      v = 0.05F + 0.8F * static_cast<float>((row * 4 + col) * 7 % 16) / 15.0F;
    }
    chart_set3(rgb, v, v, v);
    return;
  }
  if (fy < static_cast<float>(h * 58 / 100)) {
    // The colour chart: 6 x 4 patches.
    const int bottom = h * 58 / 100;
    const int left = w / 24;
    const int top = h / 30;
    const int cell_w = (w - 2 * left) / 6;
    const int cell_h = (bottom - 2 * top) / 4;
    const int patch_w = cell_w - cell_w / 6;
    const int patch_h = cell_h - cell_h / 6;
    chart_set3(rgb, 0.05F, 0.05F, 0.05F);
    const float px = fx - static_cast<float>(left + s.shift);
    const float py = fy - static_cast<float>(top);
    if (px >= 0.0F && py >= 0.0F) {
      const int col = static_cast<int>(px) / cell_w;
      const int row = static_cast<int>(py) / cell_h;
      if (col < 6 && row < 4 && px - static_cast<float>(col * cell_w) < static_cast<float>(patch_w) &&
          py - static_cast<float>(row * cell_h) < static_cast<float>(patch_h)) {
        const float* p = kChartPatch[row * 6 + col];
        chart_set3(rgb, p[0], p[1], p[2]);
      }
    }
  } else if (fy < static_cast<float>(h * 70 / 100)) {
    const float v = 0.01F + 0.89F * fx / static_cast<float>(w - 1);
    chart_set3(rgb, v, v, v);
  } else if (fy < static_cast<float>(h * 82 / 100)) {
    float hue = 6.0F * fx / static_cast<float>(w);
    hue = hue < 0.0F ? 0.0F : hue;
    const int sector = static_cast<int>(hue) % 6;
    const float frac = hue - static_cast<float>(static_cast<int>(hue));
    const float lo = 0.05F;
    const float hi = 0.7F;
    const float up = lo + (hi - lo) * frac;
    const float down = hi - (hi - lo) * frac;
    switch (sector) {
      case 0: chart_set3(rgb, hi, up, lo); break;
      case 1: chart_set3(rgb, down, hi, lo); break;
      case 2: chart_set3(rgb, lo, hi, up); break;
      case 3: chart_set3(rgb, lo, down, hi); break;
      case 4: chart_set3(rgb, up, lo, hi); break;
      default: chart_set3(rgb, hi, lo, down); break;
    }
  } else if (fx < static_cast<float>(w / 2)) {
    const float third = static_cast<float>(w / 2 > 2 ? w / 2 : 2) / 3.0F;
    const int region = fx < third ? 0 : fx < 2.0F * third ? 1 : 2;
    const float bar = static_cast<float>(4 << region);
    const float v = (static_cast<int>(std::floor(fx / bar)) & 1) != 0 ? 0.85F : 0.04F;
    chart_set3(rgb, v, v, v);
  } else {
    const int top = h * 82 / 100;
    const float dx = fx - static_cast<float>(3 * w / 4 + s.shift);
    const float dy = fy - static_cast<float>((top + h) / 2);
    const float radius = static_cast<float>((h - top) / 3);
    if (dx * dx + dy * dy < radius * radius) {
      chart_set3(rgb, 0.6F, 0.08F, 0.06F);
    } else {
      const float v = 5.0F * dx + dy > 0.0F ? 0.75F : 0.12F;
      chart_set3(rgb, v, v, v);
    }
  }
}

// The mean of a pixel's 4 x 4 points, summed row by row as they come
// (synthetic_scene_rgb, synthetic_raw.hpp:241-253), times the exposure.
ISP_HD void chart_blur(const Chart& s, const float sum[3], float rgb[3]) {
  for (int k = 0; k < 3; ++k) rgb[k] = sum[k] * 0.0625F * s.exposure;
}

// The lens-blurred scene at pixel (x, y), evaluating all 16 points.
ISP_HD void chart_scene(const Chart& s, int x, int y, float rgb[3]) {
  float sum[3] = {0.0F, 0.0F, 0.0F};
  for (int j = 0; j < 4; ++j) {
    for (int i = 0; i < 4; ++i) {
      float p[3];
      chart_point(s, chart_coord(x, i), chart_coord(y, j), p);
      for (int k = 0; k < 3; ++k) sum[k] = sum[k] + p[k];
    }
  }
  chart_blur(s, sum, rgb);
}

// Vignetting falloff at (x, y) (synthetic_falloff, synthetic_raw.hpp:255-262).
ISP_HD float chart_falloff(const Chart& s, int x, int y) {
  const float dx = static_cast<float>(x) - s.cx;
  const float dy = static_cast<float>(y) - s.cy;
  return 1.0F - s.vignetting * ((dx * dx * s.inv_cx2 + dy * dy * s.inv_cy2) * 0.5F);
}

// Colour k of the camera's response to scene colour `rgb` (the inverse colour
// matrix, clipped at 0; synthetic_camera_rgb, synthetic_raw.hpp:264-274).
ISP_HD float chart_camera(const Chart& s, const float rgb[3], int k) {
  const float v = s.inv_ccm[3 * k] * rgb[0] + s.inv_ccm[3 * k + 1] * rgb[1] + s.inv_ccm[3 * k + 2] * rgb[2];
  return v > 0.0F ? v : 0.0F;
}

// The sensor's code at (x, y) for scene colour `rgb` there (synthetic_raw_sample,
// synthetic_raw.hpp:276-289).
ISP_HD int chart_sample_of(const Chart& s, const float rgb[3], int x, int y) {
  const int channel = s.cfa[((y & 1) << 1) | (x & 1)];
  const int colour = chart_colour(channel);
  float signal = chart_camera(s, rgb, colour) * chart_falloff(s, x, y) * s.inv_wb[colour];
  signal = signal < 1.0F ? signal : 1.0F;
  const float counts = s.black[channel] + signal * (s.white - s.black[channel]);
  const int value = static_cast<int>(counts + 0.5F);
  const int white = static_cast<int>(s.white);
  return value < 0 ? 0 : value > white ? white : value;
}

// What an ideal ISP outputs at (x, y) before its tone curve, for scene colour
// `rgb` there (synthetic_truth, synthetic_raw.hpp:292-310).
ISP_HD void chart_truth_of(const Chart& s, const float rgb[3], int x, int y, float linear[3]) {
  float cam[3];
  for (int k = 0; k < 3; ++k) cam[k] = chart_camera(s, rgb, k);
  const float fall = chart_falloff(s, x, y);
  float balanced[3];
  for (int k = 0; k < 3; ++k) {
    float signal = cam[k] * fall * s.inv_wb[k];
    signal = signal < 1.0F ? signal : 1.0F;
    const float v = signal / fall * s.wb[k];
    balanced[k] = v < 1.0F ? v : 1.0F;
  }
  for (int k = 0; k < 3; ++k) {
    const float v = s.ccm[3 * k] * balanced[0] + s.ccm[3 * k + 1] * balanced[1] + s.ccm[3 * k + 2] * balanced[2];
    linear[k] = v < 0.0F ? 0.0F : v > 1.0F ? 1.0F : v;
  }
}

// One packed byte: byte b of a row whose samples are samples[0 ..] (the MIPI
// CSI-2 packing of synthetic_write_unit, synthetic_raw.hpp:314-347, read byte
// by byte). Samples past the row's width must read 0.
template <typename Samples>
ISP_HD unsigned int chart_packed_byte(int format, const Samples& samples, int b) {
  switch (format) {
    case ISP_RAW_U8:
      return static_cast<unsigned int>(samples(b)) & 0xFFU;
    case ISP_RAW_U16:
      return (static_cast<unsigned int>(samples(b >> 1)) >> (8 * (b & 1))) & 0xFFU;
    case ISP_RAW_RAW10: {
      const int unit = b / 5;
      const int k = b - 5 * unit;
      if (k < 4) return (static_cast<unsigned int>(samples(4 * unit + k)) >> 2) & 0xFFU;
      unsigned int low = 0U;
      for (int q = 0; q < 4; ++q) low |= (static_cast<unsigned int>(samples(4 * unit + q)) & 3U) << (2 * q);
      return low;
    }
    default: {
      const int unit = b / 3;
      const int k = b - 3 * unit;
      if (k < 2) return (static_cast<unsigned int>(samples(2 * unit + k)) >> 4) & 0xFFU;
      return (static_cast<unsigned int>(samples(2 * unit)) & 15U) |
             ((static_cast<unsigned int>(samples(2 * unit + 1)) & 15U) << 4);
    }
  }
}

}  // namespace isp_amd
