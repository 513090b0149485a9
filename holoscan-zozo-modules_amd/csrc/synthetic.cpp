// synthetic.cpp - host-side test and benchmark inputs (include/isp_synthetic.h).
//
// 1. The synthetic raw camera with a closed-form truth. Its scene, lens blur,
//    vignetting, colour matrix, quantisation and MIPI packing follow the
//    reference's definition step for step (synthetic_raw.hpp:102-354,
//    synthetic_raw.cpp:46-154) so that its frames are byte-identical: tests
//    compare them with the reference's own renderer.
// 2. The seeded noisy scene: a smooth colour gradient, random axis-aligned
//    rectangles of uniform random reflectance, a per-colour sensor response,
//    cos^4 vignetting ((f^2 / (f^2 + r^2))^2, f set by the corner value),
//    Poisson shot noise (gain DN per electron) and Gaussian read noise, then the
//    black level, rounding and clipping to [0, 2^bits - 1]. Every pixel draws
//    from its own SplitMix64 stream keyed by (seed, x, y), so a frame is the same
//    whatever the thread count.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/isp_synthetic.h"
#include "chart.hpp"
#include "isp_internal.hpp"

namespace isp_amd {
namespace {

constexpr int kCfaTable[4][4] = {
    {ISP_CHANNEL_R, ISP_CHANNEL_GR, ISP_CHANNEL_GB, ISP_CHANNEL_B},
    {ISP_CHANNEL_B, ISP_CHANNEL_GB, ISP_CHANNEL_GR, ISP_CHANNEL_R},
    {ISP_CHANNEL_GR, ISP_CHANNEL_R, ISP_CHANNEL_B, ISP_CHANNEL_GB},
    {ISP_CHANNEL_GB, ISP_CHANNEL_B, ISP_CHANNEL_R, ISP_CHANNEL_GR},
};

int colour_of_channel(int channel) { return channel == ISP_CHANNEL_R ? 0 : channel == ISP_CHANNEL_B ? 2 : 1; }

int format_bits(int format) {
  return format == ISP_RAW_U8 ? 8 : format == ISP_RAW_RAW10 ? 10 : format == ISP_RAW_RAW12 ? 12 : 16;
}

// Runs body(y) for every row on up to hardware_concurrency threads.
template <typename Body>
void parallel_rows(int rows, Body body) {
  const int threads = std::max(1, std::min<int>(static_cast<int>(std::thread::hardware_concurrency()), 16));
  if (threads == 1 || rows < 64) {
    for (int y = 0; y < rows; ++y) body(y);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (int y = t; y < rows; y += threads) body(y);
    });
  }
  for (std::thread& th : pool) th.join();
}

// Writes one packing unit of samples `v` (count: 1, 4 for RAW10, 2 for RAW12)
// into `row` (synthetic_write_unit, synthetic_raw.hpp:314-347).
void pack_unit(int format, int unit, const int* v, int count, std::uint8_t* row) {
  switch (format) {
    case ISP_RAW_U8:
      row[unit] = static_cast<std::uint8_t>(v[0]);
      break;
    case ISP_RAW_U16: {
      const std::uint16_t s = static_cast<std::uint16_t>(v[0]);
      std::memcpy(row + 2 * static_cast<std::size_t>(unit), &s, 2);
      break;
    }
    case ISP_RAW_RAW10: {
      std::uint8_t* g = row + 5 * static_cast<std::size_t>(unit);
      g[4] = 0;
      for (int k = 0; k < 4; ++k) {
        const int s = k < count ? v[k] : 0;
        g[k] = static_cast<std::uint8_t>(s >> 2);
        g[4] = static_cast<std::uint8_t>(g[4] | ((s & 3) << (2 * k)));
      }
      break;
    }
    default: {
      std::uint8_t* g = row + 3 * static_cast<std::size_t>(unit);
      g[2] = 0;
      for (int k = 0; k < 2; ++k) {
        const int s = k < count ? v[k] : 0;
        g[k] = static_cast<std::uint8_t>(s >> 4);
        g[2] = static_cast<std::uint8_t>(g[2] | ((s & 15) << (4 * k)));
      }
      break;
    }
  }
}

// Packs one row of samples.
void pack_row(int format, int width, const int* samples, std::uint8_t* row) {
  const int per = format == ISP_RAW_RAW10 ? 4 : format == ISP_RAW_RAW12 ? 2 : 1;
  const int units = (width + per - 1) / per;
  for (int u = 0; u < units; ++u) {
    const int first = u * per;
    pack_unit(format, u, samples + first, std::min(per, width - first), row);
  }
}

// ---- the chart camera ----------------------------------------------------

int chart_sample(const Chart& s, int x, int y) {
  float rgb[3];
  chart_scene(s, x, y, rgb);
  return chart_sample_of(s, rgb, x, y);
}

void chart_truth(const Chart& s, int x, int y, float linear[3]) {
  float rgb[3];
  chart_scene(s, x, y, rgb);
  chart_truth_of(s, rgb, x, y, linear);
}

// ---- the noisy scene ------------------------------------------------------

std::uint64_t splitmix(std::uint64_t& state) {
  std::uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// A double in (0, 1].
double uniform(std::uint64_t& state) { return (static_cast<double>(splitmix(state) >> 11) + 1.0) * 0x1.0p-53; }

struct Rect {
  int x0, y0, x1, y1;
  double refl[3];
};

struct Scene {
  isp_scene_config c;
  std::vector<Rect> rects;
  double f2;  // squared focal length of the cos^4 falloff, in pixels
  double cx, cy;
  int cfa[4];
  double white;
};

Scene make_scene(const isp_scene_config& c) {
  if (c.width < 2 || c.height < 2 || c.width > 16384 || c.height > 16384)
    throw invalid_argument("scene must be 2x2 to 16384x16384");
  if (c.format > ISP_RAW_RAW12 || c.pattern > ISP_GBRG) throw invalid_argument("scene format or pattern unknown");
  const bool packed = c.format == ISP_RAW_RAW10 || c.format == ISP_RAW_RAW12;
  if (c.bit_depth < 2 || c.bit_depth > format_bits(c.format) || (packed && c.bit_depth != format_bits(c.format)))
    throw invalid_argument("scene bit depth does not fit its sample format");
  if (!(c.corner_falloff > 0.0F && c.corner_falloff <= 1.0F)) throw invalid_argument("corner_falloff must be in (0, 1]");
  if (!(c.noise_gain > 0.0F) || !(c.read_sigma >= 0.0F)) throw invalid_argument("noise_gain > 0, read_sigma >= 0");
  if (c.rectangles < 0 || c.rectangles > 100000) throw invalid_argument("rectangles must be 0..100000");
  Scene s;
  s.c = c;
  s.white = static_cast<double>((1U << static_cast<unsigned>(c.bit_depth)) - 1U);
  if (!(c.black_level >= 0.0F && c.black_level < s.white)) throw invalid_argument("scene black level must be below white");
  for (int k = 0; k < 4; ++k) s.cfa[k] = kCfaTable[c.pattern][k];
  s.cx = (c.width - 1) / 2.0;
  s.cy = (c.height - 1) / 2.0;
  // cos^4(atan(r / f)) = (f^2 / (f^2 + r^2))^2 equals corner_falloff at the corners.
  const double r2c = s.cx * s.cx + s.cy * s.cy;
  const double root = std::sqrt(static_cast<double>(c.corner_falloff));  // f^2 / (f^2 + r2c)
  s.f2 = root >= 1.0 ? 1e300 : root * r2c / (1.0 - root);
  std::uint64_t state = c.seed * 0x2545F4914F6CDD1DULL + 0x1234567ULL;
  for (int i = 0; i < c.rectangles; ++i) {
    Rect r{};
    int xa = static_cast<int>(splitmix(state) % static_cast<std::uint64_t>(c.width));
    int xb = static_cast<int>(splitmix(state) % static_cast<std::uint64_t>(c.width));
    int ya = static_cast<int>(splitmix(state) % static_cast<std::uint64_t>(c.height));
    int yb = static_cast<int>(splitmix(state) % static_cast<std::uint64_t>(c.height));
    r.x0 = std::min(xa, xb);
    r.x1 = std::max(xa, xb);
    r.y0 = std::min(ya, yb);
    r.y1 = std::max(ya, yb);
    for (double& v : r.refl) v = static_cast<double>(splitmix(state) >> 11) * 0x1.0p-53;
    s.rects.push_back(r);
  }
  return s;
}

// The scene's reflectance of colour `k` at (x, y): the last rectangle covering
// the pixel, else the gradient.
double reflectance(const Scene& s, int x, int y, int k) {
  for (auto it = s.rects.rbegin(); it != s.rects.rend(); ++it) {
    if (x >= it->x0 && x <= it->x1 && y >= it->y0 && y <= it->y1) return it->refl[k];
  }
  const double u = s.c.width > 1 ? static_cast<double>(x) / (s.c.width - 1) : 0.0;
  const double v = s.c.height > 1 ? static_cast<double>(y) / (s.c.height - 1) : 0.0;
  return k == 0 ? 0.1 + 0.6 * u : k == 1 ? 0.1 + 0.3 * (u + v) : 0.1 + 0.6 * v;
}

double scene_falloff(const Scene& s, int x, int y) {
  const double dx = x - s.cx;
  const double dy = y - s.cy;
  const double q = s.f2 / (s.f2 + dx * dx + dy * dy);
  return q * q;
}

int noisy_sample(const Scene& s, int x, int y) {
  const int channel = s.cfa[((y & 1) << 1) | (x & 1)];
  const int k = colour_of_channel(channel);
  const double range = s.white - s.c.black_level;
  const double signal = reflectance(s, x, y, k) * s.c.illumination[k] * scene_falloff(s, x, y);
  const double lambda = std::max(0.0, signal * range / s.c.noise_gain);  // electrons
  std::uint64_t state = s.c.seed ^ (static_cast<std::uint64_t>(y) * 0x9E3779B97F4A7C15ULL) ^
                        (static_cast<std::uint64_t>(x) * 0xC2B2AE3D27D4EB4FULL);
  splitmix(state);
  double electrons;
  if (lambda < 30.0) {
    // Knuth's multiplication method.
    const double limit = std::exp(-lambda);
    double p = 1.0;
    int n = -1;
    do {
      ++n;
      p *= uniform(state);
    } while (p > limit && n < 1000);
    electrons = n;
  } else {
    const double u1 = uniform(state);
    const double u2 = uniform(state);
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    electrons = std::max(0.0, std::floor(lambda + std::sqrt(lambda) * z + 0.5));
  }
  const double u3 = uniform(state);
  const double u4 = uniform(state);
  const double read = std::sqrt(-2.0 * std::log(u3)) * std::cos(6.283185307179586 * u4) * s.c.read_sigma;
  const double dn = electrons * s.c.noise_gain + read + s.c.black_level;
  const double rounded = std::floor(dn + 0.5);
  return static_cast<int>(rounded < 0.0 ? 0.0 : rounded > s.white ? s.white : rounded);
}

template <typename F>
isp_status guarded(F&& body) {
  try {
    body();
    clear_error();
    return ISP_OK;
  } catch (const invalid_argument& e) {
    set_error(e.what());
    return ISP_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    set_error(e.what());
    return ISP_FAILURE;
  }
}

}  // namespace

Chart make_chart(const isp_synthetic_config& c) {
  if (c.width < 16 || c.height < 16 || c.width > 16384 || c.height > 16384)
    throw invalid_argument("synthetic frames must be 16x16 to 16384x16384");
  if (c.format > ISP_RAW_RAW12 || c.pattern > ISP_GBRG) throw invalid_argument("synthetic format or pattern unknown");
  const bool packed = c.format == ISP_RAW_RAW10 || c.format == ISP_RAW_RAW12;
  if (c.bit_depth < 2 || c.bit_depth > format_bits(c.format) || (packed && c.bit_depth != format_bits(c.format)))
    throw invalid_argument("synthetic bit depth does not fit its sample format");
  if (!(c.vignetting >= 0.0F && c.vignetting <= 0.9F)) throw invalid_argument("synthetic vignetting must be 0..0.9");
  if (!(std::isfinite(c.exposure) && c.exposure > 0.0F && c.exposure <= 64.0F))
    throw invalid_argument("synthetic exposure must be in (0, 64]");
  Chart s{};
  s.width = c.width;
  s.height = c.height;
  s.kind = c.kind;
  s.format = c.format;
  for (int k = 0; k < 4; ++k) s.cfa[k] = kCfaTable[c.pattern][k];
  s.white = static_cast<float>((1U << static_cast<unsigned>(c.bit_depth)) - 1U);
  for (int k = 0; k < 4; ++k) {
    if (!(c.black_levels[k] >= 0.0F && c.black_levels[k] < s.white / 2.0F))
      throw invalid_argument("synthetic black levels must be below half the white level");
    s.black[k] = c.black_levels[k];
  }
  for (int k = 0; k < 3; ++k) {
    const float g = c.wb_gains[k];
    if (!(std::isfinite(g) && g >= 0.25F && g <= 8.0F)) throw invalid_argument("synthetic wb_gains must be 0.25..8");
    s.wb[k] = g;
    s.inv_wb[k] = 1.0F / g;
  }
  // The matrix inverse by cofactors in double (synthetic_raw.cpp:21-35).
  const double a = c.ccm[0], b = c.ccm[1], cc = c.ccm[2], d = c.ccm[3], e = c.ccm[4], f = c.ccm[5],
               g = c.ccm[6], h = c.ccm[7], i = c.ccm[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + cc * (d * h - e * g);
  if (!(std::abs(det) > 1e-6)) throw invalid_argument("synthetic ccm must be invertible");
  const double cof[9] = {(e * i - f * h),  -(b * i - cc * h), (b * f - cc * e),
                         -(d * i - f * g), (a * i - cc * g),  -(a * f - cc * d),
                         (d * h - e * g),  -(a * h - b * g),  (a * e - b * d)};
  for (int k = 0; k < 9; ++k) {
    s.ccm[k] = c.ccm[k];
    s.inv_ccm[k] = static_cast<float>(cof[k] / det);
  }
  s.vignetting = c.vignetting;
  s.exposure = c.exposure;
  s.cx = static_cast<float>(c.width - 1) / 2.0F;
  s.cy = static_cast<float>(c.height - 1) / 2.0F;
  s.inv_cx2 = 1.0F / (s.cx * s.cx);
  s.inv_cy2 = 1.0F / (s.cy * s.cy);
  const unsigned span = static_cast<unsigned>(c.width / 32 > 1 ? c.width / 32 : 1);
  const unsigned phase = c.frame % (2U * span);
  s.shift = 2 * static_cast<int>(phase < span ? phase : 2U * span - phase);
  return s;
}

}  // namespace isp_amd

using namespace isp_amd;

extern "C" {

void isp_synthetic_config_init(isp_synthetic_config* c) {
  if (c == nullptr) return;
  std::memset(c, 0, sizeof(*c));
  c->width = 640;
  c->height = 480;
  c->pattern = ISP_RGGB;
  c->format = ISP_RAW_U16;
  c->bit_depth = 12;
  for (float& b : c->black_levels) b = 256.0F;
  c->wb_gains[0] = 1.9F;
  c->wb_gains[1] = 1.0F;
  c->wb_gains[2] = 1.6F;
  const float ccm[9] = {1.6F, -0.4F, -0.2F, -0.3F, 1.5F, -0.2F, -0.1F, -0.5F, 1.6F};
  std::memcpy(c->ccm, ccm, sizeof(ccm));
  c->vignetting = 0.0F;
  c->kind = ISP_SCENE_CHART;
  c->exposure = 1.0F;
  c->frame = 0;
}

isp_status isp_synthetic_render(const isp_synthetic_config* config, void* data, size_t pitch) {
  return guarded([&] {
    if (config == nullptr || data == nullptr) throw invalid_argument("config or data is null");
    const Chart s = make_chart(*config);
    if (pitch < raw_row_bytes(s.format, s.width)) throw invalid_argument("pitch is shorter than a row");
    parallel_rows(s.height, [&](int y) {
      std::vector<int> samples(static_cast<std::size_t>(s.width));
      for (int x = 0; x < s.width; ++x) samples[static_cast<std::size_t>(x)] = chart_sample(s, x, y);
      pack_row(s.format, s.width, samples.data(), static_cast<std::uint8_t*>(data) + static_cast<std::size_t>(y) * pitch);
    });
  });
}

isp_status isp_synthetic_params(const isp_synthetic_config* config, int32_t grid, isp_params* params,
                                float* lsc_gains) {
  return guarded([&] {
    if (config == nullptr || params == nullptr) throw invalid_argument("config or params is null");
    const Chart s = make_chart(*config);
    isp_params_init(params);
    params->has_channel_black_levels = 1;
    for (int k = 0; k < 4; ++k) params->channel_black_levels[k] = config->black_levels[k];
    params->white_level = s.white;
    for (int k = 0; k < 3; ++k) params->wb_gains[k] = config->wb_gains[k];
    for (int k = 0; k < 9; ++k) params->ccm[k] = config->ccm[k];
    if (config->vignetting > 0.0F) {
      if (grid < 2 || grid > 256) throw invalid_argument("lens shading grid must be 2..256 points");
      if (lsc_gains == nullptr) throw invalid_argument("lsc_gains is null");
      for (int gy = 0; gy < grid; ++gy) {
        for (int gx = 0; gx < grid; ++gx) {
          const double x = static_cast<double>(gx) * (config->width - 1) / (grid - 1);
          const double y = static_cast<double>(gy) * (config->height - 1) / (grid - 1);
          const double dx = x - s.cx;
          const double dy = y - s.cy;
          const double fall =
              1.0 - config->vignetting * 0.5 * (dx * dx / (s.cx * s.cx) + dy * dy / (s.cy * s.cy));
          lsc_gains[gy * grid + gx] = static_cast<float>(1.0 / fall);
        }
      }
      params->has_lens_shading = 1;
      params->lens_shading.width = grid;
      params->lens_shading.height = grid;
      params->lens_shading.channels = 1;
      params->lens_shading.gains = lsc_gains;
      params->lens_shading.gains_count = static_cast<int64_t>(grid) * grid;
    }
  });
}

isp_status isp_synthetic_truth(const isp_synthetic_config* config, int32_t x, int32_t y, float linear[3]) {
  return guarded([&] {
    if (config == nullptr || linear == nullptr) throw invalid_argument("config or linear is null");
    chart_truth(make_chart(*config), x, y, linear);
  });
}

isp_status isp_synthetic_truth_image(const isp_synthetic_config* config, float* linear) {
  return guarded([&] {
    if (config == nullptr || linear == nullptr) throw invalid_argument("config or linear is null");
    const Chart s = make_chart(*config);
    parallel_rows(s.height, [&](int y) {
      for (int x = 0; x < s.width; ++x) chart_truth(s, x, y, linear + (static_cast<std::size_t>(y) * s.width + x) * 3U);
    });
  });
}

void isp_scene_config_init(isp_scene_config* c) {
  if (c == nullptr) return;
  std::memset(c, 0, sizeof(*c));
  c->width = 1920;
  c->height = 1080;
  c->pattern = ISP_RGGB;
  c->format = ISP_RAW_U16;
  c->bit_depth = 12;
  c->black_level = 256.0F;
  c->seed = 1;
  c->rectangles = 40;
  c->corner_falloff = 0.6F;
  c->noise_gain = 0.5F;
  c->read_sigma = 4.0F;
  c->illumination[0] = 0.5F;
  c->illumination[1] = 1.0F;
  c->illumination[2] = 0.625F;
}

isp_status isp_scene_render(const isp_scene_config* config, void* data, size_t pitch) {
  return guarded([&] {
    if (config == nullptr || data == nullptr) throw invalid_argument("config or data is null");
    const Scene s = make_scene(*config);
    if (pitch < raw_row_bytes(config->format, config->width)) throw invalid_argument("pitch is shorter than a row");
    parallel_rows(config->height, [&](int y) {
      std::vector<int> samples(static_cast<std::size_t>(config->width));
      for (int x = 0; x < config->width; ++x) samples[static_cast<std::size_t>(x)] = noisy_sample(s, x, y);
      pack_row(config->format, config->width, samples.data(),
               static_cast<std::uint8_t*>(data) + static_cast<std::size_t>(y) * pitch);
    });
  });
}

isp_status isp_scene_lens_shading(const isp_scene_config* config, int32_t gw, int32_t gh, float* gains) {
  return guarded([&] {
    if (config == nullptr || gains == nullptr) throw invalid_argument("config or gains is null");
    if (gw < 2 || gh < 2 || gw > 256 || gh > 256) throw invalid_argument("lens shading grid must be 2..256 points");
    const Scene s = make_scene(*config);
    for (int gy = 0; gy < gh; ++gy) {
      for (int gx = 0; gx < gw; ++gx) {
        const double x = static_cast<double>(gx) * (config->width - 1) / (gw - 1);
        const double y = static_cast<double>(gy) * (config->height - 1) / (gh - 1);
        const double dx = x - s.cx;
        const double dy = y - s.cy;
        const double q = s.f2 / (s.f2 + dx * dx + dy * dy);
        gains[gy * gw + gx] = static_cast<float>(1.0 / (q * q));
      }
    }
  });
}

}  // extern "C"
