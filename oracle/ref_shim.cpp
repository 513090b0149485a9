// ref_shim.cpp - TEST INFRASTRUCTURE ONLY: a C entry into the reference's own
// CPU backend, compiled from the reference's sources where they lie
// (modules/isp/src/cpu_pipeline.cpp, isp_params.cpp, synthetic_raw.cpp) into
// oracle/_ref/libisp_ref.so by oracle/Makefile. It translates this repository's
// C structs (include/isp_c.h) into the reference's holoscan::isp types and
// repeats what IspPipeline::process does for the CPU backend
// (modules/isp/src/isp_pipeline.cpp:113-160): validate, build the tone table,
// make_frame_params, cpu_process. Used only to pin the oracle and to write the
// golden fixtures; never shipped in the product.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include <holoscan/isp/isp_params.hpp>
#include <holoscan/isp/synthetic_raw.hpp>

#include "cpu_pipeline.hpp"
#include "../include/isp_c.h"
#include "../include/isp_synthetic.h"

namespace isp = holoscan::isp;

namespace {

thread_local std::string g_error;

isp::IspParams to_params(const isp_params* c) {
  isp::IspParams p;
  p.black_level = c->black_level;
  if (c->has_channel_black_levels) {
    p.channel_black_levels = std::array<float, 4>{c->channel_black_levels[0], c->channel_black_levels[1],
                                                  c->channel_black_levels[2], c->channel_black_levels[3]};
  }
  p.white_level = c->white_level;
  if (c->has_lens_shading) {
    isp::LensShading s;
    s.width = c->lens_shading.width;
    s.height = c->lens_shading.height;
    s.channels = c->lens_shading.channels;
    s.gains.assign(c->lens_shading.gains, c->lens_shading.gains + c->lens_shading.gains_count);
    p.lens_shading = std::move(s);
  }
  p.white_balance = static_cast<isp::WhiteBalanceMode>(c->white_balance);
  for (int k = 0; k < 3; ++k) p.wb_gains[k] = c->wb_gains[k];
  p.awb_smoothing = c->awb_smoothing;
  p.white_patch_percentile = c->white_patch_percentile;
  p.demosaic = static_cast<isp::DemosaicMethod>(c->demosaic);
  for (int k = 0; k < 9; ++k) p.ccm[k] = c->ccm[k];
  p.tone_curve = static_cast<isp::ToneCurve>(c->tone_curve);
  p.gamma = c->gamma;
  if (c->lut != nullptr && c->lut_size > 0) p.lut.assign(c->lut, c->lut + c->lut_size);
  p.sharpen = c->sharpen;
  p.auto_exposure.target = c->auto_exposure.target;
  p.auto_exposure.damping = c->auto_exposure.damping;
  p.auto_exposure.min_exposure_ns = c->auto_exposure.min_exposure_ns;
  p.auto_exposure.max_exposure_ns = c->auto_exposure.max_exposure_ns;
  p.auto_exposure.min_gain = c->auto_exposure.min_gain;
  p.auto_exposure.max_gain = c->auto_exposure.max_gain;
  return p;
}

isp::RawImageView to_raw(const isp_raw_view* r) {
  return isp::RawImageView{r->data, r->width, r->height, r->pitch,
                           static_cast<isp::RawFormat>(r->format), r->bit_depth,
                           static_cast<isp::BayerPattern>(r->pattern)};
}

isp::SyntheticRawConfig to_synthetic(const isp_synthetic_config* c) {
  isp::SyntheticRawConfig s;
  s.width = c->width;
  s.height = c->height;
  s.pattern = static_cast<isp::BayerPattern>(c->pattern);
  s.format = static_cast<isp::RawFormat>(c->format);
  s.bit_depth = c->bit_depth;
  for (int k = 0; k < 4; ++k) s.black_levels[k] = c->black_levels[k];
  for (int k = 0; k < 3; ++k) s.wb_gains[k] = c->wb_gains[k];
  for (int k = 0; k < 9; ++k) s.ccm[k] = c->ccm[k];
  s.vignetting = c->vignetting;
  s.kind = static_cast<isp::SceneKind>(c->kind);
  s.exposure = c->exposure;
  s.frame = c->frame;
  return s;
}

struct RefPipeline {
  isp::IspParams params;
  std::vector<float> lut;
  isp::detail::AwbState awb{};
  isp::detail::CpuScratch scratch;
};

int fail(const std::exception& e, int code) {
  g_error = e.what();
  return code;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

int ref_validate_params(const isp_params* c) {
  try {
    isp::validate_params(to_params(c));
    return 0;
  } catch (const std::invalid_argument& e) {
    return fail(e, 1);
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

int ref_validate_raw(const isp_raw_view* raw, const isp_params* c) {
  try {
    isp::validate_raw(to_raw(raw), to_params(c));
    return 0;
  } catch (const std::invalid_argument& e) {
    return fail(e, 1);
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

size_t ref_raw_row_bytes(int format, int width) {
  return isp::raw_row_bytes(static_cast<isp::RawFormat>(format), width);
}

int ref_tone_lut(const isp_params* c, float* out, int capacity) {
  const std::vector<float> lut = isp::tone_lut(to_params(c));
  for (int i = 0; i < static_cast<int>(lut.size()) && i < capacity; ++i) out[i] = lut[i];
  return static_cast<int>(lut.size());
}

void* ref_create(const isp_params* c) {
  try {
    auto* r = new RefPipeline;
    r->params = to_params(c);
    isp::validate_params(r->params);
    r->lut = isp::tone_lut(r->params);
    r->awb = isp::detail::initial_white_balance(r->params);
    return r;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_destroy(void* handle) { delete static_cast<RefPipeline*>(handle); }

void ref_reset(void* handle) {
  auto* r = static_cast<RefPipeline*>(handle);
  r->awb = isp::detail::initial_white_balance(r->params);
}

// IspPipeline::process for the CPU backend (isp_pipeline.cpp:141-160).
int ref_process(void* handle, const isp_raw_view* raw, const isp_rgb_view* out,
                const isp_stats_out* stats, const isp_frame_exposure* exposure) {
  auto* r = static_cast<RefPipeline*>(handle);
  try {
    const isp::RawImageView view = to_raw(raw);
    isp::validate_raw(view, r->params);
    const isp::RgbImageView rgb{out->data, out->pitch, static_cast<isp::OutputFormat>(out->format)};
    if (rgb.data == nullptr) throw std::invalid_argument("output image has no data");
    const std::size_t row = static_cast<std::size_t>(view.width) * isp::output_channels(rgb.format) *
                            isp::output_channel_bytes(rgb.format);
    if (rgb.pitch < row || rgb.pitch % static_cast<std::size_t>(isp::output_channel_bytes(rgb.format))) {
      throw std::invalid_argument("output image pitch is shorter than a row or misaligned");
    }
    const isp::detail::FrameParams p = isp::detail::make_frame_params(
        r->params, view, rgb, r->lut.data(), static_cast<int>(r->lut.size()),
        r->params.lens_shading ? r->params.lens_shading->gains.data() : nullptr);
    isp::StatsOutput s{};
    if (stats != nullptr) {
      s.histogram = stats->histogram;
      s.stats = stats->stats;
    }
    isp::FrameExposure e{};
    if (exposure != nullptr) {
      e.exposure_ns = exposure->exposure_ns;
      e.gain = exposure->gain;
    }
    isp::detail::cpu_process(p, r->params, s, e, r->awb, r->scratch);
    return 0;
  } catch (const std::invalid_argument& e) {
    return fail(e, 1);
  } catch (const std::exception& e) {
    return fail(e, 2);
  }
}

void ref_awb_state(void* handle, float gains[3], int* valid) {
  auto* r = static_cast<RefPipeline*>(handle);
  for (int k = 0; k < 3; ++k) gains[k] = r->awb.gains[k];
  *valid = r->awb.valid;
}

int ref_synthetic_render(const isp_synthetic_config* c, void* data, size_t pitch) {
  try {
    isp::render_synthetic_raw(isp::make_synthetic_scene(to_synthetic(c)), data, pitch);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 1);
  }
}

// synthetic_isp_params into a C struct; `lsc` receives grid*grid gains.
int ref_synthetic_params(const isp_synthetic_config* c, int grid, isp_params* out, float* lsc) {
  try {
    const isp::IspParams p = isp::synthetic_isp_params(to_synthetic(c), grid);
    out->black_level = p.black_level;
    out->has_channel_black_levels = p.channel_black_levels.has_value();
    if (p.channel_black_levels)
      for (int k = 0; k < 4; ++k) out->channel_black_levels[k] = (*p.channel_black_levels)[k];
    out->white_level = p.white_level;
    for (int k = 0; k < 3; ++k) out->wb_gains[k] = p.wb_gains[k];
    for (int k = 0; k < 9; ++k) out->ccm[k] = p.ccm[k];
    out->has_lens_shading = p.lens_shading.has_value();
    if (p.lens_shading) {
      std::memcpy(lsc, p.lens_shading->gains.data(), p.lens_shading->gains.size() * sizeof(float));
      out->lens_shading.width = p.lens_shading->width;
      out->lens_shading.height = p.lens_shading->height;
      out->lens_shading.channels = p.lens_shading->channels;
      out->lens_shading.gains = lsc;
      out->lens_shading.gains_count = static_cast<int64_t>(p.lens_shading->gains.size());
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 1);
  }
}

int ref_synthetic_truth(const isp_synthetic_config* c, int x, int y, float linear[3]) {
  try {
    isp::synthetic_truth(isp::make_synthetic_scene(to_synthetic(c)), x, y, linear);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, 1);
  }
}

}  // extern "C"
