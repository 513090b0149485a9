// isp_params.cpp - settings: defaults, validation, the tone table, row sizes,
// names, and the thread-local error message of the C ABI.
//
// Validation follows the reference's rules and messages field by field
// (modules/isp/src/isp_params.cpp:85-164 validate_params, :199-237
// validate_raw, isp_pipeline.cpp:145-152 the output view), so a caller sees the
// same first-bad-field message; tone_lut follows isp_params.cpp:166-182 (double
// pow on the host, once per pipeline).
#include <cmath>
#include <cstring>
#include <string>

#include "isp_internal.hpp"

namespace isp_amd {
namespace {

thread_local std::string g_error;

bool finite_in(float value, float low, float high) {
  return std::isfinite(value) && value >= low && value <= high;
}

struct Name {
  std::uint8_t value;
  const char* name;
};
constexpr Name kBackends[] = {{ISP_BACKEND_GPU, "gpu"}, {ISP_BACKEND_CPU, "cpu"}, {ISP_BACKEND_AUTO, "auto"}};
constexpr Name kDemosaics[] = {{ISP_DEMOSAIC_BILINEAR, "bilinear"}, {ISP_DEMOSAIC_MALVAR, "malvar"}};
constexpr Name kPatterns[] = {{ISP_RGGB, "RGGB"}, {ISP_BGGR, "BGGR"}, {ISP_GRBG, "GRBG"}, {ISP_GBRG, "GBRG"}};
constexpr Name kWhiteBalances[] = {
    {ISP_WB_MANUAL, "manual"}, {ISP_WB_GRAY_WORLD, "gray-world"}, {ISP_WB_WHITE_PATCH, "white-patch"}};

template <std::size_t N>
const char* name_of(const Name (&names)[N], std::uint8_t value) {
  for (const Name& n : names)
    if (n.value == value) return n.name;
  return "unknown";
}
template <std::size_t N>
int parse_of(const Name (&names)[N], const char* text, std::uint8_t* value) {
  if (text == nullptr) return 0;
  for (const Name& n : names) {
    if (std::strcmp(n.name, text) == 0) {
      if (value != nullptr) *value = n.value;
      return 1;
    }
  }
  return 0;
}

int max_bits(int format) {
  return format == ISP_RAW_U8 ? 8 : format == ISP_RAW_RAW10 ? 10 : format == ISP_RAW_RAW12 ? 12 : 16;
}

}  // namespace

void set_error(const std::string& message) { g_error = message; }
void clear_error() { g_error.clear(); }

void validate_params(const isp_params& p) {
  if (!finite_in(p.black_level, 0.0F, 65535.0F)) throw invalid_argument("black_level must be 0..65535");
  if (p.has_channel_black_levels) {
    for (const float level : p.channel_black_levels)
      if (!finite_in(level, 0.0F, 65535.0F))
        throw invalid_argument("channel_black_levels must be 0..65535");
  }
  if (!finite_in(p.white_level, 0.0F, 65535.0F))
    throw invalid_argument("white_level must be 0 (from the bit depth) or up to 65535");
  if (p.has_lens_shading) {
    const isp_lens_shading& s = p.lens_shading;
    if (s.width < 2 || s.height < 2 || s.width > 256 || s.height > 256)
      throw invalid_argument("lens_shading grid must be 2x2 to 256x256 points");
    if (s.channels != 1 && s.channels != 4)
      throw invalid_argument("lens_shading channels must be 1 or 4");
    if (s.gains == nullptr ||
        s.gains_count != static_cast<std::int64_t>(s.width) * s.height * s.channels)
      throw invalid_argument("lens_shading gains must hold width x height x channels values");
    for (std::int64_t i = 0; i < s.gains_count; ++i) {
      const float gain = s.gains[i];
      if (!(std::isfinite(gain) && gain > 0.0F && gain <= 16.0F))
        throw invalid_argument("lens_shading gains must be in (0, 16]");
    }
  }
  for (const float gain : p.wb_gains)
    if (!(std::isfinite(gain) && gain > 0.0F && gain <= 64.0F))
      throw invalid_argument("wb_gains must be in (0, 64]");
  if (!(std::isfinite(p.awb_smoothing) && p.awb_smoothing > 0.0F && p.awb_smoothing <= 1.0F))
    throw invalid_argument("awb_smoothing must be in (0, 1]");
  if (!(std::isfinite(p.white_patch_percentile) && p.white_patch_percentile >= 0.5F &&
        p.white_patch_percentile < 1.0F))
    throw invalid_argument("white_patch_percentile must be in [0.5, 1)");
  for (const float value : p.ccm)
    if (!finite_in(value, -16.0F, 16.0F)) throw invalid_argument("ccm entries must be finite, -16..16");
  if (p.white_balance > ISP_WB_WHITE_PATCH) throw invalid_argument("white_balance is not a WhiteBalanceMode");
  if (p.demosaic > ISP_DEMOSAIC_MALVAR) throw invalid_argument("demosaic is not a DemosaicMethod");
  if (p.tone_curve > ISP_TONE_LUT) throw invalid_argument("tone_curve is not a ToneCurve");
  if (p.tone_curve == ISP_TONE_GAMMA && !finite_in(p.gamma, 0.1F, 10.0F))
    throw invalid_argument("gamma must be 0.1..10");
  if (p.tone_curve == ISP_TONE_LUT) {
    if (p.lut == nullptr || p.lut_size < 2 || p.lut_size > 65536)
      throw invalid_argument("lut must have 2..65536 entries");
    for (int i = 0; i < p.lut_size; ++i)
      if (!finite_in(p.lut[i], 0.0F, 1.0F)) throw invalid_argument("lut entries must be 0..1");
  }
  if (!finite_in(p.sharpen, 0.0F, 4.0F)) throw invalid_argument("sharpen must be 0..4");
  const isp_auto_exposure& ae = p.auto_exposure;
  if (!(std::isfinite(ae.target) && ae.target > 0.0F && ae.target < 1.0F))
    throw invalid_argument("auto_exposure.target must be in (0, 1)");
  if (!finite_in(ae.damping, 0.0F, 1.0F)) throw invalid_argument("auto_exposure.damping must be 0..1");
  if (ae.min_exposure_ns <= 0 || ae.max_exposure_ns < ae.min_exposure_ns)
    throw invalid_argument("auto_exposure needs 0 < min_exposure_ns <= max_exposure_ns");
  if (!(std::isfinite(ae.min_gain) && std::isfinite(ae.max_gain) && ae.min_gain > 0.0F &&
        ae.max_gain >= ae.min_gain))
    throw invalid_argument("auto_exposure needs 0 < min_gain <= max_gain");
  if (p.histogram_bins != 0 && p.histogram_bins != ISP_HISTOGRAM_BINS &&
      p.histogram_bins != ISP_HISTOGRAM_BINS_WIDE)
    throw invalid_argument("histogram_bins must be 256 or 1024");
}

Params copy_params(const isp_params* params) {
  if (params == nullptr) throw invalid_argument("params is null");
  validate_params(*params);
  Params out;
  out.c = *params;
  if (params->has_lens_shading) {
    out.lsc.assign(params->lens_shading.gains, params->lens_shading.gains + params->lens_shading.gains_count);
    out.c.lens_shading.gains = out.lsc.data();
  } else {
    out.c.lens_shading = isp_lens_shading{};
  }
  if (params->tone_curve == ISP_TONE_LUT) {
    out.lut.assign(params->lut, params->lut + params->lut_size);
    out.c.lut = out.lut.data();
  } else {
    out.c.lut = nullptr;
    out.c.lut_size = 0;
  }
  if (out.c.histogram_bins == 0) out.c.histogram_bins = ISP_HISTOGRAM_BINS;
  return out;
}

std::vector<float> tone_lut(const isp_params& p) {
  if (p.tone_curve == ISP_TONE_LUT) return std::vector<float>(p.lut, p.lut + p.lut_size);
  std::vector<float> lut(ISP_TONE_LUT_SIZE);
  for (int i = 0; i < ISP_TONE_LUT_SIZE; ++i) {
    const double x = static_cast<double>(i) / (ISP_TONE_LUT_SIZE - 1);
    double y = x;
    if (p.tone_curve == ISP_TONE_SRGB) {
      y = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
    } else if (p.tone_curve == ISP_TONE_GAMMA) {
      y = std::pow(x, 1.0 / static_cast<double>(p.gamma));
    }
    lut[static_cast<std::size_t>(i)] = static_cast<float>(y < 0.0 ? 0.0 : y > 1.0 ? 1.0 : y);
  }
  return lut;
}

std::size_t raw_row_bytes(int format, int width) {
  const auto w = static_cast<std::size_t>(width < 0 ? 0 : width);
  switch (format) {
    case ISP_RAW_U8:
      return w;
    case ISP_RAW_U16:
      return 2U * w;
    case ISP_RAW_RAW10:
      return 5U * ((w + 3U) / 4U);
    case ISP_RAW_RAW12:
      return 3U * ((w + 1U) / 2U);
    default:
      return 0;
  }
}

void validate_raw(const isp_raw_view& raw, const isp_params& params) {
  if (raw.data == nullptr) throw invalid_argument("raw image has no data");
  if (raw.format > ISP_RAW_RAW12) throw invalid_argument("raw format is not a RawFormat");
  if (raw.pattern > ISP_GBRG) throw invalid_argument("raw pattern is not a BayerPattern");
  if (raw.width < 2 || raw.height < 2 || raw.width > 16384 || raw.height > 16384)
    throw invalid_argument("raw image must be 2x2 to 16384x16384");
  if (raw.pitch < raw_row_bytes(raw.format, raw.width))
    throw invalid_argument("raw image pitch is shorter than a row");
  if (raw.format == ISP_RAW_U16 &&
      (raw.pitch % 2U != 0U || reinterpret_cast<std::uintptr_t>(raw.data) % 2U != 0U))
    throw invalid_argument("uint16 raw samples must be aligned to 2 bytes, rows too");
  const int bits = max_bits(raw.format);
  const bool packed = raw.format == ISP_RAW_RAW10 || raw.format == ISP_RAW_RAW12;
  if (raw.bit_depth < 1 || raw.bit_depth > bits || (packed && raw.bit_depth != bits))
    throw invalid_argument("raw bit depth " + std::to_string(raw.bit_depth) +
                           " does not fit its sample format");
  const float full = static_cast<float>((1U << static_cast<unsigned>(raw.bit_depth)) - 1U);
  const float white = params.white_level > 0.0F ? params.white_level : full;
  if (white > full) throw invalid_argument("white_level exceeds the raw bit depth");
  for (int k = 0; k < 4; ++k) {
    const float level = params.has_channel_black_levels ? params.channel_black_levels[k] : params.black_level;
    if (!(level < white)) throw invalid_argument("black levels must be below the white level");
  }
}

void validate_out(const isp_raw_view& raw, const isp_rgb_view& out, const isp_params& params) {
  if (out.data == nullptr) throw invalid_argument("output image has no data");
  if (out.format > ISP_OUT_RGB32F) throw invalid_argument("output format is not an OutputFormat");
  const std::size_t channel = static_cast<std::size_t>(isp_output_channel_bytes(out.format));
  const std::size_t row = static_cast<std::size_t>(raw.width) * isp_output_channels(out.format) * channel;
  if (out.pitch < row || out.pitch % channel != 0U)
    throw invalid_argument("output image pitch is shorter than a row or misaligned");
  if (reinterpret_cast<std::uintptr_t>(out.data) % channel != 0U)
    throw invalid_argument("output image data is misaligned for its channel size");
  if (params.histogram_bins == ISP_HISTOGRAM_BINS_WIDE && out.format != ISP_OUT_RGB16 &&
      out.format != ISP_OUT_RGB32F)
    throw invalid_argument("1024-bin histograms need RGB16 or RGB32F output");
}

}  // namespace isp_amd

using namespace isp_amd;

namespace {
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

extern "C" {

void isp_params_init(isp_params* p) {
  if (p == nullptr) return;
  std::memset(p, 0, sizeof(*p));
  p->black_level = 0.0F;
  p->white_level = 0.0F;
  p->white_balance = ISP_WB_MANUAL;
  p->wb_gains[0] = p->wb_gains[1] = p->wb_gains[2] = 1.0F;
  p->awb_smoothing = 0.25F;
  p->white_patch_percentile = 0.99F;
  p->demosaic = ISP_DEMOSAIC_MALVAR;
  const float identity[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(p->ccm, identity, sizeof(identity));
  p->tone_curve = ISP_TONE_SRGB;
  p->gamma = 2.2F;
  p->sharpen = 0.0F;
  p->auto_exposure.target = 0.18F;
  p->auto_exposure.damping = 0.5F;
  p->auto_exposure.min_exposure_ns = 10'000;
  p->auto_exposure.max_exposure_ns = 33'000'000;
  p->auto_exposure.min_gain = 1.0F;
  p->auto_exposure.max_gain = 16.0F;
  p->histogram_bins = ISP_HISTOGRAM_BINS;
}

isp_status isp_validate_params(const isp_params* params) {
  return guarded([&] {
    if (params == nullptr) throw invalid_argument("params is null");
    validate_params(*params);
  });
}

isp_status isp_validate_raw(const isp_raw_view* raw, const isp_params* params) {
  return guarded([&] {
    if (raw == nullptr || params == nullptr) throw invalid_argument("raw view or params is null");
    validate_raw(*raw, *params);
  });
}

size_t isp_raw_row_bytes(uint8_t format, int32_t width) { return raw_row_bytes(format, width); }

isp_status isp_tone_lut(const isp_params* params, float* out, int32_t capacity, int32_t* size) {
  return guarded([&] {
    if (params == nullptr) throw invalid_argument("params is null");
    validate_params(*params);
    const std::vector<float> lut = tone_lut(*params);
    if (size != nullptr) *size = static_cast<int32_t>(lut.size());
    if (out != nullptr) {
      if (capacity < static_cast<int32_t>(lut.size())) throw invalid_argument("tone table capacity is too small");
      std::memcpy(out, lut.data(), lut.size() * sizeof(float));
    }
  });
}

int32_t isp_output_channels(uint8_t format) { return format == ISP_OUT_RGBA8 ? 4 : 3; }
int32_t isp_output_channel_bytes(uint8_t format) {
  return format == ISP_OUT_RGB16 ? 2 : format == ISP_OUT_RGB32F ? 4 : 1;
}

const char* isp_backend_name(uint8_t b) { return name_of(kBackends, b); }
int isp_parse_backend(const char* s, uint8_t* v) { return parse_of(kBackends, s, v); }
const char* isp_demosaic_name(uint8_t m) { return name_of(kDemosaics, m); }
int isp_parse_demosaic(const char* s, uint8_t* v) { return parse_of(kDemosaics, s, v); }
const char* isp_pattern_name(uint8_t p) { return name_of(kPatterns, p); }
int isp_parse_pattern(const char* s, uint8_t* v) { return parse_of(kPatterns, s, v); }
const char* isp_white_balance_name(uint8_t w) { return name_of(kWhiteBalances, w); }
int isp_parse_white_balance(const char* s, uint8_t* v) { return parse_of(kWhiteBalances, s, v); }

uint8_t isp_resolve_auto(int gpu_device) { return gpu_device ? ISP_BACKEND_GPU : ISP_BACKEND_CPU; }

const char* isp_last_error(void) { return g_error.c_str(); }

}  // extern "C"

extern "C" {

// The host evaluation of the kernels' statistics epilogue (isp_ae.hpp), for
// verification against the host libm.
double isp_epilogue_log2(double x) { return isp_amd::isp_log2(x); }
double isp_epilogue_exp2(double y) { return isp_amd::isp_exp2(y); }
void isp_epilogue_finish_stats(const uint64_t sums[4], int64_t pixels, const float gains[3],
                               const isp_auto_exposure* ae, const isp_frame_exposure* exposure, float* out) {
  const unsigned long long s[4] = {sums[0], sums[1], sums[2], sums[3]};
  const isp_amd::AeParams p{ae->target, ae->damping, ae->min_exposure_ns, ae->max_exposure_ns, ae->min_gain, ae->max_gain};
  isp_amd::finish_stats(s, pixels, gains, p, exposure != nullptr ? exposure->exposure_ns : 0,
                        exposure != nullptr ? exposure->gain : 0.0F, out);
}

}  // extern "C"
