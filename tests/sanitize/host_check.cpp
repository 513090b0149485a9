// host_check.cpp - the host code under AddressSanitizer + UndefinedBehaviorSanitizer
// (and, built apart, ThreadSanitizer): settings validation, the tone table, the
// synthetic chart camera and the seeded noisy scene (renderers that split rows
// over threads), the statistics epilogue's host evaluation, and the CPU oracle
// (row bands on threads) over every raw format, Bayer layout, demosaic,
// sharpening, white balance mode and output format at small and odd sizes,
// the oracle's frame checker and colour conversion. The reference runs its
// pipeline test under compute-sanitizer's memcheck, racecheck, initcheck and
// synccheck (modules/isp/tests/CMakeLists.txt:37-50); this is the host half of
// that tier: the GPU kernels are checked by their own bounds logic and tests.
//
// Exit status 0 with "host_check: N cases, 0 failures" on success; any
// sanitizer report aborts (built with -fno-sanitize-recover=all).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/isp_c.h"
#include "../../include/isp_color.h"
#include "../../include/isp_source.h"
#include "../../include/isp_synthetic.h"
#include "../../oracle/isp_oracle.h"

namespace {

int g_cases = 0;
int g_failures = 0;

void expect(bool ok, const char* what) {
  ++g_cases;
  if (!ok) {
    ++g_failures;
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, isp_last_error());
  }
}

// Settings: every rule of validate_params on one bad field at a time, the
// defaults, and the tone table at its size limits.
void check_params() {
  isp_params p;
  isp_params_init(&p);
  expect(isp_validate_params(&p) == ISP_OK, "default params validate");
  expect(isp_validate_params(nullptr) == ISP_INVALID_ARGUMENT, "null params");
  const float nan = std::nanf("");
  struct Bad {
    const char* what;
    void (*set)(isp_params*, float);
  };
  const Bad bad[] = {
      {"black level", [](isp_params* q, float v) { q->black_level = v; }},
      {"white level", [](isp_params* q, float v) { q->white_level = v; }},
      {"wb gain", [](isp_params* q, float v) { q->wb_gains[1] = v; }},
      {"smoothing", [](isp_params* q, float v) { q->awb_smoothing = v; }},
      {"percentile", [](isp_params* q, float v) { q->white_patch_percentile = v; }},
      {"ccm", [](isp_params* q, float v) { q->ccm[4] = v; }},
      {"gamma", [](isp_params* q, float v) { q->gamma = v; }},
      {"sharpen", [](isp_params* q, float v) { q->sharpen = v; }},
      {"ae target", [](isp_params* q, float v) { q->auto_exposure.target = v; }},
      {"ae damping", [](isp_params* q, float v) { q->auto_exposure.damping = v; }},
      {"ae gain", [](isp_params* q, float v) { q->auto_exposure.max_gain = v; }},
  };
  for (const Bad& b : bad) {
    for (float v : {nan, -1.0F, 1e30F, -INFINITY}) {
      isp_params q = p;
      b.set(&q, v);
      const isp_status s = isp_validate_params(&q);
      expect(s == ISP_OK || s == ISP_INVALID_ARGUMENT, b.what);
      expect(s == ISP_OK || std::strlen(isp_last_error()) > 0, "an error has a message");
    }
  }
  // Lens shading grids: missing, too small, too large, the largest accepted.
  std::vector<float> grid(256 * 256 * 4, 1.0F);
  for (int w : {0, 1, 2, 17, 256, 257}) {
    for (int ch : {1, 3, 4}) {
      isp_params q = p;
      q.has_lens_shading = 1;
      q.lens_shading.width = w;
      q.lens_shading.height = w;
      q.lens_shading.channels = ch;
      q.lens_shading.gains = w >= 0 && w <= 256 ? grid.data() : nullptr;
      q.lens_shading.gains_count = w >= 0 && w <= 256 ? static_cast<int64_t>(w) * w * ch : 0;
      const isp_status s = isp_validate_params(&q);
      expect(s == ISP_OK || s == ISP_INVALID_ARGUMENT, "lens shading grid");
    }
  }
  // Tone tables: each curve, user tables of 2 and 65,536 entries, too small a capacity.
  std::vector<float> user(65536);
  for (size_t i = 0; i < user.size(); ++i) user[i] = static_cast<float>(i) / 65535.0F;
  for (int curve : {ISP_TONE_LINEAR, ISP_TONE_SRGB, ISP_TONE_GAMMA, ISP_TONE_LUT}) {
    for (int n : {2, 65536}) {
      isp_params q = p;
      q.tone_curve = static_cast<uint8_t>(curve);
      q.lut = user.data();
      q.lut_size = n;
      int32_t size = 0;
      expect(isp_tone_lut(&q, nullptr, 0, &size) == ISP_OK && size >= 2, "tone table size");
      std::vector<float> out(static_cast<size_t>(size));
      expect(isp_tone_lut(&q, out.data(), size, &size) == ISP_OK, "tone table");
      expect(isp_tone_lut(&q, out.data(), size - 1, &size) == ISP_INVALID_ARGUMENT, "tone table capacity");
      int32_t osize = oracle_tone_lut(&q, nullptr, 0);
      std::vector<float> o(static_cast<size_t>(osize > 0 ? osize : 1));
      oracle_tone_lut(&q, o.data(), osize);
      expect(osize == size && std::memcmp(o.data(), out.data(), sizeof(float) * static_cast<size_t>(size)) == 0,
             "tone table equals the oracle's");
    }
  }
  // Raw views: every format at odd widths, pitches below and at the packed row.
  for (uint8_t fmt : {ISP_RAW_U8, ISP_RAW_U16, ISP_RAW_RAW10, ISP_RAW_RAW12}) {
    for (int w : {1, 2, 3, 5, 7, 641}) {
      const size_t row = isp_raw_row_bytes(fmt, w);
      for (size_t pitch : {row - 1, row, row + 13}) {
        isp_raw_view r{grid.data(), w, 3, pitch, fmt, fmt == ISP_RAW_U8 ? 8 : fmt == ISP_RAW_RAW10 ? 10 : 12,
                       ISP_RGGB};
        const isp_status s = isp_validate_raw(&r, &p);
        expect(s == ISP_OK || s == ISP_INVALID_ARGUMENT, "raw view");
      }
    }
  }
  for (uint8_t v = 0; v < 8; ++v) {
    uint8_t back = 0;
    const char* names[] = {isp_backend_name(v), isp_demosaic_name(v), isp_pattern_name(v), isp_white_balance_name(v)};
    for (const char* n : names) expect(n != nullptr, "a name for every value");
    if (std::strcmp(isp_pattern_name(v), "?") != 0 && v < 4) {
      expect(isp_parse_pattern(isp_pattern_name(v), &back) == 1 && back == v, "pattern names parse back");
    }
  }
}

// The statistics epilogue's host evaluation against the oracle's.
void check_epilogue() {
  std::mt19937_64 rng(5);
  isp_auto_exposure ae{};
  isp_params p;
  isp_params_init(&p);
  ae = p.auto_exposure;
  for (int i = 0; i < 2000; ++i) {
    const int64_t pixels = 1 + static_cast<int64_t>(rng() % 50'000'000ULL);
    uint64_t sums[4];
    for (int k = 0; k < 3; ++k) sums[k] = rng() % (static_cast<uint64_t>(pixels) << 20);
    sums[3] = rng() % static_cast<uint64_t>(pixels);
    const float gains[3] = {1.0F + (rng() % 1000) / 500.0F, 1.0F, 0.5F + (rng() % 1000) / 700.0F};
    const isp_frame_exposure e{static_cast<int64_t>(rng() % 40'000'000ULL), 1.0F + (rng() % 15)};
    float a[ISP_STATS_SIZE], b[ISP_STATS_SIZE];
    isp_epilogue_finish_stats(sums, pixels, gains, &ae, &e, a);
    oracle_finish_stats(sums, pixels, gains, &ae, &e, b);
    expect(std::memcmp(a, b, sizeof(a)) == 0, "epilogue equals the oracle's");
  }
}

// The chart camera and the noisy scene at odd sizes in every format, and the
// chart's truth image.
void check_renderers() {
  for (uint8_t fmt : {ISP_RAW_U8, ISP_RAW_U16, ISP_RAW_RAW10, ISP_RAW_RAW12}) {
    for (int w : {16, 17, 23, 64}) {
      isp_synthetic_config c;
      isp_synthetic_config_init(&c);
      c.width = w;
      c.height = 19;
      c.format = fmt;
      c.bit_depth = fmt == ISP_RAW_U8 ? 8 : fmt == ISP_RAW_RAW10 ? 10 : 12;
      c.vignetting = 0.3F;
      for (float& b : c.black_levels) b = 8.0F;
      const size_t pitch = isp_raw_row_bytes(fmt, w);
      std::vector<uint8_t> data(pitch * static_cast<size_t>(c.height));
      expect(isp_synthetic_render(&c, data.data(), pitch) == ISP_OK, "chart render");
      std::vector<float> truth(static_cast<size_t>(w) * c.height * 3);
      expect(isp_synthetic_truth_image(&c, truth.data()) == ISP_OK, "chart truth");
      isp_params p;
      std::vector<float> lsc(9 * 9);
      expect(isp_synthetic_params(&c, 9, &p, lsc.data()) == ISP_OK, "chart params");

      isp_scene_config s;
      isp_scene_config_init(&s);
      s.width = w + 1;
      s.height = 7;
      s.format = fmt;
      s.bit_depth = c.bit_depth;
      s.black_level = 16.0F;
      const size_t sp = isp_raw_row_bytes(fmt, s.width);
      std::vector<uint8_t> sd(sp * static_cast<size_t>(s.height));
      expect(isp_scene_render(&s, sd.data(), sp) == ISP_OK, "scene render");
      std::vector<float> g(5 * 4);
      expect(isp_scene_lens_shading(&s, 5, 4, g.data()) == ISP_OK, "scene lens shading");
    }
  }
}

// Frames of 64 rows and more are rendered on several threads.
void check_threaded_renders() {
  for (uint8_t fmt : {ISP_RAW_U16, ISP_RAW_RAW10, ISP_RAW_RAW12}) {
    isp_synthetic_config c;
    isp_synthetic_config_init(&c);
    c.width = 130;
    c.height = 70;
    c.format = fmt;
    c.bit_depth = fmt == ISP_RAW_RAW10 ? 10 : 12;
    const size_t pitch = isp_raw_row_bytes(fmt, c.width);
    std::vector<uint8_t> a(pitch * static_cast<size_t>(c.height)), b(a.size());
    expect(isp_synthetic_render(&c, a.data(), pitch) == ISP_OK, "threaded chart render");
    std::vector<float> truth(static_cast<size_t>(c.width) * c.height * 3);
    expect(isp_synthetic_truth_image(&c, truth.data()) == ISP_OK, "threaded chart truth");
    isp_scene_config s;
    isp_scene_config_init(&s);
    s.width = 130;
    s.height = 66;
    s.format = fmt;
    s.bit_depth = c.bit_depth;
    s.black_level = 16.0F;
    const size_t sp = isp_raw_row_bytes(fmt, s.width);
    std::vector<uint8_t> sa(sp * static_cast<size_t>(s.height)), sb(sa.size());
    expect(isp_scene_render(&s, sa.data(), sp) == ISP_OK && isp_scene_render(&s, sb.data(), sp) == ISP_OK,
           "threaded scene render");
    expect(sa == sb, "a seeded scene renders the same bytes twice");
    // The oracle on 8 threads over this frame.
    isp_params p;
    isp_params_init(&p);
    p.black_level = 16.0F;
    p.white_balance = ISP_WB_WHITE_PATCH;
    p.sharpen = 0.5F;
    oracle_pipeline* o = nullptr;
    if (oracle_create(&p, &o) != 0) {
      expect(false, "oracle_create");
      continue;
    }
    isp_raw_view r{sa.data(), s.width, s.height, sp, fmt, s.bit_depth, s.pattern};
    std::vector<uint8_t> out(static_cast<size_t>(s.width) * 3 * s.height);
    std::vector<uint32_t> hist(4 * 256);
    float st[ISP_STATS_SIZE];
    isp_rgb_view ov{out.data(), static_cast<size_t>(s.width) * 3, ISP_OUT_RGB8};
    isp_stats_out so{hist.data(), st};
    expect(oracle_process(o, &r, &ov, &so, nullptr, 8) == 0, "oracle on 8 threads");
    oracle_destroy(o);
  }
}

// The oracle over every combination at small and odd sizes, two frames each
// (the white balance state carries over), on one and on several threads.
void check_oracle() {
  int combo = 0;
  for (uint8_t fmt : {ISP_RAW_U8, ISP_RAW_U16, ISP_RAW_RAW10, ISP_RAW_RAW12}) {
    for (uint8_t pattern : {ISP_RGGB, ISP_BGGR, ISP_GRBG, ISP_GBRG}) {
      for (uint8_t wb : {ISP_WB_MANUAL, ISP_WB_GRAY_WORLD, ISP_WB_WHITE_PATCH}) {
        for (uint8_t out_fmt : {ISP_OUT_RGB8, ISP_OUT_RGBA8, ISP_OUT_RGB16, ISP_OUT_RGB32F}) {
          ++combo;
          const int sizes[][2] = {{2, 2}, {3, 5}, {7, 3}, {33, 17}};
          const int* size = sizes[combo % 4];
          const int w = size[0], h = size[1];
          isp_scene_config s;
          isp_scene_config_init(&s);
          s.width = w;
          s.height = h;
          s.format = fmt;
          s.pattern = pattern;
          s.bit_depth = fmt == ISP_RAW_U8 ? 8 : fmt == ISP_RAW_RAW10 ? 10 : 12;
          s.seed = static_cast<uint64_t>(combo);
          s.black_level = 4.0F;
          const size_t pitch = isp_raw_row_bytes(fmt, w) + 4;
          std::vector<uint8_t> raw(pitch * static_cast<size_t>(h));
          if (isp_scene_render(&s, raw.data(), pitch) != ISP_OK) {
            expect(false, "scene render");
            continue;
          }
          isp_params p;
          isp_params_init(&p);
          p.black_level = 4.0F;
          p.white_balance = wb;
          p.demosaic = combo % 2 ? ISP_DEMOSAIC_MALVAR : ISP_DEMOSAIC_BILINEAR;
          p.sharpen = combo % 3 ? 0.5F : 0.0F;
          p.histogram_bins = combo % 5 == 0 ? ISP_HISTOGRAM_BINS_WIDE : ISP_HISTOGRAM_BINS;
          std::vector<float> grid(5 * 4 * 4, 1.25F);
          if (combo % 2) {
            p.has_lens_shading = 1;
            p.lens_shading.width = 5;
            p.lens_shading.height = 4;
            p.lens_shading.channels = combo % 4 ? 4 : 1;
            p.lens_shading.gains = grid.data();
            p.lens_shading.gains_count = 5 * 4 * p.lens_shading.channels;
          }
          if (isp_validate_params(&p) != ISP_OK) std::fprintf(stderr, "%s\n", isp_last_error());
          expect(isp_validate_params(&p) == ISP_OK, "oracle case params");
          isp_raw_view r{raw.data(), w, h, pitch, fmt, s.bit_depth, pattern};
          expect(isp_validate_raw(&r, &p) == ISP_OK, "oracle case raw view");
          oracle_pipeline* o = nullptr;
          if (oracle_create(&p, &o) != 0) {
            expect(false, "oracle_create");
            continue;
          }
          const int bins = p.histogram_bins;
          const size_t out_pitch = static_cast<size_t>(w) * isp_output_channels(out_fmt) *
                                   isp_output_channel_bytes(out_fmt);
          std::vector<uint8_t> a(out_pitch * static_cast<size_t>(h)), b(a.size());
          std::vector<uint32_t> ha(static_cast<size_t>(4 * bins)), hb(ha.size());
          float sa[ISP_STATS_SIZE], sb[ISP_STATS_SIZE];
          const isp_frame_exposure e{10'000'000, 2.0F};
          for (int frame = 0; frame < 2; ++frame) {
            isp_rgb_view va{a.data(), out_pitch, out_fmt};
            isp_stats_out so{ha.data(), sa};
            expect(oracle_process(o, &r, &va, &so, &e, 1) == 0, "oracle_process");
          }
          // The same two frames on three threads: the same bytes.
          oracle_reset(o);
          for (int frame = 0; frame < 2; ++frame) {
            isp_rgb_view vb{b.data(), out_pitch, out_fmt};
            isp_stats_out so{hb.data(), sb};
            expect(oracle_process(o, &r, &vb, &so, &e, 3) == 0, "oracle_process threads");
          }
          expect(a == b && ha == hb && std::memcmp(sa, sb, sizeof(sa)) == 0, "threads give the same bytes");
          // The white balance replay leaves the state the frames leave.
          float g1[3], g2[3];
          int v1 = 0, v2 = 0;
          oracle_awb_state(o, g1, &v1);
          oracle_reset(o);
          oracle_awb_stats* st = oracle_awb_measure(o, &r, 2);
          expect(st != nullptr, "oracle_awb_measure");
          if (st != nullptr) {
            oracle_awb_apply(o, st);
            oracle_awb_apply(o, st);
            oracle_awb_free(st);
          }
          oracle_awb_state(o, g2, &v2);
          expect(v1 == v2 && std::memcmp(g1, g2, sizeof(g1)) == 0, "white balance replay");
          oracle_destroy(o);
        }
      }
    }
  }
}

// The frame checker on a chart frame and the colour conversion at odd sizes.
void check_checker_and_colour() {
  isp_synthetic_config c;
  isp_synthetic_config_init(&c);
  c.width = 33;
  c.height = 17;
  std::vector<float> truth(static_cast<size_t>(c.width) * c.height * 3);
  expect(isp_synthetic_truth_image(&c, truth.data()) == ISP_OK, "truth image");
  std::vector<float> lut(4096);
  for (size_t i = 0; i < lut.size(); ++i) lut[i] = std::sqrt(static_cast<float>(i) / 4095.0F);
  std::vector<uint8_t> rgb(static_cast<size_t>(c.width) * c.height * 3, 77);
  std::vector<uint32_t> hist(4 * 256, 0);
  isp_frame_check_result res{};
  oracle_frame_check(truth.data(), c.width, c.height, lut.data(), static_cast<int>(lut.size()), rgb.data(),
                     static_cast<size_t>(c.width) * 3, hist.data(), nullptr, &res);
  expect(res.histogram_mismatches > 0, "an empty histogram mismatches");
  expect(std::isfinite(oracle_frame_psnr(&res, c.width, c.height)), "psnr");

  const uint8_t sources[] = {ISP_PIX_RGB8, ISP_PIX_BGR8, ISP_PIX_RGBA8, ISP_PIX_BGRA8};
  const uint8_t targets[] = {ISP_PIX_NV12, ISP_PIX_NV16, ISP_PIX_I420, ISP_PIX_YUYV, ISP_PIX_UYVY, ISP_PIX_MONO8};
  for (int w : {1, 2, 3, 17}) {
    for (int h : {1, 2, 5}) {
      for (uint8_t sf : sources) {
        const int bpp = sf == ISP_PIX_RGB8 || sf == ISP_PIX_BGR8 ? 3 : 4;
        std::vector<uint8_t> src(static_cast<size_t>(w) * h * bpp);
        for (size_t i = 0; i < src.size(); ++i) src[i] = static_cast<uint8_t>(i * 37 + w);
        isp_image_view sv{sf, w, h, {{src.data(), static_cast<size_t>(w) * bpp}, {}, {}}};
        for (uint8_t df : targets) {
          const int cw = (w + 1) / 2, ch = (h + 1) / 2;
          std::vector<uint8_t> y(static_cast<size_t>(w + 1) * 2 * h), u(static_cast<size_t>(w + 2) * h),
              v(static_cast<size_t>(cw) * ch + 1);
          isp_image_view dv{df, w, h, {{y.data(), static_cast<size_t>(w + 1) * 2}, {u.data(), static_cast<size_t>(w + 2)},
                                       {v.data(), static_cast<size_t>(cw)}}};
          const int rc = oracle_color_convert(&sv, &dv, ISP_MATRIX_BT709, ISP_RANGE_LIMITED);
          expect(rc == 0 || rc == -1, "colour conversion");
        }
      }
    }
  }
}

}  // namespace

int main() {
  check_params();
  check_epilogue();
  check_renderers();
  check_threaded_renders();
  check_oracle();
  check_checker_and_colour();
  std::printf("host_check: %d cases, %d failures\n", g_cases, g_failures);
  return g_failures == 0 ? 0 : 1;
}
