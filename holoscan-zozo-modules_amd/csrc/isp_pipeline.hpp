// isp_pipeline.hpp - the pipeline object behind the C ABI (include/isp_c.h),
// shared by the per-frame entry points (isp_pipeline.hip) and the batched and
// served forms (isp_batch.hip, isp_server.hip), which run a pipeline's frames
// with its tables and its white balance state.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "isp_internal.hpp"

namespace isp_amd {

inline void check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw runtime_failure(std::string("IspPipeline (hip): ") + what + ": " + hipGetErrorString(status));
  }
}

// The events that order a pipeline's frames (and a batch's, and a served slot's
// launches) after each other, recorded once per frame: without timing, and
// without the system-scope fence an event performs by default when it is
// recorded. Their consumers are on this device (streams that wait on them,
// and host threads that then copy with the HIP runtime, whose copies order
// themselves); the fence, a cache writeback and invalidate between every two
// frames, cost 1.8 us of a 1080p frame's 14.6.
constexpr unsigned kOrderingEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

// kCfa (reference cpu_pipeline.cpp:15-20): channel at quad position (y & 1) * 2 + (x & 1).
inline constexpr int kCfa[4][4] = {
    {ISP_CHANNEL_R, ISP_CHANNEL_GR, ISP_CHANNEL_GB, ISP_CHANNEL_B},
    {ISP_CHANNEL_B, ISP_CHANNEL_GB, ISP_CHANNEL_GR, ISP_CHANNEL_R},
    {ISP_CHANNEL_GR, ISP_CHANNEL_R, ISP_CHANNEL_B, ISP_CHANNEL_GB},
    {ISP_CHANNEL_GB, ISP_CHANNEL_B, ISP_CHANNEL_R, ISP_CHANNEL_GR},
};

constexpr bool cfa_codes_match_table() {
  for (int p = 0; p < 4; ++p) {
    if (kCfaCodes[p] != cfa_code(kCfa[p][0], kCfa[p][1], kCfa[p][2], kCfa[p][3])) return false;
  }
  return true;
}
static_assert(cfa_codes_match_table(), "the kernels' Bayer codes follow the CFA table");

// Runs `body`, mapping its exceptions to status codes and the thread's last error.
template <typename F>
isp_status guarded(F&& body) {
  try {
    body();
    clear_error();
    return ISP_OK;
  } catch (const invalid_argument& e) {
    set_error(e.what());
    return ISP_INVALID_ARGUMENT;
  } catch (const out_of_budget& e) {
    set_error(e.what());
    return ISP_OUT_OF_MEMORY;
  } catch (const std::bad_alloc&) {
    set_error("IspPipeline (hip): out of host memory");
    return ISP_FAILURE;
  } catch (const std::exception& e) {
    set_error(e.what());
    return ISP_FAILURE;
  }
}

// Copies host memory to the device on a stream of its own and waits for it
// (the frames' streams would not otherwise wait for a pageable copy).
inline void upload(const void* source, std::size_t bytes, void* destination) {
  hipStream_t stream = nullptr;
  check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  const hipError_t copied = hipMemcpyAsync(destination, source, bytes, hipMemcpyHostToDevice, stream);
  const hipError_t done = copied == hipSuccess ? hipStreamSynchronize(stream) : copied;
  static_cast<void>(hipStreamDestroy(stream));
  check(done, "upload");
}

}  // namespace isp_amd

// One pipeline = the uploaded tone table and lens shading grid, one device
// scratch (white balance sums and state, statistics sums, histogram partials,
// completion counters) and one event.
struct isp_pipeline {
  using FrameDesc = isp_amd::FrameDesc;
  using Params = isp_amd::Params;
  using LaunchConfig = isp_amd::LaunchConfig;
  using DeviceScratch = isp_amd::DeviceScratch;
  Params params;
  std::vector<float> lut;
  int device = 0;
  LaunchConfig launch{};
  float* d_lut = nullptr;
  float* d_lsc = nullptr;
  DeviceScratch* d_scratch = nullptr;
  isp_amd::AeState* d_ae = nullptr;  // device exposure feedback state, for the pipeline's lifetime
  bool ae_enabled = false;           // the host's copy of d_ae->enabled
  hipEvent_t done = nullptr;
  bool queued = false;
  hipStream_t last_stream = nullptr;  // the stream `done` was last recorded on
  // Optional kernel timing: events around each launch, on the launch stream.
  bool timing = false;
  std::vector<hipEvent_t> marks;  // awb begin, main begin, main end, per frame
  std::vector<int> mark_kind;
  std::size_t next_mark = 0;

  ~isp_pipeline() { release(); }

  hipEvent_t mark(hipStream_t stream) {
    if (next_mark == marks.size()) {
      hipEvent_t e = nullptr;
      isp_amd::check(hipEventCreate(&e), "hipEventCreate");
      marks.push_back(e);
    }
    hipEvent_t e = marks[next_mark++];
    isp_amd::check(hipEventRecord(e, stream), "hipEventRecord");
    return e;
  }

  void release() noexcept {
    if (queued && done != nullptr) static_cast<void>(hipEventSynchronize(done));
    if (d_lut != nullptr) static_cast<void>(hipFree(d_lut));
    if (d_lsc != nullptr) static_cast<void>(hipFree(d_lsc));
    if (d_scratch != nullptr) static_cast<void>(hipFree(d_scratch));
    if (d_ae != nullptr) static_cast<void>(hipFree(d_ae));
    d_ae = nullptr;
    if (done != nullptr) static_cast<void>(hipEventDestroy(done));
    for (hipEvent_t e : marks) static_cast<void>(hipEventDestroy(e));
    marks.clear();
    d_lut = d_lsc = nullptr;
    d_scratch = nullptr;
    done = nullptr;
    queued = false;
  }

  // initial_white_balance (cpu_pipeline.cpp:90-95) into the device state.
  void reset() {
    if (queued) isp_amd::check(hipEventSynchronize(done), "hipEventSynchronize");
    isp_amd::WbState state{};
    for (int k = 0; k < 3; ++k) state.gains[k] = params.c.wb_gains[k];
    state.valid = 0;
    isp_amd::upload(&state, sizeof(state), &d_scratch->wb);
  }

  // make_frame_params (cpu_pipeline.cpp:50-88) with the device tables.
  FrameDesc frame(const isp_raw_view& raw, const isp_rgb_view& out, const isp_stats_out* stats,
                  const isp_frame_exposure* exposure) const {
    const isp_params& p = params.c;
    FrameDesc f{};
    f.width = raw.width;
    f.height = raw.height;
    f.raw = static_cast<const std::uint8_t*>(raw.data);
    f.raw_pitch = static_cast<long long>(raw.pitch);
    f.raw_format = raw.format;
    f.cfa = isp_amd::kCfaCodes[raw.pattern];
    const float white = p.white_level > 0.0F ? p.white_level
                                             : static_cast<float>((1U << static_cast<unsigned>(raw.bit_depth)) - 1U);
    for (int k = 0; k < 4; ++k) {
      f.black[k] = p.has_channel_black_levels ? p.channel_black_levels[k] : p.black_level;
      f.inv_range[k] = 1.0F / (white - f.black[k]);
    }
    f.white = static_cast<int>(std::ceil(white));
    for (int k = 0; k < 3; ++k) f.wb[k] = p.wb_gains[k];
    if (p.has_lens_shading) {
      f.lsc = d_lsc;
      f.lsc_width = p.lens_shading.width;
      f.lsc_height = p.lens_shading.height;
      f.lsc_channels = p.lens_shading.channels;
      f.lsc_sx = static_cast<float>(p.lens_shading.width - 1) / static_cast<float>(raw.width - 1);
      f.lsc_sy = static_cast<float>(p.lens_shading.height - 1) / static_cast<float>(raw.height - 1);
    }
    for (int k = 0; k < 9; ++k) f.ccm[k] = p.ccm[k];
    f.lut = d_lut;
    f.lut_size = static_cast<int>(lut.size());
    f.sharpen = p.sharpen;
    f.out = static_cast<std::uint8_t*>(out.data);
    f.out_pitch = static_cast<long long>(out.pitch);
    f.out_format = out.format;
    f.histogram = stats != nullptr ? stats->histogram : nullptr;
    f.stats = stats != nullptr ? stats->stats : nullptr;
    f.bins = params.bins();
    f.exposure_ns = exposure != nullptr ? exposure->exposure_ns : 0;
    f.exposure_gain = exposure != nullptr ? exposure->gain : 0.0F;
    f.ae = isp_amd::AeParams{p.auto_exposure.target,          p.auto_exposure.damping,  p.auto_exposure.min_exposure_ns,
                    p.auto_exposure.max_exposure_ns, p.auto_exposure.min_gain, p.auto_exposure.max_gain};
    f.wb_mode = p.white_balance;
    f.percentile = p.white_patch_percentile;
    f.smoothing = p.awb_smoothing;
    f.ae_state = d_ae;
    f.zero_sign = lut.front() == 0.0F && std::signbit(lut.front()) ? 1 : 0;
    return f;
  }
};
