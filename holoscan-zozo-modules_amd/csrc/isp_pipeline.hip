// isp_pipeline.hip - the pipeline object behind the C ABI (include/isp_c.h):
// IspPipeline's life cycle and process() (reference isp_pipeline.cpp:37-90,
// cuda_pipeline.cu:226-364) on the HIP runtime.
//
// One pipeline = the uploaded tone table and lens shading grid, one device
// scratch (white balance sums and state, statistics sums, histogram
// partials, completion counters) and one event. Frames share the scratch and
// the white balance state, so each frame's GPU work waits for the previous
// frame's (hipStreamWaitEvent, never a host wait) and frames run in call order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "isp_internal.hpp"

using namespace isp_amd;

namespace {

void check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw runtime_failure(std::string("IspPipeline (hip): ") + what + ": " + hipGetErrorString(status));
  }
}

// kCfa (reference cpu_pipeline.cpp:15-20): channel at quad position (y & 1) * 2 + (x & 1).
constexpr int kCfa[4][4] = {
    {ISP_CHANNEL_R, ISP_CHANNEL_GR, ISP_CHANNEL_GB, ISP_CHANNEL_B},
    {ISP_CHANNEL_B, ISP_CHANNEL_GB, ISP_CHANNEL_GR, ISP_CHANNEL_R},
    {ISP_CHANNEL_GR, ISP_CHANNEL_R, ISP_CHANNEL_B, ISP_CHANNEL_GB},
    {ISP_CHANNEL_GB, ISP_CHANNEL_B, ISP_CHANNEL_R, ISP_CHANNEL_GR},
};

template <typename F>
isp_status guarded(F&& body) {
  try {
    body();
    clear_error();
    return ISP_OK;
  } catch (const invalid_argument& e) {
    set_error(e.what());
    return ISP_INVALID_ARGUMENT;
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
void upload(const void* source, std::size_t bytes, void* destination) {
  hipStream_t stream = nullptr;
  check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  const hipError_t copied = hipMemcpyAsync(destination, source, bytes, hipMemcpyHostToDevice, stream);
  const hipError_t done = copied == hipSuccess ? hipStreamSynchronize(stream) : copied;
  static_cast<void>(hipStreamDestroy(stream));
  check(done, "upload");
}

}  // namespace

struct isp_pipeline {
  Params params;
  std::vector<float> lut;
  int device = 0;
  LaunchConfig launch{};
  float* d_lut = nullptr;
  float* d_lsc = nullptr;
  DeviceScratch* d_scratch = nullptr;
  hipEvent_t done = nullptr;
  bool queued = false;
  // Optional kernel timing: events around each launch, on the launch stream.
  bool timing = false;
  std::vector<hipEvent_t> marks;  // awb begin, main begin, main end, per frame
  std::vector<int> mark_kind;
  std::size_t next_mark = 0;

  ~isp_pipeline() { release(); }

  hipEvent_t mark(hipStream_t stream) {
    if (next_mark == marks.size()) {
      hipEvent_t e = nullptr;
      check(hipEventCreate(&e), "hipEventCreate");
      marks.push_back(e);
    }
    hipEvent_t e = marks[next_mark++];
    check(hipEventRecord(e, stream), "hipEventRecord");
    return e;
  }

  void release() noexcept {
    if (queued && done != nullptr) static_cast<void>(hipEventSynchronize(done));
    if (d_lut != nullptr) static_cast<void>(hipFree(d_lut));
    if (d_lsc != nullptr) static_cast<void>(hipFree(d_lsc));
    if (d_scratch != nullptr) static_cast<void>(hipFree(d_scratch));
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
    if (queued) check(hipEventSynchronize(done), "hipEventSynchronize");
    WbState state{};
    for (int k = 0; k < 3; ++k) state.gains[k] = params.c.wb_gains[k];
    state.valid = 0;
    upload(&state, sizeof(state), &d_scratch->wb);
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
    f.cfa = 0;
    for (int k = 0; k < 4; ++k) f.cfa |= kCfa[raw.pattern][k] << (2 * k);
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
    f.ae = AeParams{p.auto_exposure.target,          p.auto_exposure.damping,  p.auto_exposure.min_exposure_ns,
                    p.auto_exposure.max_exposure_ns, p.auto_exposure.min_gain, p.auto_exposure.max_gain};
    f.wb_mode = p.white_balance;
    f.percentile = p.white_patch_percentile;
    f.smoothing = p.awb_smoothing;
    return f;
  }
};

extern "C" {

int isp_probe_platform(void) {
  int devices = 0;
  return hipGetDeviceCount(&devices) == hipSuccess && devices > 0 ? 1 : 0;
}

uint8_t isp_resolve_backend(uint8_t backend) {
  return backend == ISP_BACKEND_AUTO ? isp_resolve_auto(isp_probe_platform()) : backend;
}

isp_status isp_create(const isp_params* params, uint8_t backend, isp_pipeline** out) {
  return guarded([&] {
    if (out == nullptr) throw invalid_argument("out is null");
    *out = nullptr;
    Params copied = copy_params(params);
    if (backend > ISP_BACKEND_AUTO) throw invalid_argument("backend is not a Backend");
    const uint8_t resolved = isp_resolve_backend(backend);
    if (resolved != ISP_BACKEND_GPU) {
      throw invalid_argument(
          "backend cpu: this library runs the pipeline on the GPU only (no usable HIP device, or the cpu "
          "backend was asked for)");
    }
    auto p = std::make_unique<isp_pipeline>();
    p->params = std::move(copied);
    p->lut = tone_lut(p->params.c);
    check(hipGetDevice(&p->device), "hipGetDevice");
    check(hipDeviceGetAttribute(&p->launch.cu_count, hipDeviceAttributeMultiprocessorCount, p->device),
          "hipDeviceGetAttribute");
    check(hipMalloc(&p->d_lut, p->lut.size() * sizeof(float)), "hipMalloc");
    if (p->params.c.has_lens_shading) {
      check(hipMalloc(&p->d_lsc, p->params.lsc.size() * sizeof(float)), "hipMalloc");
      upload(p->params.lsc.data(), p->params.lsc.size() * sizeof(float), p->d_lsc);
    }
    check(hipMalloc(&p->d_scratch, sizeof(DeviceScratch)), "hipMalloc");
    check(hipMemset(p->d_scratch, 0, sizeof(DeviceScratch)), "hipMemset");
    check(hipEventCreateWithFlags(&p->done, hipEventDisableTiming), "hipEventCreate");
    upload(p->lut.data(), p->lut.size() * sizeof(float), p->d_lut);
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    p->reset();
    *out = p.release();
  });
}

isp_status isp_destroy(isp_pipeline* pipeline) {
  return guarded([&] { delete pipeline; });
}

uint8_t isp_backend(const isp_pipeline* pipeline) {
  (void)pipeline;
  return ISP_BACKEND_GPU;
}

uint8_t isp_location(const isp_pipeline* pipeline) {
  (void)pipeline;
  return ISP_LOCATION_DEVICE;
}

isp_status isp_reset(isp_pipeline* pipeline) {
  return guarded([&] {
    if (pipeline == nullptr) throw invalid_argument("pipeline is null");
    pipeline->reset();
  });
}

isp_status isp_process(isp_pipeline* pipeline, const isp_raw_view* raw, const isp_rgb_view* out,
                       const isp_stats_out* stats, const isp_frame_exposure* exposure, isp_stream_t stream) {
  return guarded([&] {
    if (pipeline == nullptr || raw == nullptr || out == nullptr) throw invalid_argument("pipeline, raw or out is null");
    validate_raw(*raw, pipeline->params.c);
    validate_out(*raw, *out, pipeline->params.c);
    const FrameDesc f = pipeline->frame(*raw, *out, stats, exposure);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // Clear an error an earlier, unrelated call left behind (cuda_pipeline.cu:301-303).
    static_cast<void>(hipGetLastError());
    if (pipeline->queued) check(hipStreamWaitEvent(s, pipeline->done, 0), "hipStreamWaitEvent");
    const isp_params& p = pipeline->params.c;
    const bool timing = pipeline->timing;
    const bool awb = p.white_balance != ISP_WB_MANUAL;
    if (timing && awb) pipeline->mark(s);
    if (awb) launch_awb(f, pipeline->d_scratch, pipeline->launch, s);
    if (timing) pipeline->mark(s);
    launch_main(f, pipeline->d_scratch, p.demosaic == ISP_DEMOSAIC_MALVAR, p.sharpen > 0.0F, pipeline->launch, s);
    if (timing) pipeline->mark(s);
    check(hipEventRecord(pipeline->done, s), "hipEventRecord");
    pipeline->queued = true;
  });
}

isp_status isp_timing_begin(isp_pipeline* pipeline) {
  return guarded([&] {
    if (pipeline == nullptr) throw invalid_argument("pipeline is null");
    pipeline->timing = true;
    pipeline->next_mark = 0;
  });
}

isp_status isp_timing_end(isp_pipeline* pipeline, double* awb_ms, double* main_ms, int32_t* frames) {
  return guarded([&] {
    if (pipeline == nullptr) throw invalid_argument("pipeline is null");
    pipeline->timing = false;
    const bool awb = pipeline->params.c.white_balance != ISP_WB_MANUAL;
    const std::size_t per = awb ? 3 : 2;
    const std::size_t n = pipeline->next_mark / per;
    double a = 0.0;
    double m = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
      hipEvent_t* e = &pipeline->marks[i * per];
      check(hipEventSynchronize(e[per - 1]), "hipEventSynchronize");
      float ms = 0.0F;
      if (awb) {
        check(hipEventElapsedTime(&ms, e[0], e[1]), "hipEventElapsedTime");
        a += ms;
      }
      check(hipEventElapsedTime(&ms, e[per - 2], e[per - 1]), "hipEventElapsedTime");
      m += ms;
    }
    if (awb_ms != nullptr) *awb_ms = a;
    if (main_ms != nullptr) *main_ms = m;
    if (frames != nullptr) *frames = static_cast<int32_t>(n);
    pipeline->next_mark = 0;
  });
}

}  // extern "C"
