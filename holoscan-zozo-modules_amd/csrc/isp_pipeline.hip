// isp_pipeline.hip - the pipeline object behind the C ABI (include/isp_c.h):
// IspPipeline's life cycle and process() (reference isp_pipeline.cpp:37-90,
// cuda_pipeline.cu:226-364) on the HIP runtime.
//
// One pipeline = the uploaded tone table and lens shading grid, one device
// scratch (white balance sums and state, statistics sums, histogram
// partials, completion counters) and one event. Frames share the scratch and
// the white balance state, so each frame's GPU work waits for the previous
// frame's (hipStreamWaitEvent, never a host wait) and frames run in call order.
#include "isp_pipeline.hpp"

using namespace isp_amd;


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
    // The table and one entry more, below its last (see tone() in isp_kernels.hip).
    check(hipMalloc(&p->d_lut, (p->lut.size() + 1) * sizeof(float)), "hipMalloc");
    if (p->params.c.has_lens_shading) {
      check(hipMalloc(&p->d_lsc, p->params.lsc.size() * sizeof(float)), "hipMalloc");
      upload(p->params.lsc.data(), p->params.lsc.size() * sizeof(float), p->d_lsc);
    }
    check(hipMalloc(&p->d_scratch, sizeof(DeviceScratch)), "hipMalloc");
    check(hipMemset(p->d_scratch, 0, sizeof(DeviceScratch)), "hipMemset");
    // The exposure state lives as long as the pipeline (a batch table built
    // earlier keeps its address); feedback starts switched off.
    check(hipMalloc(&p->d_ae, sizeof(AeState)), "hipMalloc");
    check(hipMemset(p->d_ae, 0, sizeof(AeState)), "hipMemset");
    check(hipEventCreateWithFlags(&p->done, kOrderingEventFlags), "hipEventCreate");
    std::vector<float> padded(p->lut);
    padded.push_back(p->lut.back() - 1.0F);
    upload(padded.data(), padded.size() * sizeof(float), p->d_lut);
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
    // The previous frame first (cuda_pipeline.cu:303-309); on the stream it was
    // queued on, the stream's own order already gives that, without a wait packet.
    if (pipeline->queued && s != pipeline->last_stream) {
      check(hipStreamWaitEvent(s, pipeline->done, 0), "hipStreamWaitEvent");
    }
    const isp_params& p = pipeline->params.c;
    const bool timing = pipeline->timing;
    const bool awb = p.white_balance != ISP_WB_MANUAL;
    if (timing && awb) pipeline->mark(s);
    if (awb) launch_awb(f, pipeline->d_scratch, pipeline->launch, s);
    if (timing) pipeline->mark(s);
    // Nothing to finish after the tiles: the main pass can do without the
    // histograms' LDS (more blocks a CU).
    const bool plain = f.histogram == nullptr && f.stats == nullptr && !pipeline->ae_enabled &&
                       p.white_balance != ISP_WB_GRAY_WORLD;
    // The frame-ordering event completes with the main kernel's dispatch
    // (cuda_pipeline.cu:359 records it after the last kernel instead); a stream
    // being captured into a graph gets the event record of its own, as there.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    check(hipStreamIsCapturing(s, &capture), "hipStreamIsCapturing");
    const bool bind = capture == hipStreamCaptureStatusNone;
    launch_main(f, pipeline->d_scratch, p.demosaic == ISP_DEMOSAIC_MALVAR, p.sharpen > 0.0F, pipeline->launch, s,
                bind ? pipeline->done : nullptr, plain);
    if (!bind) check(hipEventRecord(pipeline->done, s), "hipEventRecord");
    if (timing) pipeline->mark(s);
    pipeline->queued = true;
    pipeline->last_stream = s;
  });
}

isp_status isp_set_exposure_feedback(isp_pipeline* pipeline, int32_t enable, const isp_frame_exposure* initial) {
  return guarded([&] {
    if (pipeline == nullptr) throw invalid_argument("pipeline is null");
    if (enable != 0) {
      if (initial == nullptr || initial->exposure_ns <= 0) {
        throw invalid_argument("exposure feedback needs an initial exposure_ns > 0");
      }
      if (!(initial->gain >= 0.0F) || !std::isfinite(initial->gain)) {
        throw invalid_argument("exposure feedback needs a finite initial gain >= 0");
      }
    }
    // Queued frames finish with the state they were queued with.
    if (pipeline->queued) check(hipEventSynchronize(pipeline->done), "hipEventSynchronize");
    const AeState state = enable != 0 ? AeState{initial->exposure_ns, initial->gain, 1, 1} : AeState{0, 0.0F, 0, 0};
    upload(&state, sizeof(state), pipeline->d_ae);
    pipeline->ae_enabled = enable != 0;
  });
}

isp_status isp_get_exposure_state(isp_pipeline* pipeline, isp_frame_exposure* state) {
  return guarded([&] {
    if (pipeline == nullptr || state == nullptr) throw invalid_argument("pipeline or state is null");
    if (!pipeline->ae_enabled) throw invalid_argument("exposure feedback is not enabled");
    if (pipeline->queued) check(hipEventSynchronize(pipeline->done), "hipEventSynchronize");
    AeState host{};
    check(hipMemcpy(&host, pipeline->d_ae, sizeof(host), hipMemcpyDeviceToHost), "hipMemcpy");
    state->exposure_ns = host.exposure_ns;
    state->gain = host.gain;
  });
}

isp_status isp_epilogue_finish_stats_device(const isp_epilogue_case* cases, float* out, int32_t count,
                                            isp_stream_t stream) {
  return guarded([&] {
    if (count < 0) throw invalid_argument("count must be >= 0");
    if (count == 0) return;
    if (cases == nullptr || out == nullptr) throw invalid_argument("cases or out is null");
    launch_epilogue(cases, out, count, stream);
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
