// isp_op.hip - the IspOp operator contract over the C ABI (include/isp_op.h):
// what the reference's operator (modules/isp/src/isp_op.cpp) does around
// IspPipeline, for a runtime that lends and takes tensors through callbacks.
//
// compute(): reads the raw ImageT's tensor as the reference's describe() does
// (encodings, element types, shapes, packing, bit depth, sizes, pitch), stages a
// host frame into a device buffer on the operator's stream with the reference's
// waits and copies, borrows the outputs from the runtime, queues the frame on
// the pipeline and emits image, histogram and stats with the descriptor and
// sample metadata passed through.
#include <atomic>
#include <cstdio>

#include "../../include/isp_op.h"
#include "isp_pipeline.hpp"

using namespace isp_amd;

namespace {

std::atomic<std::uint64_t> g_host_waits{0};
std::atomic<std::uint64_t> g_copies{0};

void count_host_wait() { g_host_waits.fetch_add(1, std::memory_order_relaxed); }
void count_copy() { g_copies.fetch_add(1, std::memory_order_relaxed); }

// The operator's two error kinds (isp_op.cpp:40-46), as exceptions mapped by guarded().
[[noreturn]] void invalid(const std::string& message) { throw invalid_argument("IspOp: " + message); }
[[noreturn]] void failure(const std::string& message) { throw runtime_failure("IspOp: " + message); }

void hip_check(hipError_t status, const char* what) {
  if (status != hipSuccess) failure(std::string(what) + ": " + hipGetErrorString(status));
}

const char* encoding_name(int e) {
  static const char* const names[] = {"UNKNOWN",    "RGB8",       "RGBA8",      "RGB16",
                                      "RGB32F",     "BAYER_RGGB", "BAYER_BGGR", "BAYER_GRBG",
                                      "BAYER_GBRG", "MONO8",      "MONO16"};
  return e >= 0 && e <= ISP_ENCODING_MONO16 ? names[e] : "UNKNOWN";
}

int encoding_pattern(int e) {
  switch (e) {
    case ISP_ENCODING_BAYER_RGGB:
      return ISP_RGGB;
    case ISP_ENCODING_BAYER_BGGR:
      return ISP_BGGR;
    case ISP_ENCODING_BAYER_GRBG:
      return ISP_GRBG;
    case ISP_ENCODING_BAYER_GBRG:
      return ISP_GBRG;
    default:
      return -1;
  }
}

int output_encoding(int format) {
  return format == ISP_OUT_RGBA8 ? ISP_ENCODING_RGBA8
         : format == ISP_OUT_RGB16 ? ISP_ENCODING_RGB16
         : format == ISP_OUT_RGB32F ? ISP_ENCODING_RGB32F
                                    : ISP_ENCODING_RGB8;
}

void output_dtype(int format, uint8_t* code, uint8_t* bits) {
  *code = format == ISP_OUT_RGB32F ? ISP_DTYPE_FLOAT : ISP_DTYPE_UINT;
  *bits = format == ISP_OUT_RGB16 ? 16 : format == ISP_OUT_RGB32F ? 32 : 8;
}

bool is_type(const isp_tensor& t, int code, int bits) { return t.dtype_code == code && t.dtype_bits == bits; }

std::string pattern_text(int pattern) { return isp_pattern_name(static_cast<uint8_t>(pattern)); }

// A raw frame: its view (data still where the tensor is), the bytes its rows span.
struct Raw {
  isp_raw_view view{};
  std::size_t span = 0;
};

// describe() (isp_op.cpp:122-232) on our tensor conventions (strides in elements).
Raw describe(const isp_image& image, const isp_op_config& config) {
  const isp_tensor& t = image.data;
  if (t.memory > ISP_MEMORY_DEVICE) invalid("raw images must be in host or CUDA memory");
  const int pattern = encoding_pattern(image.encoding);
  const bool mono = image.encoding == ISP_ENCODING_MONO8 || image.encoding == ISP_ENCODING_MONO16;
  if (pattern < 0 && !mono) {
    invalid(std::string("unsupported raw encoding ") + encoding_name(image.encoding) +
            "; use BAYER_RGGB, BAYER_BGGR, BAYER_GRBG, BAYER_GBRG, MONO8 or MONO16");
  }
  if (mono && !config.has_pattern) invalid("a MONO8 or MONO16 raw image needs the Bayer pattern in the configuration");
  if (pattern >= 0 && config.has_pattern && pattern != config.pattern) {
    invalid("the image's pattern " + pattern_text(pattern) + " disagrees with the configured " +
            pattern_text(config.pattern));
  }
  Raw raw;
  isp_raw_view& v = raw.view;
  v.pattern = static_cast<uint8_t>(pattern >= 0 ? pattern : config.pattern);
  v.data = t.data;
  int container_bits = 0;
  long long row_bytes = 0;
  long long element = 1;
  if (t.ndim < 1 || t.ndim > 4) invalid("raw tensor must be [H, W] or [H, W, 1]");
  if (config.packing == ISP_PACKING_NONE) {
    const bool u8 = is_type(t, ISP_DTYPE_UINT, 8);
    if (!(u8 || is_type(t, ISP_DTYPE_UINT, 16))) invalid("raw samples must be uint8 or uint16");
    if ((image.encoding == ISP_ENCODING_MONO8 && !u8) || (image.encoding == ISP_ENCODING_MONO16 && u8)) {
      invalid("raw element type disagrees with its encoding");
    }
    if (!(t.ndim == 2 || (t.ndim == 3 && t.shape[2] == 1))) invalid("raw tensor must be [H, W] or [H, W, 1]");
    v.height = static_cast<int32_t>(t.shape[0]);
    v.width = static_cast<int32_t>(t.shape[1]);
    v.format = u8 ? ISP_RAW_U8 : ISP_RAW_U16;
    container_bits = u8 ? 8 : 16;
    element = u8 ? 1 : 2;
    if (v.width > 1 && t.strides[1] != 1) invalid("raw samples within a row must be packed");
    row_bytes = element * v.width;
    if ((image.width != 0 && image.width != v.width) || (image.height != 0 && image.height != v.height)) {
      invalid("the descriptor's size disagrees with its tensor");
    }
  } else {
    if (!is_type(t, ISP_DTYPE_UINT, 8) || t.ndim != 2) invalid("packed raw images must be uint8 [H, row bytes]");
    if (t.shape[1] > 1 && t.strides[1] != 1) invalid("packed raw rows must be contiguous");
    v.height = static_cast<int32_t>(t.shape[0]);
    v.width = image.width;
    v.format = config.packing == ISP_PACKING_MIPI_RAW10 ? ISP_RAW_RAW10 : ISP_RAW_RAW12;
    container_bits = config.packing == ISP_PACKING_MIPI_RAW10 ? 10 : 12;
    row_bytes = t.shape[1];
    if (image.height != 0 && image.height != v.height) invalid("the descriptor's height disagrees with its tensor");
    if (v.width < 0 || static_cast<std::size_t>(row_bytes) < raw_row_bytes(v.format, v.width)) {
      invalid("packed rows of " + std::to_string(row_bytes) + " bytes cannot hold " + std::to_string(v.width) +
              " samples");
    }
  }
  if (v.width < 2 || v.height < 2 || v.width > config.max_width || v.height > config.max_height) {
    invalid("raw image " + std::to_string(v.width) + "x" + std::to_string(v.height) +
            " is outside 2x2 to the configured " + std::to_string(config.max_width) + "x" +
            std::to_string(config.max_height));
  }
  const long long pitch = v.height > 1 ? t.strides[0] * element : row_bytes;
  if (pitch < row_bytes) invalid("raw rows overlap");
  v.pitch = static_cast<std::size_t>(pitch);
  if (v.format == ISP_RAW_U16 && ((reinterpret_cast<std::uintptr_t>(v.data) | v.pitch) & 1U) != 0U) {
    invalid("uint16 raw data must be aligned to 2 bytes");
  }
  const bool known_bits = image.significant_bits != 0 && image.significant_bits != ISP_UNKNOWN_SIGNIFICANT_BITS;
  v.bit_depth = config.bit_depth > 0                                         ? config.bit_depth
                : known_bits && config.packing == ISP_PACKING_NONE ? static_cast<int32_t>(image.significant_bits)
                                                                             : container_bits;
  raw.span = v.pitch * static_cast<std::size_t>(v.height - 1) + raw_row_bytes(v.format, v.width);
  return raw;
}

void validate_op_config(const isp_op_config& c) {
  validate_params(c.params);
  if (c.max_width < 2 || c.max_height < 2 || c.max_width > 16384 || c.max_height > 16384) {
    throw invalid_argument("max_width and max_height must be 2..16384");
  }
  if (c.bit_depth < 0 || c.bit_depth > 16) throw invalid_argument("bit_depth must be 0 (from the image) or 1..16");
  if ((c.packing == ISP_PACKING_MIPI_RAW10 && c.bit_depth != 0 && c.bit_depth != 10) ||
      (c.packing == ISP_PACKING_MIPI_RAW12 && c.bit_depth != 0 && c.bit_depth != 12)) {
    throw invalid_argument("packed RAW10 and RAW12 samples have 10 and 12 bits");
  }
  if (c.packing > ISP_PACKING_MIPI_RAW12) throw invalid_argument("packing is not a Packing");
  if (c.output_format > ISP_OUT_RGB32F) throw invalid_argument("output_format is not an OutputFormat");
  if (c.has_pattern && c.pattern > ISP_GBRG) throw invalid_argument("pattern is not a BayerPattern");
  if (c.backend > ISP_BACKEND_AUTO) throw invalid_argument("backend is not a Backend");
}

}  // namespace

struct isp_op {
  isp_op_config config{};
  Params params;  // owns the tables config.params points into
  uint8_t backend = ISP_BACKEND_GPU;
  isp_pipeline* pipeline = nullptr;
  std::string start_error;
  void* staging = nullptr;  // device copy of host input, grown as needed
  std::size_t staging_bytes = 0;
  bool copy_logged = false;

  ~isp_op() { release(); }

  void release() noexcept {
    if (pipeline != nullptr) static_cast<void>(isp_destroy(pipeline));
    pipeline = nullptr;
    if (staging != nullptr) static_cast<void>(hipFree(staging));  // waits for work that reads it
    staging = nullptr;
    staging_bytes = 0;
  }

  void compute(const isp_image& image, const isp_sample_metadata& metadata, hipStream_t stream,
               const isp_op_runtime& rt) {
    if (pipeline == nullptr) failure("the pipeline could not start: " + start_error);
    if (image.data.data == nullptr) invalid("raw image has no tensor");
    Raw raw = describe(image, config);
    const isp_tensor& t = image.data;

    // The raw data where the GPU reads it (isp_op.cpp:392-466).
    if (t.memory == ISP_MEMORY_DEVICE) {
      if (t.ready != nullptr) hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(t.ready), 0), "hipStreamWaitEvent");
    } else {
      if (t.memory == ISP_MEMORY_PAGEABLE) {
        // Pageable memory is not ordered on our stream: wait for its producer.
        count_host_wait();
        if (t.ready != nullptr) hip_check(hipEventSynchronize(static_cast<hipEvent_t>(t.ready)), "hipEventSynchronize");
      } else if (t.ready != nullptr) {
        hip_check(hipStreamWaitEvent(stream, static_cast<hipEvent_t>(t.ready), 0), "hipStreamWaitEvent");
      }
      if (staging_bytes < raw.span) {
        if (staging != nullptr) {
          // The old buffer may still be read by work queued on the stream.
          count_host_wait();
          hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
          const hipError_t freed = hipFree(staging);
          staging = nullptr;
          staging_bytes = 0;
          hip_check(freed, "hipFree");
        }
        if (hipMalloc(&staging, raw.span) != hipSuccess) {
          staging = nullptr;
          failure(std::string("hipMalloc: ") + hipGetErrorString(hipGetLastError()));
        }
        staging_bytes = raw.span;
      }
      hip_check(hipMemcpyAsync(staging, raw.view.data, raw.span, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
      count_copy();
      // A pageable copy is staged by the time the call returns (one wait); a
      // pinned source must outlive the copy, and nothing here can hold it, so
      // wait for the copy.
      count_host_wait();
      if (t.memory == ISP_MEMORY_PINNED) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      if (!copy_logged) {
        copy_logged = true;
        std::fprintf(stderr,
                     "IspOp (gpu backend): the raw images are in host memory and the backend reads device memory, "
                     "so each frame is copied host to device, %zu bytes (%.1f MB) per frame; to avoid the copy, "
                     "send device images\n",
                     raw.span, static_cast<double>(raw.span) / 1e6);
      }
      raw.view.data = staging;
    }

    // The outputs, borrowed from the runtime (isp_op.cpp:468-556).
    const int width = raw.view.width;
    const int height = raw.view.height;
    const int channels = isp_output_channels(config.output_format);
    const int bins = params.bins();
    isp_tensor request{};
    request.memory = ISP_MEMORY_DEVICE;
    output_dtype(config.output_format, &request.dtype_code, &request.dtype_bits);
    request.ndim = 3;
    request.shape[0] = height;
    request.shape[1] = width;
    request.shape[2] = channels;
    request.strides[0] = static_cast<int64_t>(width) * channels;
    request.strides[1] = channels;
    request.strides[2] = 1;
    isp_tensor image_out{};
    if (rt.allocate(rt.user, "image", &request, &image_out) != 0 || image_out.data == nullptr) {
      failure("the runtime could not lend the image output");
    }
    isp_tensor hist_out{};
    isp_tensor stats_out{};
    if (config.stats) {
      isp_tensor h{};
      h.memory = ISP_MEMORY_DEVICE;
      h.dtype_code = ISP_DTYPE_UINT;
      h.dtype_bits = 32;
      h.ndim = 2;
      h.shape[0] = ISP_HISTOGRAM_CHANNELS;
      h.shape[1] = bins;
      h.strides[0] = bins;
      h.strides[1] = 1;
      if (rt.allocate(rt.user, "histogram", &h, &hist_out) != 0 || hist_out.data == nullptr) {
        failure("the runtime could not lend the histogram output");
      }
      isp_tensor s{};
      s.memory = ISP_MEMORY_DEVICE;
      s.dtype_code = ISP_DTYPE_FLOAT;
      s.dtype_bits = 32;
      s.ndim = 1;
      s.shape[0] = ISP_STATS_SIZE;
      s.strides[0] = 1;
      if (rt.allocate(rt.user, "stats", &s, &stats_out) != 0 || stats_out.data == nullptr) {
        failure("the runtime could not lend the stats output");
      }
    }
    const std::size_t pitch = static_cast<std::size_t>(width) * channels *
                              static_cast<std::size_t>(isp_output_channel_bytes(config.output_format));
    const isp_rgb_view out{image_out.data, pitch, config.output_format};
    const isp_stats_out so{config.stats ? static_cast<uint32_t*>(hist_out.data) : nullptr,
                           config.stats ? static_cast<float*>(stats_out.data) : nullptr};
    const isp_frame_exposure exposure{image.exposure_time_ns, image.gain};
    const isp_status status = isp_process(pipeline, &raw.view, &out, &so, &exposure, stream);
    if (status == ISP_INVALID_ARGUMENT) invalid(isp_last_error());
    if (status != ISP_OK) failure(isp_last_error());

    // The descriptor: the raw image's header, region offsets, exposure and gain;
    // no colour space claim (isp_op.cpp:566-581). Emit order: image, histogram, stats.
    isp_image descriptor{};
    descriptor.width = width;
    descriptor.height = height;
    descriptor.encoding = static_cast<uint8_t>(output_encoding(config.output_format));
    descriptor.color_space = ISP_COLOR_SPACE_UNSPECIFIED;
    descriptor.roi_offset_x = image.roi_offset_x;
    descriptor.roi_offset_y = image.roi_offset_y;
    descriptor.header = image.header;
    descriptor.exposure_time_ns = image.exposure_time_ns;
    descriptor.gain = image.gain;
    descriptor.data = image_out;
    if (rt.emit(rt.user, "image", &descriptor, &image_out, &metadata) != 0) failure("emit of image failed");
    if (config.stats) {
      if (rt.emit(rt.user, "histogram", nullptr, &hist_out, &metadata) != 0) failure("emit of histogram failed");
      if (rt.emit(rt.user, "stats", nullptr, &stats_out, &metadata) != 0) failure("emit of stats failed");
    }
  }
};

extern "C" {

void isp_op_config_init(isp_op_config* config) {
  if (config == nullptr) return;
  std::memset(config, 0, sizeof(*config));
  config->backend = ISP_BACKEND_GPU;
  isp_params_init(&config->params);
  config->output_format = ISP_OUT_RGB8;
  config->has_pattern = 0;
  config->bit_depth = 0;
  config->packing = ISP_PACKING_NONE;
  config->stats = 0;
  config->max_width = 3840;
  config->max_height = 2160;
}

isp_status isp_op_validate_config(const isp_op_config* config) {
  return guarded([&] {
    if (config == nullptr) throw invalid_argument("config is null");
    validate_op_config(*config);
  });
}

isp_status isp_op_create(const isp_op_config* config, isp_op** out) {
  return guarded([&] {
    if (config == nullptr || out == nullptr) throw invalid_argument("config or out is null");
    *out = nullptr;
    validate_op_config(*config);
    auto op = std::make_unique<isp_op>();
    op->config = *config;
    op->params = copy_params(&config->params);
    // The copied tables, not the caller's.
    op->config.params = op->params.c;
    op->config.params.lut = op->params.lut.empty() ? nullptr : op->params.lut.data();
    op->config.params.lens_shading.gains = op->params.lsc.empty() ? nullptr : op->params.lsc.data();
    op->backend = isp_resolve_backend(config->backend);
    *out = op.release();
  });
}

isp_status isp_op_destroy(isp_op* op) {
  return guarded([&] { delete op; });
}

uint8_t isp_op_backend(const isp_op* op) { return op != nullptr ? op->backend : ISP_BACKEND_GPU; }

isp_status isp_op_ports(const isp_op* op, isp_op_port* ports, int32_t capacity, int32_t* count) {
  return guarded([&] {
    if (op == nullptr || count == nullptr) throw invalid_argument("op or count is null");
    const isp_op_config& c = op->config;
    const uint8_t memory = op->backend == ISP_BACKEND_CPU ? ISP_MEMORY_PAGEABLE : ISP_MEMORY_DEVICE;
    const uint64_t pixels = static_cast<uint64_t>(c.max_width) * static_cast<uint64_t>(c.max_height);
    const uint64_t pixel_bytes = static_cast<uint64_t>(isp_output_channels(c.output_format)) *
                                 static_cast<uint64_t>(isp_output_channel_bytes(c.output_format));
    std::vector<isp_op_port> list;
    // Two slots: one frame in compute() while the next one arrives.
    list.push_back(isp_op_port{"raw", 1, 2U, 0, 0, 0, 0, 0});
    isp_op_port image{"image", 0, 0U, memory, 0, 0, 3, pixels * pixel_bytes};
    output_dtype(c.output_format, &image.dtype_code, &image.dtype_bits);
    list.push_back(image);
    if (c.stats) {
      list.push_back(isp_op_port{"histogram", 0, 0U, memory, ISP_DTYPE_UINT, 32, 2,
                                 4ULL * ISP_HISTOGRAM_CHANNELS * static_cast<uint64_t>(op->params.bins())});
      list.push_back(isp_op_port{"stats", 0, 0U, memory, ISP_DTYPE_FLOAT, 32, 1, 4ULL * ISP_STATS_SIZE});
    }
    *count = static_cast<int32_t>(list.size());
    if (ports != nullptr) {
      for (int32_t i = 0; i < std::min<int32_t>(capacity, *count); ++i) ports[i] = list[i];
    }
  });
}

isp_status isp_op_start(isp_op* op) {
  return guarded([&] {
    if (op == nullptr) throw invalid_argument("op is null");
    op->release();
    op->start_error.clear();
    op->copy_logged = false;
    // A start failure is reported by compute(), which can fail the frame.
    if (isp_create(&op->config.params, op->backend, &op->pipeline) != ISP_OK) {
      op->pipeline = nullptr;
      op->start_error = isp_last_error();
    }
  });
}

isp_status isp_op_stop(isp_op* op) {
  return guarded([&] {
    if (op == nullptr) throw invalid_argument("op is null");
    op->release();
  });
}

isp_status isp_op_compute(isp_op* op, const isp_image* raw, const isp_sample_metadata* metadata,
                          isp_stream_t stream, const isp_op_runtime* runtime) {
  return guarded([&] {
    if (op == nullptr || raw == nullptr || runtime == nullptr || runtime->allocate == nullptr ||
        runtime->emit == nullptr) {
      throw invalid_argument("op, raw or runtime is null");
    }
    const isp_sample_metadata md = metadata != nullptr ? *metadata : isp_sample_metadata{};
    op->compute(*raw, md, static_cast<hipStream_t>(stream), *runtime);
  });
}

isp_op_statistics_t isp_op_statistics(void) {
  return isp_op_statistics_t{g_host_waits.load(std::memory_order_relaxed), g_copies.load(std::memory_order_relaxed)};
}

}  // extern "C"
