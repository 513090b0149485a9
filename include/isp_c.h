/*
 * isp_c.h - the C ABI of the MI355X image signal processor.
 *
 * Raw Bayer mosaic in, display-ready RGB out, with optional histograms, a
 * statistics vector and an auto-exposure suggestion, computed by hand-written
 * gfx950 HIP kernels on caller-owned device buffers.
 *
 * Every entry point below replaces one entry of the reference's C++ interface
 * (holoscan::isp, modules/isp/include/holoscan/isp/ headers); the comment on each
 * names the reference declaration (file:line) it stands in for. The types keep
 * the reference's field names, meanings and defaults: enums are uint8_t,
 * std::optional<T> becomes a presence flag beside a T, std::vector<float>
 * becomes a pointer plus a count that the library copies at create time.
 *
 * Errors: no exceptions cross this ABI. Every function that can fail returns
 * an isp_status; isp_last_error() then names the first unusable field with the
 * reference's own wording (std::invalid_argument -> ISP_INVALID_ARGUMENT,
 * std::runtime_error -> ISP_FAILURE).
 *
 * Threading: one host thread per pipeline, as in the reference.
 */
#ifndef ISP_C_H_
#define ISP_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */

typedef enum isp_status {
  ISP_OK = 0,
  ISP_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
  ISP_FAILURE = 2,          /* std::runtime_error ("IspPipeline (cuda): ...") */
  ISP_OUT_OF_MEMORY = 3,    /* admission refused: the HBM budget is spent (serving) */
  ISP_NOT_READY = 4         /* serving: no finished frame to collect yet */
} isp_status;

/* ---- enums (isp_params.hpp:21-61) --------------------------------------- */

enum { ISP_RGGB = 0, ISP_BGGR = 1, ISP_GRBG = 2, ISP_GBRG = 3 };            /* BayerPattern */
enum { ISP_RAW_U8 = 0, ISP_RAW_U16 = 1, ISP_RAW_RAW10 = 2, ISP_RAW_RAW12 = 3 }; /* RawFormat */
enum { ISP_DEMOSAIC_BILINEAR = 0, ISP_DEMOSAIC_MALVAR = 1 };                 /* DemosaicMethod */
enum { ISP_WB_MANUAL = 0, ISP_WB_GRAY_WORLD = 1, ISP_WB_WHITE_PATCH = 2 };   /* WhiteBalanceMode */
enum { ISP_TONE_SRGB = 0, ISP_TONE_GAMMA = 1, ISP_TONE_LINEAR = 2, ISP_TONE_LUT = 3 }; /* ToneCurve */
enum { ISP_OUT_RGB8 = 0, ISP_OUT_RGBA8 = 1, ISP_OUT_RGB16 = 2, ISP_OUT_RGB32F = 3 };  /* OutputFormat */
/* Backend (isp_params.hpp:57-61): ISP_BACKEND_GPU is the reference's kCuda. */
enum { ISP_BACKEND_GPU = 0, ISP_BACKEND_CPU = 1, ISP_BACKEND_AUTO = 2 };
/* MemoryLocation (isp_pipeline.hpp:32) */
enum { ISP_LOCATION_HOST = 0, ISP_LOCATION_DEVICE = 1 };

/* Channels of the mosaic (isp_params.hpp:65-68). */
enum { ISP_CHANNEL_R = 0, ISP_CHANNEL_GR = 1, ISP_CHANNEL_GB = 2, ISP_CHANNEL_B = 3 };

/* Tone table size of the built-in curves (isp_params.hpp:128). */
#define ISP_TONE_LUT_SIZE 4097
/* Histogram layout (isp_params.hpp:166-171): luma, red, green, blue rows. */
#define ISP_HISTOGRAM_BINS 256
#define ISP_HISTOGRAM_CHANNELS 4
/* Extension: 1024-bin histograms of 10-bit output codes (RGB16 / RGB32F outputs only). */
#define ISP_HISTOGRAM_BINS_WIDE 1024
/* Statistics vector, layout version 1 (isp_params.hpp:178-197). */
#define ISP_STATS_LAYOUT_VERSION 1
#define ISP_STATS_SIZE 12
enum {
  ISP_STAT_MEAN_R = 0, ISP_STAT_MEAN_G = 1, ISP_STAT_MEAN_B = 2, ISP_STAT_MEAN_LUMA = 3,
  ISP_STAT_AE_ERROR_EV = 4, ISP_STAT_EXPOSURE_SCALE = 5, ISP_STAT_SUGGESTED_EXPOSURE_NS = 6,
  ISP_STAT_SUGGESTED_GAIN = 7, ISP_STAT_WB_GAIN_R = 8, ISP_STAT_WB_GAIN_G = 9,
  ISP_STAT_WB_GAIN_B = 10, ISP_STAT_SATURATED_FRACTION = 11
};

/* ---- settings (isp_params.hpp:73-121) ----------------------------------- */

/* LensShading (isp_params.hpp:73-79): gains [height][width][channels], each in (0, 16]. */
typedef struct isp_lens_shading {
  int32_t width;       /* grid points across, 2..256 */
  int32_t height;      /* grid points down, 2..256 */
  int32_t channels;    /* 1 or 4 (R, Gr, Gb, B) */
  const float* gains;  /* width * height * channels values, copied at create */
  int64_t gains_count; /* the vector's size, checked against width * height * channels */
} isp_lens_shading;

/* AutoExposureParams (isp_params.hpp:82-93). */
typedef struct isp_auto_exposure {
  float target;            /* 0.18 */
  float damping;           /* 0.5 */
  int64_t min_exposure_ns; /* 10'000 */
  int64_t max_exposure_ns; /* 33'000'000 */
  float min_gain;          /* 1 */
  float max_gain;          /* 16 */
} isp_auto_exposure;

/* IspParams (isp_params.hpp:95-121). Fill with isp_params_init() first. */
typedef struct isp_params {
  float black_level;
  uint8_t has_channel_black_levels; /* std::optional<std::array<float, 4>> */
  float channel_black_levels[4];
  float white_level; /* 0: 2^bit_depth - 1 */
  uint8_t has_lens_shading; /* std::optional<LensShading> */
  isp_lens_shading lens_shading;
  uint8_t white_balance; /* ISP_WB_* */
  float wb_gains[3];
  float awb_smoothing;
  float white_patch_percentile;
  uint8_t demosaic; /* ISP_DEMOSAIC_* */
  float ccm[9];
  uint8_t tone_curve; /* ISP_TONE_* */
  float gamma;
  const float* lut; /* ISP_TONE_LUT: 2..65536 entries, copied at create */
  int32_t lut_size;
  float sharpen;
  isp_auto_exposure auto_exposure;
  /* Extension, no reference counterpart: histogram bins, 256 (the reference's
   * fixed layout; 0 means 256) or ISP_HISTOGRAM_BINS_WIDE. */
  int32_t histogram_bins;
} isp_params;

/* RawImageView (isp_params.hpp:132-140). `pitch` is the row stride in bytes. */
typedef struct isp_raw_view {
  const void* data;
  int32_t width;  /* at least 2 */
  int32_t height; /* at least 2 */
  size_t pitch;
  uint8_t format;    /* ISP_RAW_* */
  int32_t bit_depth; /* 1..8 u8, 10 RAW10, 12 RAW12, else 1..16 */
  uint8_t pattern;   /* ISP_RGGB.. */
} isp_raw_view;

/* RgbImageView (isp_params.hpp:143-147). */
typedef struct isp_rgb_view {
  void* data;
  size_t pitch;
  uint8_t format; /* ISP_OUT_* */
} isp_rgb_view;

/* StatsOutput (isp_params.hpp:151-154); either pointer may be null. */
typedef struct isp_stats_out {
  uint32_t* histogram; /* ISP_HISTOGRAM_CHANNELS x bins */
  float* stats;        /* ISP_STATS_SIZE */
} isp_stats_out;

/* FrameExposure (isp_params.hpp:157-160); 0 is unknown. */
typedef struct isp_frame_exposure {
  int64_t exposure_ns;
  float gain;
} isp_frame_exposure;

/* A hipStream_t, passed as a plain pointer; NULL is the null stream. */
typedef void* isp_stream_t;

typedef struct isp_pipeline isp_pipeline;

/* ---- settings helpers --------------------------------------------------- */

/* The default IspParams{} (isp_params.hpp:95-121). */
void isp_params_init(isp_params* params);
/* validate_params (isp_params.hpp:124, isp_params.cpp:85-164). */
isp_status isp_validate_params(const isp_params* params);
/* validate_raw (isp_params.hpp:236, isp_params.cpp:199-237). */
isp_status isp_validate_raw(const isp_raw_view* raw, const isp_params* params);
/* raw_row_bytes (isp_params.hpp:232, isp_params.cpp:184-197). */
size_t isp_raw_row_bytes(uint8_t format, int32_t width);
/* tone_lut (isp_params.hpp:129, isp_params.cpp:166-182): writes the table into
 * `out` (capacity entries) and its size into *size; with out == NULL only the size. */
isp_status isp_tone_lut(const isp_params* params, float* out, int32_t capacity, int32_t* size);
/* output_channels / output_channel_bytes (isp_params.hpp:224-229). */
int32_t isp_output_channels(uint8_t format);
int32_t isp_output_channel_bytes(uint8_t format);
/* backend_name .. parse_white_balance (isp_params.hpp:200-211): names are static
 * strings; parse returns 1 and sets *value, or 0 for an unknown name. */
const char* isp_backend_name(uint8_t backend);
int isp_parse_backend(const char* name, uint8_t* value);
const char* isp_demosaic_name(uint8_t method);
int isp_parse_demosaic(const char* name, uint8_t* value);
const char* isp_pattern_name(uint8_t pattern);
int isp_parse_pattern(const char* name, uint8_t* value);
const char* isp_white_balance_name(uint8_t mode);
int isp_parse_white_balance(const char* name, uint8_t* value);
/* probe_platform (isp_pipeline.hpp:35): 1 when a HIP device is usable. */
int isp_probe_platform(void);
/* resolve_auto / resolve_backend (isp_params.hpp:219-221, isp_pipeline.hpp:37). */
uint8_t isp_resolve_auto(int gpu_device);
uint8_t isp_resolve_backend(uint8_t backend);

/* The message of the last failed call on this thread (thread-local), "" if none. */
const char* isp_last_error(void);

/* Verification (no reference counterpart): the host evaluation of the arithmetic
 * the kernels' statistics epilogue runs (finish_stats, isp_math.hpp:387-425, with
 * a double-double log2 / exp2), for comparison with the host libm's. */
double isp_epilogue_log2(double x);
double isp_epilogue_exp2(double y);
void isp_epilogue_finish_stats(const uint64_t sums[4], int64_t pixels, const float gains[3],
                               const isp_auto_exposure* ae, const isp_frame_exposure* exposure,
                               float* out);
/* One case of the statistics epilogue: the frame's four fixed-point sums (R, G,
 * B, saturated samples), its pixel count, white balance gains, auto-exposure
 * settings and capture exposure. */
typedef struct isp_epilogue_case {
  uint64_t sums[4];
  int64_t pixels;
  float gains[3];
  isp_auto_exposure ae;
  isp_frame_exposure exposure;
} isp_epilogue_case;
/* Verification (no reference counterpart): the same epilogue evaluated on the
 * device, by the code the main pass's last block runs, one case per thread:
 * `cases` (count entries) and `out` (count x ISP_STATS_SIZE floats) are device
 * memory; queued on `stream`. */
isp_status isp_epilogue_finish_stats_device(const isp_epilogue_case* cases, float* out, int32_t count,
                                            isp_stream_t stream);

/* ---- the pipeline (isp_pipeline.hpp:39-74) ------------------------------ */

/* IspPipeline(Backend, IspParams) (isp_pipeline.hpp:44, isp_pipeline.cpp:37-46):
 * validates, resolves ISP_BACKEND_AUTO, builds the tone table, uploads it and the
 * lens shading grid, and allocates the white balance and statistics scratch on
 * the current HIP device. This library holds only the GPU backend: ISP_BACKEND_CPU,
 * or ISP_BACKEND_AUTO without a usable device, is ISP_INVALID_ARGUMENT. */
isp_status isp_create(const isp_params* params, uint8_t backend, isp_pipeline** out);
/* ~IspPipeline: waits for the frames already queued, then frees. */
isp_status isp_destroy(isp_pipeline* pipeline);
/* backend() / location() (isp_pipeline.hpp:50-52). */
uint8_t isp_backend(const isp_pipeline* pipeline);
uint8_t isp_location(const isp_pipeline* pipeline);
/* reset() (isp_pipeline.hpp:58, isp_pipeline.cpp:58-63): waits for the frames
 * queued, then forgets the white balance estimate. */
isp_status isp_reset(isp_pipeline* pipeline);
/* process(raw, out, stats, exposure, stream) (isp_pipeline.hpp:68-69,
 * isp_pipeline.cpp:65-90): queues one frame on `stream`; all buffers in device
 * memory; the frame's GPU work waits for the previous frame's on any stream.
 * `stats` and `exposure` may be NULL ({} in the reference). */
isp_status isp_process(isp_pipeline* pipeline, const isp_raw_view* raw, const isp_rgb_view* out,
                       const isp_stats_out* stats, const isp_frame_exposure* exposure,
                       isp_stream_t stream);

/* Kernel timing (no reference counterpart; the reference times whole frames with
 * CUDA events in its benchmark, tests/benchmark.cpp:137-142): between begin and
 * end, every process() records hipEvents on its stream around the white balance
 * statistics kernel and the main kernel; end waits for them and returns the
 * summed kernel times in milliseconds and the number of frames. */
isp_status isp_timing_begin(isp_pipeline* pipeline);
isp_status isp_timing_end(isp_pipeline* pipeline, double* awb_ms, double* main_ms,
                          int32_t* frames);

/* Device exposure feedback (no reference counterpart; the reference reports the
 * auto-exposure suggestion and never applies it, isp_params.hpp:80-81). With
 * `enable`, every frame of this pipeline runs at the exposure and gain its
 * previous frame suggested (statistics entries 6 and 7, the exposure rounded to
 * whole nanoseconds), the first at `initial`; the state stays on the device.
 * Waits for the frames already queued. */
isp_status isp_set_exposure_feedback(isp_pipeline* pipeline, int32_t enable, const isp_frame_exposure* initial);
/* The exposure the pipeline's next frame will run at; waits for queued frames. */
isp_status isp_get_exposure_state(isp_pipeline* pipeline, isp_frame_exposure* state);

/* ---- the batched form (no reference counterpart) -------------------------
 * Many camera streams in one persistent-kernel launch per pass (the white
 * balance statistics of the auto-balanced frames, then the main pass of all). Each frame of a batch
 * belongs to a pipeline (its stream: parameters, tables, white balance and
 * exposure state), at most one frame per pipeline, and its results are the
 * ones isp_process gives for that pipeline. Frames of any raw format, size,
 * pattern and parameters mix. The reference processes one frame per call
 * (isp_pipeline.hpp:68-69). */
#define ISP_BATCH_MAX_FRAMES 64

typedef struct isp_batch isp_batch;

typedef struct isp_batch_frame {
  isp_pipeline* pipeline;
  isp_raw_view raw;
  isp_rgb_view out;
  isp_stats_out stats; /* either pointer may be null */
  isp_frame_exposure exposure;
} isp_batch_frame;

enum { ISP_BATCH_ORDER_LARGEST_FIRST = 0, ISP_BATCH_ORDER_AS_GIVEN = 1 };
/* How the blocks of a launch share its work items: an atomic queue (dynamic
 * balance), a fixed round robin, or a fixed split into one contiguous range per
 * XCD (neighbouring tiles meet in one L2). */
enum { ISP_BATCH_QUEUE_DYNAMIC = 0, ISP_BATCH_QUEUE_STATIC = 1, ISP_BATCH_QUEUE_STATIC_XCD = 2 };

/* How a batch runs. The per-format schedule picks the main pass's tile height
 * and work-item size for each raw format and demosaic; the rest applies to the
 * whole launch. */
typedef struct isp_batch_schedule {
  int32_t tile_height[4][2];  /* [ISP_RAW_*][ISP_DEMOSAIC_*]: 16 or 32 output rows per tile */
  int32_t tiles_per_item[4][2]; /* [ISP_RAW_*][ISP_DEMOSAIC_*]: main-pass tiles per work item, 1..32 */
  int32_t awb_tiles_per_item; /* white-balance 64 x 64 tiles per work item, 1..64 */
  int32_t blocks_per_cu;      /* resident blocks per CU; 0 = as many as fit */
  int32_t order;              /* ISP_BATCH_ORDER_*: which frames' work the queue lists first */
  int32_t queue;              /* ISP_BATCH_QUEUE_* */
  int32_t graph;              /* 1: capture the launches as a hipGraph once and replay it */
  int32_t per_frame;          /* 1: every frame runs the per-frame kernels, one frame after another
                                 in the batch's launch (and graph); the other fields then do not apply */
} isp_batch_schedule;

/* The default schedule: 32-row tiles, 4 tiles per item, 8 white-balance tiles
 * per item, full occupancy, largest frames first, the dynamic queue, graph replay,
 * the batch kernels. */
void isp_batch_schedule_init(isp_batch_schedule* schedule);
isp_status isp_batch_create(isp_batch** out);
/* Waits for the batch's launches, then frees. */
isp_status isp_batch_destroy(isp_batch* batch);
isp_status isp_batch_set_schedule(isp_batch* batch, const isp_batch_schedule* schedule);
isp_status isp_batch_get_schedule(const isp_batch* batch, isp_batch_schedule* schedule);
/* Binds `count` frames (1..ISP_BATCH_MAX_FRAMES), validated as isp_process
 * validates them. The binding holds until the next call; every launch processes
 * the frames the buffers hold then. Waits for the batch's previous launch. */
isp_status isp_batch_set_frames(isp_batch* batch, const isp_batch_frame* frames, int32_t count);
/* Queues one launch of the bound frames on `stream`, after every member
 * pipeline's queued frames; each pipeline's later frames wait for it. */
isp_status isp_batch_launch(isp_batch* batch, isp_stream_t stream);
/* Waits for the batch's launches. */
isp_status isp_batch_synchronize(isp_batch* batch);
/* Autotuner: times candidate schedules on the bound frames (`launches` timed
 * launches each) and keeps the fastest; with `cache_path`, a schedule found for
 * the same device and frame set is reused without timing, and new results are
 * appended. The member pipelines' white balance and exposure state is restored
 * afterwards. *best_ms: the chosen schedule's time per launch (0 from the cache). */
isp_status isp_batch_tune(isp_batch* batch, isp_stream_t stream, const char* cache_path, int32_t launches,
                          isp_batch_schedule* chosen, double* best_ms);

/* ---- the served form (no reference counterpart) ---------------------------
 * A request loop on a thread of its own: frames from many cameras are submitted
 * as they arrive and admitted first come, first served, as launch slots and HBM
 * budget allow. Admission takes the device memory the frame needs (its raw copy
 * and its outputs) from an arena of a fixed size. Whenever a launch slot is
 * free, the admitted frames (at most one per camera and max_batch in all) run
 * in one batched launch (isp_batch_*), overlapping the launches already on the
 * GPU; requests arriving meanwhile join the next launch. Each camera is a
 * pipeline: its white balance and exposure state stay on the device, and a
 * camera's frames run in submission order. Results stay in the arena until
 * released. The reference processes one frame per call (isp_op.cpp:559-566). */
typedef struct isp_server isp_server;

typedef struct isp_server_config {
  uint64_t hbm_budget;         /* bytes of device memory for admitted frames */
  int32_t max_batch;           /* frames per launch, 1..ISP_BATCH_MAX_FRAMES */
  int32_t launch_slots;        /* launches in flight at once, each on a stream of its own, 1..8 */
  int32_t per_frame;           /* 1: one per-frame launch (isp_process) per frame instead, as a baseline */
  isp_batch_schedule schedule; /* of the batched launches */
} isp_server_config;

typedef struct isp_camera_config {
  isp_params params;
  int32_t width;
  int32_t height;
  uint8_t format;  /* ISP_RAW_* */
  int32_t bit_depth;
  uint8_t pattern; /* ISP_RGGB.. */
  uint8_t out_format;
  int32_t stats;              /* 1: histograms and the statistics vector per frame */
  isp_frame_exposure initial; /* exposure_ns > 0: device exposure feedback starting here */
} isp_camera_config;

typedef struct isp_server_result {
  int64_t request;
  int32_t camera;
  uint64_t tag;
  isp_status status;
  void* out; /* device buffers, valid until isp_server_release */
  size_t out_pitch;
  uint32_t* histogram;
  float* stats;
  int64_t submit_ns; /* isp_server_now_ns() clock */
  int64_t launch_ns;
  int64_t done_ns; /* when the loop saw the launch finish */
  int32_t batch_frames;
} isp_server_result;

typedef struct isp_server_counters {
  int64_t submitted;
  int64_t completed;
  int64_t launches;
  int64_t queued;        /* waiting for admission now */
  uint64_t budget_used;  /* bytes now */
  uint64_t budget_peak;  /* the most bytes ever admitted at once */
  int32_t max_in_flight; /* the most launches ever on the GPU at once */
} isp_server_counters;

/* 256 MiB budget, 32 frames per launch, 2 launch slots, batched, the default
 * schedule with the static queue and no graph (every launch differs). A launch
 * of a single frame runs the per-frame kernels. */
void isp_server_config_init(isp_server_config* config);
isp_status isp_server_create(const isp_server_config* config, isp_server** out);
/* Finishes every submitted frame, then frees everything. */
isp_status isp_server_destroy(isp_server* server);
/* A camera whose frames are width x height in `format`; ISP_OUT_OF_MEMORY if one
 * frame cannot fit the budget. */
isp_status isp_server_add_camera(isp_server* server, const isp_camera_config* camera, int32_t* id);
/* One frame of `camera`: its raw rows at `raw` (ISP_LOCATION_HOST: copied into
 * the arena on admission, pinned memory for an asynchronous copy, and kept valid
 * until the result is polled; ISP_LOCATION_DEVICE: read in place, valid until
 * the result is polled). */
isp_status isp_server_submit(isp_server* server, int32_t camera, const void* raw, size_t pitch, uint8_t location,
                             uint64_t tag, int64_t* request);
/* The next finished frame, waiting up to timeout_us; ISP_NOT_READY if none. */
isp_status isp_server_poll(isp_server* server, int64_t timeout_us, isp_server_result* result);
/* Returns a finished frame's arena memory. */
isp_status isp_server_release(isp_server* server, int64_t request);
isp_status isp_server_counters_get(isp_server* server, isp_server_counters* counters);
/* The steady clock of the result timestamps, in nanoseconds. */
int64_t isp_server_now_ns(void);

#ifdef __cplusplus
}
#endif

#endif /* ISP_C_H_ */
