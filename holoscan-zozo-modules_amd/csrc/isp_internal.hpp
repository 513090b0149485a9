// isp_internal.hpp - the library's internal C++ types, shared by the host code
// and the HIP kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/isp_c.h"
#include "isp_ae.hpp"

namespace isp_amd {

inline constexpr int kAwbBins = 1024;      // white-patch histogram bins (isp_math.hpp:32)
inline constexpr int kMaxHistBins = 1024;  // output histograms: 256 or the 1024-bin extension

// A validated IspParams with owned tables (the reference copies IspParams, isp_pipeline.cpp:118).
struct Params {
  isp_params c{};
  std::vector<float> lsc;  // lens shading gains, when c.has_lens_shading
  std::vector<float> lut;  // the user's table, for ISP_TONE_LUT
  int bins() const { return c.histogram_bins == ISP_HISTOGRAM_BINS_WIDE ? ISP_HISTOGRAM_BINS_WIDE : ISP_HISTOGRAM_BINS; }
};

struct invalid_argument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct runtime_failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct out_of_budget : std::runtime_error {  // ISP_OUT_OF_MEMORY
  using std::runtime_error::runtime_error;
};

// Host-side helpers (isp_params.cpp).
Params copy_params(const isp_params* params);  // validates
void validate_params(const isp_params& params);
void validate_raw(const isp_raw_view& raw, const isp_params& params);
void validate_out(const isp_raw_view& raw, const isp_rgb_view& out, const isp_params& params);
std::vector<float> tone_lut(const isp_params& params);
std::size_t raw_row_bytes(int format, int width);
void set_error(const std::string& message);
void clear_error();

// The Bayer layout as the kernels take it: 2 bits per quad position
// p = (y & 1) * 2 + (x & 1), holding the channel at that position (kCfa,
// reference cpu_pipeline.cpp:15-20), one code per BayerPattern. The kernels
// specialise their demosaic on these four codes.
constexpr int cfa_code(int c0, int c1, int c2, int c3) { return c0 | (c1 << 2) | (c2 << 4) | (c3 << 6); }
inline constexpr int kCfaCodes[4] = {
    cfa_code(ISP_CHANNEL_R, ISP_CHANNEL_GR, ISP_CHANNEL_GB, ISP_CHANNEL_B),   // RGGB
    cfa_code(ISP_CHANNEL_B, ISP_CHANNEL_GB, ISP_CHANNEL_GR, ISP_CHANNEL_R),   // BGGR
    cfa_code(ISP_CHANNEL_GR, ISP_CHANNEL_R, ISP_CHANNEL_B, ISP_CHANNEL_GB),   // GRBG
    cfa_code(ISP_CHANNEL_GB, ISP_CHANNEL_B, ISP_CHANNEL_R, ISP_CHANNEL_GR)};  // GBRG

// A camera's exposure carried on the device from one frame to the next (served
// form): the next frame runs at the previous frame's suggestion. Every pipeline
// holds one from creation to destruction, so a frame table built earlier (the
// batched form) never points at freed memory; `enabled` switches the feedback,
// and the kernels read it when the frame runs.
struct AeState {
  long long exposure_ns;
  float gain;
  int valid;    // 0: use the frame's own exposure
  int enabled;  // 0: no feedback; the frame's own exposure, no state written
};

// Everything one frame's kernels read, passed by value (FrameParams,
// isp_math.hpp:35-60, plus the outputs and the statistics switches).
struct FrameDesc {
  int width;
  int height;
  const std::uint8_t* raw;
  long long raw_pitch;
  int raw_format;
  int cfa;  // 4 x 2 bits: channel at quad position (y & 1) * 2 + (x & 1)
  float black[4];
  float inv_range[4];
  int white;
  float wb[3];
  const float* lsc;  // device grid, or null
  int lsc_width;
  int lsc_height;
  int lsc_channels;
  float lsc_sx;
  float lsc_sy;
  float ccm[9];
  const float* lut;  // device table: lut_size entries, then lut[lut_size - 1] - 1
  int lut_size;
  float sharpen;
  std::uint8_t* out;
  long long out_pitch;
  int out_format;
  std::uint32_t* histogram;  // user output or null
  float* stats;              // user output or null
  int bins;
  // The frame's capture settings and the auto-exposure bounds, for finish_stats.
  long long exposure_ns;
  float exposure_gain;
  AeParams ae;
  // White balance: mode, and the finish's settings.
  int wb_mode;
  float percentile;
  float smoothing;
  // The pipeline's device exposure state: while enabled, read instead of
  // exposure_ns / exposure_gain once valid, and replaced by this frame's
  // suggestion.
  AeState* ae_state;
  // 1 when the tone table starts at -0.0: the main pass then keeps the sign of
  // every zero before the table (the kernels' kZeroSign form).
  int zero_sign;
};

// Per-pipeline device scratch, zero between frames (each pass's last block clears
// what it consumed). The blocks of a launch add into kSlots copies of every
// accumulator, block b into copy b % kSlots, so that no address takes more than
// a kSlots-th of the blocks' device-scope atomics (same-address atomics
// serialise where they are performed); the last block adds the copies up. The
// sum copies and the completion counters each sit alone on a 256-byte line:
// atomics to one line queue behind each other even at different addresses.
inline constexpr int kSlots = 8;

struct alignas(256) LineCounter {
  unsigned int n;
};
struct alignas(256) LineSums {
  unsigned long long v[4];
};
inline constexpr std::size_t kLineSumsStride = sizeof(LineSums) / sizeof(unsigned long long);

struct AwbScratch {
  LineSums sum[kSlots];  // R, G, B fixed sums; complete quads
  unsigned int hist[kSlots][3][kAwbBins];
  LineCounter done[kSlots];  // blocks finished, per slot
  LineCounter top;           // slots finished
};
struct WbState {
  float gains[3];
  int valid;
};
struct MainScratch {
  LineSums sum[kSlots];  // R, G, B fixed sums; saturated samples
  unsigned int hist[kSlots][4][kMaxHistBins];
  LineCounter done[kSlots];
  LineCounter top;
};
struct DeviceScratch {
  AwbScratch awb;
  WbState wb;
  MainScratch main;
};

// ---- the batched form: many frames in one persistent launch ----------------

inline constexpr int kMaxBatch = 64;      // frames per launch
inline constexpr int kBatchTileW = 64;    // main-pass tile width
inline constexpr int kBatchAwbTile = 64;  // white-balance tile (samples, square)

// One frame of a batch. The variant selects the tile body: raw format (bits 0-1),
// Malvar (bit 2), sharpening (bit 3), 16-row tiles instead of 32 (bit 4).
struct BatchFrame {
  FrameDesc f;
  DeviceScratch* s;  // the frame's stream: white balance state and scratch
  int variant;
  int tiles_x;  // main pass: tiles across, tiles, tiles per work item, work items
  int tiles;
  int tiles_per_item;
  int items;
  int awb_tiles_x;  // white balance pass; 0 items when the balance is manual
  int awb_tiles;
  int awb_tiles_per_item;
  int awb_items;
  int fast_rows;
  int out_aligned;
};

// The work queues of a batch's two passes: the white balance pass lists the
// auto-balanced frames' items in awb_order, the main pass every frame's items in
// main_order. start[k] is the first item of the k-th frame in that order.
struct BatchTable {
  int count;
  int awb_items;
  int main_items;
  int queue;  // ISP_BATCH_QUEUE_*
  int awb_start[kMaxBatch];
  int awb_order[kMaxBatch];
  int main_start[kMaxBatch];
  int main_order[kMaxBatch];
  BatchFrame frames[kMaxBatch];
};

// Device control block of a batch, zero at creation; every launch leaves it
// ready for the next (the last block of a pass resets its queue).
struct BatchControl {
  unsigned int awb_next;
  unsigned int awb_exited;
  unsigned int main_next;
  unsigned int main_exited;
};

// Kernel launchers (isp_kernels.hip). `stream` is a hipStream_t.
struct LaunchConfig {
  int cu_count;
  int tile_w;
  int tile_h;
  int awb_blocks = 3;  // white balance kernel: resident blocks per CU at most
};
void launch_awb(const FrameDesc& f, DeviceScratch* scratch, const LaunchConfig& cfg, void* stream);
// `done`, when not null, is a hipEvent_t that completes with the main kernel's dispatch.
// `plain`: the frame has nothing to finish (no histograms, statistics, exposure
// state or gray-world state), which lets the kernel do without their LDS.
void launch_main(const FrameDesc& f, DeviceScratch* scratch, bool malvar, bool sharpen,
                 const LaunchConfig& cfg, void* stream, void* done = nullptr, bool plain = false);
// The statistics epilogue over `count` cases in device memory (verification).
void launch_epilogue(const isp_epilogue_case* cases, float* out, int count, void* stream);
// The batch kernels' resident blocks per CU and their launches (grid blocks);
// the main-pass kernel for batches with sharpened frames differs from the one
// without.
int batch_awb_blocks_per_cu();
// wide: the 512-thread kernels, for batches whose frames all run 32-row tiles.
// uniform >= 0: every frame runs that variant (BatchFrame::variant, 32-row tiles),
// and a kernel with that tile body alone runs the batch.
int batch_main_blocks_per_cu(bool sharpen, bool wide, int uniform = -1);
void launch_batch_awb(const BatchTable* table, BatchControl* control, int grid, void* stream);
void launch_batch_main(const BatchTable* table, BatchControl* control, int grid, bool sharpen, bool wide,
                       int uniform, void* stream);

}  // namespace isp_amd
