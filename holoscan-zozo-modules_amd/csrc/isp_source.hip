// isp_source.hip - the device-side synthetic raw source and the GPU frame
// checker (include/isp_source.h).
//
// Both kernels evaluate the synthetic chart camera (chart.hpp) tile by tile.
// A pixel's value is the mean of a 4 x 4 set of scene points half a pixel
// apart, so neighbouring pixels share most of their points: a 64 x 8 tile
// needs 130 x 18 distinct points instead of 512 x 16. Each block evaluates
// those points once into LDS (chart_point), then every pixel adds its 16 points
// from LDS in the reference's order, so each sum is bit for bit the host's.
//
//  render_kernel  one block per tile: lattice, then one raw sample per pixel
//                 into LDS, then the tile's packed bytes (u8, u16 little
//                 endian, MIPI RAW10, MIPI RAW12) written a dword per thread,
//                 byte by byte only at an unaligned or partial row end.
//                 (render_kernel, synthetic_raw.cu:11-21, writes one packing
//                 unit per thread and evaluates all 16 points per sample.)
//  check_kernel   persistent blocks walk the tiles: per pixel the truth codes
//                 (synthetic_truth, then the tone table and rounding of
//                 truth_codes, frame_check.hpp:54-62), the squared error
//                 against the image, and the four histogram rows of its codes
//                 in LDS; then, with the fence-free completion of the ISP's
//                 kernels, the last block compares the rebuilt histograms with
//                 the sent ones and writes the whole result. One launch per
//                 frame where the reference uses two memsets and two kernels
//                 (frame_check.cu:14-105).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/isp_source.h"
#include "chart.hpp"
#include "isp_device.hpp"
#include "isp_pipeline.hpp"

namespace isp_amd {
namespace {

constexpr int kSrcThreads = 256;
constexpr int kSrcTileW = 64;
constexpr int kSrcTileH = 8;
constexpr int kLatW = 2 * kSrcTileW + 2;  // lattice columns of a tile: points 2 x0 .. 2 (x0 + 63) + 3
constexpr int kLatH = 2 * kSrcTileH + 2;
constexpr int kLatPoints = kLatW * kLatH;
constexpr int kCheckEntries = ISP_HISTOGRAM_CHANNELS * ISP_HISTOGRAM_BINS;
// The block's LDS histograms in interleaved copies (`entry * kCheckCopies +
// lane % kCheckCopies`): the chart's flat patches give many lanes one code, and
// copies spread their atomics over addresses. Measured at 4K: one copy 244 us,
// two 206, four 303, eight 230 (two blocks a CU).
constexpr int kCheckCopies = 2;

struct alignas(16) P4 {
  float v[4];
};

// The scratch of isp_check_frame_device: the rebuilt histograms and the error
// sum in kSlots copies each (block b adds into copy b % kSlots; each error sum
// alone on its line), and the completion counters. Zero between checks.
struct alignas(256) LineTotal {
  unsigned long long v;
};
struct CheckScratch {
  LineTotal err[kSlots];
  unsigned int hist[kSlots][kCheckEntries];
  LineCounter done[kSlots];
  LineCounter top;
};
static_assert(sizeof(CheckScratch) == ISP_FRAME_CHECK_SCRATCH_BYTES, "scratch size is part of the ABI");

// The tile's scene points: lattice column c of the tile at x0 is the point
// coordinate chart_coord(x, i) of every (x, i) with 2 x + i = 2 x0 + c. Only
// the points some pixel of the frame uses are evaluated.
__device__ __forceinline__ void fill_lattice(const Chart& s, int x0, int y0, P4* lat) {
  const int cols = min(kLatW, 2 * (s.width - x0) + 2);
  const int rows = min(kLatH, 2 * (s.height - y0) + 2);
  for (int idx = static_cast<int>(threadIdx.x); idx < kLatPoints; idx += kSrcThreads) {
    const int r = idx / kLatW;
    const int c = idx - r * kLatW;
    if (r < rows && c < cols) {
      const int m = 2 * x0 + c;
      const int n = 2 * y0 + r;
      P4 p;
      chart_point(s, chart_coord(m >> 1, m & 1), chart_coord(n >> 1, n & 1), p.v);
      p.v[3] = 0.0F;
      lat[idx] = p;
    }
  }
}

// The blurred scene colour of tile pixel (px, py): its 16 lattice points added
// row by row, as chart_scene adds them.
__device__ __forceinline__ void lattice_scene(const Chart& s, const P4* lat, int px, int py, float rgb[3]) {
  float sum[3] = {0.0F, 0.0F, 0.0F};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const P4 p = lat[(2 * py + j) * kLatW + 2 * px + i];
#pragma unroll
      for (int k = 0; k < 3; ++k) sum[k] = sum[k] + p.v[k];
    }
  }
  chart_blur(s, sum, rgb);
}

// Bytes of one tile row of each format: 64 samples in whole packing units.
__device__ __forceinline__ int tile_row_bytes(int format) {
  return format == ISP_RAW_U8 ? 64 : format == ISP_RAW_U16 ? 128 : format == ISP_RAW_RAW10 ? 80 : 96;
}

struct RenderLds {
  P4 lat[kLatPoints];
  unsigned short samples[kSrcTileH][kSrcTileW];
};

__global__ void __launch_bounds__(kSrcThreads) render_kernel(const Chart s, std::uint8_t* data, std::size_t pitch,
                                                             int tiles_x, int row_bytes) {
  __shared__ RenderLds lds;
  const int tile = static_cast<int>(blockIdx.x);
  const int ty = tile / tiles_x;
  const int tx = tile - ty * tiles_x;
  const int x0 = tx * kSrcTileW;
  const int y0 = ty * kSrcTileH;
  fill_lattice(s, x0, y0, lds.lat);
  __syncthreads();
  for (int p = static_cast<int>(threadIdx.x); p < kSrcTileW * kSrcTileH; p += kSrcThreads) {
    const int px = p % kSrcTileW;
    const int py = p / kSrcTileW;
    const int x = x0 + px;
    const int y = y0 + py;
    int v = 0;  // samples past the width pack as 0
    if (x < s.width && y < s.height) {
      float rgb[3];
      lattice_scene(s, lds.lat, px, py, rgb);
      v = chart_sample_of(s, rgb, x, y);
    }
    lds.samples[py][px] = static_cast<unsigned short>(v);
  }
  __syncthreads();
  // The tile's bytes of each row, a dword per thread (at most 8 x 128 / 4).
  const int tb = tile_row_bytes(s.format);
  const int words = tb / 4;
  const int idx = static_cast<int>(threadIdx.x);
  if (idx >= words * kSrcTileH) return;
  const int row = idx / words;
  const int y = y0 + row;
  if (y >= s.height) return;
  const int b = 4 * (idx - row * words);  // byte within the tile row
  const int o = tx * tb + b;              // byte within the frame row
  if (o >= row_bytes) return;
  const unsigned short* line = lds.samples[row];
  const auto sample = [line](int i) { return static_cast<int>(line[i]); };
  unsigned int w = 0U;
#pragma unroll
  for (int e = 0; e < 4; ++e) w |= chart_packed_byte(s.format, sample, b + e) << (8 * e);
  std::uint8_t* dst = data + static_cast<std::size_t>(y) * pitch + static_cast<std::size_t>(o);
  if ((reinterpret_cast<std::uintptr_t>(dst) & 3U) == 0U && o + 4 <= row_bytes) {
    *reinterpret_cast<unsigned int*>(dst) = w;
  } else {
    for (int e = 0; e < 4 && o + e < row_bytes; ++e) dst[e] = static_cast<std::uint8_t>(w >> (8 * e));
  }
}

// tone (frame_check.hpp:45-51): the table at v in 0..1, interpolated linearly.
__device__ __forceinline__ float check_tone(const float* lut, int size, float v) {
  const float f = v * static_cast<float>(size - 1);
  const int i = static_cast<int>(f);
  return i >= size - 1 ? lut[size - 1] : lut[i] + (f - static_cast<float>(i)) * (lut[i + 1] - lut[i]);
}

// The 8-bit code an ideal ISP outputs (truth_codes, frame_check.hpp:54-62).
__device__ __forceinline__ int truth_code(const float* lut, int size, float linear) {
  return static_cast<int>(round(255.0 * static_cast<double>(check_tone(lut, size, linear))));
}

struct CheckLds {
  P4 lat[kLatPoints];
  unsigned int bins[kCheckEntries * kCheckCopies];
  unsigned long long wave_err[kSrcThreads / 64];
  unsigned int wave_miss[kSrcThreads / 64];
  int flag;
};

__global__ void __launch_bounds__(kSrcThreads) check_kernel(const Chart s, const float* __restrict__ lut, int lut_size,
                                                            const std::uint8_t* __restrict__ rgb, std::size_t pitch,
                                                            const unsigned int* __restrict__ histogram,
                                                            const float* __restrict__ stats, CheckScratch* work,
                                                            isp_frame_check_result* result, int tiles_x, int tiles) {
  __shared__ CheckLds lds;
  const int tid = static_cast<int>(threadIdx.x);
  for (int i = tid; i < kCheckEntries * kCheckCopies; i += kSrcThreads) lds.bins[i] = 0U;
  const int copy = tid & (kCheckCopies - 1);
  unsigned long long err = 0ULL;
  for (int tile = static_cast<int>(blockIdx.x); tile < tiles; tile += static_cast<int>(gridDim.x)) {
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x0 = tx * kSrcTileW;
    const int y0 = ty * kSrcTileH;
    __syncthreads();  // the previous tile's lattice reads are done (and the bins cleared)
    fill_lattice(s, x0, y0, lds.lat);
    __syncthreads();
    for (int p = tid; p < kSrcTileW * kSrcTileH; p += kSrcThreads) {
      const int px = p % kSrcTileW;
      const int py = p / kSrcTileW;
      const int x = x0 + px;
      const int y = y0 + py;
      if (x >= s.width || y >= s.height) continue;
      float scene[3];
      lattice_scene(s, lds.lat, px, py, scene);
      float linear[3];
      chart_truth_of(s, scene, x, y, linear);
      const std::uint8_t* pixel = rgb + static_cast<std::size_t>(y) * pitch + 3 * static_cast<std::size_t>(x);
      const int q[3] = {pixel[0], pixel[1], pixel[2]};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int d = q[k] - truth_code(lut, lut_size, linear[k]);
        err += static_cast<unsigned long long>(d * d);
      }
      const int luma = (2126 * q[0] + 7152 * q[1] + 722 * q[2] + 5000) / 10000;
      atomicAdd(&lds.bins[luma * kCheckCopies + copy], 1U);
      atomicAdd(&lds.bins[(ISP_HISTOGRAM_BINS + q[0]) * kCheckCopies + copy], 1U);
      atomicAdd(&lds.bins[(2 * ISP_HISTOGRAM_BINS + q[1]) * kCheckCopies + copy], 1U);
      atomicAdd(&lds.bins[(3 * ISP_HISTOGRAM_BINS + q[2]) * kCheckCopies + copy], 1U);
    }
  }
  // The block's share into its slot of the scratch; the returned values are
  // consumed (token) so that every atomic has completed before the count.
  err = wave_sum(err);
  if ((tid & 63) == 0) lds.wave_err[tid >> 6] = err;
  __syncthreads();
  const int slot = static_cast<int>(blockIdx.x) % kSlots;
  unsigned int token = 0U;
  if (tid == 0) {
    unsigned long long block = 0ULL;
    for (int w = 0; w < kSrcThreads / 64; ++w) block += lds.wave_err[w];
    if (block != 0ULL) token ^= fold(atomicAdd(&work->err[slot].v, block));
  }
  for (int i = tid; i < kCheckEntries; i += kSrcThreads) {
    unsigned int c = 0U;
#pragma unroll
    for (int k = 0; k < kCheckCopies; ++k) c += lds.bins[i * kCheckCopies + k];
    if (c != 0U) token ^= atomicAdd(&work->hist[slot][i], c);
  }
  if (!complete_grid(work->done, &work->top, token, &lds.flag)) return;
  // The last block: the rebuilt histograms against the sent ones (compare_kernel,
  // frame_check.cu:58-70), and the result.
  unsigned int miss = 0U;
  for (int i = tid; i < kCheckEntries; i += kSrcThreads) {
    const unsigned int rebuilt = take_slots(&work->hist[0][i], kCheckEntries);
    miss += rebuilt != histogram[i] ? 1U : 0U;
  }
  for (int offset = 32; offset > 0; offset >>= 1) miss += __shfl_down(miss, offset, 64);
  if ((tid & 63) == 0) lds.wave_miss[tid >> 6] = miss;
  __syncthreads();
  if (tid == 0) {
    unsigned int total = 0U;
    for (int w = 0; w < kSrcThreads / 64; ++w) total += lds.wave_miss[w];
    result->squared_error = take_slots(&work->err[0].v, sizeof(LineTotal) / sizeof(unsigned long long));
    result->histogram_mismatches = total;
    result->reserved = 0U;
  }
  if (tid < ISP_STATS_SIZE) result->stats[tid] = stats != nullptr ? stats[tid] : 0.0F;
  reset_grid(work->done, &work->top);
}

// Zeroes `words` 32-bit words at `p` (a fill kernel, not a memset node, so that
// a captured launch replays it correctly).
__global__ void __launch_bounds__(kSrcThreads) clear_kernel(unsigned int* __restrict__ p, int words) {
  const int i = static_cast<int>(blockIdx.x) * kSrcThreads + static_cast<int>(threadIdx.x);
  if (i < words) p[i] = 0U;
}

int check_blocks_per_cu() {
  static const int blocks = [] {
    int n = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, check_kernel, kSrcThreads, 0),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return n > 0 ? n : 1;
  }();
  return blocks;
}

}  // namespace

// The scratch of the allocation-free check (isp_check_frame_device_static): a
// file-scope device array of one CheckScratch (ISP_FRAME_CHECK_SCRATCH_BYTES,
// 37,120 bytes), of which every GPU the library runs on holds its own copy
// until the process ends. Host code never uses the array's name as an address
// (on the host it is only a stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[sizeof(CheckScratch)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
CheckScratch* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current. The first lookup also queues a clear of the array on
// `stream` ahead of the check: the check leaves the array zeroed for the next
// one, but nothing has zeroed it before the first.
CheckScratch* static_scratch(hipStream_t stream) {
  int device = 0;
  check(hipGetDevice(&device), "hipGetDevice");
  if (device < 0 || device >= kMaxDevices) throw runtime_failure("IspPipeline (hip): device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device] != nullptr) return g_scratch_ptr[device];
  void* ptr = nullptr;
  check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (frame check scratch)");
  if (ptr == nullptr) throw runtime_failure("IspPipeline (hip): hipGetSymbolAddress returned no address");
  constexpr int kWords = static_cast<int>(sizeof(CheckScratch) / sizeof(unsigned int));
  static_cast<void>(hipGetLastError());
  clear_kernel<<<dim3((kWords + kSrcThreads - 1) / kSrcThreads), dim3(kSrcThreads), 0, stream>>>(
      static_cast<unsigned int*>(ptr), kWords);
  check(hipGetLastError(), "clear_kernel");
  g_scratch_ptr[device] = static_cast<CheckScratch*>(ptr);
  return g_scratch_ptr[device];
}

// The launch of one check with the workspace `work` (validated arguments).
void launch_check(const isp_synthetic_config* camera, const float* lut, int32_t lut_size, const uint8_t* rgb,
                  size_t pitch, const uint32_t* histogram, const float* stats, CheckScratch* work,
                  isp_frame_check_result* result, hipStream_t stream) {
  const Chart s = make_chart(*camera);
  if (lut_size < 2) throw invalid_argument("the tone table needs at least 2 entries");
  const std::size_t packed = 3 * static_cast<std::size_t>(s.width);
  if (pitch == 0) pitch = packed;
  if (pitch < packed) throw invalid_argument("pitch is shorter than an RGB8 row");
  int device = 0;
  int cus = 1;
  check(hipGetDevice(&device), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
  const int tiles_x = (s.width + kSrcTileW - 1) / kSrcTileW;
  const int tiles = tiles_x * ((s.height + kSrcTileH - 1) / kSrcTileH);
  const int grid = std::min(tiles, cus * check_blocks_per_cu());
  static_cast<void>(hipGetLastError());
  check_kernel<<<dim3(static_cast<unsigned int>(grid)), dim3(kSrcThreads), 0, stream>>>(
      s, lut, lut_size, rgb, pitch, histogram, stats, work, result, tiles_x, tiles);
  check(hipGetLastError(), "check_kernel");
}

}  // namespace
}  // namespace isp_amd

using namespace isp_amd;

extern "C" {

isp_status isp_synthetic_render_device(const isp_synthetic_config* config, void* data, size_t pitch,
                                       isp_stream_t stream) {
  return guarded([&] {
    if (config == nullptr || data == nullptr) throw invalid_argument("config or data is null");
    const Chart s = make_chart(*config);
    const std::size_t row = raw_row_bytes(s.format, s.width);
    if (pitch < row) throw invalid_argument("pitch is shorter than a row");
    const int tiles_x = (s.width + kSrcTileW - 1) / kSrcTileW;
    const int tiles_y = (s.height + kSrcTileH - 1) / kSrcTileH;
    // Clear an error an earlier, unrelated call left behind (synthetic_raw.cu:29-31).
    static_cast<void>(hipGetLastError());
    render_kernel<<<dim3(static_cast<unsigned int>(tiles_x * tiles_y)), dim3(kSrcThreads), 0,
                    static_cast<hipStream_t>(stream)>>>(s, static_cast<std::uint8_t*>(data), pitch, tiles_x,
                                                        static_cast<int>(row));
    check(hipGetLastError(), "render_kernel");
  });
}

isp_status isp_check_frame_device(const isp_synthetic_config* camera, const float* lut, int32_t lut_size,
                                  const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                                  const float* stats, void* work, isp_frame_check_result* result,
                                  isp_stream_t stream) {
  return guarded([&] {
    if (camera == nullptr || lut == nullptr || rgb == nullptr || histogram == nullptr || work == nullptr ||
        result == nullptr) {
      throw invalid_argument("camera, lut, rgb, histogram, scratch or result is null");
    }
    launch_check(camera, lut, lut_size, rgb, pitch, histogram, stats, static_cast<CheckScratch*>(work), result,
                 static_cast<hipStream_t>(stream));
  });
}

isp_status isp_check_frame_device_static(const isp_synthetic_config* camera, const float* lut, int32_t lut_size,
                                         const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                                         const float* stats, isp_frame_check_result* result,
                                         isp_stream_t stream) {
  return guarded([&] {
    if (camera == nullptr || lut == nullptr || rgb == nullptr || histogram == nullptr || result == nullptr) {
      throw invalid_argument("camera, lut, rgb, histogram or result is null");
    }
    // The array's 32-bit bin counters take up to a whole frame per slot.
    const std::uint64_t pixels = static_cast<std::uint64_t>(camera->width > 0 ? camera->width : 0) *
                                 static_cast<std::uint64_t>(camera->height > 0 ? camera->height : 0);
    if (pixels > ISP_FRAME_CHECK_STATIC_MAX_PIXELS) {
      throw invalid_argument("frame of " + std::to_string(pixels) +
                             " pixels is beyond the static frame check scratch's maximum of " +
                             std::to_string(static_cast<unsigned long long>(ISP_FRAME_CHECK_STATIC_MAX_PIXELS)) +
                             " pixels");
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    launch_check(camera, lut, lut_size, rgb, pitch, histogram, stats, static_scratch(st), result, st);
  });
}

double isp_frame_check_psnr(const isp_frame_check_result* result, int32_t width, int32_t height) {
  if (result == nullptr) return 0.0;
  const double squared = static_cast<double>(result->squared_error);
  return squared == 0.0 ? 99.0
                        : 10.0 * std::log10(255.0 * 255.0 * 3.0 * static_cast<double>(width) *
                                            static_cast<double>(height) / squared);
}

}  // extern "C"
