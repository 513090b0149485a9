// isp_kernels.hip - the ISP's gfx950 kernels.
//
// Two launches per frame, one when the white balance is manual:
//
//  awb_kernel   white balance statistics over complete unsaturated 2x2 quads
//               (reference: white_balance_quad isp_math.hpp:311-332, awb_kernel
//               cuda_pipeline.cu:119-158) with an LDS-privatised 3 x 1024
//               white-patch histogram; the last block to finish turns the sums
//               into gains (finish_white_balance, isp_math.hpp:357-384) in the
//               pipeline's device state and clears the sums for the next frame.
//
//  main_kernel  persistent blocks walk the frame's tiles: unpack + black level +
//               lens shading + white balance into an LDS tile with a mirrored
//               halo, demosaic (bilinear or Malvar-He-Cutler, only the two
//               formulas each CFA site needs), unsharp mask over an LDS ring of
//               demosaiced RGB, colour matrix, tone table, store, and the
//               statistics (2^-20 fixed sums, saturated count) and 4 x bins
//               histograms of the written codes in registers and LDS. The last
//               block writes the histograms and the statistics vector
//               (finish_stats, isp_math.hpp:387-425, with the exact log2/exp2
//               of isp_ae.hpp) and clears the scratch.
//
// Every float expression keeps the reference's operation order
// (isp_math.hpp:121-305) and the library is built with -ffp-contract=off, so
// the output is bit for bit the reference CPU backend's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "isp_internal.hpp"

namespace isp_amd {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int mirror_any(int v, int n) {
  const int period = 2 * (n - 1);
  v %= period;
  if (v < 0) v += period;
  return v < n ? v : period - v;
}
// mirror (isp_math.hpp:82-91), with the in-range case first.
__device__ __forceinline__ int mirror_near(int v, int n) {
  if (static_cast<unsigned>(v) < static_cast<unsigned>(n)) return v;
  if (v < 0 && -v < n) return -v;
  if (v >= n && v < 2 * n - 1) return 2 * (n - 1) - v;
  return mirror_any(v, n);
}

__device__ __forceinline__ int channel_at(int cfa, int x, int y) {
  return (cfa >> (2 * (((y & 1) << 1) | (x & 1)))) & 3;
}
__device__ __forceinline__ int color_of(int channel) {
  return channel == ISP_CHANNEL_R ? 0 : channel == ISP_CHANNEL_B ? 2 : 1;
}
__device__ __forceinline__ float clamp01(float v) { return v < 0.0F ? 0.0F : v > 1.0F ? 1.0F : v; }

// load_raw (isp_math.hpp:101-119).
__device__ __forceinline__ int load_raw(const FrameDesc& f, int x, int y) {
  const std::uint8_t* row = f.raw + static_cast<long long>(y) * f.raw_pitch;
  switch (f.raw_format) {
    case ISP_RAW_U8:
      return row[x];
    case ISP_RAW_U16:
      return reinterpret_cast<const std::uint16_t*>(row)[x];
    case ISP_RAW_RAW10: {
      const std::uint8_t* g = row + 5 * (x >> 2);
      const int k = x & 3;
      return (static_cast<int>(g[k]) << 2) | ((g[4] >> (2 * k)) & 3);
    }
    default: {
      const std::uint8_t* g = row + 3 * (x >> 1);
      const int k = x & 1;
      return (static_cast<int>(g[k]) << 4) | ((g[2] >> (4 * k)) & 15);
    }
  }
}

// lens_shading_gain (isp_math.hpp:121-140); f.lsc is not null.
__device__ __forceinline__ float lsc_gain(const FrameDesc& f, int x, int y, int channel) {
  const float fx = static_cast<float>(x) * f.lsc_sx;
  const float fy = static_cast<float>(y) * f.lsc_sy;
  int ix = static_cast<int>(fx);
  int iy = static_cast<int>(fy);
  ix = ix > f.lsc_width - 2 ? f.lsc_width - 2 : ix;
  iy = iy > f.lsc_height - 2 ? f.lsc_height - 2 : iy;
  const float tx = fx - static_cast<float>(ix);
  const float ty = fy - static_cast<float>(iy);
  const int step = f.lsc_channels;
  const int row = f.lsc_width * step;
  const float* g = f.lsc + (iy * f.lsc_width + ix) * step + (step == 4 ? channel : 0);
  const float top = g[0] + (g[step] - g[0]) * tx;
  const float bottom = g[row] + (g[row + step] - g[row]) * tx;
  return top + (bottom - top) * ty;
}

// linearize (isp_math.hpp:144-148).
__device__ __forceinline__ float linearize(const FrameDesc& f, int raw, int x, int y, int channel) {
  float v = (static_cast<float>(raw) - f.black[channel]) * f.inv_range[channel];
  v = v > 0.0F ? v : 0.0F;
  if (f.lsc != nullptr) v = v * lsc_gain(f, x, y, channel);
  return v;
}

// Wave64 sum of a 64-bit value.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
  return v;
}

__device__ __forceinline__ unsigned int atomic_load_u32(unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_load_u64(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Whether this block is the last of the grid to reach this point; publishes
// the block's earlier atomics first.
__device__ __forceinline__ bool last_block(unsigned int* done) {
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int ticket = atomicAdd(done, 1U);
    is_last = ticket == gridDim.x - 1 ? 1 : 0;
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last != 0;
}

// ---- white balance statistics -------------------------------------------

// histogram_percentile (isp_math.hpp:336-352) over hist[1024], by one wave:
// each lane owns 16 consecutive bins; an exclusive prefix sum finds the bin.
__device__ float percentile_wave(const unsigned int* hist, float percentile) {
  const int lane = threadIdx.x & 63;
  unsigned long long mine = 0;
  for (int b = 0; b < 16; ++b) mine += hist[lane * 16 + b];
  // inclusive scan over lanes
  unsigned long long scan = mine;
  for (int offset = 1; offset < 64; offset <<= 1) {
    const unsigned long long up = __shfl_up(scan, offset, 64);
    if (lane >= offset) scan += up;
  }
  const unsigned long long total = __shfl(scan, 63, 64);
  if (total == 0) return 0.0F;
  const double target = static_cast<double>(percentile) * static_cast<double>(total);
  unsigned long long below = scan - mine;
  int found = 1 << 30;
  double within = 0.0;
  for (int b = 0; b < 16; ++b) {
    const unsigned int h = hist[lane * 16 + b];
    if (found == (1 << 30) && h != 0U && static_cast<double>(below + h) >= target) {
      found = lane * 16 + b;
      within = (target - static_cast<double>(below)) / static_cast<double>(h);
    }
    below += h;
  }
  // the lowest lane with a hit holds the first such bin
  const unsigned long long hits = __ballot(found != (1 << 30));
  if (hits == 0) return 1.0F;
  const int first = __ffsll(static_cast<long long>(hits)) - 1;
  const int bin = __shfl(found, first, 64);
  const double w = __shfl(within, first, 64);
  return static_cast<float>((static_cast<double>(bin) + w) / kAwbBins);
}

__global__ void __launch_bounds__(kThreads) awb_kernel(FrameDesc f, DeviceScratch* s) {
  __shared__ unsigned int hist[3 * kAwbBins];
  __shared__ unsigned long long part[kThreads / 64][4];
  const bool patch = f.wb_mode == ISP_WB_WHITE_PATCH;
  const int tid = threadIdx.x;
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) hist[i] = 0U;
  }
  __syncthreads();
  const int qw = f.width / 2;
  const int qh = f.height / 2;
  unsigned long long local[4] = {0, 0, 0, 0};
  for (int qy = blockIdx.x; qy < qh; qy += gridDim.x) {
    for (int qx = tid; qx < qw; qx += kThreads) {
      int raw[4];
      bool ok = true;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        raw[k] = load_raw(f, 2 * qx + (k & 1), 2 * qy + (k >> 1));
        ok = ok && raw[k] < f.white;
      }
      if (!ok) continue;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int x = 2 * qx + (k & 1);
        const int y = 2 * qy + (k >> 1);
        const int channel = channel_at(f.cfa, x, y);
        float v = linearize(f, raw[k], x, y, channel);
        v = v < 4.0F ? v : 4.0F;
        const int color = color_of(channel);
        local[color] += static_cast<unsigned int>(v * 1048576.0F + 0.5F);
        if (patch) {
          const int b = static_cast<int>(v * static_cast<float>(kAwbBins));
          atomicAdd(&hist[color * kAwbBins + (b < kAwbBins - 1 ? b : kAwbBins - 1)], 1U);
        }
      }
      local[3] += 1;
    }
  }
  const int wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned long long v = wave_sum(local[k]);
    if ((tid & 63) == 0) part[wave][k] = v;
  }
  __syncthreads();
  if (tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += part[w][tid];
    if (total != 0) atomicAdd(tid < 3 ? &s->awb.sum[tid] : &s->awb.quads, total);
  }
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) {
      if (hist[i] != 0U) atomicAdd(&s->awb.hist[i / kAwbBins][i % kAwbBins], hist[i]);
    }
  }
  if (!last_block(&s->awb.done)) return;

  // The last block: finish_white_balance (isp_math.hpp:357-384), then clear.
  __shared__ float level_s[3];
  __shared__ unsigned long long quads_s;
  if (tid == 0) quads_s = atomic_load_u64(&s->awb.quads);
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) hist[i] = atomic_load_u32(&s->awb.hist[i / kAwbBins][i % kAwbBins]);
  }
  __syncthreads();
  if (quads_s != 0) {
    if (patch) {
      if (tid < 64) {
        for (int k = 0; k < 3; ++k) {
          const float level = percentile_wave(&hist[k * kAwbBins], f.percentile);
          if (tid == 0) level_s[k] = level;
        }
      }
    } else if (tid == 0) {
      const double quads = static_cast<double>(quads_s);
      level_s[0] = static_cast<float>(static_cast<double>(atomic_load_u64(&s->awb.sum[0])) / quads);
      level_s[1] = static_cast<float>(static_cast<double>(atomic_load_u64(&s->awb.sum[1])) / (2.0 * quads));
      level_s[2] = static_cast<float>(static_cast<double>(atomic_load_u64(&s->awb.sum[2])) / quads);
    }
  }
  __syncthreads();
  if (tid == 0 && quads_s != 0) {
    const float level[3] = {level_s[0], level_s[1], level_s[2]};
    if (level[0] > 0.0F && level[1] > 0.0F && level[2] > 0.0F) {
      float target[3] = {level[1] / level[0], 1.0F, level[1] / level[2]};
      for (int k = 0; k < 3; ++k) {
        target[k] = target[k] < 0.0625F ? 0.0625F : target[k] > 16.0F ? 16.0F : target[k];
      }
      WbState& st = s->wb;
      for (int k = 0; k < 3; ++k) {
        st.gains[k] = st.valid != 0 && f.smoothing < 1.0F ? st.gains[k] + f.smoothing * (target[k] - st.gains[k])
                                                          : target[k];
      }
      st.valid = 1;
    }
  }
  // Clear for the next frame.
  if (tid < 3) s->awb.sum[tid] = 0;
  if (tid == 3) s->awb.quads = 0;
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) s->awb.hist[i / kAwbBins][i % kAwbBins] = 0U;
  }
  if (tid == 0) s->awb.done = 0U;
}

// ---- the main pass ----------------------------------------------------------

// Demosaic (isp_math.hpp:161-215) of the sample at p (an LDS pointer into a tile
// of row stride `sw`) whose channel is `channel` (uniform across the wave).
template <bool kMalvar>
__device__ __forceinline__ void demosaic_at(const float* p, int sw, int channel, float rgb[3]) {
  const float c = p[0];
  const float n12 = p[-sw];
  const float n32 = p[sw];
  const float n21 = p[-1];
  const float n23 = p[1];
  const bool rb = channel == ISP_CHANNEL_R || channel == ISP_CHANNEL_B;
  float first;   // at_green (R/B sites) or across (G sites)
  float second;  // diagonal (R/B sites) or down (G sites)
  if constexpr (!kMalvar) {
    if (rb) {
      first = (n12 + n32 + n21 + n23) * 0.25F;
      second = (p[-sw - 1] + p[-sw + 1] + p[sw - 1] + p[sw + 1]) * 0.25F;
    } else {
      first = (n21 + n23) * 0.5F;
      second = (n12 + n32) * 0.5F;
    }
  } else {
    const float n02 = p[-2 * sw];
    const float n42 = p[2 * sw];
    const float n20 = p[-2];
    const float n24 = p[2];
    const float vertical2 = n02 + n42;
    if (rb) {
      const float axial1 = n12 + n32 + n21 + n23;
      const float axial2 = n02 + n42 + n20 + n24;
      const float diagonal1 = p[-sw - 1] + p[-sw + 1] + p[sw - 1] + p[sw + 1];
      first = clamp01((4.0F * c + 2.0F * axial1 - axial2) * 0.125F);
      second = clamp01((6.0F * c + 2.0F * diagonal1 - 1.5F * axial2) * 0.125F);
    } else {
      const float diagonal1 = p[-sw - 1] + p[-sw + 1] + p[sw - 1] + p[sw + 1];
      const float horizontal2 = n20 + n24;
      first = clamp01((5.0F * c + 4.0F * (n21 + n23) - horizontal2 - diagonal1 + 0.5F * vertical2) * 0.125F);
      second = clamp01((5.0F * c + 4.0F * (n12 + n32) - vertical2 - diagonal1 + 0.5F * horizontal2) * 0.125F);
    }
  }
  switch (channel) {
    case ISP_CHANNEL_R:
      rgb[0] = c;
      rgb[1] = first;
      rgb[2] = second;
      break;
    case ISP_CHANNEL_B:
      rgb[0] = second;
      rgb[1] = first;
      rgb[2] = c;
      break;
    case ISP_CHANNEL_GR:
      rgb[0] = first;
      rgb[1] = c;
      rgb[2] = second;
      break;
    default:
      rgb[0] = second;
      rgb[1] = c;
      rgb[2] = first;
      break;
  }
}

// tone (isp_math.hpp:235-243).
__device__ __forceinline__ float tone(const float* lut, int size, float x) {
  const float f = x * static_cast<float>(size - 1);
  const int i = static_cast<int>(f);
  if (i >= size - 1) return lut[size - 1];
  const float t = f - static_cast<float>(i);
  return lut[i] + t * (lut[i + 1] - lut[i]);
}

template <bool kMalvar, bool kSharpen, int kTileW, int kTileH>
__global__ void __launch_bounds__(kThreads) main_kernel(FrameDesc f, DeviceScratch* s, int tiles_x, int tiles) {
  constexpr int kRadius = kMalvar ? 2 : 1;
  constexpr int kHalo = kRadius + (kSharpen ? 1 : 0);
  constexpr int kSW = kTileW + 2 * kHalo;
  constexpr int kSH = kTileH + 2 * kHalo;
  constexpr int kRW = kSharpen ? kTileW + 2 : 1;
  constexpr int kRH = kSharpen ? kTileH + 2 : 1;
  __shared__ float samp[kSH * kSW];
  __shared__ float ring[3][kRH * kRW];
  __shared__ unsigned int hist[4 * kMaxHistBins];
  __shared__ unsigned long long part[kThreads / 64][4];

  const int tid = threadIdx.x;
  const bool want_hist = f.histogram != nullptr;
  const bool want_stats = f.stats != nullptr;
  const int bins = f.bins;
  const bool aligned4 = ((reinterpret_cast<unsigned long long>(f.out) | static_cast<unsigned long long>(f.out_pitch)) & 3ULL) == 0;
  float gains[3];
  if (f.wb_mode != ISP_WB_MANUAL) {
    for (int k = 0; k < 3; ++k) gains[k] = s->wb.gains[k];
  } else {
    for (int k = 0; k < 3; ++k) gains[k] = f.wb[k];
  }
  if (want_hist) {
    for (int i = tid; i < 4 * bins; i += kThreads) hist[i] = 0U;
  }
  unsigned long long acc[4] = {0, 0, 0, 0};
  const int W = f.width;
  const int H = f.height;

  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x0 = tx * kTileW;
    const int y0 = ty * kTileH;
    __syncthreads();
    // 1. Samples of the tile and its halo, reflected at the frame's edges,
    // preprocessed once (preprocess, isp_math.hpp:151-155).
    unsigned int saturated = 0;
    for (int i = tid; i < kSW * kSH; i += kThreads) {
      const int sy = i / kSW;
      const int sx = i - sy * kSW;
      const int gx = x0 + sx - kHalo;
      const int gy = y0 + sy - kHalo;
      const int x = mirror_near(gx, W);
      const int y = mirror_near(gy, H);
      const int raw = load_raw(f, x, y);
      const int channel = channel_at(f.cfa, x, y);
      float v = linearize(f, raw, x, y, channel) * gains[color_of(channel)];
      v = v < 1.0F ? v : 1.0F;
      samp[i] = v;
      if (sx >= kHalo && sx < kHalo + kTileW && sy >= kHalo && sy < kHalo + kTileH && gx < W && gy < H) {
        saturated += raw >= f.white ? 1U : 0U;
      }
    }
    acc[3] += saturated;
    __syncthreads();
    // 2. With sharpening, the RGB of the tile and a one-pixel ring, in 2x2
    // quads starting at virtual pixel (x0 - 1, y0 - 1).
    if constexpr (kSharpen) {
      constexpr int kQW = kRW / 2;
      constexpr int kQuads = kQW * (kRH / 2);
      for (int qi = tid; qi < kQuads; qi += kThreads) {
        const int qy = qi / kQW;
        const int qx = qi - qy * kQW;
        const int rx = 2 * qx;
        const int ry = 2 * qy;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
#pragma unroll
          for (int dx = 0; dx < 2; ++dx) {
            // virtual pixel (x0 + rx + dx - 1, y0 + ry + dy - 1): parity (1 - dx, 1 - dy)
            const int channel = (f.cfa >> (2 * (((1 - dy) << 1) | (1 - dx)))) & 3;
            float rgb[3];
            demosaic_at<kMalvar>(&samp[(ry + dy + kHalo - 1) * kSW + rx + dx + kHalo - 1], kSW, channel, rgb);
            const int r = (ry + dy) * kRW + rx + dx;
            ring[0][r] = rgb[0];
            ring[1][r] = rgb[1];
            ring[2][r] = rgb[2];
          }
        }
      }
      __syncthreads();
    }
    // 3. Each thread: a 4 x 2 block of output pixels.
    constexpr int kBW = kTileW / 4;
    constexpr int kBlocks = kBW * (kTileH / 2);
    for (int b = tid; b < kBlocks; b += kThreads) {
      const int by = b / kBW;
      const int bx = b - by * kBW;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int py = 2 * by + r;
        const int y = y0 + py;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int px = 4 * bx + c;
          const int x = x0 + px;
          float rgb[3];
          if constexpr (kSharpen) {
            const int q = (py + 1) * kRW + px + 1;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const float* m = &ring[k][q];
              const float cc = m[0];
              const float blur = (4.0F * cc + 2.0F * (m[-kRW] + m[kRW] + m[-1] + m[1]) +
                                  (m[-kRW - 1] + m[-kRW + 1] + m[kRW - 1] + m[kRW + 1])) *
                                 0.0625F;
              rgb[k] = clamp01(cc + f.sharpen * (cc - blur));
            }
          } else {
            const int channel = (f.cfa >> (2 * ((r << 1) | (c & 1)))) & 3;
            demosaic_at<kMalvar>(&samp[(py + kHalo) * kSW + px + kHalo], kSW, channel, rgb);
          }
          if (x >= W || y >= H) continue;
          // correct_color (isp_math.hpp:229-233)
          float lin[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            lin[k] = clamp01(f.ccm[3 * k] * rgb[0] + f.ccm[3 * k + 1] * rgb[1] + f.ccm[3 * k + 2] * rgb[2]);
          }
          if (want_stats) {
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += static_cast<unsigned int>(lin[k] * 1048576.0F + 0.5F);
          }
          float t[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) t[k] = tone(f.lut, f.lut_size, lin[k]);
          // store (isp_math.hpp:250-276) and output_codes (:279-301)
          std::uint8_t* row = f.out + static_cast<long long>(y) * f.out_pitch;
          int q[3];
          switch (f.out_format) {
            case ISP_OUT_RGB8:
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<unsigned char>(t[k] * 255.0F + 0.5F);
                row[3 * x + k] = static_cast<std::uint8_t>(v);
                q[k] = static_cast<int>(v);
              }
              break;
            case ISP_OUT_RGBA8: {
              unsigned int packed = 0xFF000000U;
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<unsigned char>(t[k] * 255.0F + 0.5F);
                packed |= v << (8 * k);
                q[k] = static_cast<int>(v);
              }
              if (aligned4) {
                reinterpret_cast<unsigned int*>(row)[x] = packed;
              } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) row[4 * x + k] = static_cast<std::uint8_t>(packed >> (8 * k));
              }
              break;
            }
            case ISP_OUT_RGB16:
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<std::uint16_t>(t[k] * 65535.0F + 0.5F);
                reinterpret_cast<std::uint16_t*>(row)[3 * x + k] = static_cast<std::uint16_t>(v);
                q[k] = static_cast<int>((v * static_cast<unsigned int>(bins - 1) + 32767U) / 65535U);
              }
              break;
            default:
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                reinterpret_cast<float*>(row)[3 * x + k] = t[k];
                q[k] = static_cast<int>(clamp01(t[k]) * static_cast<float>(bins - 1) + 0.5F);
              }
              break;
          }
          if (want_hist) {
            const int luma = (2126 * q[0] + 7152 * q[1] + 722 * q[2] + 5000) / 10000;
            atomicAdd(&hist[luma], 1U);
            atomicAdd(&hist[bins + q[0]], 1U);
            atomicAdd(&hist[2 * bins + q[1]], 1U);
            atomicAdd(&hist[3 * bins + q[2]], 1U);
          }
        }
      }
    }
  }

  // Flush: the block's sums and histograms into the frame's scratch.
  const int wave = tid >> 6;
  if (want_stats) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned long long v = wave_sum(acc[k]);
      if ((tid & 63) == 0) part[wave][k] = v;
    }
  }
  __syncthreads();
  if (want_stats && tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += part[w][tid];
    if (total != 0) atomicAdd(&s->main.sum[tid], total);
  }
  if (want_hist) {
    for (int i = tid; i < 4 * bins; i += kThreads) {
      if (hist[i] != 0U) atomicAdd(&s->main.hist[i / bins][i % bins], hist[i]);
    }
  }
  if (!last_block(&s->main.done)) return;

  // The last block: the histograms and the statistics vector, then clear.
  if (want_hist) {
    for (int i = tid; i < 4 * bins; i += kThreads) {
      unsigned int* cell = &s->main.hist[i / bins][i % bins];
      f.histogram[i] = atomic_load_u32(cell);
      *cell = 0U;
    }
  }
  if (want_stats && tid == 0) {
    unsigned long long sums[4];
    for (int k = 0; k < 4; ++k) sums[k] = atomic_load_u64(&s->main.sum[k]);
    finish_stats(sums, static_cast<long long>(W) * H, gains, f.ae, f.exposure_ns, f.exposure_gain, f.stats);
  }
  if (tid < 4) s->main.sum[tid] = 0;
  if (tid == 0) s->main.done = 0U;
}

void check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw runtime_failure(std::string("IspPipeline (hip): ") + what + ": " + hipGetErrorString(status));
  }
}

template <typename Kernel>
int blocks_per_cu(Kernel kernel) {
  static std::mutex mutex;
  static std::map<const void*, int> cache;
  std::lock_guard<std::mutex> lock(mutex);
  const auto it = cache.find(reinterpret_cast<const void*>(kernel));
  if (it != cache.end()) return it->second;
  int blocks = 0;
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kThreads, 0), "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  blocks = std::max(blocks, 1);
  cache[reinterpret_cast<const void*>(kernel)] = blocks;
  return blocks;
}

template <bool kMalvar, bool kSharpen>
void launch_main_t(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream) {
  constexpr int kTileW = 64;
  constexpr int kTileH = 32;
  auto kernel = main_kernel<kMalvar, kSharpen, kTileW, kTileH>;
  const int tiles_x = (f.width + kTileW - 1) / kTileW;
  const int tiles_y = (f.height + kTileH - 1) / kTileH;
  const int tiles = tiles_x * tiles_y;
  const int grid = std::max(1, std::min(tiles, cfg.cu_count * blocks_per_cu(kernel)));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, f, s, tiles_x, tiles);
  check(hipGetLastError(), "main_kernel launch");
}

}  // namespace

void launch_awb(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, void* stream) {
  const int rows = f.height / 2;
  const int grid = std::max(1, std::min(rows, cfg.cu_count * 4));
  hipLaunchKernelGGL(awb_kernel, dim3(grid), dim3(kThreads), 0, static_cast<hipStream_t>(stream), f, s);
  check(hipGetLastError(), "awb_kernel launch");
}

void launch_main(const FrameDesc& f, DeviceScratch* s, bool malvar, bool sharpen, const LaunchConfig& cfg,
                 void* stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (malvar) {
    sharpen ? launch_main_t<true, true>(f, s, cfg, st) : launch_main_t<true, false>(f, s, cfg, st);
  } else {
    sharpen ? launch_main_t<false, true>(f, s, cfg, st) : launch_main_t<false, false>(f, s, cfg, st);
  }
}

}  // namespace isp_amd
