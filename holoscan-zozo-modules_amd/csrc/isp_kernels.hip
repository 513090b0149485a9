// isp_kernels.hip - the ISP's gfx950 kernels.
//
// Two launches per frame, one when the white balance is manual:
//
//  awb_kernel   white balance statistics (reference: white_balance_quad,
//               isp_math.hpp:311-332; awb_kernel, cuda_pipeline.cu:119-158).
//               Persistent blocks walk 64x64 tiles; each thread takes 4
//               complete 2x2 quads from two coalesced row-chunk loads, skips
//               quads with a saturated sample, and sums the linearised values
//               (2^-20 units) per colour in registers; white patch adds an
//               LDS-privatised 3 x 1024 histogram. The last block to finish
//               turns the sums into gains (finish_white_balance,
//               isp_math.hpp:357-384) in the pipeline's device state and clears
//               the sums for the next frame.
//
//  main_kernel  persistent blocks walk 64x64 output tiles:
//               1. coalesced chunk loads of the raw rows (RAW12: 8 samples from
//                  3 dwords, RAW10: 16 from 5, u16: 8 from 4, u8: 8 from 2),
//                  unpack, black level, lens shading, white balance and clip
//                  into an LDS tile with a mirrored halo;
//               2. with sharpening, Malvar / bilinear demosaic of the tile and
//                  a one-pixel ring into LDS (4x2 pixels per thread, windows
//                  read as 128-bit LDS loads, only the two formulas each CFA
//                  site needs);
//               3. per thread a 4x4 block of output pixels: unsharp mask (or
//                  the demosaic itself), colour matrix, tone table, one vector
//                  store per row, the statistics sums in registers and the
//                  4 x bins histograms of the written codes in LDS.
//               The last block writes the histograms and the statistics vector
//               (finish_stats, isp_math.hpp:387-425, with the exact log2/exp2
//               of isp_ae.hpp) and clears the scratch.
//
// Lens shading: the reference's per-sample bilinear interpolation
// (lens_shading_gain, isp_math.hpp:121-140) is top + (bottom - top) * ty with
// top and bottom depending only on the column and the grid row. Each tile
// tabulates top and bottom - top per column for the (at most two) grid rows its
// sample rows fall in, so a sample costs one multiply-add pair and gets the
// same float result.
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
constexpr int kTile = 64;  // output tile edge (main and white balance passes)

struct alignas(4) W3 {
  unsigned int v[3];
};
struct alignas(4) W5 {
  unsigned int v[5];
};
struct alignas(8) W2 {
  unsigned int v[2];
};
struct alignas(16) W4 {
  unsigned int v[4];
};
struct alignas(16) F4 {
  float v[4];
};
struct alignas(8) F2 {
  float v[2];
};

__device__ __forceinline__ int mirror_any(int v, int n) {
  const int period = 2 * (n - 1);
  v %= period;
  if (v < 0) v += period;
  return v < n ? v : period - v;
}
// mirror (isp_math.hpp:82-91), with the in-range and one-reflection cases first.
__device__ __forceinline__ int mirror_near(int v, int n) {
  if (static_cast<unsigned>(v) < static_cast<unsigned>(n)) return v;
  if (v < 0 && -v < n) return -v;
  if (v >= n && v < 2 * n - 1) return 2 * (n - 1) - v;
  return mirror_any(v, n);
}

__device__ __forceinline__ int cfa_channel(int cfa, int p) { return (cfa >> (2 * p)) & 3; }
__device__ __forceinline__ int color_of(int channel) {
  return channel == ISP_CHANNEL_R ? 0 : channel == ISP_CHANNEL_B ? 2 : 1;
}
__device__ __forceinline__ float clamp01(float v) { return v < 0.0F ? 0.0F : v > 1.0F ? 1.0F : v; }

// load_raw (isp_math.hpp:101-119), byte-wise: the edges and unaligned frames.
__device__ __forceinline__ int load_raw(const std::uint8_t* row, int format, int x) {
  switch (format) {
    case ISP_RAW_U8:
      return row[x];
    case ISP_RAW_U16:
      return row[2 * x] | (row[2 * x + 1] << 8);
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

template <int kFmt>
struct Chunk {
  // Samples per aligned chunk load.
  static constexpr int kSamples = kFmt == ISP_RAW_RAW10 ? 16 : 8;
};

__device__ __forceinline__ unsigned int byte_of(const unsigned int* w, int i) {
  return (w[i >> 2] >> (8 * (i & 3))) & 0xFFU;
}

// kSamples samples at x (a multiple of kSamples) of an aligned row.
template <int kFmt>
__device__ __forceinline__ void load_chunk(const std::uint8_t* row, int x, int* s) {
  if constexpr (kFmt == ISP_RAW_U8) {
    const W2 w = *reinterpret_cast<const W2*>(row + x);
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = static_cast<int>(byte_of(w.v, k));
  } else if constexpr (kFmt == ISP_RAW_U16) {
    const W4 w = *reinterpret_cast<const W4*>(row + 2 * x);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      s[2 * k] = static_cast<int>(w.v[k] & 0xFFFFU);
      s[2 * k + 1] = static_cast<int>(w.v[k] >> 16);
    }
  } else if constexpr (kFmt == ISP_RAW_RAW12) {
    const W3 w = *reinterpret_cast<const W3*>(row + 3 * (x >> 1));
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const unsigned int lo = byte_of(w.v, 3 * g + 2);
      s[2 * g] = static_cast<int>((byte_of(w.v, 3 * g) << 4) | (lo & 15U));
      s[2 * g + 1] = static_cast<int>((byte_of(w.v, 3 * g + 1) << 4) | (lo >> 4));
    }
  } else {
    const W5 w = *reinterpret_cast<const W5*>(row + 5 * (x >> 2));
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const unsigned int lo = byte_of(w.v, 5 * g + 4);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[4 * g + k] = static_cast<int>((byte_of(w.v, 5 * g + k) << 2) | ((lo >> (2 * k)) & 3U));
      }
    }
  }
}

// Wave64 sum of a 64-bit value.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset, 64);
  return v;
}

// Grid-wide completion without cache-wide fences. Blocks publish their
// partial sums with device-scope atomics whose results each thread consumes
// (folded into `token`), so every atomic has been performed at the point of
// coherence before the thread reaches the barrier; only then does one thread
// take a ticket. The last block reads the totals back with atomic exchanges,
// which also clear them for the next frame. (An agent-scope fence here would
// write back and invalidate the whole L2 of the block's XCD, once per block.)
__device__ __forceinline__ bool last_block(unsigned int* done, unsigned int token) {
  __shared__ int is_last;
  // Consuming the token makes the atomics return their values and the wave wait
  // for them here (the asm reads a register and touches no memory).
  __asm__ volatile("" : : "v"(token));
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int ticket = atomicAdd(done, 1U);
    is_last = ticket == gridDim.x - 1 ? 1 : 0;
  }
  __syncthreads();
  return is_last != 0;
}

__device__ __forceinline__ unsigned int fold(unsigned long long v) {
  return static_cast<unsigned int>(v) ^ static_cast<unsigned int>(v >> 32);
}

// ---- lens shading ---------------------------------------------------------

// lens_shading_gain (isp_math.hpp:121-140), per sample: the slow path for
// tiles whose rows span more than two grid rows.
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

__device__ __forceinline__ int lsc_row(const FrameDesc& f, int y, float* ty) {
  const float fy = static_cast<float>(y) * f.lsc_sy;
  int iy = static_cast<int>(fy);
  iy = iy > f.lsc_height - 2 ? f.lsc_height - 2 : iy;
  *ty = fy - static_cast<float>(iy);
  return iy;
}

// The per-tile lens shading table: for sample columns cx0 .. cx0 + kCols - 1,
// both row parities and grid rows iy0, iy0 + 1: (top, bottom - top).
template <int kCols>
struct LscTable {
  F2 e[2][2][kCols];  // [grid row - iy0][row parity][column]
};

template <int kCols>
__device__ void build_lsc_table(const FrameDesc& f, int cx0, int iy0, LscTable<kCols>& t) {
  const int step = f.lsc_channels;
  const int row = f.lsc_width * step;
  for (int i = threadIdx.x; i < 4 * kCols; i += blockDim.x) {
    const int c = i % kCols;
    const int par = (i / kCols) & 1;
    const int seg = i / (2 * kCols);
    const int x = cx0 + c;
    if (x < 0 || x >= f.width) continue;
    const int iy = iy0 + seg > f.lsc_height - 2 ? f.lsc_height - 2 : iy0 + seg;
    const float fx = static_cast<float>(x) * f.lsc_sx;
    int ix = static_cast<int>(fx);
    ix = ix > f.lsc_width - 2 ? f.lsc_width - 2 : ix;
    const float tx = fx - static_cast<float>(ix);
    const int channel = cfa_channel(f.cfa, (par << 1) | (x & 1));
    const float* g = f.lsc + (iy * f.lsc_width + ix) * step + (step == 4 ? channel : 0);
    const float top = g[0] + (g[step] - g[0]) * tx;
    const float bottom = g[row] + (g[row + step] - g[row]) * tx;
    t.e[seg][par][c].v[0] = top;
    t.e[seg][par][c].v[1] = bottom - top;
  }
}

// The range of grid rows the tile's (mirrored) sample rows [r0, r0 + n) fall in.
__device__ void lsc_row_range(const FrameDesc& f, int r0, int n, int* lo, int* hi) {
  __shared__ int s_lo;
  __shared__ int s_hi;
  if (threadIdx.x == 0) {
    s_lo = 1 << 30;
    s_hi = -1;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float ty;
    const int iy = lsc_row(f, mirror_near(r0 + i, f.height), &ty);
    atomicMin(&s_lo, iy);
    atomicMax(&s_hi, iy);
  }
  __syncthreads();
  *lo = s_lo;
  *hi = s_hi;
}

// Per quad position p = (y & 1) * 2 + (x & 1): black, 1 / range, colour gain.
struct QuadConst {
  float black[4];
  float inv[4];
  float gain[4];
};

__device__ __forceinline__ QuadConst quad_consts(const FrameDesc& f, const float gains[3]) {
  QuadConst q;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int ch = cfa_channel(f.cfa, p);
    q.black[p] = f.black[ch];
    q.inv[p] = f.inv_range[ch];
    q.gain[p] = gains[color_of(ch)];
  }
  return q;
}

// ---- white balance statistics -------------------------------------------

// histogram_percentile (isp_math.hpp:336-352) over hist[1024], by one wave:
// each lane owns 16 consecutive bins; a prefix sum finds the bin.
__device__ float percentile_wave(const unsigned int* hist, float percentile) {
  const int lane = threadIdx.x & 63;
  unsigned long long mine = 0;
  for (int b = 0; b < 16; ++b) mine += hist[lane * 16 + b];
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
  const unsigned long long hits = __ballot(found != (1 << 30));
  if (hits == 0) return 1.0F;
  const int first = __ffsll(static_cast<long long>(hits)) - 1;
  const int bin = __shfl(found, first, 64);
  const double w = __shfl(within, first, 64);
  return static_cast<float>((static_cast<double>(bin) + w) / kAwbBins);
}

template <int kFmt>
__global__ void __launch_bounds__(kThreads) awb_kernel(FrameDesc f, DeviceScratch* s, int tiles_x, int tiles,
                                                       int fast_rows) {
  constexpr int kC = Chunk<kFmt>::kSamples;
  constexpr int kCols = kTile;
  __shared__ unsigned int hist[3 * kAwbBins];
  __shared__ unsigned long long part[kThreads / 64][4];
  __shared__ LscTable<kCols> lsc_tab;
  const bool patch = f.wb_mode == ISP_WB_WHITE_PATCH;
  const int tid = threadIdx.x;
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) hist[i] = 0U;
  }
  QuadConst qc;
  {
    const float ones[3] = {1.0F, 1.0F, 1.0F};
    qc = quad_consts(f, ones);
  }
  int colour[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) colour[p] = color_of(cfa_channel(f.cfa, p));
  const int qw = f.width / 2;
  const int qh = f.height / 2;
  unsigned long long local[4] = {0, 0, 0, 0};

  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int tyi = tile / tiles_x;
    const int x0 = (tile - tyi * tiles_x) * kTile;
    const int y0 = tyi * kTile;
    bool table = false;
    int iy0 = 0;
    __syncthreads();  // the previous tile's table readers are done
    if (f.lsc != nullptr) {
      int lo, hi;
      lsc_row_range(f, y0, min(kTile, f.height - y0), &lo, &hi);
      table = hi - lo <= 1;
      iy0 = lo;
      if (table) build_lsc_table<kCols>(f, x0, iy0, lsc_tab);
      __syncthreads();
    }
    // Each task: kC columns x 2 rows = kC / 2 quads.
    constexpr int kTasksX = kTile / kC;
    constexpr int kTasks = kTasksX * (kTile / 2);
    for (int task = tid; task < kTasks; task += kThreads) {
      const int qr = task / kTasksX;
      const int cx = x0 + (task - qr * kTasksX) * kC;
      const int qy = (y0 >> 1) + qr;
      if (qy >= qh || cx >= 2 * qw) continue;
      int raw[2][kC];
      const std::uint8_t* rows[2] = {f.raw + static_cast<long long>(2 * qy) * f.raw_pitch,
                                     f.raw + static_cast<long long>(2 * qy + 1) * f.raw_pitch};
      const bool whole = fast_rows != 0 && cx + kC <= f.width;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        if (whole) {
          load_chunk<kFmt>(rows[r], cx, raw[r]);
        } else {
#pragma unroll
          for (int k = 0; k < kC; ++k) raw[r][k] = cx + k < 2 * qw ? load_raw(rows[r], kFmt, cx + k) : 0;
        }
      }
      float tyr[2] = {0.0F, 0.0F};
      int seg[2] = {0, 0};
      if (table) {
#pragma unroll
        for (int r = 0; r < 2; ++r) seg[r] = lsc_row(f, 2 * qy + r, &tyr[r]) - iy0;
      }
      unsigned int sum[3] = {0U, 0U, 0U};
      unsigned int quads = 0U;
#pragma unroll
      for (int q = 0; q < kC / 2; ++q) {
        const int qx = (cx >> 1) + q;
        if (qx >= qw) continue;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) ok = ok && raw[k >> 1][2 * q + (k & 1)] < f.white;
        if (!ok) continue;
        quads += 1U;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int r = k >> 1;
          const int c = 2 * q + (k & 1);
          const int p = k;  // quad position: (y & 1) * 2 + (x & 1); quads start at even (x, y)
          float v = (static_cast<float>(raw[r][c]) - qc.black[p]) * qc.inv[p];
          v = v > 0.0F ? v : 0.0F;
          if (f.lsc != nullptr) {
            float gain;
            if (table) {
              const F2 e = lsc_tab.e[seg[r]][r][cx - x0 + c];
              gain = e.v[0] + e.v[1] * tyr[r];
            } else {
              gain = lsc_gain(f, cx + c, 2 * qy + r, cfa_channel(f.cfa, p));
            }
            v = v * gain;
          }
          v = v < 4.0F ? v : 4.0F;
          const int col = colour[p];
          const unsigned int fixed = static_cast<unsigned int>(v * 1048576.0F + 0.5F);
          sum[col] += fixed;
          if (patch) {
            const int b = static_cast<int>(v * static_cast<float>(kAwbBins));
            atomicAdd(&hist[col * kAwbBins + (b < kAwbBins - 1 ? b : kAwbBins - 1)], 1U);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) local[k] += sum[k];
      local[3] += quads;
    }
  }
  const int wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned long long v = wave_sum(local[k]);
    if ((tid & 63) == 0) part[wave][k] = v;
  }
  __syncthreads();
  unsigned int token = 0;
  if (tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += part[w][tid];
    if (total != 0) token ^= fold(atomicAdd(tid < 3 ? &s->awb.sum[tid] : &s->awb.quads, total));
  }
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) {
      if (hist[i] != 0U) token ^= atomicAdd(&s->awb.hist[i / kAwbBins][i % kAwbBins], hist[i]);
    }
  }
  if (!last_block(&s->awb.done, token)) return;

  // The last block: finish_white_balance (isp_math.hpp:357-384); the reads clear.
  __shared__ float level_s[3];
  __shared__ unsigned long long sums_s[4];
  if (tid < 4) sums_s[tid] = atomicExch(tid < 3 ? &s->awb.sum[tid] : &s->awb.quads, 0ULL);
  if (patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) hist[i] = atomicExch(&s->awb.hist[i / kAwbBins][i % kAwbBins], 0U);
  }
  __syncthreads();
  const unsigned long long quads_s = sums_s[3];
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
      level_s[0] = static_cast<float>(static_cast<double>(sums_s[0]) / quads);
      level_s[1] = static_cast<float>(static_cast<double>(sums_s[1]) / (2.0 * quads));
      level_s[2] = static_cast<float>(static_cast<double>(sums_s[2]) / quads);
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
  if (tid == 0) atomicExch(&s->awb.done, 0U);
}

// ---- the main pass ----------------------------------------------------------

// The two demosaic formulas (isp_math.hpp:161-215) a CFA site needs, from a
// window w[row][col] whose centre is w[cy][cx]: at an R or B site green and the
// diagonal colour, at a G site the colour across and the colour down.
template <bool kMalvar, int kRows, int kCols>
__device__ __forceinline__ void demosaic_pair(const float (&w)[kRows][kCols], int cy, int cx, bool rb,
                                              float* first, float* second) {
  const float c = w[cy][cx];
  const float n12 = w[cy - 1][cx];
  const float n32 = w[cy + 1][cx];
  const float n21 = w[cy][cx - 1];
  const float n23 = w[cy][cx + 1];
  if constexpr (!kMalvar) {
    if (rb) {
      *first = (n12 + n32 + n21 + n23) * 0.25F;
      *second = (w[cy - 1][cx - 1] + w[cy - 1][cx + 1] + w[cy + 1][cx - 1] + w[cy + 1][cx + 1]) * 0.25F;
    } else {
      *first = (n21 + n23) * 0.5F;
      *second = (n12 + n32) * 0.5F;
    }
  } else {
    const float n02 = w[cy - 2][cx];
    const float n42 = w[cy + 2][cx];
    const float n20 = w[cy][cx - 2];
    const float n24 = w[cy][cx + 2];
    const float diagonal1 = w[cy - 1][cx - 1] + w[cy - 1][cx + 1] + w[cy + 1][cx - 1] + w[cy + 1][cx + 1];
    const float vertical2 = n02 + n42;
    if (rb) {
      const float axial1 = n12 + n32 + n21 + n23;
      const float axial2 = n02 + n42 + n20 + n24;
      *first = clamp01((4.0F * c + 2.0F * axial1 - axial2) * 0.125F);
      *second = clamp01((6.0F * c + 2.0F * diagonal1 - 1.5F * axial2) * 0.125F);
    } else {
      const float horizontal2 = n20 + n24;
      *first = clamp01((5.0F * c + 4.0F * (n21 + n23) - horizontal2 - diagonal1 + 0.5F * vertical2) * 0.125F);
      *second = clamp01((5.0F * c + 4.0F * (n12 + n32) - vertical2 - diagonal1 + 0.5F * horizontal2) * 0.125F);
    }
  }
}

// RGB of a site of channel `ch` (uniform) from its sample c and the pair above.
__device__ __forceinline__ void assign_rgb(int ch, float c, float first, float second, float* rgb) {
  switch (ch) {
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

// tone (isp_math.hpp:235-243) on a table in global memory.
__device__ __forceinline__ float tone(const float* __restrict__ lut, int size, float x) {
  const float f = x * static_cast<float>(size - 1);
  const int i = static_cast<int>(f);
  if (i >= size - 1) return lut[size - 1];
  const float t = f - static_cast<float>(i);
  const float a = lut[i];
  return a + t * (lut[i + 1] - a);
}

template <int kFmt, bool kMalvar, bool kSharpen>
struct MainGeom {
  static constexpr int kRadius = kMalvar ? 2 : 1;
  static constexpr int kHalo = kRadius + (kSharpen ? 1 : 0);
  static constexpr int kC = Chunk<kFmt>::kSamples;
  // Sample tile: frame columns x0 - kHalo .. x0 + kTile + kHalo - 1 from LDS
  // column 0; the stride leaves room for the 8-wide windows of the last blocks
  // and keeps every window read 16-byte aligned.
  static constexpr int kSW = (kTile + 2 * kHalo + 2 + 3) / 4 * 4;
  static constexpr int kSH = kTile + 2 * kHalo;
  // Ring of demosaiced RGB (tile plus one pixel), 4-wide blocks.
  static constexpr int kRW = kSharpen ? (kTile + 2 + 3) / 4 * 4 : 4;
  static constexpr int kRH = kSharpen ? kTile + 2 : 2;
  // Lens shading table columns: the chunk-aligned span covering the sample columns.
  static constexpr int kLscCols = (kTile + 2 * kHalo + 2 * kC - 2) / kC * kC;
};

template <int kFmt, bool kMalvar, bool kSharpen>
__global__ void __launch_bounds__(kThreads) main_kernel(FrameDesc f, DeviceScratch* s, int tiles_x, int tiles,
                                                        int fast_rows, int out_aligned) {
  using G = MainGeom<kFmt, kMalvar, kSharpen>;
  constexpr int kR = G::kRadius;
  constexpr int kH = G::kHalo;
  constexpr int kC = G::kC;
  constexpr int kSW = G::kSW;
  constexpr int kSH = G::kSH;
  constexpr int kRW = G::kRW;
  constexpr int kRH = G::kRH;
  constexpr int kLscCols = G::kLscCols;
  __shared__ __attribute__((aligned(16))) float samp[kSH * kSW];
  __shared__ __attribute__((aligned(16))) float ring[3][kRH * kRW];
  __shared__ unsigned int hist[4 * ISP_HISTOGRAM_BINS];
  __shared__ LscTable<kLscCols> lsc_tab;
  __shared__ unsigned long long part[kThreads / 64][4];

  const int tid = threadIdx.x;
  const bool want_hist = f.histogram != nullptr;
  const bool want_stats = f.stats != nullptr;
  const int bins = f.bins;
  const bool wide = bins != ISP_HISTOGRAM_BINS;  // 1024-bin extension: global atomics
  float gains[3];
  if (f.wb_mode != ISP_WB_MANUAL) {
#pragma unroll
    for (int k = 0; k < 3; ++k) gains[k] = s->wb.gains[k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) gains[k] = f.wb[k];
  }
  const QuadConst qc = quad_consts(f, gains);
  if (want_hist && !wide) {
    for (int i = tid; i < 4 * ISP_HISTOGRAM_BINS; i += kThreads) hist[i] = 0U;
  }
  unsigned long long acc[4] = {0, 0, 0, 0};
  unsigned int wide_token = 0;
  const int W = f.width;
  const int H = f.height;
  const float* __restrict__ lut = f.lut;
  const int lut_size = f.lut_size;
  const float ccm[9] = {f.ccm[0], f.ccm[1], f.ccm[2], f.ccm[3], f.ccm[4], f.ccm[5], f.ccm[6], f.ccm[7], f.ccm[8]};

  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int tyi = tile / tiles_x;
    const int x0 = (tile - tyi * tiles_x) * kTile;
    const int y0 = tyi * kTile;
    const int sx0 = x0 - kH;  // frame column of sample column 0
    const int sy0 = y0 - kH;
    const int cx0 = (sx0 >= 0 ? sx0 / kC : -((-sx0 + kC - 1) / kC)) * kC;  // chunk-aligned start
    const bool edge_x = sx0 < 0 || x0 + kTile + kH > W;
    __syncthreads();  // the previous tile's readers are done
    bool table = false;
    int iy0 = 0;
    if (f.lsc != nullptr) {
      int lo, hi;
      lsc_row_range(f, sy0, kSH, &lo, &hi);
      table = hi - lo <= 1;
      iy0 = lo;
      if (table) build_lsc_table<kLscCols>(f, cx0, iy0, lsc_tab);
      __syncthreads();
    }

    // 1. Samples (preprocess, isp_math.hpp:151-155).
    unsigned int saturated = 0;
    const int n_chunks = (x0 + kTile + kH - cx0 + kC - 1) / kC;
    for (int task = tid; task < kSH * n_chunks; task += kThreads) {
      const int sr = task / n_chunks;
      const int cx = cx0 + (task - sr * n_chunks) * kC;
      const int yv = sy0 + sr;  // virtual row
      const int y = mirror_near(yv, H);
      const std::uint8_t* row = f.raw + static_cast<long long>(y) * f.raw_pitch;
      int raw[kC];
      if (fast_rows != 0 && cx >= 0 && cx + kC <= W) {
        load_chunk<kFmt>(row, cx, raw);
      } else {
#pragma unroll
        for (int k = 0; k < kC; ++k) raw[k] = cx + k >= 0 && cx + k < W ? load_raw(row, kFmt, cx + k) : 0;
      }
      const int par = y & 1;
      const float b0 = par ? qc.black[2] : qc.black[0];
      const float b1 = par ? qc.black[3] : qc.black[1];
      const float i0 = par ? qc.inv[2] : qc.inv[0];
      const float i1 = par ? qc.inv[3] : qc.inv[1];
      const float g0 = par ? qc.gain[2] : qc.gain[0];
      const float g1 = par ? qc.gain[3] : qc.gain[1];
      float tyr = 0.0F;
      int seg = 0;
      if (table) seg = lsc_row(f, y, &tyr) - iy0;
      const bool count = want_stats && yv >= y0 && yv < y0 + kTile && yv < H;
#pragma unroll
      for (int k = 0; k < kC; ++k) {
        const int x = cx + k;
        const int col = x - sx0;
        if (col < 0 || col >= kSW || x < 0 || x >= W) continue;
        const bool odd = (k & 1) != 0;  // cx is even
        float v = (static_cast<float>(raw[k]) - (odd ? b1 : b0)) * (odd ? i1 : i0);
        v = v > 0.0F ? v : 0.0F;
        if (f.lsc != nullptr) {
          float gain;
          if (table) {
            const F2 e = lsc_tab.e[seg][par][x - cx0];
            gain = e.v[0] + e.v[1] * tyr;
          } else {
            gain = lsc_gain(f, x, y, cfa_channel(f.cfa, (par << 1) | (x & 1)));
          }
          v = v * gain;
        }
        v = v * (odd ? g1 : g0);
        v = v < 1.0F ? v : 1.0F;
        samp[sr * kSW + col] = v;
        if (count && x >= x0 && x < x0 + kTile) saturated += raw[k] >= f.white ? 1U : 0U;
      }
    }
    acc[3] += saturated;
    if (edge_x) {
      // Columns outside the frame take the sample of their reflection.
      __syncthreads();
      for (int i = tid; i < kSH * kSW; i += kThreads) {
        const int sr = i / kSW;
        const int col = i - sr * kSW;
        const int x = sx0 + col;
        if (x >= 0 && x < W) continue;
        const int src = mirror_near(x, W) - sx0;
        samp[i] = src >= 0 && src < kSW && src + sx0 < W ? samp[sr * kSW + src] : 0.0F;
      }
    }
    __syncthreads();

    // 2. With sharpening: the demosaiced RGB of the tile and a one-pixel ring,
    // ring pixel (rc, rr) = virtual pixel (x0 - 1 + rc, y0 - 1 + rr), in 4x2 blocks.
    if constexpr (kSharpen) {
      constexpr int kBX = kRW / 4;
      constexpr int kBlocks = kBX * (kRH / 2);
      for (int b = tid; b < kBlocks; b += kThreads) {
        const int by = b / kBX;
        const int bx = b - by * kBX;
        const int rc = 4 * bx;
        const int rr = 2 * by;
        // window: sample rows rr + kH - 1 - kR .. + 1 + kR, columns rc + kH - 1 - kR, 8 wide
        constexpr int kWR = 2 + 2 * kR;
        constexpr int kWC = 8;
        float w[kWR][kWC];
        const int base = (rr + kH - 1 - kR) * kSW + rc + kH - 1 - kR;
#pragma unroll
        for (int r = 0; r < kWR; ++r) {
          const F4 a = *reinterpret_cast<const F4*>(&samp[base + r * kSW]);
          const F4 c = *reinterpret_cast<const F4*>(&samp[base + r * kSW + 4]);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            w[r][k] = a.v[k];
            w[r][4 + k] = c.v[k];
          }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          float rgbs[4][3];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            // virtual pixel parity: ((1 + i) & 1, (1 + j) & 1)
            const int ch = cfa_channel(f.cfa, (((1 + j) & 1) << 1) | ((1 + i) & 1));
            const bool rb = ch == ISP_CHANNEL_R || ch == ISP_CHANNEL_B;
            float first, second;
            demosaic_pair<kMalvar>(w, kR + j, kR + i, rb, &first, &second);
            assign_rgb(ch, w[kR + j][kR + i], first, second, rgbs[i]);
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            F4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) o.v[i] = rgbs[i][k];
            *reinterpret_cast<F4*>(&ring[k][(rr + j) * kRW + rc]) = o;
          }
        }
      }
      __syncthreads();
    }

    // 3. A 4x4 block of output pixels per thread.
    {
      const int bx = tid & 15;
      const int by = tid >> 4;
      const int px = 4 * bx;
      const int py = 4 * by;
      const int xo = x0 + px;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int yo = y0 + py + j;
        float rgb[4][3];
        if constexpr (kSharpen) {
          // ring rows py + j .. + 2, columns px .. px + 5
          float m[3][3][6];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const float* src = &ring[k][(py + j + r) * kRW + px];
              const F4 a = *reinterpret_cast<const F4*>(src);
              const F2 b2 = *reinterpret_cast<const F2*>(src + 4);
              m[k][r][0] = a.v[0];
              m[k][r][1] = a.v[1];
              m[k][r][2] = a.v[2];
              m[k][r][3] = a.v[3];
              m[k][r][4] = b2.v[0];
              m[k][r][5] = b2.v[1];
            }
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              // sharpen (isp_math.hpp:219-227)
              const float c = m[k][1][i + 1];
              const float blur = (4.0F * c + 2.0F * (m[k][0][i + 1] + m[k][2][i + 1] + m[k][1][i] + m[k][1][i + 2]) +
                                  (m[k][0][i] + m[k][0][i + 2] + m[k][2][i] + m[k][2][i + 2])) *
                                 0.0625F;
              rgb[i][k] = clamp01(c + f.sharpen * (c - blur));
            }
          }
        } else {
          constexpr int kWR = 1 + 2 * kR;
          constexpr int kWC = 8;
          float w[kWR][kWC];
          const int base = (py + j + kH - kR) * kSW + px + kH - kR;
#pragma unroll
          for (int r = 0; r < kWR; ++r) {
            const F4 a = *reinterpret_cast<const F4*>(&samp[base + r * kSW]);
            const F4 c = *reinterpret_cast<const F4*>(&samp[base + r * kSW + 4]);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              w[r][k] = a.v[k];
              w[r][4 + k] = c.v[k];
            }
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int ch = cfa_channel(f.cfa, ((j & 1) << 1) | (i & 1));
            const bool rb = ch == ISP_CHANNEL_R || ch == ISP_CHANNEL_B;
            float first, second;
            demosaic_pair<kMalvar>(w, kR, kR + i, rb, &first, &second);
            assign_rgb(ch, w[kR][kR + i], first, second, rgb[i]);
          }
        }
        if (yo >= H) continue;
        // correct_color, to_fixed sums, tone, store, output codes (isp_math.hpp:229-305)
        float t[4][3];
        int q[4][3];
        const bool valid[4] = {xo < W, xo + 1 < W, xo + 2 < W, xo + 3 < W};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float lin[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            lin[k] = clamp01(ccm[3 * k] * rgb[i][0] + ccm[3 * k + 1] * rgb[i][1] + ccm[3 * k + 2] * rgb[i][2]);
          }
          if (want_stats && valid[i]) {
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += static_cast<unsigned int>(lin[k] * 1048576.0F + 0.5F);
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) t[i][k] = tone(lut, lut_size, lin[k]);
        }
        std::uint8_t* orow = f.out + static_cast<long long>(yo) * f.out_pitch;
        const bool whole = out_aligned != 0 && xo + 4 <= W;
        switch (f.out_format) {
          case ISP_OUT_RGB8: {
            unsigned int byte[12];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<unsigned char>(t[i][k] * 255.0F + 0.5F);
                q[i][k] = static_cast<int>(v);
                byte[3 * i + k] = v;
              }
            }
            if (whole) {
              W3 o;
#pragma unroll
              for (int w = 0; w < 3; ++w) {
                o.v[w] = byte[4 * w] | (byte[4 * w + 1] << 8) | (byte[4 * w + 2] << 16) | (byte[4 * w + 3] << 24);
              }
              *reinterpret_cast<W3*>(orow + 3 * xo) = o;
            } else {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                if (!valid[i]) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) orow[3 * (xo + i) + k] = static_cast<std::uint8_t>(byte[3 * i + k]);
              }
            }
            break;
          }
          case ISP_OUT_RGBA8: {
            unsigned int word[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              word[i] = 0xFF000000U;
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<unsigned char>(t[i][k] * 255.0F + 0.5F);
                q[i][k] = static_cast<int>(v);
                word[i] |= v << (8 * k);
              }
            }
            if (whole) {
              W4 o;
#pragma unroll
              for (int i = 0; i < 4; ++i) o.v[i] = word[i];
              *reinterpret_cast<W4*>(orow + 4 * xo) = o;
            } else {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                if (!valid[i]) continue;
#pragma unroll
                for (int k = 0; k < 4; ++k) orow[4 * (xo + i) + k] = static_cast<std::uint8_t>(word[i] >> (8 * k));
              }
            }
            break;
          }
          case ISP_OUT_RGB16: {
            unsigned int half[12];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const unsigned int v = static_cast<std::uint16_t>(t[i][k] * 65535.0F + 0.5F);
                half[3 * i + k] = v;
                q[i][k] = static_cast<int>((v * static_cast<unsigned int>(bins - 1) + 32767U) / 65535U);
              }
            }
            if (whole) {
#pragma unroll
              for (int w = 0; w < 3; ++w) {
                W2 o;
                o.v[0] = half[4 * w] | (half[4 * w + 1] << 16);
                o.v[1] = half[4 * w + 2] | (half[4 * w + 3] << 16);
                *reinterpret_cast<W2*>(orow + 6 * xo + 8 * w) = o;
              }
            } else {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                if (!valid[i]) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                  const std::uint16_t v = static_cast<std::uint16_t>(half[3 * i + k]);
                  orow[6 * (xo + i) + 2 * k] = static_cast<std::uint8_t>(v & 0xFF);
                  orow[6 * (xo + i) + 2 * k + 1] = static_cast<std::uint8_t>(v >> 8);
                }
              }
            }
            break;
          }
          default: {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                q[i][k] = static_cast<int>(clamp01(t[i][k]) * static_cast<float>(bins - 1) + 0.5F);
              }
            }
            if (whole) {
#pragma unroll
              for (int w = 0; w < 3; ++w) {
                F4 o;
#pragma unroll
                for (int k = 0; k < 4; ++k) o.v[k] = t[(4 * w + k) / 3][(4 * w + k) % 3];
                *reinterpret_cast<F4*>(orow + 12 * xo + 16 * w) = o;
              }
            } else {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                if (!valid[i]) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) reinterpret_cast<float*>(orow)[3 * (xo + i) + k] = t[i][k];
              }
            }
            break;
          }
        }
        if (want_hist) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (!valid[i]) continue;
            const int luma = (2126 * q[i][0] + 7152 * q[i][1] + 722 * q[i][2] + 5000) / 10000;
            if (!wide) {
              atomicAdd(&hist[luma], 1U);
              atomicAdd(&hist[ISP_HISTOGRAM_BINS + q[i][0]], 1U);
              atomicAdd(&hist[2 * ISP_HISTOGRAM_BINS + q[i][1]], 1U);
              atomicAdd(&hist[3 * ISP_HISTOGRAM_BINS + q[i][2]], 1U);
            } else {
              wide_token ^= atomicAdd(&s->main.hist[0][luma], 1U);
              wide_token ^= atomicAdd(&s->main.hist[1][q[i][0]], 1U);
              wide_token ^= atomicAdd(&s->main.hist[2][q[i][1]], 1U);
              wide_token ^= atomicAdd(&s->main.hist[3][q[i][2]], 1U);
            }
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
  unsigned int token = wide_token;
  if (want_stats && tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += part[w][tid];
    if (total != 0) token ^= fold(atomicAdd(&s->main.sum[tid], total));
  }
  if (want_hist && !wide) {
    for (int i = tid; i < 4 * ISP_HISTOGRAM_BINS; i += kThreads) {
      if (hist[i] != 0U) token ^= atomicAdd(&s->main.hist[i / ISP_HISTOGRAM_BINS][i % ISP_HISTOGRAM_BINS], hist[i]);
    }
  }
  if (!last_block(&s->main.done, token)) return;

  // The last block: the histograms and the statistics vector; the reads clear.
  if (want_hist) {
    for (int i = tid; i < 4 * bins; i += kThreads) f.histogram[i] = atomicExch(&s->main.hist[i / bins][i % bins], 0U);
  }
  __shared__ unsigned long long sums_s[4];
  if (tid < 4) sums_s[tid] = atomicExch(&s->main.sum[tid], 0ULL);
  __syncthreads();
  if (want_stats && tid == 0) {
    const unsigned long long sums[4] = {sums_s[0], sums_s[1], sums_s[2], sums_s[3]};
    finish_stats(sums, static_cast<long long>(W) * H, gains, f.ae, f.exposure_ns, f.exposure_gain, f.stats);
  }
  if (tid == 0) atomicExch(&s->main.done, 0U);
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
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kThreads, 0),
        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  blocks = std::max(blocks, 1);
  cache[reinterpret_cast<const void*>(kernel)] = blocks;
  return blocks;
}

// Rows can be read with aligned chunk loads.
bool rows_aligned(const FrameDesc& f) {
  const auto p = reinterpret_cast<unsigned long long>(f.raw);
  return ((p | static_cast<unsigned long long>(f.raw_pitch)) & 15ULL) == 0;
}

bool out_aligned(const FrameDesc& f) {
  const auto p = reinterpret_cast<unsigned long long>(f.out);
  return ((p | static_cast<unsigned long long>(f.out_pitch)) & 15ULL) == 0;
}

template <int kFmt, bool kMalvar, bool kSharpen>
void launch_main_t(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream) {
  auto kernel = main_kernel<kFmt, kMalvar, kSharpen>;
  const int tiles_x = (f.width + kTile - 1) / kTile;
  const int tiles_y = (f.height + kTile - 1) / kTile;
  const int tiles = tiles_x * tiles_y;
  const int grid = std::max(1, std::min(tiles, cfg.cu_count * blocks_per_cu(kernel)));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, f, s, tiles_x, tiles, rows_aligned(f) ? 1 : 0,
                     out_aligned(f) ? 1 : 0);
  check(hipGetLastError(), "main_kernel launch");
}

template <int kFmt>
void launch_main_f(const FrameDesc& f, DeviceScratch* s, bool malvar, bool sharpen, const LaunchConfig& cfg,
                   hipStream_t st) {
  if (malvar) {
    sharpen ? launch_main_t<kFmt, true, true>(f, s, cfg, st) : launch_main_t<kFmt, true, false>(f, s, cfg, st);
  } else {
    sharpen ? launch_main_t<kFmt, false, true>(f, s, cfg, st) : launch_main_t<kFmt, false, false>(f, s, cfg, st);
  }
}

template <int kFmt>
void launch_awb_f(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream) {
  auto kernel = awb_kernel<kFmt>;
  const int tiles_x = (f.width + kTile - 1) / kTile;
  const int tiles_y = (f.height + kTile - 1) / kTile;
  const int tiles = tiles_x * tiles_y;
  const int grid = std::max(1, std::min(tiles, cfg.cu_count * blocks_per_cu(kernel)));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), 0, stream, f, s, tiles_x, tiles, rows_aligned(f) ? 1 : 0);
  check(hipGetLastError(), "awb_kernel launch");
}

}  // namespace

void launch_awb(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, void* stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  switch (f.raw_format) {
    case ISP_RAW_U8:
      launch_awb_f<ISP_RAW_U8>(f, s, cfg, st);
      break;
    case ISP_RAW_U16:
      launch_awb_f<ISP_RAW_U16>(f, s, cfg, st);
      break;
    case ISP_RAW_RAW10:
      launch_awb_f<ISP_RAW_RAW10>(f, s, cfg, st);
      break;
    default:
      launch_awb_f<ISP_RAW_RAW12>(f, s, cfg, st);
      break;
  }
}

void launch_main(const FrameDesc& f, DeviceScratch* s, bool malvar, bool sharpen, const LaunchConfig& cfg,
                 void* stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  switch (f.raw_format) {
    case ISP_RAW_U8:
      launch_main_f<ISP_RAW_U8>(f, s, malvar, sharpen, cfg, st);
      break;
    case ISP_RAW_U16:
      launch_main_f<ISP_RAW_U16>(f, s, malvar, sharpen, cfg, st);
      break;
    case ISP_RAW_RAW10:
      launch_main_f<ISP_RAW_RAW10>(f, s, malvar, sharpen, cfg, st);
      break;
    default:
      launch_main_f<ISP_RAW_RAW12>(f, s, malvar, sharpen, cfg, st);
      break;
  }
}

}  // namespace isp_amd
