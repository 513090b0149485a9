// isp_kernels.hip - the ISP's gfx950 kernels.
//
// Per-frame form, two launches per frame (one when the white balance is manual):
//
//  awb_kernel   white balance statistics (reference: white_balance_quad,
//               isp_math.hpp:311-332; awb_kernel, cuda_pipeline.cu:119-158).
//               Persistent blocks walk 64x64 tiles; each thread takes kC/2
//               complete 2x2 quads from two coalesced row-chunk loads, skips
//               quads with a saturated sample, and sums the linearised values
//               (2^-20 units) per quad position in registers, in straight-line
//               code for whole chunks; white patch adds an LDS-privatised
//               3 x 1024 histogram. Each block gives the positions their
//               colours once, when it adds its sums to the frame's scratch.
//               The last block to finish turns the sums into gains
//               (finish_white_balance, isp_math.hpp:357-384) in the
//               pipeline's device state.
//
//  main_kernel  persistent blocks walk 64 x 32 output tiles:
//               1. every thread first issues its coalesced row-chunk loads
//                  (RAW12: 8 samples from 3 dwords, RAW10: 16 from 5, u16: 8
//                  from 4, u8: 8 from 2), then unpacks them and applies black
//                  level, lens shading, white balance and the clip, writing
//                  four samples per 128-bit LDS store into a tile with a
//                  mirrored halo;
//               2. with sharpening, Malvar / bilinear demosaic of the tile and
//                  a one-pixel ring into LDS (8 x 2 pixels per thread, windows
//                  read as 128-bit LDS loads, only the two formulas each CFA
//                  site needs);
//               3. per thread a 4 x 2 block of output pixels: unsharp mask (or
//                  the demosaic itself), colour matrix, tone table (through
//                  the caches), one nontemporal vector store per row, the
//                  statistics sums in registers and the 4 x bins histograms of
//                  the written codes in eight lane-interleaved LDS copies (fewer
//                  same-address conflicts).
//               The last block writes the histograms and the statistics vector
//               (finish_stats, isp_math.hpp:387-425, with the exact log2/exp2
//               of isp_ae.hpp).
//
// Batched form, many frames in two persistent launches (batch_awb_kernel, then
// batch_main_kernel): each pass's grid takes work items (a few tiles of one
// frame) from a device queue that lists the frames largest first. The same
// tile bodies run per item, chosen per frame among all raw formats, both
// demosaics, with and without sharpening and two tile heights. A block keeps
// one frame's partial sums and histograms in LDS while it takes consecutive
// items of that frame and adds them to the frame's scratch when it moves on;
// the block whose contribution completes a frame's item count runs that
// frame's epilogue. The kernel boundary between the passes orders each frame's
// white balance before its main pass.
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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <map>
#include <type_traits>
#include <utility>
#include <mutex>
#include <stdexcept>
#include <string>

#include "isp_device.hpp"
#include "isp_internal.hpp"

namespace isp_amd {
namespace {

constexpr int kThreads = 256;

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

// 16 bytes of LDS read as a native 16-byte-aligned vector, which becomes one
// ds_read_b128 where all four elements are used. gfx950's LDS serves a wave's
// 128-bit reads at a 16-byte lane stride without bank conflicts; the same bytes
// as ds_read2_b64 (what the compiler makes of a struct copy it can only prove
// 8-byte aligned) cost several conflict cycles.
typedef float V4 __attribute__((ext_vector_type(4)));
// Output words as vectors for nontemporal stores (dword-aligned where the type says so).
typedef unsigned int U4 __attribute__((ext_vector_type(4)));
typedef unsigned int U3 __attribute__((ext_vector_type(3), aligned(4)));
typedef unsigned int U2 __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ V4 lds_v4(const float* p) { return *reinterpret_cast<const V4*>(p); }

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
// clamp01 (isp_math.hpp:157), bit for bit for every finite v, -0 included (kept,
// as the reference keeps it): the upper cut as a v_min_f32, the lower as a
// select, one operation fewer than two selects.
__device__ __forceinline__ float clamp01(float v) { return v < 0.0F ? 0.0F : fminf(v, 1.0F); }
// The sign of a zero before the tone table reaches an output only through the
// table's first entry: tone(+-0) = lut[0] + (+-0) (lut[1] - lut[0]), which is
// +0 or lut[0] itself for either sign unless lut[0] is -0.0 (a user table;
// validation reads it as 0). Every other use of these values (products, sums,
// to_fixed, the output codes) gives the same result for +0 and -0, and no
// operation of the path turns a zero's sign into a nonzero value. So for a
// table that does not start at -0.0 (kZeroSign false) a clamp may give either
// zero for -0: clamp_unit is then med3(v, 0, 1), which the compiler folds into
// the producing instruction's clamp modifier, and the tone fraction is
// v_fract_f32. With kZeroSign the reference's forms keep every sign.
template <bool kZeroSign>
__device__ __forceinline__ float clamp_unit(float v) {
  if constexpr (kZeroSign) {
    return clamp01(v);
  } else {
    return __builtin_amdgcn_fmed3f(v, 0.0F, 1.0F);
  }
}

// p * a + b for p = 2 or 4, in one instruction. Scaling up by a power of two
// is exact for every float short of overflow (the values here are at most a few
// units), subnormals included, so the single rounding of the fused operation is
// the rounding of the reference's sum: bit for bit the unfused p * a + b.
// (Halving is not exact for a subnormal with its last bit set, so products by
// 0.5 stay unfused.)
__device__ __forceinline__ float exact_fma(float p, float a, float b) { return __builtin_fmaf(p, a, b); }

// to_fixed (isp_math.hpp:245-247). (Folding 2^20 into an exact_fma holds the
// constant in a scalar register, which in the main pass costs more in scalar
// spills than the product saves.)
__device__ __forceinline__ unsigned int to_fixed(float v) {
  return static_cast<unsigned int>(v * 1048576.0F + 0.5F);
}

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

// The white balance state is written with device-scope atomics (performed at
// the point of coherence), as the statistics scratch is.
__device__ __forceinline__ unsigned int coherent_write(float* p, float v) {
  return atomicExch(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}

// ---- lens shading ---------------------------------------------------------

constexpr int kMaxGrid = 1024;  // lens shading grid values staged in LDS when they fit

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

// The grid row of frame row y and the weight ty within it.
__device__ __forceinline__ int lsc_row(const FrameDesc& f, int y, float* ty) {
  const float fy = static_cast<float>(y) * f.lsc_sy;
  int iy = static_cast<int>(fy);
  iy = iy > f.lsc_height - 2 ? f.lsc_height - 2 : iy;
  *ty = fy - static_cast<float>(iy);
  return iy;
}
__device__ __forceinline__ int lsc_row(const FrameDesc& f, int y) {
  float ty;
  return lsc_row(f, y, &ty);
}

// Per-tile table for sample columns cx0 .. cx0 + kCols - 1, both row parities
// and grid rows iy0, iy0 + 1: (top, bottom - top) of lens_shading_gain.
template <int kCols>
struct LscTable {
  F2 e[2][2][kCols];  // [grid row - iy0][row parity][column]
};

// `grid` is the lens shading grid (LDS or global memory).
template <int kCols>
__device__ __forceinline__ void build_lsc_table(const FrameDesc& f, const float* grid, int cx0, int iy0,
                                                LscTable<kCols>& t) {
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
    const float* g = grid + (iy * f.lsc_width + ix) * step + (step == 4 ? channel : 0);
    const float top = g[0] + (g[step] - g[0]) * tx;
    const float bottom = g[row] + (g[row + step] - g[row]) * tx;
    t.e[seg][par][c].v[0] = top;
    t.e[seg][par][c].v[1] = bottom - top;
  }
}

// Copies the lens shading grid into LDS when it fits; returns where to read it.
__device__ __forceinline__ const float* stage_grid(const FrameDesc& f, float* grid_s) {
  if (f.lsc == nullptr) return nullptr;
  const int n = f.lsc_width * f.lsc_height * f.lsc_channels;
  if (n > kMaxGrid) return f.lsc;
  for (int i = threadIdx.x; i < n; i += blockDim.x) grid_s[i] = f.lsc[i];
  return grid_s;
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

// Raw words of one chunk (kSamples samples), as loaded.
template <int kFmt>
struct Packed {
  static constexpr int kWords = kFmt == ISP_RAW_U8 ? 2 : kFmt == ISP_RAW_U16 ? 4 : kFmt == ISP_RAW_RAW12 ? 3 : 5;
  unsigned int w[kWords];
};

template <int kFmt>
__device__ __forceinline__ void load_packed(const std::uint8_t* row, int x, Packed<kFmt>& p) {
  if constexpr (kFmt == ISP_RAW_U8) {
    const W2 w = *reinterpret_cast<const W2*>(row + x);
    p.w[0] = w.v[0];
    p.w[1] = w.v[1];
  } else if constexpr (kFmt == ISP_RAW_U16) {
    const W4 w = *reinterpret_cast<const W4*>(row + 2 * x);
#pragma unroll
    for (int k = 0; k < 4; ++k) p.w[k] = w.v[k];
  } else if constexpr (kFmt == ISP_RAW_RAW12) {
    const W3 w = *reinterpret_cast<const W3*>(row + 3 * (x >> 1));
#pragma unroll
    for (int k = 0; k < 3; ++k) p.w[k] = w.v[k];
  } else {
    const W5 w = *reinterpret_cast<const W5*>(row + 5 * (x >> 2));
#pragma unroll
    for (int k = 0; k < 5; ++k) p.w[k] = w.v[k];
  }
}

template <int kFmt>
__device__ __forceinline__ void unpack(const Packed<kFmt>& p, int* s) {
  if constexpr (kFmt == ISP_RAW_U8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = static_cast<int>(byte_of(p.w, k));
  } else if constexpr (kFmt == ISP_RAW_U16) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      s[2 * k] = static_cast<int>(p.w[k] & 0xFFFFU);
      s[2 * k + 1] = static_cast<int>(p.w[k] >> 16);
    }
  } else if constexpr (kFmt == ISP_RAW_RAW12) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const unsigned int lo = byte_of(p.w, 3 * g + 2);
      s[2 * g] = static_cast<int>((byte_of(p.w, 3 * g) << 4) | (lo & 15U));
      s[2 * g + 1] = static_cast<int>((byte_of(p.w, 3 * g + 1) << 4) | (lo >> 4));
    }
  } else {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const unsigned int lo = byte_of(p.w, 5 * g + 4);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[4 * g + k] = static_cast<int>((byte_of(p.w, 5 * g + k) << 2) | ((lo >> (2 * k)) & 3U));
      }
    }
  }
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

constexpr int kAwbTile = kBatchAwbTile;
// A thread's white balance accumulators: the fixed-point sums of the four quad
// positions, then the count of complete quads.
constexpr int kAwbAcc = 5;

// The white balance pass's LDS histograms (white patch) in interleaved copies,
// `bin * kAwbCopies + lane % kAwbCopies`, which spread the atomics of lanes
// that share a bin (flat regions) over addresses. After the flush, the last
// block reuses the first 3 x kAwbBins entries for the frame's totals.
constexpr int kAwbCopies = 2;

// LDS of the white balance pass: per frame, and per tile.
struct AwbFrameLds {
  unsigned int hist[3 * kAwbBins * kAwbCopies];
  float grid[kMaxGrid];
  unsigned long long part[kThreads / 64][4];
  unsigned long long sums[4];
  float level[3];
  int flag;
};
struct AwbTileLds {
  LscTable<kAwbTile> lsc;
};

// A block's constants for one frame of the white balance pass.
struct AwbCtx {
  QuadConst qc;
  int colour[4];
  const float* grid;
  bool patch;
  bool lsc;
};

// Starts a block's work on a frame. The tile's first barrier orders these LDS
// writes before any reader.
__device__ __forceinline__ void awb_begin(const FrameDesc& f, AwbFrameLds& fl, AwbCtx& c) {
  c.patch = f.wb_mode == ISP_WB_WHITE_PATCH;
  if (c.patch) {
    for (int i = threadIdx.x; i < 3 * kAwbBins * kAwbCopies; i += kThreads) fl.hist[i] = 0U;
  }
  c.grid = stage_grid(f, fl.grid);
  const float ones[3] = {1.0F, 1.0F, 1.0F};
  c.qc = quad_consts(f, ones);
#pragma unroll
  for (int p = 0; p < 4; ++p) c.colour[p] = color_of(cfa_channel(f.cfa, p));
  c.lsc = f.lsc != nullptr;
}

// A thread's task in a 64 x 64 sample tile: one row chunk of kC samples in
// each of the two rows of a quad row (kC / 2 complete quads).
template <int kFmt>
struct AwbTask {
  const std::uint8_t* row0;
  int x0, y0, cx, qy;
  bool active, whole;
};

template <int kFmt>
__device__ __forceinline__ AwbTask<kFmt> awb_task(const FrameDesc& f, int tile, int tiles_x, int fast_rows) {
  constexpr int kC = Chunk<kFmt>::kSamples;
  constexpr int kTasksX = kAwbTile / kC;
  constexpr int kTasks = kTasksX * (kAwbTile / 2);
  static_assert(kTasks <= kThreads, "one task per thread");
  const int tid = threadIdx.x;
  const int qw = f.width / 2;
  const int qh = f.height / 2;
  AwbTask<kFmt> t;
  const int tyi = tile / tiles_x;
  t.x0 = (tile - tyi * tiles_x) * kAwbTile;
  t.y0 = tyi * kAwbTile;
  const int qr = tid / kTasksX;
  t.cx = t.x0 + (tid - qr * kTasksX) * kC;
  t.qy = (t.y0 >> 1) + qr;
  t.active = tid < kTasks && t.qy < qh && t.cx < 2 * qw;
  t.whole = t.active && fast_rows != 0 && t.cx + kC <= f.width;
  t.row0 = f.raw + static_cast<long long>(2 * t.qy) * f.raw_pitch;
  return t;
}

// A tile's coalesced chunk loads (whole chunks only; the edges read byte by
// byte in awb_tile).
template <int kFmt>
__device__ __forceinline__ void awb_load(const FrameDesc& f, const AwbTask<kFmt>& t, Packed<kFmt> (&pk)[2]) {
  if (t.whole) {
    load_packed<kFmt>(t.row0, t.cx, pk[0]);
    load_packed<kFmt>(t.row0 + f.raw_pitch, t.cx, pk[1]);
  }
}

// The common case of a tile's task: a whole chunk, gray-world sums, lens
// shading from the tile's table (or none). Straight-line code: quad position p =
// 2 r + (column & 1) is known per sample (chunks start at even columns), the
// chunk's table entries come in 128-bit LDS reads, and a whole chunk's quads
// all lie inside the frame. Same arithmetic as awb_tile's general path.
template <int kFmt, bool kLsc>
__device__ __forceinline__ void awb_chunk_fast(const FrameDesc& f, const AwbCtx& c, AwbTileLds& tl,
                                               const AwbTask<kFmt>& t, const Packed<kFmt> (&pk)[2], int lo,
                                               unsigned long long local[kAwbAcc]) {
  constexpr int kC = Chunk<kFmt>::kSamples;
  int raw[2][kC];
  unpack<kFmt>(pk[0], raw[0]);
  unpack<kFmt>(pk[1], raw[1]);
  float gain[2][kC];
  if constexpr (kLsc) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      float ty;
      const int seg = lsc_row(f, 2 * t.qy + r, &ty) - lo;
      const F4* e = reinterpret_cast<const F4*>(&tl.lsc.e[seg][r][t.cx - t.x0]);
#pragma unroll
      for (int h = 0; h < kC / 2; ++h) {
        const F4 v = e[h];
        gain[r][2 * h] = v.v[0] + v.v[1] * ty;
        gain[r][2 * h + 1] = v.v[2] + v.v[3] * ty;
      }
    }
  }
  unsigned int pos[4] = {0U, 0U, 0U, 0U};
  unsigned int quads = 0U;
#pragma unroll
  for (int q = 0; q < kC / 2; ++q) {
    const bool ok = raw[0][2 * q] < f.white && raw[0][2 * q + 1] < f.white && raw[1][2 * q] < f.white &&
                    raw[1][2 * q + 1] < f.white;
    unsigned int fx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = k >> 1;
      const int col = 2 * q + (k & 1);
      float v = (static_cast<float>(raw[r][col]) - c.qc.black[k]) * c.qc.inv[k];
      v = v > 0.0F ? v : 0.0F;
      if constexpr (kLsc) v = v * gain[r][col];
      v = v < 4.0F ? v : 4.0F;
      fx[k] = to_fixed(v);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) pos[k] += ok ? fx[k] : 0U;
    quads += ok ? 1U : 0U;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) local[k] += pos[k];
  local[4] += quads;
}

// One 64 x 64 sample tile, its chunks loaded by awb_load: white_balance_quad
// (isp_math.hpp:311-332) over its complete quads: the fixed-point values add up
// per quad position into local[0..3] (exact integer sums; awb_flush gives them
// their colours), complete quads into local[4].
template <int kFmt>
__device__ __forceinline__ void awb_tile(const FrameDesc& f, const AwbCtx& c, AwbFrameLds& fl, AwbTileLds& tl,
                                         const AwbTask<kFmt>& t, const Packed<kFmt> (&pk)[2],
                                         unsigned long long local[kAwbAcc]) {
  constexpr int kC = Chunk<kFmt>::kSamples;
  const int H = f.height;
  const int qw = f.width / 2;
  int lo = 0;
  bool table = false;
  __syncthreads();  // the previous tile's table readers are done
  if (c.lsc) {
    lo = lsc_row(f, t.y0);
    table = lsc_row(f, min(t.y0 + kAwbTile - 1, H - 1)) - lo <= 1;
    if (table) build_lsc_table<kAwbTile>(f, c.grid, t.x0, lo, tl.lsc);
  }
  __syncthreads();
  if (!t.active) return;
  // Every tile has the table where the frame's grid rows allow it; the test is
  // made uniform so that the choice below is a scalar branch.
  table = __builtin_amdgcn_readfirstlane(static_cast<int>(table)) != 0;
  if (t.whole && (table || !c.lsc) && !c.patch) {
    if (c.lsc) {
      awb_chunk_fast<kFmt, true>(f, c, tl, t, pk, lo, local);
    } else {
      awb_chunk_fast<kFmt, false>(f, c, tl, t, pk, lo, local);
    }
    return;
  }
  const std::uint8_t* row1 = t.row0 + f.raw_pitch;
  int raw[2][kC];
  if (t.whole) {
    unpack<kFmt>(pk[0], raw[0]);
    unpack<kFmt>(pk[1], raw[1]);
  } else {
#pragma unroll
    for (int k = 0; k < kC; ++k) {
      raw[0][k] = t.cx + k < 2 * qw ? load_raw(t.row0, kFmt, t.cx + k) : 0;
      raw[1][k] = t.cx + k < 2 * qw ? load_raw(row1, kFmt, t.cx + k) : 0;
    }
  }
  float tyr[2] = {0.0F, 0.0F};
  int seg[2] = {0, 0};
  if (table) {
#pragma unroll
    for (int r = 0; r < 2; ++r) seg[r] = lsc_row(f, 2 * t.qy + r, &tyr[r]) - lo;
  }
  unsigned int pos[4] = {0U, 0U, 0U, 0U};
  unsigned int quads = 0U;
#pragma unroll
  for (int q = 0; q < kC / 2; ++q) {
    if ((t.cx >> 1) + q >= qw) continue;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) ok = ok && raw[k >> 1][2 * q + (k & 1)] < f.white;
    if (!ok) continue;
    quads += 1U;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = k >> 1;
      const int col = 2 * q + (k & 1);
      // white_balance_quad (isp_math.hpp:311-332): linearize, cap at 4, fixed point, bin
      float v = (static_cast<float>(raw[r][col]) - c.qc.black[k]) * c.qc.inv[k];
      v = v > 0.0F ? v : 0.0F;
      if (c.lsc) {
        float gain;
        if (table) {
          const F2 e = tl.lsc.e[seg[r]][r][t.cx - t.x0 + col];
          gain = e.v[0] + e.v[1] * tyr[r];
        } else {
          gain = lsc_gain(f, t.cx + col, 2 * t.qy + r, cfa_channel(f.cfa, k));
        }
        v = v * gain;
      }
      v = v < 4.0F ? v : 4.0F;
      pos[k] += to_fixed(v);
      if (c.patch) {
        const int b = static_cast<int>(v * static_cast<float>(kAwbBins));
        const int bin = c.colour[k] * kAwbBins + (b < kAwbBins - 1 ? b : kAwbBins - 1);
        atomicAdd(&fl.hist[bin * kAwbCopies + (threadIdx.x & (kAwbCopies - 1))], 1U);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) local[k] += pos[k];
  local[4] += quads;
}

// The tiles tile, tile + stride, ... below `end` of one frame. (Loading the
// next tile's chunks while one is summed measured no faster.)
template <int kFmt>
__device__ __forceinline__ void awb_tiles(const FrameDesc& f, const AwbCtx& c, AwbFrameLds& fl, AwbTileLds& tl,
                                          int tile, int end, int stride, int tiles_x, int fast_rows,
                                          unsigned long long local[kAwbAcc]) {
  for (; tile < end; tile += stride) {
    const AwbTask<kFmt> t = awb_task<kFmt>(f, tile, tiles_x, fast_rows);
    Packed<kFmt> pk[2];
    awb_load<kFmt>(f, t, pk);
    awb_tile<kFmt>(f, c, fl, tl, t, pk, local);
  }
}

// Adds the block's partial sums and histograms to the frame's scratch; returns
// the token of the atomics (see complete_grid()).
__device__ __forceinline__ unsigned int awb_flush(AwbScratch* a, AwbFrameLds& fl, const AwbCtx& c,
                                                  const unsigned long long local[kAwbAcc]) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  // The quad positions' sums by colour (R, G, B), and the quad count.
  unsigned long long sums[4] = {0, 0, 0, local[4]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    sums[0] += c.colour[k] == 0 ? local[k] : 0ULL;
    sums[1] += c.colour[k] == 1 ? local[k] : 0ULL;
    sums[2] += c.colour[k] == 2 ? local[k] : 0ULL;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned long long v = wave_sum(sums[k]);
    if ((tid & 63) == 0) fl.part[wave][k] = v;
  }
  __syncthreads();
  unsigned int token = 0;
  if (tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kThreads / 64; ++w) total += fl.part[w][tid];
    if (total != 0) token ^= fold(atomicAdd(&a->sum[blockIdx.x % kSlots].v[tid], total));
  }
  if (c.patch) {
    const int x = blockIdx.x % kSlots;
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) {
      unsigned int total = 0U;
#pragma unroll
      for (int k = 0; k < kAwbCopies; ++k) total += fl.hist[i * kAwbCopies + k];
      if (total != 0U) token ^= atomicAdd(&a->hist[x][i / kAwbBins][i % kAwbBins], total);
    }
  }
  return token;
}

// Gray-world levels from the sums (finish_white_balance, isp_math.hpp:366-370):
// the means in double, G over both green samples of a quad.
__device__ __forceinline__ void gray_world_levels(const unsigned long long sums[4], float level[3]) {
  const double quads = static_cast<double>(sums[3]);
  level[0] = static_cast<float>(static_cast<double>(sums[0]) / quads);
  level[1] = static_cast<float>(static_cast<double>(sums[1]) / (2.0 * quads));
  level[2] = static_cast<float>(static_cast<double>(sums[2]) / quads);
}

// The rest of finish_white_balance (isp_math.hpp:371-384): the clamped target
// gains, smoothed from the state `old` unless it is unset or smoothing is 1.
// False (state kept) when a level is not positive.
__device__ __forceinline__ bool wb_target(const float level[3], const float old_gains[3], int old_valid,
                                          float smoothing, float out[3]) {
  if (!(level[0] > 0.0F && level[1] > 0.0F && level[2] > 0.0F)) return false;
  float target[3] = {level[1] / level[0], 1.0F, level[1] / level[2]};
  for (int k = 0; k < 3; ++k) target[k] = target[k] < 0.0625F ? 0.0625F : target[k] > 16.0F ? 16.0F : target[k];
  const bool smooth = old_valid != 0 && smoothing < 1.0F;
  for (int k = 0; k < 3; ++k) out[k] = smooth ? old_gains[k] + smoothing * (target[k] - old_gains[k]) : target[k];
  return true;
}

// The frame's last contributor: finish_white_balance (isp_math.hpp:357-384)
// into the stream's white balance state; the reads clear the scratch.
__device__ __forceinline__ void awb_finish(const FrameDesc& f, DeviceScratch* s, AwbFrameLds& fl, const AwbCtx& c) {
  const int tid = threadIdx.x;
  if (tid < 4) fl.sums[tid] = take_slots(&s->awb.sum[0].v[tid], kLineSumsStride);
  if (c.patch) {
    for (int i = tid; i < 3 * kAwbBins; i += kThreads) {
      fl.hist[i] = take_slots(&s->awb.hist[0][i / kAwbBins][i % kAwbBins], 3 * kAwbBins);
    }
  }
  __syncthreads();
  const unsigned long long quads_s = fl.sums[3];
  if (quads_s != 0) {
    if (c.patch) {
      if (tid < 64) {
        for (int k = 0; k < 3; ++k) {
          const float level = percentile_wave(&fl.hist[k * kAwbBins], f.percentile);
          if (tid == 0) fl.level[k] = level;
        }
      }
    } else if (tid == 0) {
      const unsigned long long sums[4] = {fl.sums[0], fl.sums[1], fl.sums[2], fl.sums[3]};
      gray_world_levels(sums, fl.level);
    }
  }
  __syncthreads();
  if (tid == 0) {
    unsigned int token = 0;
    if (quads_s != 0) {
      const float level[3] = {fl.level[0], fl.level[1], fl.level[2]};
      // The state was last written by an earlier launch.
      WbState& st = s->wb;
      const float old[3] = {st.gains[0], st.gains[1], st.gains[2]};
      float g[3];
      if (wb_target(level, old, st.valid, f.smoothing, g)) {
        for (int k = 0; k < 3; ++k) token ^= coherent_write(&st.gains[k], g[k]);
        token ^= atomicExch(reinterpret_cast<unsigned int*>(&st.valid), 1U);
      }
    }
    __asm__ volatile("" : : "v"(token));  // the state is written before the counters clear
  }
}

template <int kFmt>
__global__ void __launch_bounds__(kThreads) awb_kernel(FrameDesc f, DeviceScratch* s, int tiles_x, int tiles,
                                                       int fast_rows) {
  __shared__ AwbFrameLds fl;
  __shared__ __attribute__((aligned(16))) AwbTileLds tl;
  AwbCtx c;
  awb_begin(f, fl, c);
  unsigned long long local[kAwbAcc] = {0, 0, 0, 0, 0};
  awb_tiles<kFmt>(f, c, fl, tl, blockIdx.x, tiles, gridDim.x, tiles_x, fast_rows, local);
  const unsigned int token = awb_flush(&s->awb, fl, c, local);
  // Gray world: the sums stay in the scratch, and the main pass that follows on
  // the stream turns them into gains (main_gray_world_gains); no completion here.
  if (!c.patch) return;
  if (!complete_grid(s->awb.done, &s->awb.top, token, &fl.flag)) return;
  awb_finish(f, s, fl, c);
  reset_grid(s->awb.done, &s->awb.top);
}

// ---- the main pass ----------------------------------------------------------

// The two demosaic formulas (isp_math.hpp:161-215) a CFA site needs, from a
// window w[row][col] whose centre is w[cy][cx]: at an R or B site green and the
// diagonal colour, at a G site the colour across and the colour down.
template <bool kMalvar, bool kZeroSign, int kRows, int kCols>
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
    // The reference's sums in its order, each product by 2 or 4 folded into its
    // sum (exact_fma); 4c + 2 axial1 is 2 (2c + axial1), its rounding the double
    // of that sum's (both exact when the sum is subnormal).
    if (rb) {
      const float axial1 = n12 + n32 + n21 + n23;
      const float axial2 = n02 + n42 + n20 + n24;
      *first = clamp_unit<kZeroSign>(exact_fma(2.0F, exact_fma(2.0F, c, axial1), -axial2) * 0.125F);
      *second = clamp_unit<kZeroSign>((exact_fma(2.0F, diagonal1, 6.0F * c) - 1.5F * axial2) * 0.125F);
    } else {
      const float horizontal2 = n20 + n24;
      *first = clamp_unit<kZeroSign>(
          (exact_fma(4.0F, n21 + n23, 5.0F * c) - horizontal2 - diagonal1 + 0.5F * vertical2) * 0.125F);
      *second = clamp_unit<kZeroSign>(
          (exact_fma(4.0F, n12 + n32, 5.0F * c) - vertical2 - diagonal1 + 0.5F * horizontal2) * 0.125F);
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

// tone (isp_math.hpp:235-243). The fraction is the reference's f - (float)i:
// v_fract_f32 gives the same for every f >= 0 but not for f = -0, which the
// colour matrix's clamp passes on (a row of negative weights on a black pixel)
// and which keeps its sign through a table whose first entry is -0. The table
// carries one entry more, lut[n1 + 1] = lut[n1] - 1: at x = 1 (the only x with
// i = n1, since x <= 1), t = +0 and b - a < 0, so a + t * (b - a) is lut[n1]
// plus -0, which is lut[n1] exactly, -0 included, as the reference returns it;
// below, the entries read are the reference's. No clamp of the index, no
// select. The table is read from HBM through the caches: its 16 KB stay in L1
// and L2, and staging it into each block's LDS measured slower (the copy at
// every block's start, and the random reads' bank conflicts in LDS).
// The entries are addressed by a 32-bit byte offset from the table's (uniform)
// base, which a load takes as is; a 64-bit element index would cost a sign
// extension and a 64-bit add per lookup.
template <bool kZeroSign>
__device__ __forceinline__ float tone(const float* lut, int n1, float x) {
  const float f = x * static_cast<float>(n1);
  const int i = static_cast<int>(f);
  // f >= 0 (or -0): f - (float)i is exact, and v_fract_f32 equals it but at -0.
  float t;
  if constexpr (kZeroSign) {
    t = f - static_cast<float>(i);
  } else {
    t = __builtin_amdgcn_fractf(f);
  }
  const float* e = reinterpret_cast<const float*>(reinterpret_cast<const char*>(lut) + (static_cast<unsigned>(i) << 2));
  const float a = e[0];
  const float b = e[1];
  return a + t * (b - a);
}

// tone of a thread's 4 x 3 values, its twelve table reads issued before any is
// used (each read waits on the caches; issued one pair at a time they would
// queue behind each other).
template <bool kZeroSign>
__device__ __forceinline__ void tone12(const float* lut, int n1, const float (&x)[4][3], float (&out)[4][3]) {
  // In two groups of six reads: all twelve in flight at once would hold 24
  // registers of results beside the pass's other live values.
#pragma unroll
  for (int g = 0; g < 12; g += 6) {
    float frac[6];
    unsigned int off[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const float f = x[(g + m) / 3][(g + m) % 3] * static_cast<float>(n1);
      const int i = static_cast<int>(f);
      if constexpr (kZeroSign) {
        frac[m] = f - static_cast<float>(i);
      } else {
        frac[m] = __builtin_amdgcn_fractf(f);
      }
      off[m] = static_cast<unsigned int>(i) << 2;
    }
    float a[6];
    float b[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const float* e = reinterpret_cast<const float*>(reinterpret_cast<const char*>(lut) + off[m]);
      a[m] = e[0];
      b[m] = e[1];
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) out[(g + m) / 3][(g + m) % 3] = a[m] + frac[m] * (b[m] - a[m]);
  }
}

// luma_code (isp_math.hpp:290-295) of output codes below 1024:
// (2126 R + 7152 G + 722 B + 5000) / 10000 in full-rate 24-bit multiplies. The
// sum n is below 10,235,001 < 2^24, and n / 10000 rounded down is the high
// word of n * ceil(2^35 / 10000) shifted right by 3, exactly for every such n
// (the error term n * 0.1632 / 2^35 stays below 1 / 10000).
__device__ __forceinline__ unsigned int luma_code(unsigned int r, unsigned int g, unsigned int b) {
  const unsigned int n = __umul24(r, 2126U) + __umul24(g, 7152U) + __umul24(b, 722U) + 5000U;
  unsigned int hi;
  __asm__("v_mul_hi_u32_u24 %0, 0x346dc6, %1" : "=v"(hi) : "v"(n));  // 0x346dc6 = ceil(2^35 / 10000)
  return hi >> 3;
}

template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock = kThreads>
struct MainGeom {
  static constexpr int kR = kMalvar ? 2 : 1;
  static constexpr int kH = kR + (kSharpen ? 1 : 0);
  static constexpr int kC = Chunk<kFmt>::kSamples;
  // Sample tile: LDS column = frame x - x0 + kPad, columns x0 - 4 .. x0 + 71,
  // rows y0 - kH .. y0 + kTileH + kH - 1; every 4-column group is 16-byte aligned.
  static constexpr int kPad = 4;
  static constexpr int kSW = 76;
  static constexpr int kSH = kTileH + 2 * kH;
  // Row chunks from x0 - kC cover the sample columns.
  static constexpr int kNC = (kSW - kPad + kC + kC - 1) / kC;
  static constexpr int kTasks = kSH * kNC;
  static constexpr int kTPT = (kTasks + kBlock - 1) / kBlock;
  // Ring of demosaiced RGB: ring column rc = frame x - x0 + 2, row rr = y - y0 + 1;
  // blocks of kRBW columns and 2 rows, one per thread (8 columns for 256-thread
  // blocks, 4 for 512).
  static constexpr int kRW = 72;
  static constexpr int kRH = kSharpen ? kTileH + 2 : 2;
  static constexpr int kRBW = kBlock >= 512 ? 4 : 8;
  static constexpr int kRingBlocks = (kRW / kRBW) * (kRH / 2);
  // Output rows per thread (a 4 x kRows3 block).
  static constexpr int kRows3 = kTileH / (kBlock / 16);
  static constexpr int kLscCols = kNC * kC;
  static_assert(kRingBlocks <= kBlock || !kSharpen, "one ring block per thread");
  static_assert(kTileH == 16 || kTileH == 32, "tile heights");
  static_assert(kRows3 >= 1, "a row per thread at least");
};

// The main pass's LDS histogram entries: 4 channels of up to 1024 bins, in as
// many interleaved copies as fit (bin-major, `bin * copies + lane % copies`):
// eight for 256 bins. Copies spread the atomics of lanes that share a code
// (flat image regions) over addresses; eight measured 2-7 % faster than four.
constexpr int kHistEntries = 8 * kMaxHistBins;

// LDS of the main pass: per frame (histograms, lens shading grid,
// reductions), and per tile (samples, demosaiced ring, lens shading table).
struct MainFrameLds {
  static constexpr bool kPlain = false;
  unsigned int hist[kHistEntries];
  float grid[kMaxGrid];
  unsigned long long part[8][4];  // per wave, up to 512 threads
  unsigned long long sums[4];
  float gains[3];
  int wb_update;  // gray world: the frame's gains replace the white balance state
  int flag;
};
// The same for a frame with nothing to finish (no histograms, statistics or
// device state): no histograms, which leaves room for more blocks a CU.
struct MainPlainFrameLds {
  static constexpr bool kPlain = true;
  unsigned int hist[1];  // never used
  float grid[kMaxGrid];
};
template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock = kThreads>
struct MainTileLds {
  using G = MainGeom<kFmt, kMalvar, kSharpen, kTileH, kBlock>;
  alignas(16) float samp[G::kSH * G::kSW];
  alignas(16) float ring[3][G::kRH * G::kRW];
  alignas(16) LscTable<G::kLscCols> lsc;
};

// A block's constants for one frame of the main pass.
struct MainCtx {
  QuadConst qc;
  float gains[3];
  const float* grid;
  int n1;
  int bins;
  int copies;      // interleaved histogram copies: 8 for 256 bins, 2 for 1024
  int copy_shift;  // log2(copies) + 2: a code's byte offset in its row is (code << copy_shift) | 4 copy
  bool lsc;
  bool want_hist;
  bool want_stats;
};

__device__ __forceinline__ void main_set_gains(const FrameDesc& f, const float gains[3], MainCtx& c) {
#pragma unroll
  for (int k = 0; k < 3; ++k) c.gains[k] = gains[k];
  c.qc = quad_consts(f, gains);
}

// Gray world in the per-frame form: the white balance kernel leaves its sums in
// the scratch, and each block of the main pass turns them into the frame's
// gains (finish_white_balance, isp_math.hpp:357-384) with the state an earlier
// frame left, all blocks alike; the last block stores the state and clears the
// sums (main_gray_world_store). No block changes either before every block has
// read them: the last block is known only once all have counted. Thread 0
// reads and computes while the other threads stage the block's tables; the
// result is read after a barrier.
__device__ __forceinline__ void main_gray_world_gains(const FrameDesc& f, const DeviceScratch* s, MainFrameLds& fl) {
  if (threadIdx.x != 0) return;
  unsigned long long sums[4] = {0, 0, 0, 0};
#pragma unroll
  for (int x = 0; x < kSlots; ++x) {
#pragma unroll
    for (int k = 0; k < 4; ++k) sums[k] += s->awb.sum[x].v[k];
  }
  const float old[3] = {s->wb.gains[0], s->wb.gains[1], s->wb.gains[2]};
  float g[3] = {old[0], old[1], old[2]};
  int update = 0;
  if (sums[3] != 0) {
    float level[3];
    gray_world_levels(sums, level);
    update = wb_target(level, old, s->wb.valid, f.smoothing, g) ? 1 : 0;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) fl.gains[k] = g[k];
  fl.wb_update = update;
}

// The last block of a gray-world main pass: the state for the next frame, and
// the white balance sums cleared.
__device__ __forceinline__ void main_gray_world_store(DeviceScratch* s, const MainFrameLds& fl) {
  const int tid = threadIdx.x;
  if (tid < 4 * kSlots) atomicExch(&s->awb.sum[tid >> 2].v[tid & 3], 0ULL);
  if (tid == 0 && fl.wb_update != 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) coherent_write(&s->wb.gains[k], fl.gains[k]);
    atomicExch(reinterpret_cast<unsigned int*>(&s->wb.valid), 1U);
  }
}

// Starts a block's work on a frame. The tile's first barrier orders these LDS
// writes before any reader.
template <int kBlock = kThreads, typename FL = MainFrameLds>
__device__ __forceinline__ void main_begin(const FrameDesc& f, const float gains[3], FL& fl, MainCtx& c) {
  const int tid = threadIdx.x;
  c.want_hist = !FL::kPlain && f.histogram != nullptr;
  c.want_stats = !FL::kPlain && (f.stats != nullptr || (f.ae_state != nullptr && f.ae_state->enabled != 0));
  c.bins = f.bins;
  c.copies = kHistEntries / (4 * c.bins);
  c.copy_shift = __builtin_ctz(static_cast<unsigned int>(c.copies)) + 2;
  main_set_gains(f, gains, c);
  if (c.want_hist) {
    for (int i = tid; i < kHistEntries; i += kBlock) fl.hist[i] = 0U;
  }
  c.n1 = f.lut_size - 1;
  c.lsc = f.lsc != nullptr;
  c.grid = stage_grid(f, fl.grid);
}

// Lens shading of a tile's samples: none, from the tile's table, or per sample
// (tiles whose rows span more than two grid rows).
constexpr int kLscNone = 0;
constexpr int kLscTable = 1;
constexpr int kLscSlow = 2;

// One row chunk's preprocessing constants: black level, 1 / range and colour
// gain of its even and odd columns, and where it lies.
struct PreRow {
  float b0, b1, i0, i1, g0, g1;
  int cx, y, par, ci;
  bool count;
};

// preprocess (isp_math.hpp:151-155) of one row chunk into sample-tile row
// `out`: (raw - black) / range, cut at 0, lens shading, white balance, cut at 1;
// saturated samples of the tile's own rows and columns counted. With the table,
// each group of 4 samples reads its entries in two 128-bit LDS loads, in
// straight-line code, so the loads are issued ahead of the arithmetic.
template <int kFmt, int kLsc, int kPad, int kSW>
__device__ __forceinline__ void pre_chunk(const FrameDesc& f, const PreRow& r, const int (&raw)[Chunk<kFmt>::kSamples],
                                          const F2* table, float tyr, float* out, unsigned int& saturated) {
  constexpr int kC = Chunk<kFmt>::kSamples;
  unsigned int sat = 0U;
#pragma unroll
  for (int h = 0; h < kC / 4; ++h) {
    const int col0 = kPad - kC + r.ci * kC + 4 * h;
    if (col0 < 0 || col0 + 4 > kSW) continue;
    float gain[4];
    if constexpr (kLsc == kLscTable) {
      const F4* e = reinterpret_cast<const F4*>(table) + 2 * h;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const F4 v = e[q];
        gain[2 * q] = v.v[0] + v.v[1] * tyr;
        gain[2 * q + 1] = v.v[2] + v.v[3] * tyr;
      }
    }
    F4 o;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int k = 4 * h + m;
      const bool odd = (m & 1) != 0;  // cx and col0 are even
      float v = (static_cast<float>(raw[k]) - (odd ? r.b1 : r.b0)) * (odd ? r.i1 : r.i0);
      v = v > 0.0F ? v : 0.0F;
      if constexpr (kLsc == kLscTable) {
        v = v * gain[m];
      } else if constexpr (kLsc == kLscSlow) {
        const int x = r.cx + k;
        const float g = x >= 0 && x < f.width ? lsc_gain(f, x, r.y, cfa_channel(f.cfa, (r.par << 1) | (x & 1))) : 1.0F;
        v = v * g;
      }
      v = v * (odd ? r.g1 : r.g0);
      v = v < 1.0F ? v : 1.0F;
      o.v[m] = v;
      sat += raw[k] >= f.white ? 1U : 0U;
    }
    *reinterpret_cast<F4*>(&out[col0]) = o;
  }
  // Only the tile's own samples count. Samples past the frame's right edge were
  // loaded as 0, below every white level (which exceeds the black level >= 0).
  saturated += r.count ? sat : 0U;
}

// The 8-bit histogram code of an RGB16 value (output_codes, isp_math.hpp:282-284):
// (v * (bins - 1) + 32767) / 65535 for v <= 65535 and bins - 1 <= 1023, in a
// 24-bit multiply and shifts: for every m < 2^26, m / 65535 rounded down is
// (m + (m >> 16) + 1) >> 16 (checked for every v at both bin counts).
__device__ __forceinline__ unsigned int code_of_rgb16(unsigned int v, unsigned int top) {
  const unsigned int m = __umul24(v, top) + 32767U;
  return (m + (m >> 16) + 1U) >> 16;
}

// Stage 2 of a tile with sharpening, for the Bayer layout kCfa (a compile-time
// channel per quad position): the demosaiced RGB of a kRBW x 2 block of the
// tile's ring (ring columns rc0.., rows rr0, rr0 + 1) into the three LDS planes
// (kPlane floats apart). Each site evaluates only its own two formulas and its
// RGB assignment is a renaming. The window of samples rows rr0 .. rr0 + 1 + 2 kR,
// columns rc0 .. rc0 + kRBW + 3 comes in 128-bit LDS loads.
template <bool kMalvar, bool kZeroSign, bool kWhole, int kCfa, int kR, int kRBW, int kRW, int kSW, int kPlane>
__device__ __forceinline__ void ring_block(const float* samp, float* ring, int rc0, int rr0) {
  constexpr int kWR = 2 + 2 * kR;
  constexpr int kWC = kRBW + 4;
  float w[kWR][kWC];
#pragma unroll
  for (int r = 0; r < kWR; ++r) {
#pragma unroll
    for (int q = 0; q < kWC / 4; ++q) {
      const float* p = &samp[(rr0 + r) * kSW + rc0 + 4 * q];
      if constexpr (kWhole) {
        const V4 a = lds_v4(p);
#pragma unroll
        for (int k = 0; k < 4; ++k) w[r][4 * q + k] = a[k];
      } else {
        const F4 a = *reinterpret_cast<const F4*>(p);
#pragma unroll
        for (int k = 0; k < 4; ++k) w[r][4 * q + k] = a.v[k];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    F4 planes[3][kRBW / 4];
#pragma unroll
    for (int i = 0; i < kRBW; ++i) {
      // ring pixel parity: x = x0 - 2 + rc0 + i, y = y0 - 1 + rr0 + j (x0, rc0, y0, rr0 even)
      const int ch = cfa_channel(kCfa, (((1 + j) & 1) << 1) | (i & 1));
      const bool rb = ch == ISP_CHANNEL_R || ch == ISP_CHANNEL_B;
      float first, second;
      demosaic_pair<kMalvar, kZeroSign>(w, kR + j, 2 + i, rb, &first, &second);
      float rgb[3];
      assign_rgb(ch, w[kR + j][2 + i], first, second, rgb);
#pragma unroll
      for (int k = 0; k < 3; ++k) planes[k][i / 4].v[i % 4] = rgb[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int q = 0; q < kRBW / 4; ++q) {
        *reinterpret_cast<F4*>(&ring[k * kPlane + (rr0 + j) * kRW + rc0 + 4 * q]) = planes[k][q];
      }
    }
  }
}

// ring_block for the Bayer code `cfa` (uniform).
template <bool kMalvar, bool kZeroSign, bool kWhole, int kR, int kRBW, int kRW, int kSW, int kPlane>
__device__ __forceinline__ void ring_block_any(int cfa, const float* samp, float* ring, int rc0, int rr0) {
  switch (cfa) {
    case kCfaCodes[ISP_RGGB]:
      ring_block<kMalvar, kZeroSign, kWhole, kCfaCodes[ISP_RGGB], kR, kRBW, kRW, kSW, kPlane>(samp, ring, rc0, rr0);
      break;
    case kCfaCodes[ISP_BGGR]:
      ring_block<kMalvar, kZeroSign, kWhole, kCfaCodes[ISP_BGGR], kR, kRBW, kRW, kSW, kPlane>(samp, ring, rc0, rr0);
      break;
    case kCfaCodes[ISP_GRBG]:
      ring_block<kMalvar, kZeroSign, kWhole, kCfaCodes[ISP_GRBG], kR, kRBW, kRW, kSW, kPlane>(samp, ring, rc0, rr0);
      break;
    default:
      ring_block<kMalvar, kZeroSign, kWhole, kCfaCodes[ISP_GBRG], kR, kRBW, kRW, kSW, kPlane>(samp, ring, rc0, rr0);
      break;
  }
}

// Demosaic of a thread's 4 x kRows3 output block without sharpening, from its
// sample window w (centre rows kR .., columns 4 ..), for the Bayer layout kCfa
// and output rows of parity kYPar + j.
template <bool kMalvar, bool kZeroSign, int kCfa, int kYPar, int kR, int kRows3, int kWR>
__device__ __forceinline__ void plain_block(const float (&w)[kWR][12], float (&rgb)[kRows3][4][3]) {
#pragma unroll
  for (int j = 0; j < kRows3; ++j) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = cfa_channel(kCfa, (((kYPar + j) & 1) << 1) | (i & 1));
      const bool rb = ch == ISP_CHANNEL_R || ch == ISP_CHANNEL_B;
      float first, second;
      demosaic_pair<kMalvar, kZeroSign>(w, kR + j, 4 + i, rb, &first, &second);
      assign_rgb(ch, w[kR + j][4 + i], first, second, rgb[j][i]);
    }
  }
}

template <bool kMalvar, bool kZeroSign, int kCfa, int kR, int kRows3, int kWR>
__device__ __forceinline__ void plain_block_par(int ypar, const float (&w)[kWR][12], float (&rgb)[kRows3][4][3]) {
  if (ypar != 0) {
    plain_block<kMalvar, kZeroSign, kCfa, 1, kR, kRows3>(w, rgb);
  } else {
    plain_block<kMalvar, kZeroSign, kCfa, 0, kR, kRows3>(w, rgb);
  }
}

// plain_block for the Bayer code `cfa` and the first row's parity `ypar` (both uniform).
template <bool kMalvar, bool kZeroSign, int kR, int kRows3, int kWR>
__device__ __forceinline__ void plain_block_any(int cfa, int ypar, const float (&w)[kWR][12],
                                                float (&rgb)[kRows3][4][3]) {
  switch (cfa) {
    case kCfaCodes[ISP_RGGB]:
      plain_block_par<kMalvar, kZeroSign, kCfaCodes[ISP_RGGB], kR, kRows3>(ypar, w, rgb);
      break;
    case kCfaCodes[ISP_BGGR]:
      plain_block_par<kMalvar, kZeroSign, kCfaCodes[ISP_BGGR], kR, kRows3>(ypar, w, rgb);
      break;
    case kCfaCodes[ISP_GRBG]:
      plain_block_par<kMalvar, kZeroSign, kCfaCodes[ISP_GRBG], kR, kRows3>(ypar, w, rgb);
      break;
    default:
      plain_block_par<kMalvar, kZeroSign, kCfaCodes[ISP_GBRG], kR, kRows3>(ypar, w, rgb);
      break;
  }
}

// The first output row (within the tile) of thread `tid`'s 4 x kRows3 block.
// With one row a thread and no sharpening, the rows are dealt so that every
// wave's four rows have one parity (row 8 (w / 2) + 2 (r % 4) + w % 2 for the
// r-th group of 16 threads, in wave w = r / 4): the demosaic's site types are
// then the same across a wave. Otherwise rows follow the threads in order.
template <bool kSharpen, int kRows3>
__device__ __forceinline__ int output_row(int tid) {
  if constexpr (!kSharpen && kRows3 == 1) {
    const int r = tid >> 4;
    const int w = r >> 2;
    return 8 * (w >> 1) + 2 * (r & 3) + (w & 1);
  } else {
    return kRows3 * (tid >> 4);
  }
}

// XCDs of the chip; blocks are dispatched to them in turn.
constexpr int kXcds = 8;
static_assert((kXcds & (kXcds - 1)) == 0, "a power of two");

// A block's share of `total` items when the grid's blocks split into one group
// per XCD (block b in group b % 8, the XCD it runs on) and each group sweeps
// one contiguous range: the block's first item, the end of its group's range,
// and the stride between its items. In unsigned 32-bit arithmetic, and with
// eight groups (every grid of eight blocks or more) the divisions are shifts:
// every wave runs this, and a 64-bit division is some 150 scalar instructions.
// total x 8 stays far below 2^32: frames are at most 16384 x 16384 samples
// (validate_raw), 2^17 tiles, and a batch holds at most 64 of them.
struct XcdShare {
  int first;
  int end;
  int stride;
};
__device__ __forceinline__ XcdShare xcd_share(int total) {
  const unsigned int g = gridDim.x;
  const unsigned int b = blockIdx.x;
  const unsigned int t = static_cast<unsigned int>(total);
  XcdShare s;
  if (g >= static_cast<unsigned int>(kXcds)) {
    constexpr unsigned int kShift = __builtin_ctz(kXcds);
    const unsigned int x = b & (kXcds - 1);
    s.first = static_cast<int>(((t * x) >> kShift) + (b >> kShift));
    s.end = static_cast<int>((t * (x + 1)) >> kShift);
    s.stride = static_cast<int>((g - x + kXcds - 1) >> kShift);
  } else {
    s.first = static_cast<int>(t * b / g);
    s.end = static_cast<int>(t * (b + 1) / g);
    s.stride = 1;
  }
  return s;
}

// Waves per SIMD the main pass's registers are held to: four (two 512-thread
// blocks a CU); six (three blocks) without sharpening for u8 and u16 samples,
// whose tile LDS (52 KB) lets three blocks in and whose bodies fit 80 VGPRs
// without spilling (the packed formats' unpacking would spill). Three blocks
// made the 1080p plain frame 8 % faster and the 4K one 7-9 %. The plain kernel
// (256-thread blocks of 16-row tiles, no histograms) runs at kPlainWaves: six
// (80 VGPRs, every format without spilling) or eight (64 VGPRs, all but RAW10).
template <int kFmt, bool kSharpen, bool kPlain = false, int kPlainWaves = 6>
constexpr int main_waves_per_eu() {
  return kPlain ? kPlainWaves : !kSharpen && (kFmt == ISP_RAW_U8 || kFmt == ISP_RAW_U16) ? 6 : 4;
}

// One kTileH x 64 output tile; statistics sums into acc[0..2], saturated
// samples into acc[3].
// kSpare: a per-frame kernel at four waves a SIMD, with registers to spare for
// the tone table's twelve reads in flight and whole 128-bit window loads (the
// batch kernels, with every tile body in one function, and the six-wave kernels
// have none).
template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock = kThreads, bool kZeroSign = false,
          bool kSpare = false, typename FL = MainFrameLds>
__device__ __forceinline__ void main_tile(const FrameDesc& f, const MainCtx& cx_, FL& fl,
                                          MainTileLds<kFmt, kMalvar, kSharpen, kTileH, kBlock>& tl, int tile,
                                          int tiles_x, int fast_rows, int out_aligned, unsigned long long acc[4]) {
  using G = MainGeom<kFmt, kMalvar, kSharpen, kTileH, kBlock>;
  constexpr int kR = G::kR;
  constexpr int kH = G::kH;
  constexpr int kC = G::kC;
  constexpr int kPad = G::kPad;
  constexpr int kSW = G::kSW;
  constexpr int kSH = G::kSH;
  constexpr int kNC = G::kNC;
  constexpr int kTasks = G::kTasks;
  constexpr int kTPT = G::kTPT;
  constexpr int kRW = G::kRW;
  constexpr int kRows3 = G::kRows3;
  constexpr int kLscCols = G::kLscCols;
  const MainCtx& c = cx_;
  float* samp = tl.samp;
  const int tid = threadIdx.x;
  const int W = f.width;
  const int H = f.height;
  const int bins = c.bins;
  const int copies = c.copies;
  const unsigned int copy4 = static_cast<unsigned int>((tid & 63) & (copies - 1)) << 2;  // this lane's copy, in bytes
  const int copy_shift = c.copy_shift;
  const int tyi = tile / tiles_x;
  const int x0 = (tile - tyi * tiles_x) * 64;
  const int y0 = tyi * kTileH;
  const int sy0 = y0 - kH;  // frame row of sample row 0
  const int cx0 = x0 - kC;  // frame column of chunk 0
  const bool interior_x = cx0 >= 0 && cx0 + kNC * kC <= W;

  // 1a. Issue this thread's raw chunk loads.
  Packed<kFmt> pk[kTPT];
#pragma unroll
  for (int t = 0; t < kTPT; ++t) {
    const int task = tid + t * kBlock;
    if (task >= kTasks) break;
    const int sr = task / kNC;
    const int ci = task - sr * kNC;
    const int cx = cx0 + ci * kC;
    if (fast_rows != 0 && cx >= 0 && cx + kC <= W) {
      const int y = mirror_near(sy0 + sr, H);
      load_packed<kFmt>(f.raw + static_cast<long long>(y) * f.raw_pitch, cx, pk[t]);
    }
  }
  // The previous tile's readers are done before this tile writes LDS: with
  // lens shading the table's barrier below comes first; with sharpening the
  // last readers of the sample tile (the ring) passed a barrier already, and
  // the ring's readers pass the one after this tile's preprocessing.
  if (!kSharpen && !c.lsc) __syncthreads();
  int lo = 0;
  bool table = false;
  if (c.lsc) {
    lo = lsc_row(f, sy0 > 0 ? sy0 : 0);
    table = lsc_row(f, sy0 + kSH - 1 < H - 1 ? sy0 + kSH - 1 : H - 1) - lo <= 1;
    // The same for every thread: made uniform so that the per-sample choice
    // below is a scalar branch.
    table = __builtin_amdgcn_readfirstlane(static_cast<int>(table)) != 0;
    if (table) build_lsc_table<kLscCols>(f, c.grid, cx0, lo, tl.lsc);
    __syncthreads();
  }

  // 1b. Unpack and preprocess (isp_math.hpp:151-155) into the sample tile.
  unsigned int saturated = 0;
#pragma unroll
  for (int t = 0; t < kTPT; ++t) {
    const int task = tid + t * kBlock;
    if (task >= kTasks) break;
    const int sr = task / kNC;
    const int ci = task - sr * kNC;
    const int cx = cx0 + ci * kC;
    const int yv = sy0 + sr;
    const int y = mirror_near(yv, H);
    int raw[kC];
    if (fast_rows != 0 && cx >= 0 && cx + kC <= W) {
      unpack<kFmt>(pk[t], raw);
    } else {
      const std::uint8_t* row = f.raw + static_cast<long long>(y) * f.raw_pitch;
#pragma unroll
      for (int k = 0; k < kC; ++k) raw[k] = cx + k >= 0 && cx + k < W ? load_raw(row, kFmt, cx + k) : 0;
    }
    const int par = y & 1;
    const float b0 = par ? c.qc.black[2] : c.qc.black[0];
    const float b1 = par ? c.qc.black[3] : c.qc.black[1];
    const float i0 = par ? c.qc.inv[2] : c.qc.inv[0];
    const float i1 = par ? c.qc.inv[3] : c.qc.inv[1];
    const float g0 = par ? c.qc.gain[2] : c.qc.gain[0];
    const float g1 = par ? c.qc.gain[3] : c.qc.gain[1];
    const bool count = c.want_stats && yv >= y0 && yv < y0 + kTileH && yv < H && ci >= 1 && ci <= 64 / kC;
    const PreRow row{b0, b1, i0, i1, g0, g1, cx, y, par, ci, count};
    float* out = &samp[sr * kSW];
    if (!c.lsc) {
      pre_chunk<kFmt, kLscNone, kPad, kSW>(f, row, raw, nullptr, 0.0F, out, saturated);
    } else if (table) {
      float tyr;
      const int seg = lsc_row(f, y, &tyr) - lo;
      pre_chunk<kFmt, kLscTable, kPad, kSW>(f, row, raw, &tl.lsc.e[seg][par][ci * kC], tyr, out, saturated);
    } else {
      pre_chunk<kFmt, kLscSlow, kPad, kSW>(f, row, raw, nullptr, 0.0F, out, saturated);
    }
  }
  acc[3] += saturated;
  if (!interior_x) {
    // Columns outside the frame take the sample of their reflection.
    __syncthreads();
    for (int i = tid; i < kSH * kSW; i += kBlock) {
      const int sr = i / kSW;
      const int col = i - sr * kSW;
      const int x = x0 - kPad + col;
      if (x >= 0 && x < W) continue;
      const int src = mirror_near(x, W) - x0 + kPad;
      samp[i] = src >= 0 && src < kSW ? samp[sr * kSW + src] : 0.0F;
    }
  }
  __syncthreads();

  // 2. With sharpening: the demosaiced RGB of the tile and a one-pixel ring
  // (ring column rc = x - x0 + 2, row rr = y - y0 + 1), kRBW x 2 pixels per thread.
  if constexpr (kSharpen) {
    if (tid < G::kRingBlocks) {
      constexpr int kRBW = G::kRBW;
      const int by = tid / (kRW / kRBW);
      const int bx = tid - by * (kRW / kRBW);
      ring_block_any<kMalvar, kZeroSign, kSpare, kR, kRBW, kRW, kSW, G::kRH * kRW>(__builtin_amdgcn_readfirstlane(f.cfa), samp,
                                                                 &tl.ring[0][0], kRBW * bx, 2 * by);
    }
    __syncthreads();
  }

  // 3. A 4 x kRows3 block of output pixels per thread.
  const int px = 4 * (tid & 15);
  const int py = output_row<kSharpen, kRows3>(tid);
  const int xo = x0 + px;
  float rgb[kRows3][4][3];
  if constexpr (kSharpen) {
    // sharpen (isp_math.hpp:219-227) over ring rows py .. py + kRows3 + 1, columns px .. px + 7
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float m[kRows3 + 2][8];
#pragma unroll
      for (int r = 0; r < kRows3 + 2; ++r) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const float* p = &tl.ring[k][(py + r) * kRW + px + 4 * q];
          if constexpr (kSpare) {
            const V4 a = lds_v4(p);
#pragma unroll
            for (int e = 0; e < 4; ++e) m[r][4 * q + e] = a[e];
          } else {
            const F4 a = *reinterpret_cast<const F4*>(p);
#pragma unroll
            for (int e = 0; e < 4; ++e) m[r][4 * q + e] = a.v[e];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kRows3; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float cc = m[j + 1][i + 2];
          // (4 cc + 2 axial + diagonal) / 16 with the powers of two folded:
          // 4 cc + 2 axial is 2 (2 cc + axial), exactly.
          const float axial = m[j][i + 2] + m[j + 2][i + 2] + m[j + 1][i + 1] + m[j + 1][i + 3];
          const float diagonal = m[j][i + 1] + m[j][i + 3] + m[j + 2][i + 1] + m[j + 2][i + 3];
          const float blur = exact_fma(2.0F, exact_fma(2.0F, cc, axial), diagonal) * 0.0625F;
          rgb[j][i][k] = clamp_unit<kZeroSign>(cc + f.sharpen * (cc - blur));
        }
      }
    }
  } else {
    // demosaic from samples rows py + kH - kR .. py + kRows3 - 1 + kH + kR, columns px .. px + 11
    constexpr int kWR = kRows3 + 2 * kR;
    float w[kWR][12];
#pragma unroll
    for (int r = 0; r < kWR; ++r) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float* p = &samp[(py + kH - kR + r) * kSW + px + 4 * q];
        if constexpr (kSpare) {
          const V4 a = lds_v4(p);
#pragma unroll
          for (int k = 0; k < 4; ++k) w[r][4 * q + k] = a[k];
        } else {
          const F4 a = *reinterpret_cast<const F4*>(p);
#pragma unroll
          for (int k = 0; k < 4; ++k) w[r][4 * q + k] = a.v[k];
        }
      }
    }
    // y0 is even, and the row parity of py is the same across a wave (output_row).
    plain_block_any<kMalvar, kZeroSign, kR, kRows3>(__builtin_amdgcn_readfirstlane(f.cfa),
                                         __builtin_amdgcn_readfirstlane(py & 1), w, rgb);
  }
  const bool full_x = xo + 4 <= W;
#pragma unroll
  for (int j = 0; j < kRows3; ++j) {
    const int yo = y0 + py + j;
    if (yo >= H) break;
    // correct_color, to_fixed sums, tone, store, output codes (isp_math.hpp:229-305)
    float lin[4][3];
    float t[4][3];
    unsigned int q[4][3];
    unsigned int sums[3] = {0U, 0U, 0U};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lin[i][k] = clamp_unit<kZeroSign>(f.ccm[3 * k] * rgb[j][i][0] + f.ccm[3 * k + 1] * rgb[j][i][1] +
                            f.ccm[3 * k + 2] * rgb[j][i][2]);
      }
      if (c.want_stats && (full_x || xo + i < W)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) sums[k] += to_fixed(lin[i][k]);
      }
    }
    if constexpr (kSpare) {
      tone12<kZeroSign>(f.lut, c.n1, lin, t);
    } else {
#pragma unroll
      for (int m = 0; m < 12; ++m) t[m / 3][m % 3] = tone<kZeroSign>(f.lut, c.n1, lin[m / 3][m % 3]);
    }
    if (c.want_stats) {
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += sums[k];
    }
    std::uint8_t* orow = f.out + static_cast<long long>(yo) * f.out_pitch;
    // Whole groups go out as one nontemporal vector store a row: the image is
    // written once and read by the next stage, not here, and streaming it past
    // the caches keeps the L2 for the raw halos (4K main pass 1.7 us faster).
    const bool whole = out_aligned != 0 && full_x;
    switch (f.out_format) {
      case ISP_OUT_RGB8: {
        unsigned int byte[12];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const unsigned int v = static_cast<unsigned char>(t[i][k] * 255.0F + 0.5F);
            q[i][k] = v;
            byte[3 * i + k] = v;
          }
        }
        if (whole) {
          U3 o;
#pragma unroll
          for (int w = 0; w < 3; ++w) {
            o[w] = byte[4 * w] | (byte[4 * w + 1] << 8) | (byte[4 * w + 2] << 16) | (byte[4 * w + 3] << 24);
          }
          __builtin_nontemporal_store(o, reinterpret_cast<U3*>(orow + 3 * xo));
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (xo + i >= W) continue;
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
            q[i][k] = v;
            word[i] |= v << (8 * k);
          }
        }
        if (whole) {
          __builtin_nontemporal_store(U4{word[0], word[1], word[2], word[3]}, reinterpret_cast<U4*>(orow + 4 * xo));
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (xo + i >= W) continue;
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
            q[i][k] = code_of_rgb16(v, static_cast<unsigned int>(bins - 1));
          }
        }
        if (whole) {
#pragma unroll
          for (int w = 0; w < 3; ++w) {
            __builtin_nontemporal_store(U2{half[4 * w] | (half[4 * w + 1] << 16), half[4 * w + 2] | (half[4 * w + 3] << 16)},
                                        reinterpret_cast<U2*>(orow + 6 * xo + 8 * w));
          }
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (xo + i >= W) continue;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              orow[6 * (xo + i) + 2 * k] = static_cast<std::uint8_t>(half[3 * i + k] & 0xFFU);
              orow[6 * (xo + i) + 2 * k + 1] = static_cast<std::uint8_t>(half[3 * i + k] >> 8);
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
            q[i][k] = static_cast<unsigned int>(clamp01(t[i][k]) * static_cast<float>(bins - 1) + 0.5F);
          }
        }
        if (whole) {
#pragma unroll
          for (int w = 0; w < 3; ++w) {
            V4 o;
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = t[(4 * w + k) / 3][(4 * w + k) % 3];
            __builtin_nontemporal_store(o, reinterpret_cast<V4*>(orow + 12 * xo + 16 * w));
          }
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (xo + i >= W) continue;
#pragma unroll
            for (int k = 0; k < 3; ++k) reinterpret_cast<float*>(orow)[3 * (xo + i) + k] = t[i][k];
          }
        }
        break;
      }
    }
    if (c.want_hist) {
      // Row r of the histograms starts at entry r * kHistEntries / 4 (bins x
      // copies is kHistEntries / 4 for either bin count): a code's entry is a
      // shift and an or from the row's base, which the LDS instruction's offset holds.
      unsigned char* h = reinterpret_cast<unsigned char*>(fl.hist);
      constexpr int kRowBytes = kHistEntries;  // kHistEntries / 4 entries of 4 bytes
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (!full_x && xo + i >= W) continue;
        const unsigned int luma = luma_code(q[i][0], q[i][1], q[i][2]);
        atomicAdd(reinterpret_cast<unsigned int*>(h + ((luma << copy_shift) | copy4)), 1U);
        atomicAdd(reinterpret_cast<unsigned int*>(h + kRowBytes + ((q[i][0] << copy_shift) | copy4)), 1U);
        atomicAdd(reinterpret_cast<unsigned int*>(h + 2 * kRowBytes + ((q[i][1] << copy_shift) | copy4)), 1U);
        atomicAdd(reinterpret_cast<unsigned int*>(h + 3 * kRowBytes + ((q[i][2] << copy_shift) | copy4)), 1U);
      }
    }
  }
}

// Adds the block's sums and histograms to the frame's scratch; returns the
// token of the atomics (see complete_grid()).
template <int kBlock = kThreads>
__device__ __forceinline__ unsigned int main_flush(MainScratch* m, MainFrameLds& fl, const MainCtx& c,
                                                   const unsigned long long acc[4]) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  if (c.want_stats) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned long long v = wave_sum(acc[k]);
      if ((tid & 63) == 0) fl.part[wave][k] = v;
    }
  }
  __syncthreads();
  unsigned int token = 0;
  if (c.want_stats && tid < 4) {
    unsigned long long total = 0;
    for (int w = 0; w < kBlock / 64; ++w) total += fl.part[w][tid];
    if (total != 0) token ^= fold(atomicAdd(&m->sum[blockIdx.x % kSlots].v[tid], total));
  }
  if (c.want_hist) {
    const int bins = c.bins;
    const int x = blockIdx.x % kSlots;
    for (int i = tid; i < 4 * bins; i += kBlock) {
      unsigned int total = 0;
      for (int k = 0; k < c.copies; ++k) total += fl.hist[i * c.copies + k];
      if (total != 0U) token ^= atomicAdd(&m->hist[x][i / bins][i % bins], total);
    }
  }
  return token;
}

// The frame's last contributor: the histograms and the statistics vector
// (finish_stats, isp_math.hpp:387-425); the reads clear the scratch. With a
// device exposure state, the frame runs at the state's exposure and leaves its
// suggestion there for the stream's next frame. The first wave runs the serial
// epilogue while the other waves gather the histograms.
template <int kBlock = kThreads>
__device__ __forceinline__ void main_finish(const FrameDesc& f, MainScratch* m, MainFrameLds& fl, const MainCtx& c) {
  const int tid = threadIdx.x;
  if (tid < 4) fl.sums[tid] = take_slots(&m->sum[0].v[tid], kLineSumsStride);
  __syncthreads();
  if (c.want_hist && tid >= 64) {
    const int bins = c.bins;
    for (int i = tid - 64; i < 4 * bins; i += kBlock - 64) {
      f.histogram[i] = take_slots(&m->hist[0][i / bins][i % bins], 4 * kMaxHistBins);
    }
  }
  if (c.want_stats && tid == 0) {
    const unsigned long long sums[4] = {fl.sums[0], fl.sums[1], fl.sums[2], fl.sums[3]};
    long long exposure_ns = f.exposure_ns;
    float exposure_gain = f.exposure_gain;
    AeState* ae = f.ae_state != nullptr && f.ae_state->enabled != 0 ? f.ae_state : nullptr;
    if (ae != nullptr && ae->valid != 0) {
      exposure_ns = ae->exposure_ns;
      exposure_gain = ae->gain;
    }
    float out[ISP_STATS_SIZE];
    finish_stats(sums, static_cast<long long>(f.width) * f.height, c.gains, f.ae, exposure_ns, exposure_gain, out);
    if (f.stats != nullptr) {
      for (int k = 0; k < ISP_STATS_SIZE; ++k) f.stats[k] = out[k];
    }
    if (ae != nullptr && out[ISP_STAT_SUGGESTED_EXPOSURE_NS] > 0.0F) {
      ae->exposure_ns = static_cast<long long>(static_cast<double>(out[ISP_STAT_SUGGESTED_EXPOSURE_NS]) + 0.5);
      ae->gain = out[ISP_STAT_SUGGESTED_GAIN];
      ae->valid = 1;
    }
  }
}


// kBlock threads per block: 512-thread blocks (one output row per thread) put
// twice the waves on a CU for the same LDS, with registers held to what
// main_waves_per_eu() leaves.
//
// kPlain: a frame with nothing to finish (no histograms, statistics, exposure
// state or gray-world state; the host decides), in 256-thread blocks of 16-row
// tiles without the histograms' LDS: more, smaller blocks a CU, so that a small
// frame's tiles spread over the chip in fewer rounds.
template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock, bool kZeroSign, bool kPlain = false,
          int kPlainWaves = 6>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(kBlock >= 512 || kPlain ? main_waves_per_eu<kFmt, kSharpen, kPlain, kPlainWaves>()
                                                           : 1)))
main_kernel(FrameDesc f, DeviceScratch* s, int tiles_x, int tiles, int fast_rows, int out_aligned) {
  using FL = std::conditional_t<kPlain, MainPlainFrameLds, MainFrameLds>;
  __shared__ FL fl;
  __shared__ MainTileLds<kFmt, kMalvar, kSharpen, kTileH, kBlock> tl;
  const bool gray_world = !kPlain && f.wb_mode == ISP_WB_GRAY_WORLD;
  float gains[3];
  if (f.wb_mode != ISP_WB_MANUAL && !gray_world) {
#pragma unroll
    for (int k = 0; k < 3; ++k) gains[k] = s->wb.gains[k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) gains[k] = f.wb[k];  // gray world: replaced below
  }
  if constexpr (!kPlain) {
    if (gray_world) main_gray_world_gains(f, s, fl);
  }
  MainCtx c;
  main_begin<kBlock>(f, gains, fl, c);
  __syncthreads();  // the staged grid (and gray world's gains) before any use
  if constexpr (!kPlain) {
    if (gray_world) {
      const float g[3] = {fl.gains[0], fl.gains[1], fl.gains[2]};
      main_set_gains(f, g, c);
    }
  }
  unsigned long long acc[4] = {0, 0, 0, 0};
  // Tiles by XCD: blocks are dispatched to the eight XCDs in turn (block b to
  // XCD b % 8), and each XCD's blocks sweep one contiguous band of tiles, so
  // that neighbouring tiles, whose halos overlap, are read through one L2.
  const XcdShare share = xcd_share(tiles);
  for (int tile = share.first; tile < share.end; tile += share.stride) {
    main_tile<kFmt, kMalvar, kSharpen, kTileH, kBlock, kZeroSign,
              main_waves_per_eu<kFmt, kSharpen, kPlain, kPlainWaves>() <= 4>(
        f, c, fl, tl, tile, tiles_x, fast_rows, out_aligned, acc);
  }
  // A frame with no histograms, statistics, exposure state or gray-world state
  // to finish has nothing to count: its blocks end with their tiles (the
  // completion chain's serial device atomics take several microseconds, most
  // of a small frame's time).
  if constexpr (!kPlain) {
    if (!c.want_stats && !c.want_hist && !gray_world) return;
    const unsigned int token = main_flush<kBlock>(&s->main, fl, c, acc);
    if (!complete_grid(s->main.done, &s->main.top, token, &fl.flag)) return;
    main_finish<kBlock>(f, &s->main, fl, c);
    if (gray_world) main_gray_world_store(s, fl);
    reset_grid(s->main.done, &s->main.top);
  }
}

// ---- the batched form ----------------------------------------------------------

template <typename A, typename... B>
constexpr std::size_t max_size() {
  if constexpr (sizeof...(B) == 0) {
    return sizeof(A);
  } else {
    return sizeof(A) > max_size<B...>() ? sizeof(A) : max_size<B...>();
  }
}

// The largest tile LDS of every body a main-pass batch kernel runs: with
// sharpening, 32-row tiles dominate (RAW10's 16-sample chunks widen its lens
// shading table); without it the LDS is small enough for 3 blocks per CU.
template <bool kSharpen, int kBlock>
constexpr std::size_t batch_tile_lds() {
  if constexpr (kSharpen) {
    return max_size<MainTileLds<ISP_RAW_U8, true, true, 32, kBlock>, MainTileLds<ISP_RAW_U16, true, true, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW10, true, true, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW12, true, true, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW10, false, true, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW12, false, true, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW10, true, false, 32, kBlock>>();
  } else {
    return max_size<MainTileLds<ISP_RAW_U8, true, false, 32, kBlock>,
                    MainTileLds<ISP_RAW_U16, true, false, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW10, true, false, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW12, true, false, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW10, false, false, 32, kBlock>,
                    MainTileLds<ISP_RAW_RAW12, false, false, 32, kBlock>>();
  }
}

constexpr int variant_of(int fmt, bool malvar, bool sharpen, int tile_h) {
  return fmt | (malvar ? 4 : 0) | (sharpen ? 8 : 0) | (tile_h == 16 ? 16 : 0);
}

template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock>
__device__ __forceinline__ void batch_main_item(const BatchFrame& b, const MainCtx& c, MainFrameLds& fl,
                                                unsigned char* tile_lds, int item, unsigned long long acc[4]) {
  auto& tl = *reinterpret_cast<MainTileLds<kFmt, kMalvar, kSharpen, kTileH, kBlock>*>(tile_lds);
  const int first = item * b.tiles_per_item;
  const int end = min(first + b.tiles_per_item, b.tiles);
  for (int t = first; t < end; ++t) {
    main_tile<kFmt, kMalvar, kSharpen, kTileH, kBlock>(b.f, c, fl, tl, t, b.tiles_x, b.fast_rows, b.out_aligned,
                                                       acc);
  }
}

template <int kFmt>
__device__ __forceinline__ void batch_awb_item(const BatchFrame& b, const AwbCtx& c, AwbFrameLds& fl,
                                               AwbTileLds& tl, int item, unsigned long long acc[kAwbAcc]) {
  const int first = item * b.awb_tiles_per_item;
  const int end = min(first + b.awb_tiles_per_item, b.awb_tiles);
  awb_tiles<kFmt>(b.f, c, fl, tl, first, end, 1, b.awb_tiles_x, b.fast_rows, acc);
}

// The frame (in the pass's order) that holds item `item` of a pass, by one
// coalesced load of the pass's start table per wave; returns its index in the
// table and the item within the frame.
__device__ __forceinline__ int batch_locate(const int* start, const int* order, int count, int item, int* within) {
  const int lane = threadIdx.x & 63;
  const int s = lane < count ? start[lane] : INT_MAX;
  const unsigned long long mask = __ballot(s <= item);
  const int k = 63 - __clzll(static_cast<long long>(mask));
  const int ks = __builtin_amdgcn_readfirstlane(k);
  *within = item - start[ks];
  return order[ks];
}

// The work of one pass over a batch, by the table's queue mode:
//  dynamic     every block's first item is its own (blockIdx.x); later ones come
//              from an atomic counter, fetched while the current item runs, and
//              the last block to leave resets the counter for the next launch
//              (so a captured launch replays without a memset);
//  static      block b takes items b, b + grid, ... (no atomics);
//  static XCD  the items split into one contiguous range per XCD (blocks are
//              dispatched to the 8 XCDs in turn, block b to XCD b % 8), each
//              range taken in turn by that XCD's blocks, so neighbouring tiles,
//              whose halos overlap, meet in the same L2.

struct BatchQueue {
  unsigned int* next;
  unsigned int* exited;
  int total;
  int mode;   // ISP_BATCH_QUEUE_*
  int* slot;  // block-shared
  unsigned int pending = 0;
  int step = 0;  // static XCD: items between this block's turns, and its range's end
  int end = 0;

  __device__ __forceinline__ int first() {
    if (mode == ISP_BATCH_QUEUE_STATIC) return static_cast<int>(blockIdx.x);
    if (mode == ISP_BATCH_QUEUE_STATIC_XCD) {
      const XcdShare share = xcd_share(total);
      step = share.stride;
      end = share.end;
      return share.first < end ? share.first : total;
    }
    if (threadIdx.x == 0) *slot = static_cast<int>(blockIdx.x);
    __syncthreads();
    const int item = *slot;
    __syncthreads();
    return item;
  }
  // Dynamic: fetches the item after the current one (thread 0; read by next_item()).
  __device__ __forceinline__ void prefetch() {
    if (mode == ISP_BATCH_QUEUE_DYNAMIC && threadIdx.x == 0) pending = gridDim.x + atomicAdd(next, 1U);
  }
  __device__ __forceinline__ int next_item(int item) {
    if (mode == ISP_BATCH_QUEUE_STATIC) return item + static_cast<int>(gridDim.x);
    if (mode == ISP_BATCH_QUEUE_STATIC_XCD) return item + step < end ? item + step : total;
    if (threadIdx.x == 0) *slot = static_cast<int>(pending);
    __syncthreads();
    const int next_one = *slot;
    __syncthreads();
    return next_one;
  }
  __device__ __forceinline__ void leave() {
    if (mode == ISP_BATCH_QUEUE_DYNAMIC && threadIdx.x == 0) {
      const unsigned int ticket = atomicAdd(exited, 1U);
      if (ticket == gridDim.x - 1) {
        atomicExch(next, 0U);
        atomicExch(exited, 0U);
      }
    }
  }
};

// White balance statistics of every auto-balanced frame of a batch: the items
// of all frames in one persistent grid; each frame's last contributor turns
// its sums into the stream's gains, which the main pass (the next launch) reads.
__global__ void __launch_bounds__(kThreads) batch_awb_kernel(const BatchTable* __restrict__ table, BatchControl* ctl) {
  __shared__ AwbFrameLds fl;
  __shared__ __attribute__((aligned(16))) AwbTileLds tl;
  __shared__ int slot;
  BatchQueue q{&ctl->awb_next, &ctl->awb_exited, table->awb_items, table->queue, &slot};
  const int count = table->count;
  int cur = -1;
  unsigned int items_done = 0;
  unsigned long long acc[kAwbAcc] = {0, 0, 0, 0, 0};
  AwbCtx ac;
  for (int item = q.first(); item < q.total; item = q.next_item(item)) {
    q.prefetch();
    int local = 0;
    const int f = batch_locate(table->awb_start, table->awb_order, count, item, &local);
    const BatchFrame& b = table->frames[f];
    if (f != cur) {
      if (cur >= 0) {
        const BatchFrame& pb = table->frames[cur];
        const unsigned int token = awb_flush(&pb.s->awb, fl, ac, acc);
        if (complete_items(&pb.s->awb.done[0].n, items_done, static_cast<unsigned int>(pb.awb_items), token,
                           &fl.flag)) {
          awb_finish(pb.f, pb.s, fl, ac);
          if (threadIdx.x == 0) atomicExch(&pb.s->awb.done[0].n, 0U);
        }
        __syncthreads();  // the epilogue is done with the frame LDS
#pragma unroll
        for (int k = 0; k < kAwbAcc; ++k) acc[k] = 0;
        items_done = 0;
      }
      cur = f;
      awb_begin(b.f, fl, ac);
    }
    switch (b.f.raw_format) {
      case ISP_RAW_U8:
        batch_awb_item<ISP_RAW_U8>(b, ac, fl, tl, local, acc);
        break;
      case ISP_RAW_U16:
        batch_awb_item<ISP_RAW_U16>(b, ac, fl, tl, local, acc);
        break;
      case ISP_RAW_RAW10:
        batch_awb_item<ISP_RAW_RAW10>(b, ac, fl, tl, local, acc);
        break;
      default:
        batch_awb_item<ISP_RAW_RAW12>(b, ac, fl, tl, local, acc);
        break;
    }
    ++items_done;
  }
  if (cur >= 0) {
    const BatchFrame& pb = table->frames[cur];
    const unsigned int token = awb_flush(&pb.s->awb, fl, ac, acc);
    if (complete_items(&pb.s->awb.done[0].n, items_done, static_cast<unsigned int>(pb.awb_items), token, &fl.flag)) {
      awb_finish(pb.f, pb.s, fl, ac);
      if (threadIdx.x == 0) atomicExch(&pb.s->awb.done[0].n, 0U);
    }
  }
  q.leave();
}

#define ISP_BATCH_MAIN_CASE(FMT, MAL, SHP, TH)                                        \
  case variant_of(FMT, MAL, SHP, TH):                                                 \
    if constexpr ((kSharpen || !SHP) && (kBlock == kThreads || TH == 32)) {           \
      batch_main_item<FMT, MAL, SHP, TH, kBlock>(b, mc, fl, tile_lds, local, acc);    \
    }                                                                                 \
    break;
#define ISP_BATCH_MAIN_CASES(FMT)            \
  ISP_BATCH_MAIN_CASE(FMT, false, false, 32) \
  ISP_BATCH_MAIN_CASE(FMT, false, true, 32)  \
  ISP_BATCH_MAIN_CASE(FMT, true, false, 32)  \
  ISP_BATCH_MAIN_CASE(FMT, true, true, 32)   \
  ISP_BATCH_MAIN_CASE(FMT, false, false, 16) \
  ISP_BATCH_MAIN_CASE(FMT, false, true, 16)  \
  ISP_BATCH_MAIN_CASE(FMT, true, false, 16)  \
  ISP_BATCH_MAIN_CASE(FMT, true, true, 16)

// The main pass of every frame of a batch in one persistent grid. kSharpen: the
// batch has frames with sharpening (their tile bodies need the larger LDS); a
// batch without them runs a lighter kernel. kBlock: 512-thread blocks (16 waves
// per CU, as the per-frame kernel) run 32-row tiles only; a batch with 16-row
// tiles runs the 256-thread kernels.
// kOnly >= 0: every frame of the batch runs the tile body of variant kOnly
// (32-row tiles), which is then the only one compiled in: no switch over the
// variants, and the registers and LDS of that body alone.
template <bool kSharpen, int kBlock, int kOnly = -1>
__device__ __forceinline__ void batch_main_body(const BatchTable* __restrict__ table, BatchControl* ctl) {
  constexpr int kOnlyFmt = kOnly & 3;
  constexpr bool kOnlyMalvar = (kOnly & 4) != 0;
  __shared__ MainFrameLds fl;
  __shared__ __attribute__((aligned(16))) unsigned char
      tile_lds[kOnly >= 0 ? sizeof(MainTileLds<kOnlyFmt, kOnlyMalvar, kSharpen, 32, kBlock>)
                          : batch_tile_lds<kSharpen, kBlock>()];
  __shared__ int slot;
  BatchQueue q{&ctl->main_next, &ctl->main_exited, table->main_items, table->queue, &slot};
  const int count = table->count;
  int cur = -1;
  unsigned int items_done = 0;
  unsigned long long acc[4] = {0, 0, 0, 0};
  // The frame's constants stay in LDS, read where the tile bodies use them:
  // held in registers across the switch over all variants they would crowd
  // out the tile bodies' own.
  __shared__ MainCtx mc;
  auto close = [&]() {
    const BatchFrame& pb = table->frames[cur];
    const unsigned int token = main_flush<kBlock>(&pb.s->main, fl, mc, acc);
    if (complete_items(&pb.s->main.done[0].n, items_done, static_cast<unsigned int>(pb.items), token, &fl.flag)) {
      main_finish<kBlock>(pb.f, &pb.s->main, fl, mc);
      if (threadIdx.x == 0) atomicExch(&pb.s->main.done[0].n, 0U);
    }
  };
  for (int item = q.first(); item < q.total; item = q.next_item(item)) {
    q.prefetch();
    int local = 0;
    const int f = batch_locate(table->main_start, table->main_order, count, item, &local);
    const BatchFrame& b = table->frames[f];
    if (f != cur) {
      if (cur >= 0) {
        close();
        __syncthreads();  // the epilogue is done with the frame LDS
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] = 0;
        items_done = 0;
      }
      cur = f;
      // The white balance pass of this batch ended with the previous launch.
      float gains[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) gains[k] = b.f.wb_mode != ISP_WB_MANUAL ? b.s->wb.gains[k] : b.f.wb[k];
      MainCtx begun;
      main_begin<kBlock>(b.f, gains, fl, begun);
      if (threadIdx.x == 0) mc = begun;
      __syncthreads();
    }
    if constexpr (kOnly >= 0) {
      batch_main_item<kOnlyFmt, kOnlyMalvar, kSharpen, 32, kBlock>(b, mc, fl, tile_lds, local, acc);
    } else {
      switch (b.variant) {
        ISP_BATCH_MAIN_CASES(ISP_RAW_U8)
        ISP_BATCH_MAIN_CASES(ISP_RAW_U16)
        ISP_BATCH_MAIN_CASES(ISP_RAW_RAW10)
        ISP_BATCH_MAIN_CASES(ISP_RAW_RAW12)
        default:
          break;
      }
    }
    ++items_done;
  }
  if (cur >= 0) close();
  q.leave();
}

#undef ISP_BATCH_MAIN_CASES
#undef ISP_BATCH_MAIN_CASE

__global__ void __launch_bounds__(kThreads) batch_main_kernel_sharpen(const BatchTable* __restrict__ table,
                                                                      BatchControl* ctl) {
  batch_main_body<true, kThreads>(table, ctl);
}

// Three resident blocks per CU: the registers are held to what three waves per
// SIMD leave (the LDS of this kernel fits three blocks).
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 3)))
batch_main_kernel_plain(const BatchTable* __restrict__ table, BatchControl* ctl) {
  batch_main_body<false, kThreads>(table, ctl);
}

// 512-thread forms, registers held to what four waves per SIMD leave.
constexpr int kWideBlock = 512;
__global__ void __launch_bounds__(kWideBlock) __attribute__((amdgpu_waves_per_eu(4)))
batch_main_kernel_sharpen_wide(const BatchTable* __restrict__ table, BatchControl* ctl) {
  batch_main_body<true, kWideBlock>(table, ctl);
}
__global__ void __launch_bounds__(kWideBlock) __attribute__((amdgpu_waves_per_eu(4)))
batch_main_kernel_plain_wide(const BatchTable* __restrict__ table, BatchControl* ctl) {
  batch_main_body<false, kWideBlock>(table, ctl);
}
// A batch whose frames all run one variant with 32-row tiles (variant_of(), no
// 16-row bit): that tile body alone, at the per-frame kernel's registers.
template <int kVariant>
__global__ void __launch_bounds__(kWideBlock)
__attribute__((amdgpu_waves_per_eu(main_waves_per_eu<kVariant & 3, (kVariant & 8) != 0>())))
batch_main_kernel_uniform(const BatchTable* __restrict__ table, BatchControl* ctl) {
  batch_main_body<(kVariant & 8) != 0, kWideBlock, kVariant>(table, ctl);
}

// The uniform kernel of a variant (0..15, 32-row tiles).
using BatchKernel = void (*)(const BatchTable*, BatchControl*);
template <int... kV>
constexpr BatchKernel uniform_kernel_of(int variant, std::integer_sequence<int, kV...>) {
  BatchKernel k = nullptr;
  ((k = kV == variant ? batch_main_kernel_uniform<kV> : k), ...);
  return k;
}
BatchKernel uniform_kernel(int variant) {
  return uniform_kernel_of(variant, std::make_integer_sequence<int, 16>{});
}

// ---- verification: the statistics epilogue on the device ---------------------

// finish_stats (isp_math.hpp:387-425) over many cases, one per thread: the same
// function main_finish runs, so its double arithmetic (the double-double
// log2 / exp2 of isp_ae.hpp) is checked as the device evaluates it.
__global__ void __launch_bounds__(kThreads) epilogue_kernel(const isp_epilogue_case* __restrict__ cases,
                                                            float* __restrict__ out, int count) {
  const int i = static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  if (i >= count) return;
  const isp_epilogue_case& c = cases[i];
  const unsigned long long sums[4] = {c.sums[0], c.sums[1], c.sums[2], c.sums[3]};
  const AeParams ae{c.ae.target, c.ae.damping, c.ae.min_exposure_ns, c.ae.max_exposure_ns, c.ae.min_gain,
                    c.ae.max_gain};
  float r[ISP_STATS_SIZE];
  finish_stats(sums, c.pixels, c.gains, ae, c.exposure.exposure_ns, c.exposure.gain, r);
#pragma unroll
  for (int k = 0; k < ISP_STATS_SIZE; ++k) out[static_cast<long long>(i) * ISP_STATS_SIZE + k] = r[k];
}

// ---- launches ---------------------------------------------------------------

void check(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw runtime_failure(std::string("IspPipeline (hip): ") + what + ": " + hipGetErrorString(status));
  }
}

template <typename Kernel>
int blocks_per_cu(Kernel kernel, int threads = kThreads) {
  static std::mutex mutex;
  static std::map<const void*, int> cache;
  std::lock_guard<std::mutex> lock(mutex);
  const auto it = cache.find(reinterpret_cast<const void*>(kernel));
  if (it != cache.end()) return it->second;
  int blocks = 0;
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0),
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

constexpr int kMainTileH = 32;

// Threads per block of the per-frame main kernel: 512, two blocks per CU (16
// waves) with sharpening and without.
template <bool kSharpen>
constexpr int kMainBlock = 512;

template <int kFmt, bool kMalvar, bool kSharpen, int kTileH, int kBlock, bool kPlain, int kPlainWaves = 6>
void launch_main_k(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream,
                   hipEvent_t done) {
  auto kernel = f.zero_sign != 0 ? main_kernel<kFmt, kMalvar, kSharpen, kTileH, kBlock, true, kPlain, kPlainWaves>
                                  : main_kernel<kFmt, kMalvar, kSharpen, kTileH, kBlock, false, kPlain, kPlainWaves>;
  const int tiles_x = (f.width + 63) / 64;
  const int tiles_y = (f.height + kTileH - 1) / kTileH;
  const int tiles = tiles_x * tiles_y;
  const int grid = std::max(1, std::min(tiles, cfg.cu_count * blocks_per_cu(kernel, kBlock)));
  const int rows = rows_aligned(f) ? 1 : 0;
  const int outs = out_aligned(f) ? 1 : 0;
  if (done != nullptr) {
    // The event completes with the dispatch itself: no marker packet of its own.
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, stream, nullptr, done, 0, f, s, tiles_x, tiles, rows,
                          outs);
  } else {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, stream, f, s, tiles_x, tiles, rows, outs);
  }
  check(hipGetLastError(), "main_kernel launch");
}

// Frames with nothing to finish and no sharpening run the plain kernel: 16-row
// tiles in 256-thread blocks without the histograms' LDS. At six waves a SIMD a
// tile's stages have more registers (a frame whose tiles all fit in one round of
// the resident blocks runs about 15 % faster than at eight); at eight the chip
// holds more blocks, which wins once the tiles outnumber the six-wave kernel's
// (1080p: one round instead of two). RAW10's unpacking needs more registers
// than eight waves leave.
template <int kFmt, bool kMalvar, bool kSharpen>
void launch_main_t(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream,
                   hipEvent_t done, bool plain) {
  if constexpr (!kSharpen) {
    if (plain) {
      if constexpr (kFmt != ISP_RAW_RAW10) {
        const int tiles = (f.width + 63) / 64 * ((f.height + 15) / 16);
        const int resident = cfg.cu_count * blocks_per_cu(main_kernel<kFmt, kMalvar, false, 16, kThreads, false, true, 6>,
                                                           kThreads);
        if (tiles > resident) {
          launch_main_k<kFmt, kMalvar, false, 16, kThreads, true, 8>(f, s, cfg, stream, done);
          return;
        }
      }
      launch_main_k<kFmt, kMalvar, false, 16, kThreads, true, 6>(f, s, cfg, stream, done);
      return;
    }
  }
  launch_main_k<kFmt, kMalvar, kSharpen, kMainTileH, kMainBlock<kSharpen>, false>(f, s, cfg, stream, done);
}

template <int kFmt>
void launch_main_f(const FrameDesc& f, DeviceScratch* s, bool malvar, bool sharpen, const LaunchConfig& cfg,
                   hipStream_t st, hipEvent_t done, bool plain) {
  if (malvar) {
    sharpen ? launch_main_t<kFmt, true, true>(f, s, cfg, st, done, plain)
            : launch_main_t<kFmt, true, false>(f, s, cfg, st, done, plain);
  } else {
    sharpen ? launch_main_t<kFmt, false, true>(f, s, cfg, st, done, plain)
            : launch_main_t<kFmt, false, false>(f, s, cfg, st, done, plain);
  }
}

template <int kFmt>
void launch_awb_f(const FrameDesc& f, DeviceScratch* s, const LaunchConfig& cfg, hipStream_t stream) {
  auto kernel = awb_kernel<kFmt>;
  const int tiles_x = (f.width + kAwbTile - 1) / kAwbTile;
  const int tiles_y = (f.height + kAwbTile - 1) / kAwbTile;
  const int tiles = tiles_x * tiles_y;
  const int grid = std::max(1, std::min(tiles, cfg.cu_count * std::min(cfg.awb_blocks, blocks_per_cu(kernel))));
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
                 void* stream, void* done, bool plain) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const hipEvent_t ev = static_cast<hipEvent_t>(done);
  switch (f.raw_format) {
    case ISP_RAW_U8:
      launch_main_f<ISP_RAW_U8>(f, s, malvar, sharpen, cfg, st, ev, plain);
      break;
    case ISP_RAW_U16:
      launch_main_f<ISP_RAW_U16>(f, s, malvar, sharpen, cfg, st, ev, plain);
      break;
    case ISP_RAW_RAW10:
      launch_main_f<ISP_RAW_RAW10>(f, s, malvar, sharpen, cfg, st, ev, plain);
      break;
    default:
      launch_main_f<ISP_RAW_RAW12>(f, s, malvar, sharpen, cfg, st, ev, plain);
      break;
  }
}

void launch_epilogue(const isp_epilogue_case* cases, float* out, int count, void* stream) {
  const int grid = (count + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(epilogue_kernel, dim3(grid), dim3(kThreads), 0, static_cast<hipStream_t>(stream), cases, out,
                     count);
  check(hipGetLastError(), "epilogue_kernel launch");
}

int batch_main_blocks_per_cu(bool sharpen, bool wide, int uniform) {
  if (uniform >= 0) return blocks_per_cu(uniform_kernel(uniform), kWideBlock);
  if (wide) {
    return sharpen ? blocks_per_cu(batch_main_kernel_sharpen_wide, kWideBlock)
                   : blocks_per_cu(batch_main_kernel_plain_wide, kWideBlock);
  }
  return sharpen ? blocks_per_cu(batch_main_kernel_sharpen) : blocks_per_cu(batch_main_kernel_plain);
}

int batch_awb_blocks_per_cu() { return std::min(4, blocks_per_cu(batch_awb_kernel)); }

void launch_batch_awb(const BatchTable* table, BatchControl* control, int grid, void* stream) {
  hipLaunchKernelGGL(batch_awb_kernel, dim3(grid), dim3(kThreads), 0, static_cast<hipStream_t>(stream), table,
                     control);
  check(hipGetLastError(), "batch_awb_kernel launch");
}

void launch_batch_main(const BatchTable* table, BatchControl* control, int grid, bool sharpen, bool wide,
                       int uniform, void* stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (uniform >= 0) {
    hipLaunchKernelGGL(uniform_kernel(uniform), dim3(grid), dim3(kWideBlock), 0, st, table, control);
  } else if (wide) {
    if (sharpen) {
      hipLaunchKernelGGL(batch_main_kernel_sharpen_wide, dim3(grid), dim3(kWideBlock), 0, st, table, control);
    } else {
      hipLaunchKernelGGL(batch_main_kernel_plain_wide, dim3(grid), dim3(kWideBlock), 0, st, table, control);
    }
  } else if (sharpen) {
    hipLaunchKernelGGL(batch_main_kernel_sharpen, dim3(grid), dim3(kThreads), 0, st, table, control);
  } else {
    hipLaunchKernelGGL(batch_main_kernel_plain, dim3(grid), dim3(kThreads), 0, st, table, control);
  }
  check(hipGetLastError(), "batch_main_kernel launch");
}

}  // namespace isp_amd
