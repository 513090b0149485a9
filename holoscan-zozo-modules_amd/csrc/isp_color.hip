// isp_color.hip - colour conversion of the ISP's RGB output to the YCbCr
// layouts encoders take (include/isp_color.h).
//
// The conversion moves few bytes a pixel (RGB8 to NV12 reads 3 and writes 1.5)
// but spends about 30 float operations on each (nine products, nine sums, the
// rounding of a luma and a share of two chroma codes), so at 4K it is bound by
// the vector ALUs rather than by HBM, and the kernel is written to spend as few
// operations as the reference's arithmetic allows: one thread converts a span
// of 8 pixels (two rows of them for the 4:2:0 layouts, whose chroma sample
// covers a 2 x 2 block), converts bytes with v_cvt_f32_ubyte*, clamps with one
// median and packs each code into its output word with v_cvt_pk_u8_f32. Rows
// are loaded and planes stored as the widest vectors their alignment allows,
// consecutive threads on consecutive spans. Spans at the right or bottom edge,
// and views whose planes or pitches are not aligned for the vectors, take a
// byte-wise path with the same arithmetic.
//
// Arithmetic, as the reference's per-pixel code (kernel_math.hpp:80-361) does
// it, in float without contraction: a pixel's (Y, Cb, Cr) is m * (R, G, B) + o
// row by row ((m0 R + m1 G) + m2 B) + o0; a code is floor(v + 0.5) clamped to
// 0..255; a chroma sample is the sum of its block's Cb (Cr) values, row by row,
// from 0, divided by the block's pixel count. At one size every resampling mode
// of the reference reads each source pixel with weight 1, so the pixel is the
// source's bytes as floats. The matrix and offsets are computed in double and
// rounded to float as make_plan does (conversion.cpp:176-198, :428-436).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/isp_color.h"
#include "isp_pipeline.hpp"

namespace isp_amd {
namespace {

enum Model { kYuv, kRgb, kGray };

// format_info (pixel_format.hpp:60-86).
struct Info {
  int model;
  int planes;
  int shift_x;
  int shift_y;
  int bytes;  // per sample
  int channels;
};

Info info_of(int format) {
  switch (format) {
    case ISP_PIX_NV12:
      return {kYuv, 2, 1, 1, 1, 1};
    case ISP_PIX_NV16:
      return {kYuv, 2, 1, 0, 1, 1};
    case ISP_PIX_I420:
      return {kYuv, 3, 1, 1, 1, 1};
    case ISP_PIX_YUYV:
    case ISP_PIX_UYVY:
      return {kYuv, 1, 1, 0, 1, 2};
    case ISP_PIX_RGB8:
    case ISP_PIX_BGR8:
      return {kRgb, 1, 0, 0, 1, 3};
    case ISP_PIX_RGBA8:
    case ISP_PIX_BGRA8:
      return {kRgb, 1, 0, 0, 1, 4};
    case ISP_PIX_MONO8:
      return {kGray, 1, 0, 0, 1, 1};
    case ISP_PIX_MONO16:
      return {kGray, 1, 0, 0, 2, 1};
    default:
      return {kRgb, 1, 0, 0, 2, 3};
  }
}

const char* format_name(int format) {
  static const char* const kNames[] = {"nv12",  "nv16",  "i420",  "yuyv",  "uyvy",   "rgb8",
                                       "bgr8",  "rgba8", "bgra8", "mono8", "mono16", "rgb16"};
  return format >= 0 && format <= ISP_PIX_RGB16 ? kNames[format] : "unknown";
}

// plane_size (pixel_format.hpp:111-123).
void plane_size(int format, int width, int height, int plane, std::size_t* row_bytes, int* rows) {
  const Info info = info_of(format);
  const auto w = static_cast<std::size_t>(width);
  if (plane == 0) {
    *row_bytes = w * static_cast<std::size_t>(info.channels * info.bytes);
    *rows = height;
    return;
  }
  const std::size_t chroma_width = (w + ((1U << info.shift_x) - 1U)) >> info.shift_x;
  *rows = (height + (1 << info.shift_y) - 1) >> info.shift_y;
  *row_bytes = chroma_width * (info.planes == 2 ? 2U : 1U);
}

// check_view (conversion.cpp:223-244), with its messages.
void check_view(const isp_image_view& v, const char* which) {
  const std::string name(which);
  if (v.format > ISP_PIX_RGB16) throw invalid_argument(name + " image format is not a PixelFormat");
  if (v.width <= 0 || v.height <= 0) throw invalid_argument(name + " image size must be positive");
  if ((v.format == ISP_PIX_YUYV || v.format == ISP_PIX_UYVY) && v.width % 2 != 0) {
    throw invalid_argument(name + " image: " + format_name(v.format) + " needs an even width, not " +
                           std::to_string(v.width));
  }
  for (int plane = 0; plane < info_of(v.format).planes; ++plane) {
    std::size_t row = 0;
    int rows = 0;
    plane_size(v.format, v.width, v.height, plane, &row, &rows);
    if (v.planes[plane].data == nullptr) {
      throw invalid_argument(name + " image plane " + std::to_string(plane) + " has no data");
    }
    if (v.planes[plane].pitch < row) {
      throw invalid_argument(name + " image plane " + std::to_string(plane) + " pitch " +
                             std::to_string(v.planes[plane].pitch) + " is shorter than a row of " +
                             std::to_string(row) + " bytes");
    }
  }
}

// c = m * (R, G, B) + o.
struct Affine {
  float m[9];
  float o[3];
};

// from_rgb (conversion.cpp:176-198) for a YCbCr or gray destination, rounded
// to float as make_plan rounds its composed map (the RGB source's to_rgb is the
// identity, which the composition passes through exactly).
Affine affine_of(int dst_model, int matrix, int range) {
  double kr = 0.299;
  double kb = 0.114;
  if (matrix == ISP_MATRIX_BT709) {
    kr = 0.2126;
    kb = 0.0722;
  } else if (matrix == ISP_MATRIX_BT2020) {
    kr = 0.2627;
    kb = 0.0593;
  }
  const double kg = 1.0 - kr - kb;
  double m[3][4] = {};
  if (dst_model == kGray) {
    m[0][0] = kr;
    m[0][1] = kg;
    m[0][2] = kb;
  } else {
    const bool limited = range == ISP_RANGE_LIMITED;
    const double y_scale = limited ? 219.0 : 255.0;
    const double y_offset = limited ? 16.0 : 0.0;
    const double c_scale = limited ? 224.0 : 255.0;
    const double cb_den = 2.0 * (1.0 - kb);
    const double cr_den = 2.0 * (1.0 - kr);
    const double s = y_scale / 255.0;
    const double c = c_scale / 255.0;
    const double rows[3][4] = {
        {kr * s, kg * s, kb * s, y_offset},
        {-kr / cb_den * c, -kg / cb_den * c, (1.0 - kb) / cb_den * c, 128.0},
        {(1.0 - kr) / cr_den * c, -kg / cr_den * c, -kb / cr_den * c, 128.0},
    };
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) m[i][j] = rows[i][j];
  }
  Affine a{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) a.m[i * 3 + j] = static_cast<float>(m[i][j]);
    a.o[i] = static_cast<float>(m[i][3]);
  }
  return a;
}

struct alignas(8) V2 {
  unsigned int w[2];
};
struct alignas(16) V4 {
  unsigned int w[4];
};

struct Planes {
  std::uint8_t* data[3];
  std::size_t pitch[3];
};

struct Ycc {
  float y, cb, cr;
};

// to_u8 (kernel_math.hpp:80-83) up to the cast: floor(v + 0.5) clamped to
// 0..255, an exact integer in float. The values here are finite, so the clamp
// is one median of three.
__device__ __forceinline__ float code_f(float v) {
  return __builtin_amdgcn_fmed3f(floorf(v + 0.5F), 0.0F, 255.0F);
}
__device__ __forceinline__ unsigned int code(float v) {
  return static_cast<unsigned int>(code_f(v));
}
// Puts code_f(v) into byte `i` of `word` (one v_cvt_pk_u8_f32: the value is an
// integer in 0..255, so the conversion is exact).
__device__ __forceinline__ unsigned int put_code(float v, unsigned int i, unsigned int word) {
  return __builtin_amdgcn_cvt_pk_u8_f32(code_f(v), i, word);
}
// put_code for a luma value, which needs no clamp: its weights are positive
// and sum to at most 1 (to 219/255 over a 16 offset in limited range), so v
// lies in [0, 255] up to a few ulps and floor(v + 0.5) in 0..255.
__device__ __forceinline__ unsigned int put_luma_code(float v, unsigned int i, unsigned int word) {
  return __builtin_amdgcn_cvt_pk_u8_f32(floorf(v + 0.5F), i, word);
}

// apply (kernel_math.hpp:234-242) to the source's bytes.
__device__ __forceinline__ Ycc apply(const Affine& a, float r, float g, float b) {
  return {a.m[0] * r + a.m[1] * g + a.m[2] * b + a.o[0], a.m[3] * r + a.m[4] * g + a.m[5] * b + a.o[1],
          a.m[6] * r + a.m[7] * g + a.m[8] * b + a.o[2]};
}

template <int kSrc>
constexpr int src_bytes() {
  return kSrc == ISP_PIX_RGBA8 || kSrc == ISP_PIX_BGRA8 ? 4 : 3;
}
template <int kSrc>
constexpr bool src_bgr() {
  return kSrc == ISP_PIX_BGR8 || kSrc == ISP_PIX_BGRA8;
}

// Pixel x of source row `row` (load, kernel_math.hpp:107-156), byte-wise.
template <int kSrc>
__device__ __forceinline__ Ycc pixel_at(const Affine& a, const std::uint8_t* row, int x) {
  const std::uint8_t* p = row + src_bytes<kSrc>() * x;
  const float c0 = static_cast<float>(p[src_bgr<kSrc>() ? 2 : 0]);
  const float c1 = static_cast<float>(p[1]);
  const float c2 = static_cast<float>(p[src_bgr<kSrc>() ? 0 : 2]);
  return apply(a, c0, c1, c2);
}

// Pixel k of a span's row from its loaded words.
template <int kSrc>
__device__ __forceinline__ Ycc pixel_of(const Affine& a, const unsigned int* w, int k) {
  constexpr int kB = src_bytes<kSrc>();
  const auto byte = [w](int i) { return static_cast<float>((w[i >> 2] >> (8 * (i & 3))) & 0xFFU); };
  const float c0 = byte(kB * k + (src_bgr<kSrc>() ? 2 : 0));
  const float c1 = byte(kB * k + 1);
  const float c2 = byte(kB * k + (src_bgr<kSrc>() ? 0 : 2));
  return apply(a, c0, c1, c2);
}

// Pixels of a thread's span. 8 measures faster at 4K than 16 (more waves,
// fewer registers), and faster than 8 with several span rows a thread.
constexpr int kSpan = 8;

// The widest vector (4, 8 or 16 bytes) that a run of `bytes` bytes starting at
// a multiple of `bytes` keeps aligned.
constexpr int unit_of(int bytes) {
  return bytes % 16 == 0 ? 16 : (bytes % 8 == 0 ? 8 : 4);
}

template <int kDst>
constexpr int block_rows() {
  return kDst == ISP_PIX_NV12 || kDst == ISP_PIX_I420 ? 2 : 1;
}

// Y goes to plane 0 at (x, y) (store, kernel_math.hpp:253-300).
template <int kDst>
__device__ __forceinline__ void put_luma(const Planes& d, int x, int y, unsigned int v) {
  std::uint8_t* row = d.data[0] + static_cast<std::size_t>(y) * d.pitch[0];
  if constexpr (kDst == ISP_PIX_YUYV) {
    row[4 * (x >> 1) + 2 * (x & 1)] = static_cast<std::uint8_t>(v);
  } else if constexpr (kDst == ISP_PIX_UYVY) {
    row[4 * (x >> 1) + 1 + 2 * (x & 1)] = static_cast<std::uint8_t>(v);
  } else {
    row[x] = static_cast<std::uint8_t>(v);
  }
}

// The chroma sample (cx, cy) (store_chroma, kernel_math.hpp:303-331).
template <int kDst>
__device__ __forceinline__ void put_chroma(const Planes& d, int cx, int cy, unsigned int cb, unsigned int cr) {
  if constexpr (kDst == ISP_PIX_NV12 || kDst == ISP_PIX_NV16) {
    std::uint8_t* q = d.data[1] + static_cast<std::size_t>(cy) * d.pitch[1] + 2 * cx;
    q[0] = static_cast<std::uint8_t>(cb);
    q[1] = static_cast<std::uint8_t>(cr);
  } else if constexpr (kDst == ISP_PIX_I420) {
    d.data[1][static_cast<std::size_t>(cy) * d.pitch[1] + cx] = static_cast<std::uint8_t>(cb);
    d.data[2][static_cast<std::size_t>(cy) * d.pitch[2] + cx] = static_cast<std::uint8_t>(cr);
  } else if constexpr (kDst == ISP_PIX_YUYV) {
    std::uint8_t* q = d.data[0] + static_cast<std::size_t>(cy) * d.pitch[0] + 4 * cx;
    q[1] = static_cast<std::uint8_t>(cb);
    q[3] = static_cast<std::uint8_t>(cr);
  } else if constexpr (kDst == ISP_PIX_UYVY) {
    std::uint8_t* q = d.data[0] + static_cast<std::size_t>(cy) * d.pitch[0] + 4 * cx;
    q[0] = static_cast<std::uint8_t>(cb);
    q[2] = static_cast<std::uint8_t>(cr);
  }
}

// The byte-wise form of a span: convert_block (kernel_math.hpp:335-361) for
// each chroma block (or pixel, for MONO8) of the span inside the image.
template <int kSrc, int kDst>
__device__ void span_bytes(const Affine& a, const std::uint8_t* src, std::size_t src_pitch, const Planes& d, int W,
                           int H, int x0, int y0) {
  constexpr int kBY = block_rows<kDst>();
  const int y1 = min(y0 + kBY, H);
  if constexpr (kDst == ISP_PIX_MONO8) {
    const std::uint8_t* row = src + static_cast<std::size_t>(y0) * src_pitch;
    for (int x = x0; x < min(x0 + kSpan, W); ++x) put_luma<kDst>(d, x, y0, code(pixel_at<kSrc>(a, row, x).y));
  } else {
    for (int bx = x0; bx < min(x0 + kSpan, W); bx += 2) {
      const int bx1 = min(bx + 2, W);
      float cb = 0.0F;
      float cr = 0.0F;
      int n = 0;
      for (int y = y0; y < y1; ++y) {
        const std::uint8_t* row = src + static_cast<std::size_t>(y) * src_pitch;
        for (int x = bx; x < bx1; ++x) {
          const Ycc v = pixel_at<kSrc>(a, row, x);
          put_luma<kDst>(d, x, y, code(v.y));
          cb += v.cb;
          cr += v.cr;
          ++n;
        }
      }
      const float count = static_cast<float>(n);
      put_chroma<kDst>(d, bx >> 1, y0 / kBY, code(cb / count), code(cr / count));
    }
  }
}

// Words of one span row of the source (kSpan pixels).
template <int kSrc>
constexpr int span_words() {
  return kSpan * src_bytes<kSrc>() / 4;
}

typedef unsigned int U4 __attribute__((ext_vector_type(4)));
typedef unsigned int U2 __attribute__((ext_vector_type(2)));

// kN words from p, as vectors of kUnit bytes (p aligned to kUnit).
template <int kUnit, int kN>
__device__ __forceinline__ void load_words(const std::uint8_t* p, unsigned int (&w)[kN]) {
#pragma unroll
  for (int i = 0; i < kN; i += kUnit / 4) {
    if constexpr (kUnit == 16) {
      const V4 v = *reinterpret_cast<const V4*>(p + 4 * i);
#pragma unroll
      for (int k = 0; k < 4; ++k) w[i + k] = v.w[k];
    } else if constexpr (kUnit == 8) {
      const V2 v = *reinterpret_cast<const V2*>(p + 4 * i);
      w[i] = v.w[0];
      w[i + 1] = v.w[1];
    } else {
      w[i] = *reinterpret_cast<const unsigned int*>(p + 4 * i);
    }
  }
}

// kN words to p, as vectors of kUnit bytes (p aligned to kUnit). The stores
// are nontemporal: the planes are written once and read by the next stage,
// not by this kernel, and streaming them past the caches measures about 10 %
// faster at 4K.
template <int kUnit, int kN>
__device__ __forceinline__ void store_words(std::uint8_t* p, const unsigned int (&w)[kN]) {
#pragma unroll
  for (int i = 0; i < kN; i += kUnit / 4) {
    if constexpr (kUnit == 16) {
      __builtin_nontemporal_store(U4{w[i], w[i + 1], w[i + 2], w[i + 3]}, reinterpret_cast<U4*>(p + 4 * i));
    } else if constexpr (kUnit == 8) {
      __builtin_nontemporal_store(U2{w[i], w[i + 1]}, reinterpret_cast<U2*>(p + 4 * i));
    } else {
      __builtin_nontemporal_store(w[i], reinterpret_cast<unsigned int*>(p + 4 * i));
    }
  }
}

// Vector units of the source rows, plane 0 and the chroma planes.
template <int kSrc>
constexpr int src_unit() {
  return unit_of(kSpan * src_bytes<kSrc>());
}
template <int kDst>
constexpr int luma_unit() {
  return unit_of(kDst == ISP_PIX_YUYV || kDst == ISP_PIX_UYVY ? 2 * kSpan : kSpan);
}
template <int kDst>
constexpr int chroma_unit() {
  return unit_of(kDst == ISP_PIX_I420 ? kSpan / 2 : kSpan);
}


// One span (kSpan pixels of kBY rows at (x0, y0)) from its loaded words to
// every destination plane, with vector stores.
template <int kSrc, int kDst>
__device__ __forceinline__ void convert_span(const Affine& a, const unsigned int (&w)[block_rows<kDst>()][span_words<kSrc>()],
                                             const Planes& d, int x0, int y0) {
  constexpr int kBY = block_rows<kDst>();
  constexpr int kBlocks = kSpan / 2;  // chroma blocks of a span
  // Row by row: each pixel's luma code, put in its byte of the output word (a
  // plane-0 word of 4 codes; for packed 4:2:2 the word of its pair), and its Cb
  // and Cr added to its chroma block's sums, which so run row by row from 0 as
  // the reference's do.
  constexpr bool kPacked = kDst == ISP_PIX_YUYV || kDst == ISP_PIX_UYVY;
  unsigned int y[kBY][kPacked ? kBlocks : kSpan / 4];
  float sb[kBlocks];
  float sr[kBlocks];
#pragma unroll
  for (int r = 0; r < kBY; ++r) {
#pragma unroll
    for (int k = 0; k < kSpan; ++k) {
      const Ycc v = pixel_of<kSrc>(a, w[r], k);
      const int word = kPacked ? k >> 1 : k >> 2;
      const int byte = kPacked ? (kDst == ISP_PIX_UYVY ? 1 : 0) + 2 * (k & 1) : k & 3;
      const bool first = kPacked ? (k & 1) == 0 : (k & 3) == 0;
      y[r][word] = put_luma_code(v.y, byte, first ? 0U : y[r][word]);
      if constexpr (kDst != ISP_PIX_MONO8) {
        if (r == 0 && (k & 1) == 0) {
          sb[k >> 1] = v.cb;
          sr[k >> 1] = v.cr;
        } else {
          sb[k >> 1] += v.cb;
          sr[k >> 1] += v.cr;
        }
      }
    }
  }
  if constexpr (kDst == ISP_PIX_MONO8) {
    store_words<luma_unit<kDst>()>(d.data[0] + static_cast<std::size_t>(y0) * d.pitch[0] + x0, y[0]);
  } else {
    // Chroma means: the division by a power of two is exact as a product.
    float cb[kBlocks];
    float cr[kBlocks];
    const float count = static_cast<float>(2 * kBY);
#pragma unroll
    for (int b = 0; b < kBlocks; ++b) {
      cb[b] = sb[b] / count;
      cr[b] = sr[b] / count;
    }
    if constexpr (kDst == ISP_PIX_YUYV || kDst == ISP_PIX_UYVY) {
      unsigned int o[kBlocks];
#pragma unroll
      for (int b = 0; b < kBlocks; ++b) {
        o[b] = kDst == ISP_PIX_YUYV ? put_code(cr[b], 3, put_code(cb[b], 1, y[0][b]))
                                    : put_code(cr[b], 2, put_code(cb[b], 0, y[0][b]));
      }
      store_words<luma_unit<kDst>()>(d.data[0] + static_cast<std::size_t>(y0) * d.pitch[0] + 2 * x0, o);
    } else {
#pragma unroll
      for (int r = 0; r < kBY; ++r) store_words<luma_unit<kDst>()>(d.data[0] + static_cast<std::size_t>(y0 + r) * d.pitch[0] + x0, y[r]);
      const int cy = y0 / kBY;
      if constexpr (kDst == ISP_PIX_I420) {
        unsigned int u[kBlocks / 4];
        unsigned int v[kBlocks / 4];
#pragma unroll
        for (int i = 0; i < kBlocks / 4; ++i) {
          u[i] = 0;
          v[i] = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            u[i] = put_code(cb[4 * i + k], k, u[i]);
            v[i] = put_code(cr[4 * i + k], k, v[i]);
          }
        }
        store_words<chroma_unit<kDst>()>(d.data[1] + static_cast<std::size_t>(cy) * d.pitch[1] + (x0 >> 1), u);
        store_words<chroma_unit<kDst>()>(d.data[2] + static_cast<std::size_t>(cy) * d.pitch[2] + (x0 >> 1), v);
      } else {
        unsigned int o[kBlocks / 2];
#pragma unroll
        for (int i = 0; i < kBlocks / 2; ++i) {
          o[i] = 0;
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            o[i] = put_code(cb[2 * i + k], 2 * k, o[i]);
            o[i] = put_code(cr[2 * i + k], 2 * k + 1, o[i]);
          }
        }
        store_words<chroma_unit<kDst>()>(d.data[1] + static_cast<std::size_t>(cy) * d.pitch[1] + x0, o);
      }
    }
  }
}

// Grid: x over spans of kSpan pixels, y over span rows (of block_rows() image rows).
template <int kSrc, int kDst>
__global__ void __launch_bounds__(256) rgb_to_ycc_kernel(const std::uint8_t* __restrict__ src, std::size_t src_pitch,
                                                         Planes d, Affine a, int W, int H, int vec) {
  constexpr int kBY = block_rows<kDst>();
  const int x0 = kSpan * static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  // A wave is one row of 64 threads (the block is 64 x 4), so its span row is
  // uniform: read from the first lane, its row addresses are scalar.
  const int y0 = __builtin_amdgcn_readfirstlane(kBY * static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y));
  if (x0 >= W || y0 >= H) return;
  if (vec == 0 || x0 + kSpan > W || y0 + kBY > H) {
    span_bytes<kSrc, kDst>(a, src, src_pitch, d, W, H, x0, y0);
    return;
  }
  unsigned int w[kBY][span_words<kSrc>()];
#pragma unroll
  for (int r = 0; r < kBY; ++r) {
    load_words<src_unit<kSrc>()>(src + static_cast<std::size_t>(y0 + r) * src_pitch + src_bytes<kSrc>() * x0, w[r]);
  }
  convert_span<kSrc, kDst>(a, w, d, x0, y0);
}

bool aligned(const void* p, std::size_t pitch, std::size_t n) {
  return reinterpret_cast<std::uintptr_t>(p) % n == 0 && pitch % n == 0;
}

template <int kSrc, int kDst>
void launch(const isp_image_view& s, const isp_image_view& t, const Affine& a, hipStream_t stream) {
  constexpr int kBY = block_rows<kDst>();
  Planes d{};
  for (int i = 0; i < 3; ++i) {
    d.data[i] = t.planes[i].data;
    d.pitch[i] = t.planes[i].pitch;
  }
  // The vector form needs the source rows and each destination plane aligned
  // to its vectors.
  bool vec = aligned(s.planes[0].data, s.planes[0].pitch, src_unit<kSrc>()) &&
             aligned(d.data[0], d.pitch[0], luma_unit<kDst>());
  if constexpr (kDst == ISP_PIX_NV12 || kDst == ISP_PIX_NV16) vec = vec && aligned(d.data[1], d.pitch[1], chroma_unit<kDst>());
  if constexpr (kDst == ISP_PIX_I420) {
    vec = vec && aligned(d.data[1], d.pitch[1], chroma_unit<kDst>()) && aligned(d.data[2], d.pitch[2], chroma_unit<kDst>());
  }
  const int spans = (s.width + kSpan - 1) / kSpan;
  const int span_rows = (s.height + kBY - 1) / kBY;
  const dim3 block(64, 4);
  const dim3 grid(static_cast<unsigned int>((spans + 63) / 64), static_cast<unsigned int>((span_rows + 3) / 4));
  rgb_to_ycc_kernel<kSrc, kDst><<<grid, block, 0, stream>>>(s.planes[0].data, s.planes[0].pitch, d, a, s.width,
                                                            s.height, vec ? 1 : 0);
}

template <int kSrc>
void launch_to(const isp_image_view& s, const isp_image_view& t, const Affine& a, hipStream_t stream) {
  switch (t.format) {
    case ISP_PIX_NV12:
      return launch<kSrc, ISP_PIX_NV12>(s, t, a, stream);
    case ISP_PIX_NV16:
      return launch<kSrc, ISP_PIX_NV16>(s, t, a, stream);
    case ISP_PIX_I420:
      return launch<kSrc, ISP_PIX_I420>(s, t, a, stream);
    case ISP_PIX_YUYV:
      return launch<kSrc, ISP_PIX_YUYV>(s, t, a, stream);
    case ISP_PIX_UYVY:
      return launch<kSrc, ISP_PIX_UYVY>(s, t, a, stream);
    default:
      return launch<kSrc, ISP_PIX_MONO8>(s, t, a, stream);
  }
}

}  // namespace
}  // namespace isp_amd

using namespace isp_amd;

extern "C" {

void isp_color_params_init(isp_color_params* params) {
  if (params == nullptr) return;
  params->matrix = ISP_MATRIX_BT601;
  params->range = ISP_RANGE_LIMITED;
}

isp_status isp_color_plane_size(uint8_t format, int32_t width, int32_t height, int32_t plane, size_t* row_bytes,
                                int32_t* rows) {
  return guarded([&] {
    if (row_bytes == nullptr || rows == nullptr) throw invalid_argument("row_bytes or rows is null");
    if (format > ISP_PIX_RGB16) throw invalid_argument("format is not a PixelFormat");
    if (plane < 0 || plane >= info_of(format).planes) throw invalid_argument("the format has no such plane");
    int r = 0;
    plane_size(format, width, height, plane, row_bytes, &r);
    *rows = r;
  });
}

isp_status isp_color_convert_device(const isp_image_view* source, const isp_image_view* destination,
                                    const isp_color_params* params, isp_stream_t stream) {
  return guarded([&] {
    if (source == nullptr || destination == nullptr) throw invalid_argument("source or destination is null");
    isp_color_params p;
    isp_color_params_init(&p);
    if (params != nullptr) p = *params;
    if (p.matrix > ISP_MATRIX_BT2020) throw invalid_argument("matrix is not a ColorMatrix");
    if (p.range > ISP_RANGE_FULL) throw invalid_argument("range is not a ColorRange");
    check_view(*source, "source");
    check_view(*destination, "destination");
    const int s = source->format;
    const int t = destination->format;
    const bool rgb_in = s == ISP_PIX_RGB8 || s == ISP_PIX_BGR8 || s == ISP_PIX_RGBA8 || s == ISP_PIX_BGRA8;
    const bool ycc_out = info_of(t).model == kYuv || t == ISP_PIX_MONO8;
    if (!rgb_in || !ycc_out) {
      throw invalid_argument(std::string("this library converts rgb8, bgr8, rgba8 and bgra8 into nv12, nv16, i420, "
                                         "yuyv, uyvy and mono8, not ") +
                             format_name(s) + " into " + format_name(t));
    }
    if (source->width != destination->width || source->height != destination->height) {
      throw invalid_argument("this library converts at one size, not " + std::to_string(source->width) + "x" +
                             std::to_string(source->height) + " into " + std::to_string(destination->width) + "x" +
                             std::to_string(destination->height));
    }
    const Affine a = affine_of(info_of(t).model, p.matrix, p.range);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    static_cast<void>(hipGetLastError());
    switch (s) {
      case ISP_PIX_RGB8:
        launch_to<ISP_PIX_RGB8>(*source, *destination, a, st);
        break;
      case ISP_PIX_BGR8:
        launch_to<ISP_PIX_BGR8>(*source, *destination, a, st);
        break;
      case ISP_PIX_RGBA8:
        launch_to<ISP_PIX_RGBA8>(*source, *destination, a, st);
        break;
      default:
        launch_to<ISP_PIX_BGRA8>(*source, *destination, a, st);
        break;
    }
    check(hipGetLastError(), "rgb_to_ycc_kernel");
  });
}

}  // extern "C"
