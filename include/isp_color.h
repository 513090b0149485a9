/*
 * isp_color.h - the step after the ISP: colour conversion of its RGB output to
 * the YCbCr layouts encoders take, on the GPU, C ABI.
 *
 * A same-size slice of the reference's colour conversion module
 * (modules/color-convert): Converter::convert on the cuda backend
 * (include/holoscan/color_convert/converter.hpp:41-80) with the arithmetic of
 * its per-pixel code (src/kernel_math.hpp:80-361) and its colour matrices
 * (src/conversion.cpp:91-198, make_plan :397-446). Sources: RGB8, BGR8, RGBA8,
 * BGRA8 (alpha is not carried into YCbCr). Destinations: NV12, NV16, I420,
 * YUYV, UYVY and MONO8 (full-range luma). Matrices BT.601, BT.709 and BT.2020
 * (non-constant luminance), limited or full range. No resizing, cropping or
 * letterboxing: source and destination have one size. Every byte is the one the
 * reference's CPU converter (reference_convert, src/reference.cpp:14-27) writes.
 */
#ifndef ISP_COLOR_H_
#define ISP_COLOR_H_

#include "isp_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PixelFormat (pixel_format.hpp:16-29), same order and values. */
enum {
  ISP_PIX_NV12 = 0,
  ISP_PIX_NV16 = 1,
  ISP_PIX_I420 = 2,
  ISP_PIX_YUYV = 3,
  ISP_PIX_UYVY = 4,
  ISP_PIX_RGB8 = 5,
  ISP_PIX_BGR8 = 6,
  ISP_PIX_RGBA8 = 7,
  ISP_PIX_BGRA8 = 8,
  ISP_PIX_MONO8 = 9,
  ISP_PIX_MONO16 = 10,
  ISP_PIX_RGB16 = 11
};
/* ColorMatrix, ColorRange (pixel_format.hpp:40-44). */
enum { ISP_MATRIX_BT601 = 0, ISP_MATRIX_BT709 = 1, ISP_MATRIX_BT2020 = 2 };
enum { ISP_RANGE_LIMITED = 0, ISP_RANGE_FULL = 1 };

/* PlaneView / ImageView (image_view.hpp:19-30), device memory. */
typedef struct isp_plane_view {
  uint8_t* data;
  size_t pitch; /* bytes from one row to the next, at least the packed row */
} isp_plane_view;

typedef struct isp_image_view {
  uint8_t format; /* ISP_PIX_* */
  int32_t width;
  int32_t height;
  isp_plane_view planes[3]; /* the format's planes: 1, 2 (NV12, NV16) or 3 (I420) */
} isp_image_view;

/* The colour settings of ConversionParams (conversion.hpp:58-69); the defaults
 * are the reference's: BT.601, limited range. */
typedef struct isp_color_params {
  uint8_t matrix; /* ISP_MATRIX_* */
  uint8_t range;  /* ISP_RANGE_* */
} isp_color_params;

void isp_color_params_init(isp_color_params* params);

/* plane_size (pixel_format.hpp:111-123): bytes of a packed row and rows of
 * `plane` of a width x height image of `format`. */
isp_status isp_color_plane_size(uint8_t format, int32_t width, int32_t height, int32_t plane,
                                size_t* row_bytes, int32_t* rows);

/* Converter::convert (converter.hpp:60-67) on the cuda backend: converts
 * `source` into `destination`, both in device memory, queued on `stream` after
 * the work already there. ISP_INVALID_ARGUMENT for a pair, size, plane or pitch
 * this build does not convert, with the reason in isp_last_error(). */
isp_status isp_color_convert_device(const isp_image_view* source, const isp_image_view* destination,
                                    const isp_color_params* params, isp_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ISP_COLOR_H_ */
