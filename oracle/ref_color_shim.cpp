// ref_color_shim.cpp - TEST INFRASTRUCTURE ONLY: a C entry into the reference's
// own CPU colour converter (reference_convert, modules/color-convert/src/
// reference.cpp:14-27, with src/conversion.cpp), compiled from its sources where
// they lie into oracle/_ref/libcc_ref.so by oracle/Makefile. It maps this
// repository's isp_image_view (include/isp_color.h) onto the reference's
// ImageView and ConversionParams. Used only to pin the oracle; never shipped in
// the product.
#include <exception>
#include <string>

#include <holoscan/color_convert/conversion.hpp>
#include <holoscan/color_convert/reference.hpp>

#include "../include/isp_color.h"

namespace cc = holoscan::color_convert;

namespace {

thread_local std::string g_error;

cc::ImageView to_view(const isp_image_view& v) {
  cc::ImageView out;
  out.format = static_cast<cc::PixelFormat>(v.format);
  out.width = v.width;
  out.height = v.height;
  for (int i = 0; i < 3; ++i) out.planes[static_cast<std::size_t>(i)] = {v.planes[i].data, v.planes[i].pitch};
  out.location = cc::MemoryLocation::kHost;
  return out;
}

}  // namespace

extern "C" {

const char* ref_cc_last_error() { return g_error.c_str(); }

// 0, or -1 with the reference's message in ref_cc_last_error(). `mode` is a
// ResizeMode (0 nearest, 1 bilinear, 2 area).
int ref_cc_convert(const isp_image_view* source, const isp_image_view* destination, int matrix, int range,
                   int mode) {
  try {
    cc::ConversionParams p;
    p.matrix = static_cast<cc::ColorMatrix>(matrix);
    p.range = static_cast<cc::ColorRange>(range);
    p.mode = static_cast<cc::ResizeMode>(mode);
    cc::reference_convert(to_view(*source), to_view(*destination), p);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // extern "C"
