/*
 * isp_source.h - the device-side synthetic raw source and the GPU frame
 * checker, C ABI.
 *
 * 1. isp_synthetic_render_device: the synthetic chart camera of isp_synthetic.h
 *    rendered into device memory on a stream, byte for byte the frame that
 *    isp_synthetic_render writes on the host, in every sample format and MIPI
 *    CSI-2 packing (render_synthetic_raw_cuda,
 *    modules/isp/include/holoscan/isp/synthetic_raw_cuda.hpp:17-19,
 *    modules/isp/src/synthetic_raw.cu:24-37). A camera that renders into HBM
 *    feeds the ISP without a PCIe copy.
 * 2. isp_check_frame_device: the per-frame ground truth of the reference's
 *    demo on the GPU (check_frame_cuda, applications/isp/src/frame_check.hpp:70-76,
 *    frame_check.cu:14-105): the squared error of an RGB8 image against the
 *    camera's truth through the ISP's tone curve, the image's four histogram
 *    rows rebuilt from its pixels and compared bin by bin with the ones the ISP
 *    wrote, and the ISP's statistics, gathered into one 64-byte result in
 *    device memory that the caller reads back once per frame.
 */
#ifndef ISP_SOURCE_H_
#define ISP_SOURCE_H_

#include "isp_c.h"
#include "isp_synthetic.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FrameCheckResult (frame_check.hpp:33-41). */
typedef struct isp_frame_check_result {
  uint64_t squared_error;        /* sum over pixels and channels of (code - truth code)^2 */
  uint32_t histogram_mismatches; /* bins that differ from the counts of the image's own pixels */
  uint32_t reserved;
  float stats[ISP_STATS_SIZE]; /* the ISP's statistics, as sent */
} isp_frame_check_result;

/* Device scratch of isp_check_frame_device: zero it once when it is allocated;
 * every check leaves it zeroed again. */
#define ISP_FRAME_CHECK_SCRATCH_BYTES 37120

/* render_synthetic_raw_cuda (synthetic_raw_cuda.hpp:17-19): the frame of
 * `config` into device memory at `data`, rows `pitch` bytes apart (at least
 * isp_raw_row_bytes), queued on `stream`. Bytes past a row's packed samples are
 * not written. */
isp_status isp_synthetic_render_device(const isp_synthetic_config* config, void* data, size_t pitch,
                                       isp_stream_t stream);

/* check_frame_cuda (frame_check.hpp:70-76): checks the RGB8 image `rgb` (rows
 * `pitch` bytes apart, 0: 3 * width) of frame `camera` against the camera's
 * truth through the tone table `lut` (lut_size entries, as isp_tone_lut writes
 * it), compares `histogram` (ISP_HISTOGRAM_CHANNELS x ISP_HISTOGRAM_BINS uint32)
 * with the image's own, and copies `stats` (ISP_STATS_SIZE floats; NULL: zeros)
 * into `*result`. Every pointer is device memory; the whole check is one launch
 * queued on `stream`. Checks that share a scratch must not overlap. */
isp_status isp_check_frame_device(const isp_synthetic_config* camera, const float* lut, int32_t lut_size,
                                  const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                                  const float* stats, void* scratch, isp_frame_check_result* result,
                                  isp_stream_t stream);

/* No counterpart in the reference: isp_check_frame_device without the scratch
 * argument, for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It allocates
 * nothing: its scratch is a static device array of ISP_FRAME_CHECK_SCRATCH_BYTES
 * that each GPU holds from the library's first use on it until the process
 * ends, cleared once on that first call and left zeroed by every check. All
 * calls on one GPU share that array, so calls that overlap on one GPU race:
 * queue them on one stream, or wait between them. Frames above
 * ISP_FRAME_CHECK_STATIC_MAX_PIXELS (what the array's 32-bit bin counters hold)
 * are refused with ISP_INVALID_ARGUMENT. Otherwise as isp_check_frame_device. */
#define ISP_FRAME_CHECK_STATIC_MAX_PIXELS 4294967295ULL
isp_status isp_check_frame_device_static(const isp_synthetic_config* camera, const float* lut, int32_t lut_size,
                                         const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                                         const float* stats, isp_frame_check_result* result,
                                         isp_stream_t stream);

/* The PSNR in dB of a checked frame (check_op.cpp:226-228): 99 when the error is 0. */
double isp_frame_check_psnr(const isp_frame_check_result* result, int32_t width, int32_t height);

#ifdef __cplusplus
}
#endif

#endif /* ISP_SOURCE_H_ */
