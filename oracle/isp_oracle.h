/*
 * isp_oracle.h - TEST INFRASTRUCTURE ONLY: the CPU oracle of the ISP.
 *
 * A plain-C restatement of the reference's CPU backend
 * (modules/isp/src/cpu_pipeline.cpp, modules/isp/src/isp_math.hpp,
 * modules/isp/src/isp_params.cpp:166-182), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker of
 * the HIP path. It is never linked into, loaded by or called from the product
 * library. Each function in isp_oracle.c cites the reference lines it follows.
 *
 * Parity pins: see the header of isp_oracle.c.
 */
#ifndef ISP_ORACLE_H_
#define ISP_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/isp_c.h"
#include "../include/isp_source.h"
#include "../include/isp_color.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_pipeline oracle_pipeline;

/* IspPipeline(Backend::kCpu, params) without validation: copies the tables. */
int oracle_create(const isp_params* params, oracle_pipeline** out);
void oracle_destroy(oracle_pipeline* pipeline);
/* Forgets the white balance estimate. */
void oracle_reset(oracle_pipeline* pipeline);
/* One frame in host memory (cpu_process); `threads` > 1 splits every pass
 * into row bands (the sums are integers, so the result is the same). */
int oracle_process(oracle_pipeline* pipeline, const isp_raw_view* raw, const isp_rgb_view* out,
                   const isp_stats_out* stats, const isp_frame_exposure* exposure, int threads);
/* The white balance state after the last frame: gains[3], valid. */
void oracle_awb_state(const oracle_pipeline* pipeline, float gains[3], int* valid);

/* The white balance estimate split in two, for replaying a long sequence of
 * frames cheaply: the statistics of one frame (the white-balance pass of
 * oracle_process, cpu_pipeline.cpp:101-119), and their effect on the state
 * (finish_white_balance, cpu_pipeline.cpp:120-122). Measuring frame n and
 * applying it leaves the state oracle_process(frame n) leaves. */
typedef struct oracle_awb_stats oracle_awb_stats;
oracle_awb_stats* oracle_awb_measure(const oracle_pipeline* pipeline, const isp_raw_view* raw, int threads);
void oracle_awb_apply(oracle_pipeline* pipeline, const oracle_awb_stats* stats);
void oracle_awb_free(oracle_awb_stats* stats);

/* Pieces, for unit tests against closed forms. */
int oracle_tone_lut(const isp_params* params, float* out, int capacity);
int oracle_mirror(int v, int n);
int oracle_load_raw(const isp_raw_view* raw, int x, int y);
/* finish_stats on given sums (isp_math.hpp:387-425). */
void oracle_finish_stats(const uint64_t sums[4], int64_t pixels, const float gains[3],
                         const isp_auto_exposure* ae, const isp_frame_exposure* exposure,
                         float* out);
/* oracle_finish_stats over `count` cases; out holds ISP_STATS_SIZE floats per case. */
void oracle_finish_stats_cases(const isp_epilogue_case* cases, float* out, int count);
/* histogram_percentile (isp_math.hpp:336-352) over a 1024-bin histogram. */
float oracle_histogram_percentile(const uint32_t* hist, float percentile);

/* The frame checker of the reference's demo on the host (check_frame_host,
 * applications/isp/src/check_op.cpp:44-72): `truth` holds the camera's linear
 * truth, width*height*3 floats (synthetic_truth per pixel); `rgb` the RGB8
 * image, rows `pitch` bytes apart; `histogram` the 4 x 256 rows to compare;
 * `stats` (may be NULL: zeros) is copied into the result. */
void oracle_frame_check(const float* truth, int width, int height, const float* lut, int lut_size,
                        const uint8_t* rgb, size_t pitch, const uint32_t* histogram,
                        const float* stats, isp_frame_check_result* result);
/* The PSNR of a checked frame (check_op.cpp:226-228). */
double oracle_frame_psnr(const isp_frame_check_result* result, int width, int height);

/* The colour conversion module's CPU converter (reference_convert,
 * modules/color-convert/src/reference.cpp:14-27) at one size, from an RGB
 * layout (RGB8, BGR8, RGBA8, BGRA8) into a YCbCr one (NV12, NV16, I420, YUYV,
 * UYVY) or MONO8, views in host memory. 0, or -1 for a pair or size it does not
 * convert. */
int oracle_color_convert(const isp_image_view* source, const isp_image_view* destination, int matrix,
                         int range);

#ifdef __cplusplus
}
#endif

#endif /* ISP_ORACLE_H_ */
