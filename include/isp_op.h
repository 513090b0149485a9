/*
 * isp_op.h - the IspOp operator contract (modules/isp/include/holoscan/isp/isp_op.hpp)
 * over the C ABI, for a runtime without the Holoscan SDK.
 *
 * The operator takes one raw ImageT per compute() from its `raw` input and emits
 * an RGB ImageT on `image`, then, when configured with `stats`, a uint32
 * [4, bins] tensor on `histogram` and a float32 [12] tensor on `stats`, in that
 * order (isp_op.cpp:584-602). The runtime side is two callbacks: one lends an
 * output tensor of a requested shape and dtype (TensorOutputLoan), the other
 * emits it. Tensors follow DLPack's conventions; descriptors carry the ImageT
 * fields the operator reads and passes through (isp_op.cpp:566-583).
 *
 * Raw images may be in device, pinned host or pageable host memory; host ones
 * are staged to the device on the operator's stream with the waits and copies
 * of the reference's table (modules/isp/README.md:146-153), counted in
 * isp_op_statistics(). This library runs the GPU backend only.
 */
#ifndef ISP_OP_H_
#define ISP_OP_H_

#include "isp_c.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Packing (isp_op.hpp:32-36). */
enum { ISP_PACKING_NONE = 0, ISP_PACKING_MIPI_RAW10 = 1, ISP_PACKING_MIPI_RAW12 = 2 };

/* Where a tensor's memory is (DLPack kDLCPU, kDLCUDAHost, kDLCUDA). */
enum { ISP_MEMORY_PAGEABLE = 0, ISP_MEMORY_PINNED = 1, ISP_MEMORY_DEVICE = 2 };

/* DLPack dtype codes. */
enum { ISP_DTYPE_INT = 0, ISP_DTYPE_UINT = 1, ISP_DTYPE_FLOAT = 2 };

/* The ImageEncoding values the operator reads or writes (sensor schema). */
enum {
  ISP_ENCODING_UNKNOWN = 0,
  ISP_ENCODING_RGB8 = 1,
  ISP_ENCODING_RGBA8 = 2,
  ISP_ENCODING_RGB16 = 3,
  ISP_ENCODING_RGB32F = 4,
  ISP_ENCODING_BAYER_RGGB = 5,
  ISP_ENCODING_BAYER_BGGR = 6,
  ISP_ENCODING_BAYER_GRBG = 7,
  ISP_ENCODING_BAYER_GBRG = 8,
  ISP_ENCODING_MONO8 = 9,
  ISP_ENCODING_MONO16 = 10
};

enum { ISP_COLOR_SPACE_UNSPECIFIED = 0, ISP_COLOR_SPACE_SRGB = 1 };

/* ImageT's significant_bits when unknown. */
#define ISP_UNKNOWN_SIGNIFICANT_BITS 255

typedef struct isp_tensor {
  void* data;
  uint8_t memory; /* ISP_MEMORY_* */
  uint8_t dtype_code;
  uint8_t dtype_bits;
  int32_t ndim;        /* 1..4 */
  int64_t shape[4];
  int64_t strides[4];  /* in elements, DLPack's convention */
  void* ready;         /* optional hipEvent_t: the producer's work on this memory */
} isp_tensor;

/* The ImageT fields the operator reads and passes through. `header` is opaque
 * and passed on as it is. */
typedef struct isp_image {
  int32_t width;
  int32_t height;
  uint8_t encoding;         /* ISP_ENCODING_* */
  uint8_t significant_bits; /* 0 or ISP_UNKNOWN_SIGNIFICANT_BITS: unknown */
  uint8_t color_space;
  int32_t roi_offset_x;
  int32_t roi_offset_y;
  int64_t exposure_time_ns;
  float gain;
  const void* header;
  isp_tensor data;
} isp_image;

/* SampleMetadata: kept from the raw sample on every output (isp_op.cpp:583). */
typedef struct isp_sample_metadata {
  int64_t capture_timestamp_ns;
  uint64_t frame_id;
} isp_sample_metadata;

/* IspOpConfig (isp_op.hpp:38-55), with isp_op_config_init's defaults. */
typedef struct isp_op_config {
  uint8_t backend; /* ISP_BACKEND_GPU (the reference's kCuda), or AUTO */
  isp_params params;
  uint8_t output_format; /* ISP_OUT_RGB8 */
  uint8_t has_pattern;   /* std::optional<BayerPattern> */
  uint8_t pattern;
  int32_t bit_depth; /* 0: the image's significant bits, else the container's */
  uint8_t packing;   /* ISP_PACKING_NONE */
  uint8_t stats;     /* declare the histogram and stats outputs */
  int32_t max_width; /* 3840, 2..16384 */
  int32_t max_height; /* 2160, 2..16384 */
} isp_op_config;

/* A port as setup() declares it (isp_op.cpp:323-344). */
typedef struct isp_op_port {
  const char* name;
  uint8_t is_input;
  uint32_t queue_depth;  /* inputs */
  uint8_t memory;        /* outputs: ISP_MEMORY_DEVICE for the GPU backend */
  uint8_t dtype_code;
  uint8_t dtype_bits;
  int32_t rank;
  uint64_t max_bytes;
} isp_op_port;

/* IspOpStatistics (isp_op.hpp:62-70). */
typedef struct isp_op_statistics_t {
  uint64_t host_waits;
  uint64_t copies;
} isp_op_statistics_t;

/* The runtime's side of compute(): lend an output tensor of `request`'s shape,
 * dtype and memory (nonzero return: refused), and emit it; for `image`, with its
 * descriptor. Both return 0 on success. */
typedef struct isp_op_runtime {
  void* user;
  int (*allocate)(void* user, const char* port, const isp_tensor* request, isp_tensor* out);
  int (*emit)(void* user, const char* port, const isp_image* image, const isp_tensor* tensor,
              const isp_sample_metadata* metadata);
} isp_op_runtime;

typedef struct isp_op isp_op;

void isp_op_config_init(isp_op_config* config);
/* validate_config (isp_op.hpp:58, isp_op.cpp:267-281). */
isp_status isp_op_validate_config(const isp_op_config* config);
/* IspOp(IspOpConfig) (isp_op.cpp:315-319): validates, resolves the backend. */
isp_status isp_op_create(const isp_op_config* config, isp_op** out);
isp_status isp_op_destroy(isp_op* op);
/* backend() (isp_op.hpp:90). */
uint8_t isp_op_backend(const isp_op* op);
/* setup() (isp_op.cpp:323-344): the declared ports, at most `capacity`. */
isp_status isp_op_ports(const isp_op* op, isp_op_port* ports, int32_t capacity, int32_t* count);
/* start() / stop() (isp_op.cpp:352-368): build and release the pipeline; a start
 * failure is reported by the next compute(). */
isp_status isp_op_start(isp_op* op);
isp_status isp_op_stop(isp_op* op);
/* compute() (isp_op.cpp:370-603): one raw sample; `stream` is the operator's
 * hipStream_t. ISP_INVALID_ARGUMENT / ISP_FAILURE as kInvalidArgument /
 * kFailure, with the reference's "IspOp: ..." messages. */
isp_status isp_op_compute(isp_op* op, const isp_image* raw, const isp_sample_metadata* metadata,
                          isp_stream_t stream, const isp_op_runtime* runtime);
/* isp_op_statistics() (isp_op.hpp:71, isp_op.cpp:262-265). */
isp_op_statistics_t isp_op_statistics(void);

#ifdef __cplusplus
}
#endif

#endif /* ISP_OP_H_ */
