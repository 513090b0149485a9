"""ctypes mirror of include/isp_c.h, include/isp_synthetic.h, include/isp_source.h, include/isp_color.h
and include/isp_op.h.

The field order and types follow the C headers exactly; the names follow the
reference's holoscan::isp types (modules/isp/include/holoscan/isp/isp_params.hpp).
"""
import ctypes as C

# enums (isp_params.hpp:21-61)
RGGB, BGGR, GRBG, GBRG = 0, 1, 2, 3
RAW_U8, RAW_U16, RAW_RAW10, RAW_RAW12 = 0, 1, 2, 3
BILINEAR, MALVAR = 0, 1
WB_MANUAL, WB_GRAY_WORLD, WB_WHITE_PATCH = 0, 1, 2
TONE_SRGB, TONE_GAMMA, TONE_LINEAR, TONE_LUT = 0, 1, 2, 3
OUT_RGB8, OUT_RGBA8, OUT_RGB16, OUT_RGB32F = 0, 1, 2, 3
BACKEND_GPU, BACKEND_CPU, BACKEND_AUTO = 0, 1, 2
LOCATION_HOST, LOCATION_DEVICE = 0, 1
SCENE_CHART, SCENE_GRAY = 0, 1

TONE_LUT_SIZE = 4097
HISTOGRAM_BINS = 256
HISTOGRAM_BINS_WIDE = 1024
HISTOGRAM_CHANNELS = 4
STATS_SIZE = 12
(STAT_MEAN_R, STAT_MEAN_G, STAT_MEAN_B, STAT_MEAN_LUMA, STAT_AE_ERROR_EV, STAT_EXPOSURE_SCALE,
 STAT_SUGGESTED_EXPOSURE_NS, STAT_SUGGESTED_GAIN, STAT_WB_GAIN_R, STAT_WB_GAIN_G, STAT_WB_GAIN_B,
 STAT_SATURATED_FRACTION) = range(12)

OK, INVALID_ARGUMENT, FAILURE, OUT_OF_MEMORY, NOT_READY = 0, 1, 2, 3, 4


class LensShading(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32),
                ("gains", C.POINTER(C.c_float)), ("gains_count", C.c_int64)]


class AutoExposure(C.Structure):
    _fields_ = [("target", C.c_float), ("damping", C.c_float),
                ("min_exposure_ns", C.c_int64), ("max_exposure_ns", C.c_int64),
                ("min_gain", C.c_float), ("max_gain", C.c_float)]


class Params(C.Structure):
    _fields_ = [("black_level", C.c_float),
                ("has_channel_black_levels", C.c_uint8), ("channel_black_levels", C.c_float * 4),
                ("white_level", C.c_float),
                ("has_lens_shading", C.c_uint8), ("lens_shading", LensShading),
                ("white_balance", C.c_uint8), ("wb_gains", C.c_float * 3),
                ("awb_smoothing", C.c_float), ("white_patch_percentile", C.c_float),
                ("demosaic", C.c_uint8), ("ccm", C.c_float * 9),
                ("tone_curve", C.c_uint8), ("gamma", C.c_float),
                ("lut", C.POINTER(C.c_float)), ("lut_size", C.c_int32),
                ("sharpen", C.c_float), ("auto_exposure", AutoExposure),
                ("histogram_bins", C.c_int32)]


BATCH_MAX_FRAMES = 64
BATCH_ORDER_LARGEST_FIRST, BATCH_ORDER_AS_GIVEN = 0, 1
BATCH_QUEUE_DYNAMIC, BATCH_QUEUE_STATIC, BATCH_QUEUE_STATIC_XCD = 0, 1, 2


class RawView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("pitch", C.c_size_t), ("format", C.c_uint8), ("bit_depth", C.c_int32),
                ("pattern", C.c_uint8)]


class RgbView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("pitch", C.c_size_t), ("format", C.c_uint8)]


class StatsOut(C.Structure):
    _fields_ = [("histogram", C.c_void_p), ("stats", C.c_void_p)]


class FrameExposure(C.Structure):
    _fields_ = [("exposure_ns", C.c_int64), ("gain", C.c_float)]


class EpilogueCase(C.Structure):
    _fields_ = [("sums", C.c_uint64 * 4), ("pixels", C.c_int64), ("gains", C.c_float * 3),
                ("ae", AutoExposure), ("exposure", FrameExposure)]


class BatchFrame(C.Structure):
    _fields_ = [("pipeline", C.c_void_p), ("raw", RawView), ("out", RgbView),
                ("stats", StatsOut), ("exposure", FrameExposure)]


class BatchSchedule(C.Structure):
    _fields_ = [("tile_height", (C.c_int32 * 2) * 4), ("tiles_per_item", (C.c_int32 * 2) * 4),
                ("awb_tiles_per_item", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("order", C.c_int32), ("queue", C.c_int32), ("graph", C.c_int32), ("per_frame", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int32), ("launch_slots", C.c_int32),
                ("per_frame", C.c_int32), ("schedule", BatchSchedule)]


class CameraConfig(C.Structure):
    _fields_ = [("params", Params), ("width", C.c_int32), ("height", C.c_int32), ("format", C.c_uint8),
                ("bit_depth", C.c_int32), ("pattern", C.c_uint8), ("out_format", C.c_uint8),
                ("stats", C.c_int32), ("initial", FrameExposure)]


class ServerResult(C.Structure):
    _fields_ = [("request", C.c_int64), ("camera", C.c_int32), ("tag", C.c_uint64), ("status", C.c_int),
                ("out", C.c_void_p), ("out_pitch", C.c_size_t), ("histogram", C.c_void_p),
                ("stats", C.c_void_p), ("submit_ns", C.c_int64), ("launch_ns", C.c_int64),
                ("done_ns", C.c_int64), ("batch_frames", C.c_int32)]


class ServerCounters(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("launches", C.c_int64),
                ("queued", C.c_int64), ("budget_used", C.c_uint64), ("budget_peak", C.c_uint64),
                ("max_in_flight", C.c_int32)]


PACKING_NONE, PACKING_MIPI_RAW10, PACKING_MIPI_RAW12 = 0, 1, 2
MEMORY_PAGEABLE, MEMORY_PINNED, MEMORY_DEVICE = 0, 1, 2
DTYPE_INT, DTYPE_UINT, DTYPE_FLOAT = 0, 1, 2
(ENCODING_UNKNOWN, ENCODING_RGB8, ENCODING_RGBA8, ENCODING_RGB16, ENCODING_RGB32F, ENCODING_BAYER_RGGB,
 ENCODING_BAYER_BGGR, ENCODING_BAYER_GRBG, ENCODING_BAYER_GBRG, ENCODING_MONO8, ENCODING_MONO16) = range(11)
COLOR_SPACE_UNSPECIFIED, COLOR_SPACE_SRGB = 0, 1
UNKNOWN_SIGNIFICANT_BITS = 255


class Tensor(C.Structure):
    _fields_ = [("data", C.c_void_p), ("memory", C.c_uint8), ("dtype_code", C.c_uint8),
                ("dtype_bits", C.c_uint8), ("ndim", C.c_int32), ("shape", C.c_int64 * 4),
                ("strides", C.c_int64 * 4), ("ready", C.c_void_p)]


class Image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("encoding", C.c_uint8),
                ("significant_bits", C.c_uint8), ("color_space", C.c_uint8), ("roi_offset_x", C.c_int32),
                ("roi_offset_y", C.c_int32), ("exposure_time_ns", C.c_int64), ("gain", C.c_float),
                ("header", C.c_void_p), ("data", Tensor)]


class SampleMetadata(C.Structure):
    _fields_ = [("capture_timestamp_ns", C.c_int64), ("frame_id", C.c_uint64)]


class OpConfig(C.Structure):
    _fields_ = [("backend", C.c_uint8), ("params", Params), ("output_format", C.c_uint8),
                ("has_pattern", C.c_uint8), ("pattern", C.c_uint8), ("bit_depth", C.c_int32),
                ("packing", C.c_uint8), ("stats", C.c_uint8), ("max_width", C.c_int32),
                ("max_height", C.c_int32)]


class OpPort(C.Structure):
    _fields_ = [("name", C.c_char_p), ("is_input", C.c_uint8), ("queue_depth", C.c_uint32),
                ("memory", C.c_uint8), ("dtype_code", C.c_uint8), ("dtype_bits", C.c_uint8),
                ("rank", C.c_int32), ("max_bytes", C.c_uint64)]


class OpStatistics(C.Structure):
    _fields_ = [("host_waits", C.c_uint64), ("copies", C.c_uint64)]


ALLOCATE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.POINTER(Tensor), C.POINTER(Tensor))
EMIT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.POINTER(Image), C.POINTER(Tensor),
                      C.POINTER(SampleMetadata))


class OpRuntime(C.Structure):
    _fields_ = [("user", C.c_void_p), ("allocate", ALLOCATE_FN), ("emit", EMIT_FN)]


class SyntheticConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pattern", C.c_uint8),
                ("format", C.c_uint8), ("bit_depth", C.c_int32),
                ("black_levels", C.c_float * 4), ("wb_gains", C.c_float * 3),
                ("ccm", C.c_float * 9), ("vignetting", C.c_float), ("kind", C.c_uint8),
                ("exposure", C.c_float), ("frame", C.c_uint32)]


class SceneConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pattern", C.c_uint8),
                ("format", C.c_uint8), ("bit_depth", C.c_int32), ("black_level", C.c_float),
                ("seed", C.c_uint64), ("rectangles", C.c_int32),
                ("corner_falloff", C.c_float), ("noise_gain", C.c_float),
                ("read_sigma", C.c_float), ("illumination", C.c_float * 3)]


class FrameCheckResult(C.Structure):
    _fields_ = [("squared_error", C.c_uint64), ("histogram_mismatches", C.c_uint32),
                ("reserved", C.c_uint32), ("stats", C.c_float * STATS_SIZE)]


FRAME_CHECK_SCRATCH_BYTES = 37120

# include/isp_color.h
(PIX_NV12, PIX_NV16, PIX_I420, PIX_YUYV, PIX_UYVY, PIX_RGB8, PIX_BGR8, PIX_RGBA8, PIX_BGRA8, PIX_MONO8,
 PIX_MONO16, PIX_RGB16) = range(12)
MATRIX_BT601, MATRIX_BT709, MATRIX_BT2020 = range(3)
RANGE_LIMITED, RANGE_FULL = range(2)


class PlaneView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("pitch", C.c_size_t)]


class ImageView(C.Structure):
    _fields_ = [("format", C.c_uint8), ("width", C.c_int32), ("height", C.c_int32), ("planes", PlaneView * 3)]


class ColorParams(C.Structure):
    _fields_ = [("matrix", C.c_uint8), ("range", C.c_uint8)]

P = C.POINTER
_void = C.c_void_p

# (name, restype, argtypes) of every symbol the headers declare.
PROTOTYPES = [
    ("isp_params_init", None, [P(Params)]),
    ("isp_validate_params", C.c_int, [P(Params)]),
    ("isp_validate_raw", C.c_int, [P(RawView), P(Params)]),
    ("isp_raw_row_bytes", C.c_size_t, [C.c_uint8, C.c_int32]),
    ("isp_tone_lut", C.c_int, [P(Params), P(C.c_float), C.c_int32, P(C.c_int32)]),
    ("isp_output_channels", C.c_int32, [C.c_uint8]),
    ("isp_output_channel_bytes", C.c_int32, [C.c_uint8]),
    ("isp_backend_name", C.c_char_p, [C.c_uint8]),
    ("isp_parse_backend", C.c_int, [C.c_char_p, P(C.c_uint8)]),
    ("isp_demosaic_name", C.c_char_p, [C.c_uint8]),
    ("isp_parse_demosaic", C.c_int, [C.c_char_p, P(C.c_uint8)]),
    ("isp_pattern_name", C.c_char_p, [C.c_uint8]),
    ("isp_parse_pattern", C.c_int, [C.c_char_p, P(C.c_uint8)]),
    ("isp_white_balance_name", C.c_char_p, [C.c_uint8]),
    ("isp_parse_white_balance", C.c_int, [C.c_char_p, P(C.c_uint8)]),
    ("isp_probe_platform", C.c_int, []),
    ("isp_resolve_auto", C.c_uint8, [C.c_int]),
    ("isp_resolve_backend", C.c_uint8, [C.c_uint8]),
    ("isp_last_error", C.c_char_p, []),
    ("isp_epilogue_log2", C.c_double, [C.c_double]),
    ("isp_epilogue_exp2", C.c_double, [C.c_double]),
    ("isp_epilogue_finish_stats", None, [P(C.c_uint64), C.c_int64, P(C.c_float), P(AutoExposure),
                                         P(FrameExposure), P(C.c_float)]),
    ("isp_epilogue_finish_stats_device", C.c_int, [_void, _void, C.c_int32, _void]),
    ("isp_create", C.c_int, [P(Params), C.c_uint8, P(_void)]),
    ("isp_destroy", C.c_int, [_void]),
    ("isp_backend", C.c_uint8, [_void]),
    ("isp_location", C.c_uint8, [_void]),
    ("isp_reset", C.c_int, [_void]),
    ("isp_process", C.c_int, [_void, P(RawView), P(RgbView), P(StatsOut), P(FrameExposure), _void]),
    ("isp_timing_begin", C.c_int, [_void]),
    ("isp_timing_end", C.c_int, [_void, P(C.c_double), P(C.c_double), P(C.c_int32)]),
    ("isp_set_exposure_feedback", C.c_int, [_void, C.c_int32, P(FrameExposure)]),
    ("isp_get_exposure_state", C.c_int, [_void, P(FrameExposure)]),
    ("isp_batch_schedule_init", None, [P(BatchSchedule)]),
    ("isp_batch_create", C.c_int, [P(_void)]),
    ("isp_batch_destroy", C.c_int, [_void]),
    ("isp_batch_set_schedule", C.c_int, [_void, P(BatchSchedule)]),
    ("isp_batch_get_schedule", C.c_int, [_void, P(BatchSchedule)]),
    ("isp_batch_set_frames", C.c_int, [_void, P(BatchFrame), C.c_int32]),
    ("isp_batch_launch", C.c_int, [_void, _void]),
    ("isp_batch_synchronize", C.c_int, [_void]),
    ("isp_batch_tune", C.c_int, [_void, _void, C.c_char_p, C.c_int32, P(BatchSchedule), P(C.c_double)]),
    ("isp_server_config_init", None, [P(ServerConfig)]),
    ("isp_server_create", C.c_int, [P(ServerConfig), P(_void)]),
    ("isp_server_destroy", C.c_int, [_void]),
    ("isp_server_add_camera", C.c_int, [_void, P(CameraConfig), P(C.c_int32)]),
    ("isp_server_submit", C.c_int, [_void, C.c_int32, _void, C.c_size_t, C.c_uint8, C.c_uint64, P(C.c_int64)]),
    ("isp_server_poll", C.c_int, [_void, C.c_int64, P(ServerResult)]),
    ("isp_server_release", C.c_int, [_void, C.c_int64]),
    ("isp_server_counters_get", C.c_int, [_void, P(ServerCounters)]),
    ("isp_server_now_ns", C.c_int64, []),
    ("isp_op_config_init", None, [P(OpConfig)]),
    ("isp_op_validate_config", C.c_int, [P(OpConfig)]),
    ("isp_op_create", C.c_int, [P(OpConfig), P(_void)]),
    ("isp_op_destroy", C.c_int, [_void]),
    ("isp_op_backend", C.c_uint8, [_void]),
    ("isp_op_ports", C.c_int, [_void, P(OpPort), C.c_int32, P(C.c_int32)]),
    ("isp_op_start", C.c_int, [_void]),
    ("isp_op_stop", C.c_int, [_void]),
    ("isp_op_compute", C.c_int, [_void, P(Image), P(SampleMetadata), _void, P(OpRuntime)]),
    ("isp_op_statistics", OpStatistics, []),
    ("isp_synthetic_config_init", None, [P(SyntheticConfig)]),
    ("isp_synthetic_render", C.c_int, [P(SyntheticConfig), _void, C.c_size_t]),
    ("isp_synthetic_params", C.c_int, [P(SyntheticConfig), C.c_int32, P(Params), P(C.c_float)]),
    ("isp_synthetic_truth", C.c_int, [P(SyntheticConfig), C.c_int32, C.c_int32, P(C.c_float)]),
    ("isp_synthetic_truth_image", C.c_int, [P(SyntheticConfig), P(C.c_float)]),
    ("isp_scene_config_init", None, [P(SceneConfig)]),
    ("isp_scene_render", C.c_int, [P(SceneConfig), _void, C.c_size_t]),
    ("isp_scene_lens_shading", C.c_int, [P(SceneConfig), C.c_int32, C.c_int32, P(C.c_float)]),
    ("isp_synthetic_render_device", C.c_int, [P(SyntheticConfig), _void, C.c_size_t, _void]),
    ("isp_check_frame_device", C.c_int, [P(SyntheticConfig), _void, C.c_int32, _void, C.c_size_t, _void,
                                         _void, _void, _void, _void]),
    ("isp_check_frame_device_static", C.c_int, [P(SyntheticConfig), _void, C.c_int32, _void, C.c_size_t,
                                                _void, _void, _void, _void]),
    ("isp_frame_check_psnr", C.c_double, [P(FrameCheckResult), C.c_int32, C.c_int32]),
    ("isp_color_params_init", None, [P(ColorParams)]),
    ("isp_color_plane_size", C.c_int, [C.c_uint8, C.c_int32, C.c_int32, C.c_int32, P(C.c_size_t), P(C.c_int32)]),
    ("isp_color_convert_device", C.c_int, [P(ImageView), P(ImageView), P(ColorParams), _void]),
]


def bind(lib, prototypes=PROTOTYPES):
    """Sets restype/argtypes of every prototype on `lib`; raises if one is missing."""
    for name, restype, argtypes in prototypes:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib
