"""ctypes mirror of include/pikgpu.h and loader of the HIP library.

The product path is the C ABI in libpikgpu.so (built from pik_amd/csrc);
this module only describes its structs for Python callers. Loading fails
loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpikgpu.so")

ABI_VERSION = 3

OK = 0
EINVAL = -1
EHIP = -2
ENOMEM = -3
ESTATE = -4
EUNSUPPORTED = -5
EDATA = -6

# AC strategy raw types (pik/ac_strategy.h:55-75).
ACS_DCT = 0
ACS_IDENTITY = 1
ACS_DCT2X2 = 2
ACS_DCT4X4 = 3
ACS_DCT16X16 = 4
ACS_DCT32X32 = 5
ACS_LINES = 6
ACS_DCT_NOHF = 7
ACS_DCT4X4_NOHF = 8

# FrameHeader::flags bits (pik/headers.h:185-196).
FLAG_GRADIENT_MAP = 1
FLAG_GRAYSCALE_OPT = 2
FLAG_NOISE = 4

OUTPUT_LINEAR = 1
OUTPUT_SRGB8 = 2

GABORISH_OFF, GABORISH_500, GABORISH_750, GABORISH_875, GABORISH_1000 = range(5)

STAGE_DC_DEQUANT = 0
STAGE_DC = 1
STAGE_COEFFS = 2
STAGE_IDCT = 3
STAGE_GUIDE = 4
STAGE_EPF = 5
STAGE_AR = 6
STAGE_DOTS = 7
STAGE_GABORISH = 8
STAGE_LINEAR = 9
NUM_STAGES = 10
STAGE_NAMES = ["dc_dequant", "dc", "coeffs", "idct", "guide", "epf", "ar",
               "dots", "gaborish", "linear"]

c_float_p = ctypes.POINTER(ctypes.c_float)


class Dot(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_uint32), ("y", ctypes.c_uint32),
        ("xsize", ctypes.c_uint32), ("ysize", ctypes.c_uint32),
        ("transform", ctypes.c_int32),
        ("dx", ctypes.c_int32), ("dy", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("pixels", (ctypes.c_int8 * 64) * 3),
    ]


class FrameC(ctypes.Structure):
    _fields_ = [
        ("xsize", ctypes.c_uint32), ("ysize", ctypes.c_uint32),
        ("gaborish", ctypes.c_int32),
        ("predict_lf", ctypes.c_int32),
        ("predict_hf", ctypes.c_int32),
        ("adaptive_reconstruction", ctypes.c_int32),
        ("epf_enable_adaptive", ctypes.c_int32),
        ("epf_use_sharpened", ctypes.c_int32),
        ("global_scale", ctypes.c_int32),
        ("quant_dc", ctypes.c_int32),
        ("ac", ctypes.c_void_p * 3), ("ac_stride", ctypes.c_size_t),
        ("dc", ctypes.c_void_p * 3), ("dc_stride", ctypes.c_size_t),
        ("quant_field", ctypes.c_void_p), ("quant_field_stride", ctypes.c_size_t),
        ("ac_strategy", ctypes.c_void_p), ("ac_strategy_stride", ctypes.c_size_t),
        ("sigma_lut_ids", ctypes.c_void_p), ("sigma_lut_stride", ctypes.c_size_t),
        ("dequant_cf", ctypes.c_void_p), ("dequant_cf_stride", ctypes.c_size_t),
        ("ytox_map", ctypes.c_void_p), ("ytox_stride", ctypes.c_size_t),
        ("ytob_map", ctypes.c_void_p), ("ytob_stride", ctypes.c_size_t),
        ("ytox_dc", ctypes.c_int32), ("ytob_dc", ctypes.c_int32),
        ("dequant_map", ctypes.c_void_p),
        ("dequant_table", ctypes.c_void_p), ("dequant_table_size", ctypes.c_size_t),
        ("dequant_offsets", ctypes.c_void_p),
        ("num_dequant_offsets", ctypes.c_size_t),
        ("dots", ctypes.c_void_p), ("num_dots", ctypes.c_size_t),
        ("grayscale", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
    ]


class ConfigC(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


_lib = None


ADJ_DEVICE_POINTERS = 1


class AdjIoC(ctypes.Structure):
    """pikgpu_adj_io."""
    _fields_ = [
        ("flags", ctypes.c_uint32),
        ("qf", ctypes.c_void_p),
        ("qf_stride", ctypes.c_size_t),
        ("target", ctypes.c_void_p * 3),
        ("dl_dlinear", ctypes.c_void_p * 3),
        ("in_stride", ctypes.c_size_t),
        ("linear", ctypes.c_void_p * 3),
        ("linear_stride", ctypes.c_size_t),
        ("dl_dqf", ctypes.c_void_p),
        ("dqf_stride", ctypes.c_size_t),
        ("dl_dcoeffs", ctypes.c_void_p * 3),
        ("dcoeffs_stride", ctypes.c_size_t),
        ("loss", ctypes.c_void_p),
    ]


class SearchParamsC(ctypes.Structure):
    """pikgpu_search_params."""
    _fields_ = [
        ("steps", ctypes.c_int32),
        ("learning_rate", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("lambda_", ctypes.c_float),
        ("qf_max", ctypes.c_int32),
        ("trust_low", ctypes.c_float),
        ("trust_high", ctypes.c_float),
    ]


class SearchIoC(ctypes.Structure):
    """pikgpu_search_io."""
    _fields_ = [
        ("opsin", ctypes.c_void_p * 3),
        ("opsin_stride", ctypes.c_size_t),
        ("target", ctypes.c_void_p * 3),
        ("target_stride", ctypes.c_size_t),
        ("weight", ctypes.c_void_p),
        ("weight_stride", ctypes.c_size_t),
        ("quant_field", ctypes.c_void_p),
        ("quant_field_stride", ctypes.c_size_t),
        ("ac", ctypes.c_void_p * 3),
        ("ac_stride", ctypes.c_size_t),
        ("dc", ctypes.c_void_p * 3),
        ("dc_stride", ctypes.c_size_t),
        ("loss", ctypes.c_void_p),
        ("rate", ctypes.c_void_p),
        ("lambda_used", ctypes.c_void_p),
    ]


class AcPassC(ctypes.Structure):
    """pikgpu_ac_pass."""
    _fields_ = [
        ("xsize", ctypes.c_uint32),
        ("ysize", ctypes.c_uint32),
        ("ac_strategy", ctypes.c_void_p),
        ("ac_strategy_stride", ctypes.c_size_t),
        ("num_groups", ctypes.c_size_t),
        ("group_data", ctypes.c_void_p),
        ("group_size", ctypes.c_void_p),
        ("ac", ctypes.c_void_p * 3),
        ("ac_stride", ctypes.c_size_t),
        ("noise", ctypes.c_void_p),
        ("group_consumed", ctypes.c_void_p),
    ]


class DcSectionC(ctypes.Structure):
    """pikgpu_dc_section."""
    _fields_ = [
        ("xsize", ctypes.c_uint32),
        ("ysize", ctypes.c_uint32),
        ("data", ctypes.c_void_p),
        ("size", ctypes.c_size_t),
        ("dc", ctypes.c_void_p * 3),
        ("dc_stride", ctypes.c_size_t),
        ("ac_strategy", ctypes.c_void_p),
        ("ac_strategy_stride", ctypes.c_size_t),
        ("quant_field", ctypes.c_void_p),
        ("quant_field_stride", ctypes.c_size_t),
        ("sigma_lut_ids", ctypes.c_void_p),
        ("sigma_lut_stride", ctypes.c_size_t),
        ("consumed", ctypes.c_size_t),
    ]


BATCH_PHASES = 19
BATCH_QUEUE, BATCH_BRANCHES = 0, 1


class ScheduleC(ctypes.Structure):
    """pikgpu_batch_schedule."""
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("persist", ctypes.c_int32 * BATCH_PHASES),
    ]


class ParsedInfoC(ctypes.Structure):
    """pikgpu_parsed_info."""
    _fields_ = [
        ("frame_size", ctypes.c_uint64),
        ("num_passes", ctypes.c_uint32),
        ("num_groups", ctypes.c_uint32),
        ("epf_sigma", ctypes.c_uint32),
        ("num_downsampling", ctypes.c_uint32),
        ("downsampling", ctypes.c_void_p),
        ("noise", ctypes.c_void_p),
        ("consumed", ctypes.c_size_t),
    ]


def load():
    """Loads libpikgpu.so (the HIP path). Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libpikgpu.so is not built (%s); run __graft_entry__.build()"
            % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.pikgpu_create.restype = vp
    lib.pikgpu_create.argtypes = [ctypes.c_int, ctypes.POINTER(ConfigC)]
    lib.pikgpu_destroy.argtypes = [vp]
    lib.pikgpu_destroy.restype = None
    lib.pikgpu_last_error.argtypes = [vp]
    lib.pikgpu_last_error.restype = ctypes.c_char_p
    lib.pikgpu_frame_upload.argtypes = [vp, ctypes.POINTER(FrameC)]
    lib.pikgpu_run.argtypes = [vp, ctypes.c_int]
    lib.pikgpu_sync.argtypes = [vp]
    lib.pikgpu_download.argtypes = [vp, c_float_p * 3, ctypes.c_size_t]
    lib.pikgpu_reconstruct.argtypes = [vp, ctypes.POINTER(FrameC),
                                       c_float_p * 3, ctypes.c_size_t]
    lib.pikgpu_stage_size.argtypes = [vp, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]
    lib.pikgpu_stage_dump.argtypes = [vp, ctypes.c_int, c_float_p * 3,
                                      ctypes.c_size_t]
    lib.pikgpu_last_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p),
                                        c_float_p, ctypes.c_int]
    lib.pikgpu_stream.argtypes = [vp]
    lib.pikgpu_stream.restype = vp
    lib.pikgpu_reconstruct_noalloc.argtypes = [
        ctypes.POINTER(FrameC), c_float_p * 3, ctypes.c_size_t, vp,
        ctypes.c_char_p, ctypes.c_size_t]
    lib.pikgpu_search_quant_field.argtypes = [
        vp, ctypes.POINTER(SearchParamsC), ctypes.POINTER(SearchIoC)]
    lib.pikgpu_batch_create.restype = vp
    lib.pikgpu_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(ConfigC)]
    lib.pikgpu_batch_destroy.argtypes = [vp]
    lib.pikgpu_batch_destroy.restype = None
    lib.pikgpu_batch_last_error.argtypes = [vp]
    lib.pikgpu_batch_last_error.restype = ctypes.c_char_p
    lib.pikgpu_batch_upload.argtypes = [vp, ctypes.c_void_p, ctypes.c_int]
    lib.pikgpu_batch_run.argtypes = [vp]
    lib.pikgpu_batch_sync.argtypes = [vp]
    lib.pikgpu_batch_last_ms.argtypes = [vp]
    lib.pikgpu_batch_last_ms.restype = ctypes.c_float
    lib.pikgpu_batch_download.argtypes = [vp, ctypes.c_int, c_float_p * 3,
                                          ctypes.c_size_t]
    lib.pikgpu_reconstruct_batch.argtypes = [vp, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p]
    lib.pikgpu_set_output.argtypes = [vp, ctypes.c_int]
    lib.pikgpu_set_downsample.argtypes = [vp, ctypes.c_int]
    lib.pikgpu_quantize.argtypes = [vp, vp, ctypes.c_size_t, vp,
                                    ctypes.c_size_t, vp, ctypes.c_size_t]
    lib.pikgpu_frame_add_pass.argtypes = [vp, ctypes.c_void_p, ctypes.c_size_t]
    lib.pikgpu_download_srgb8.argtypes = [vp, ctypes.c_void_p, ctypes.c_size_t]
    lib.pikgpu_default_dequant_tables.argtypes = [
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    lib.pikgpu_validate_frame.argtypes = [ctypes.POINTER(FrameC),
                                          ctypes.c_char_p, ctypes.c_size_t]
    lib.pikgpu_fwd_adj.argtypes = [vp, ctypes.POINTER(AdjIoC)]
    lib.pikgpu_decode_ac_pass.argtypes = [ctypes.POINTER(AcPassC), ctypes.c_int,
                                          ctypes.c_char_p, ctypes.c_size_t]
    lib.pikgpu_decode_dc_groups.argtypes = [
        ctypes.POINTER(DcSectionC), ctypes.c_int, ctypes.c_char_p,
        ctypes.c_size_t]
    lib.pikgpu_batch_set_schedule.argtypes = [vp, ctypes.POINTER(ScheduleC)]
    lib.pikgpu_batch_add_pass.argtypes = [vp, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_size_t]
    lib.pikgpu_batch_get_schedule.argtypes = [vp, ctypes.POINTER(ScheduleC)]
    lib.pikgpu_batch_phase_name.argtypes = [ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int)]
    lib.pikgpu_batch_phase_name.restype = ctypes.c_char_p
    lib.pikgpu_parse_frame.argtypes = [
        vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
        ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
    lib.pikgpu_parsed_get.argtypes = [vp, ctypes.POINTER(FrameC),
                                      ctypes.POINTER(ParsedInfoC)]
    lib.pikgpu_parsed_pass_ac.argtypes = [vp, ctypes.c_uint32, vp * 3,
                                          ctypes.POINTER(ctypes.c_size_t)]
    lib.pikgpu_parsed_free.argtypes = [vp]
    lib.pikgpu_parsed_free.restype = None
    lib.pikgpu_decode_frame.argtypes = [
        vp, vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
        ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    _lib = lib
    return lib


EXPORTED_SYMBOLS = [
    "pikgpu_create", "pikgpu_destroy", "pikgpu_last_error",
    "pikgpu_frame_upload", "pikgpu_run", "pikgpu_sync", "pikgpu_download",
    "pikgpu_reconstruct", "pikgpu_stage_size", "pikgpu_stage_dump",
    "pikgpu_last_timings", "pikgpu_stream", "pikgpu_validate_frame",
    "pikgpu_default_dequant_tables", "pikgpu_set_output",
    "pikgpu_download_srgb8", "pikgpu_batch_create", "pikgpu_batch_destroy",
    "pikgpu_batch_last_error", "pikgpu_batch_upload", "pikgpu_batch_run",
    "pikgpu_batch_sync", "pikgpu_batch_last_ms", "pikgpu_batch_download",
    "pikgpu_reconstruct_batch", "pikgpu_fwd_adj", "pikgpu_set_downsample", "pikgpu_quantize",
    "pikgpu_frame_add_pass", "pikgpu_reconstruct_noalloc",
    "pikgpu_search_quant_field", "pikgpu_decode_ac_pass",
    "pikgpu_decode_dc_groups", "pikgpu_parse_frame", "pikgpu_parsed_get",
    "pikgpu_parsed_pass_ac", "pikgpu_parsed_free", "pikgpu_decode_frame",
    "pikgpu_batch_set_schedule", "pikgpu_batch_get_schedule",
    "pikgpu_batch_phase_name", "pikgpu_batch_add_pass",
]
