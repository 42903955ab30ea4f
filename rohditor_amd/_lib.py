"""ctypes binding of the librhd C ABI (include/rhd.h).

The library is loaded from this package directory only; a missing library is
an error, never a fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librhd.so"

OK, CANCELLED, UNSUPPORTED, OVER_BUDGET, INVALID_DIMENSIONS, BASE_MISMATCH, \
    UNSUPPORTED_EDITS, SYNCHRONIZATION, INVALID_INPUT, READBACK, RAW_UNSUPPORTED, \
    RAW_CORRUPT = range(12)
STATUS_NAMES = ["Ok", "Cancelled", "Unsupported", "OverBudget", "InvalidDimensions",
                "BaseMismatch", "UnsupportedEdits", "Synchronization", "InvalidInput",
                "Readback", "RawUnsupported", "RawCorrupt"]
ARW_NONE, ARW_COMPRESSED, ARW_UNCOMPRESSED16 = range(3)


class SensorDesc(ctypes.Structure):
    _fields_ = [
        ("sensor_width", c_uint32), ("sensor_height", c_uint32), ("row_stride", c_uint64),
        ("sensor_pattern", c_int32), ("crop_x", c_uint32), ("crop_y", c_uint32),
        ("crop_width", c_uint32), ("crop_height", c_uint32),
        ("black_levels", POINTER(c_float)), ("black_repeat_width", c_uint32),
        ("black_repeat_height", c_uint32), ("white_levels", POINTER(c_float)),
        ("white_count", c_uint32), ("highlight", c_int32), ("clip_levels", c_float * 3),
        ("detection_levels", c_float * 3), ("has_neutral_ceiling", c_int32),
        ("neutral_ceiling", c_float * 3), ("demosaic", c_int32),
        ("demosaic_balance", c_float * 3),
    ]


class DetailDesc(ctypes.Structure):
    _fields_ = [
        ("capture_enabled", c_int32), ("amount", c_float), ("sigma", c_float),
        ("noise_protection", c_float), ("ceilings", c_float * 3), ("iterations", c_int32),
        ("false_colour", c_int32), ("balance", c_float * 3), ("nr_chroma", c_float),
        ("nr_luminance", c_float), ("nr_luminance_detail", c_float),
    ]


class ArwLayout(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("bits_per_sample", c_uint32),
                ("compression", c_uint32), ("offset", c_uint64), ("bytes", c_uint64),
                ("curve", ctypes.c_uint16 * 4), ("kind", c_int32)]


class NoiseProfile(ctypes.Structure):
    _fields_ = [("a", c_float), ("b", c_float), ("bins", c_uint32)]


class ClipStats(ctypes.Structure):
    _fields_ = [("affected_sites", c_uint64), ("changed_sites", c_uint64),
                ("nominal_over_white_sites", c_uint64), ("affected_by_channel", c_uint64 * 3)]


class OpposedStats(ctypes.Structure):
    _fields_ = [("suspected_clipped_sites", c_uint64), ("reconstructed_sites", c_uint64),
                ("changed_sites", c_uint64), ("fallback_sites", c_uint64),
                ("fully_unsupported_sites", c_uint64), ("neutralised_sites", c_uint64),
                ("suspected_by_channel", c_uint64 * 3), ("regions", c_uint32)]


class Metrics(ctypes.Structure):
    _fields_ = [
        ("normalize_ms", c_float), ("highlight_ms", c_float), ("demosaic_ms", c_float),
        ("capture_ms", c_float), ("total_ms", c_float), ("uploaded_bytes", c_uint64),
        ("readback_bytes", c_uint64), ("estimated_gpu_bytes", c_uint64), ("launches", c_uint32),
        ("highlight_method", c_int32), ("clip", ClipStats), ("opposed", OpposedStats),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [("lanes", c_uint32), ("admission_bytes", c_uint64),
                ("base_cache_bytes", c_uint64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", c_int32), ("lane", c_uint32), ("base_cache_hit", c_int32),
                ("queued_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("megapixels", ctypes.c_double),
                ("metrics", Metrics)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("cancelled", c_uint64),
                ("coalesced", c_uint64), ("failed", c_uint64), ("base_hits", c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("megapixels", ctypes.c_double),
                ("span_ms", ctypes.c_double), ("base_cache_bytes", c_uint64),
                ("running_bytes", c_uint64)]


class OpticsDesc(ctypes.Structure):
    """rhd_optics_desc: a resolved OpticsExecution (crates/optics/src/plan.rs:36-48)."""
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("norm_scale", ctypes.c_double),
                ("norm_unscale", ctypes.c_double), ("center_x", ctypes.c_double),
                ("center_y", ctypes.c_double), ("scale", c_float), ("distortion", c_int32),
                ("distortion_k", c_float * 3), ("tca", c_int32), ("tca_red", c_float * 3),
                ("tca_blue", c_float * 3), ("vignetting", c_int32),
                ("vignetting_k", c_float * 3)]


# Every symbol include/rhd.h declares, with (restype, argtypes).
_P = c_void_p
SIGNATURES = {
    "rhd_create": (c_int, [c_int, c_uint64, POINTER(_P)]),
    "rhd_destroy": (None, [_P]),
    "rhd_set_budget": (c_int, [_P, c_uint64]),
    "rhd_reserved_bytes": (c_uint64, [_P]),
    "rhd_last_error": (c_char_p, [_P]),
    "rhd_status_name": (c_char_p, [c_int]),
    "rhd_status_is_routing": (c_int, [c_int]),
    "rhd_supports": (c_int, [POINTER(SensorDesc), POINTER(DetailDesc)]),
    "rhd_mosaic_upload": (c_int, [_P, _P, c_uint64, c_uint32, c_uint32, POINTER(_P)]),
    "rhd_mosaic_release": (None, [_P]),
    "rhd_normalize": (c_int, [_P, _P, POINTER(SensorDesc), _P, POINTER(_P), POINTER(Metrics)]),
    "rhd_highlight": (c_int, [_P, _P, POINTER(SensorDesc), _P, POINTER(Metrics)]),
    "rhd_demosaic": (c_int, [_P, _P, POINTER(SensorDesc), POINTER(DetailDesc), _P, POINTER(_P),
                             POINTER(Metrics)]),
    "rhd_capture": (c_int, [_P, _P, POINTER(DetailDesc), _P, POINTER(Metrics)]),
    "rhd_capture_noalloc": (c_int, [_P, _P, c_uint32, c_uint32, POINTER(DetailDesc)]),
    "rhd_develop_resident": (c_int, [_P, _P, POINTER(SensorDesc), POINTER(DetailDesc), _P,
                                     POINTER(_P), POINTER(Metrics)]),
    "rhd_develop": (c_int, [_P, _P, c_uint64, POINTER(SensorDesc), POINTER(DetailDesc), _P,
                            POINTER(_P), POINTER(Metrics)]),
    "rhd_plane_read": (c_int, [_P, _P, _P, c_uint64]),
    "rhd_plane_write": (c_int, [_P, _P, _P, c_uint64]),
    "rhd_plane_upload": (c_int, [_P, _P, c_uint32, c_uint32, c_int32, POINTER(_P)]),
    "rhd_plane_dims": (None, [_P, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int32)]),
    "rhd_plane_release": (None, [_P]),
    "rhd_source_read_rgb": (c_int, [_P, _P, _P, c_uint64]),
    "rhd_source_upload": (c_int, [_P, _P, c_uint32, c_uint32, POINTER(_P)]),
    "rhd_source_planes": (c_int, [_P, POINTER(_P), POINTER(_P), POINTER(_P)]),
    "rhd_source_dims": (None, [_P, POINTER(c_uint32), POINTER(c_uint32)]),
    "rhd_source_release": (None, [_P]),
    "rhd_libm_eval": (c_int, [_P, c_int32, _P, _P, c_uint64]),
    "rhd_synchronize": (c_int, [_P]),
    "rhd_profile_enable": (c_int, [_P, c_int]),
    "rhd_profile_reset": (c_int, [_P]),
    "rhd_profile_only": (c_int, [_P, c_char_p]),
    "rhd_profile_count": (c_uint32, [_P]),
    "rhd_profile_entry": (c_int, [_P, c_uint32, POINTER(c_char_p), POINTER(ctypes.c_double),
                                  POINTER(c_uint64), POINTER(c_uint64)]),
    "rhd_preview_dimensions": (c_int, [c_uint32, c_uint32, c_uint32, POINTER(c_uint32),
                                       POINTER(c_uint32)]),
    "rhd_optics_find_scale": (c_int, [POINTER(OpticsDesc), POINTER(c_float)]),
    "rhd_reduce_preview": (c_int, [_P, _P, POINTER(OpticsDesc), c_uint32, c_uint32, _P,
                                   POINTER(_P), POINTER(c_float)]),
    "rhd_optics_full": (c_int, [_P, _P, POINTER(OpticsDesc), _P, POINTER(_P), POINTER(c_float)]),
    "rhd_color_transform": (c_int, [_P, _P, POINTER(c_float), POINTER(c_float), _P,
                                    POINTER(c_float)]),
    "rhd_arw_probe": (c_int, [_P, c_uint64, POINTER(ArwLayout)]),
    "rhd_mosaic_decode_arw": (c_int, [_P, _P, c_uint64, POINTER(ArwLayout), POINTER(_P),
                                      POINTER(c_float)]),
    "rhd_mosaic_read": (c_int, [_P, _P, _P, c_uint64]),
    "rhd_noise_profiles": (c_int, [_P, _P, POINTER(c_float), POINTER(NoiseProfile),
                                   POINTER(c_int32), POINTER(c_float)]),
    "rhd_denoise_chroma": (c_int, [_P, _P, c_float, POINTER(c_float), POINTER(c_float),
                                   POINTER(NoiseProfile), POINTER(c_float)]),
    "rhd_denoise_luma": (c_int, [_P, _P, c_float, c_float, POINTER(c_float), POINTER(c_float),
                                 POINTER(NoiseProfile), POINTER(c_float)]),
    "rhd_server_create": (c_int, [_P, POINTER(ServerConfig), POINTER(_P)]),
    "rhd_server_destroy": (None, [_P]),
    "rhd_server_put_mosaic": (c_int, [_P, c_uint64, _P, c_uint64, c_uint32, c_uint32]),
    "rhd_server_put_arw": (c_int, [_P, c_uint64, _P, c_uint64, POINTER(ArwLayout)]),
    "rhd_server_drop_mosaic": (c_int, [_P, c_uint64]),
    "rhd_server_submit": (c_int, [_P, c_uint64, POINTER(SensorDesc), POINTER(DetailDesc),
                                  POINTER(c_uint64)]),
    "rhd_server_submit_spatial": (c_int, [_P, c_uint64, POINTER(SensorDesc), POINTER(DetailDesc),
                                          POINTER(OpticsDesc), c_uint32, c_uint32,
                                          POINTER(c_uint64)]),
    "rhd_server_wait": (c_int, [_P, c_uint64, c_int32, POINTER(_P), POINTER(RequestInfo)]),
    "rhd_server_read_stats": (c_int, [_P, POINTER(ServerStats)]),
    "rhd_server_reset_stats": (c_int, [_P]),
    "rhd_batch_create": (c_int, [_P, c_uint32, POINTER(_P), POINTER(SensorDesc),
                                 POINTER(DetailDesc), c_uint32, POINTER(_P)]),
    "rhd_batch_run": (c_int, [_P, POINTER(c_int32), POINTER(Metrics), POINTER(c_float)]),
    "rhd_batch_source": (_P, [_P, c_uint32]),
    "rhd_batch_lanes": (c_uint32, [_P]),
    "rhd_batch_destroy": (None, [_P]),
}

_lib = None


def load():
    """Load librhd.so from the package directory, binding every declared symbol."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"librhd.so is not built at {LIB_PATH}; run `python -c 'import __graft_entry__ as g;"
                " g.build()'` (the HIP library is required; there is no CPU fallback)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib
