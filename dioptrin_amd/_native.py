"""ctypes binding of libdioptrin_hip.so (the C ABI in include/dioptrin_hip.h).

The shared library is the only compute path: if it is missing or cannot be
loaded this module raises, it never falls back to Python or CPU code.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdioptrin_hip.so")

DIO_OK, DIO_ERR_UNAVAILABLE, DIO_ERR_UNSUPPORTED, DIO_ERR_INVALID, DIO_ERR_DEVICE = range(5)
METHODS = {"counting": 0, "pixel_split": 1, "pyfai_full": 2, "supersampled": 3}
UNITS = {"2th_rad": 0, "2th_deg": 1, "q_A^-1": 2, "q_nm^-1": 3}
PRECISIONS = {"f32": 0, "f32_compensated": 1, "f64": 2, "f32_wide": 3}
ORDERS = {"detector": 0, "flip_vertical": 1}
DTYPES = {"uint16": 0, "uint32": 1, "int32": 2, "float32": 3, "float64": 4}
WITH_ERRORS, WITH_AZIMUTHAL_STD, DEVICE_FRAMES, RAW_SUMS = 1, 2, 4, 8

# Exported symbols, in include/dioptrin_hip.h order.
EXPORTS = [
    "dio_last_error", "dio_version", "dio_open", "dio_close", "dio_device_info_get",
    "dio_plan_create", "dio_plan_destroy", "dio_plan_info_get", "dio_plan_radial",
    "dio_plan_azimuthal", "dio_plan_export_csr", "dio_plan_denominators", "dio_plan_normalization",
    "dio_plan_counts",
    "dio_integrate", "dio_integrate_device", "dio_integrate_device_noalloc",
    "dio_integrate_masked",
    "dio_integrate_masked_device", "dio_integrate_variances", "dio_multi_integrate",
    "dio_integrate_bslz4", "dio_bslz4_scan", "dio_bslz4_decode_device", "dio_bslz4_compress",
    "dio_bslz4_bound",
    "dio_plan_autotune", "dio_plan_set_tile",
    "dio_graph_create", "dio_graph_launch", "dio_graph_destroy", "dio_plan_time_gather",
    "dio_plan_time_passes", "dio_device_alloc", "dio_device_free", "dio_memcpy",
    "dio_host_alloc", "dio_host_free", "dio_synchronize",
    "dio_synthetic_poisson_u32", "dio_server_create", "dio_server_submit", "dio_server_poll",
    "dio_server_stats_get", "dio_server_destroy",
]


class DioError(RuntimeError):
    """Device or runtime failure (DIO_ERR_DEVICE)."""


class DioUnavailableError(RuntimeError):
    """No usable device or library (DIO_ERR_UNAVAILABLE)."""


class DioUnsupportedError(ValueError):
    """Configuration not supported on the device (DIO_ERR_UNSUPPORTED)."""


class Poni(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "pixel1", "pixel2", "distance", "poni1", "poni2", "rot1", "rot2", "rot3", "wavelength")]


class PlanDesc(C.Structure):
    _fields_ = [
        ("dim1", C.c_uint32), ("dim2", C.c_uint32), ("method", C.c_int32), ("n_ss", C.c_uint32),
        ("unit", C.c_int32), ("ndim", C.c_uint32), ("n_radial", C.c_uint32),
        ("n_azimuthal", C.c_uint32), ("has_radial_range", C.c_int32),
        ("radial_min", C.c_double), ("radial_max", C.c_double),
        ("has_azimuthal_range", C.c_int32), ("azimuthal_min", C.c_double),
        ("azimuthal_max", C.c_double), ("has_polarization", C.c_int32),
        ("polarization_factor", C.c_double), ("mask", C.c_void_p), ("precision", C.c_int32),
        ("input_order", C.c_int32)]


class PlanInfo(C.Structure):
    _fields_ = [
        ("nbins", C.c_uint32), ("n_radial", C.c_uint32), ("n_azimuthal", C.c_uint32),
        ("nnz", C.c_uint64), ("nnz_masked", C.c_uint64), ("weighted", C.c_uint32),
        ("n_segments", C.c_uint32), ("uncertain_angles", C.c_uint64),
        ("singular_pixels", C.c_uint32), ("build_ms", C.c_double), ("hist_min", C.c_double),
        ("hist_step", C.c_double), ("device_bytes", C.c_uint64),
        ("max_tile_groups", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_int32),
                ("xcds", C.c_int32), ("total_memory", C.c_uint64),
                ("runtime_version", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("plan_cache_entries", C.c_uint32),
                ("hbm_budget_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64), ("failed", C.c_uint64), ("frames", C.c_uint64),
        ("queued", C.c_uint64), ("inflight", C.c_uint32), ("peak_inflight", C.c_uint32),
        ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64), ("plan_builds", C.c_uint64),
        ("cache_hits", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
        ("mean_ms", C.c_double), ("max_ms", C.c_double), ("span_s", C.c_double),
        ("frames_per_s", C.c_double), ("plan_bytes", C.c_uint64),
        ("scratch_bytes", C.c_uint64), ("request_bytes", C.c_uint64),
        ("plans_resident", C.c_uint32), ("builds_pending", C.c_uint32)]


class Bslz4Block(C.Structure):
    _fields_ = [("src_off", C.c_uint32), ("comp_len", C.c_uint32), ("dst_off", C.c_uint32),
                ("uncomp_bytes", C.c_uint32)]


class RawSums(C.Structure):
    _fields_ = [("sum_signal", C.c_void_p), ("sum_norm", C.c_void_p),
                ("sum_variance", C.c_void_p), ("count", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DioUnavailableError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                f"g.build()' or make -C dioptrin_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp, u64, u32, i32, dp = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_void_p
        sig = {
            "dio_last_error": (C.c_char_p, []),
            "dio_version": (C.c_char_p, []),
            "dio_open": (i32, [i32, C.POINTER(vp)]),
            "dio_close": (i32, [vp]),
            "dio_device_info_get": (i32, [vp, C.POINTER(DeviceInfo)]),
            "dio_plan_create": (i32, [vp, C.POINTER(Poni), C.POINTER(PlanDesc), C.POINTER(vp)]),
            "dio_plan_destroy": (i32, [vp]),
            "dio_plan_info_get": (i32, [vp, C.POINTER(PlanInfo)]),
            "dio_plan_radial": (i32, [vp, dp]),
            "dio_plan_azimuthal": (i32, [vp, dp]),
            "dio_plan_export_csr": (i32, [vp, i32, vp, vp, vp, vp, u64]),
            "dio_plan_denominators": (i32, [vp, dp]),
            "dio_plan_normalization": (i32, [vp, dp]),
            "dio_plan_counts": (i32, [vp, dp]),
            "dio_integrate": (i32, [vp, vp, i32, u64, u32, dp, dp, dp]),
            "dio_integrate_device": (i32, [vp, vp, i32, u64, u32, dp, dp, dp, vp]),
            "dio_integrate_device_noalloc": (i32, [vp, vp, i32, u64, u32, dp, dp, dp, vp]),
            "dio_integrate_masked": (i32, [vp, vp, i32, u64, vp, u32, dp, dp]),
            "dio_integrate_masked_device": (i32, [vp, vp, i32, u64, vp, u32, dp, dp, vp]),
            "dio_integrate_variances": (i32, [vp, vp, i32, vp, i32, u64, u32, dp, dp]),
            "dio_multi_integrate": (i32, [u32, vp, vp, vp, vp, vp, vp, u64, u32, dp, dp,
                                          C.POINTER(RawSums)]),
            "dio_integrate_bslz4": (i32, [vp, vp, vp, u64, i32, u32, dp, dp]),
            "dio_bslz4_scan": (i32, [vp, vp, u64, u64, i32, vp, u64, C.POINTER(u64),
                                     C.POINTER(u32), C.POINTER(u32)]),
            "dio_bslz4_decode_device": (i32, [vp, vp, u64, vp, u64, u32, u32, i32, vp, u64, vp,
                                              vp]),
            "dio_bslz4_compress": (i32, [vp, u64, i32, u32, vp, u64, C.POINTER(u64)]),
            "dio_bslz4_bound": (u64, [u64, i32, u32]),
            "dio_plan_autotune": (i32, [vp, i32, u64, u32, C.POINTER(u32), C.POINTER(u32)]),
            "dio_plan_set_tile": (i32, [vp, u32, u32]),
            "dio_graph_create": (i32, [vp, vp, i32, u64, u32, dp, dp, dp, C.POINTER(vp)]),
            "dio_graph_launch": (i32, [vp, vp]),
            "dio_graph_destroy": (i32, [vp]),
            "dio_plan_time_gather": (i32, [vp, vp, i32, u64, u32, dp, dp, i32,
                                           C.POINTER(C.c_double)]),
            "dio_plan_time_passes": (i32, [vp, vp, i32, u64, u32, dp, dp, C.POINTER(C.c_float),
                                           C.POINTER(u32),
                                           u32, C.POINTER(u32)]),
            "dio_device_alloc": (i32, [vp, u64, C.POINTER(vp)]),
            "dio_host_alloc": (i32, [vp, u64, C.POINTER(vp)]),
            "dio_host_free": (i32, [vp, vp]),
            "dio_device_free": (i32, [vp, vp]),
            "dio_memcpy": (i32, [vp, vp, vp, u64, i32]),
            "dio_synchronize": (i32, [vp]),
            "dio_synthetic_poisson_u32": (i32, [vp, vp, u64, vp, u64, u64, vp]),
            "dio_server_create": (i32, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
            "dio_server_submit": (i32, [vp, C.POINTER(Poni), C.POINTER(PlanDesc), vp, i32, u64,
                                        u32, dp, dp, dp, C.POINTER(u64)]),
            "dio_server_poll": (i32, [vp, u64, C.c_double, C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]),
            "dio_server_stats_get": (i32, [vp, C.POINTER(ServerStats)]),
            "dio_server_destroy": (i32, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    """Raise the Python exception matching a status code."""
    if status == DIO_OK:
        return
    msg = lib().dio_last_error().decode(errors="replace")
    if status == DIO_ERR_UNAVAILABLE:
        raise DioUnavailableError(msg)
    if status == DIO_ERR_UNSUPPORTED:
        raise DioUnsupportedError(msg)
    if status == DIO_ERR_INVALID:
        raise ValueError(msg)
    raise DioError(msg)
