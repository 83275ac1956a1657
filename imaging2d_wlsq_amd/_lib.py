"""Loads libwlsq_hip.so and declares its C ABI (include/wlsq_hip.h) for ctypes.

There is no CPU fallback: if the library is missing or fails to load, every
entry raises. Build it with ``python -m imaging2d_wlsq_amd.build``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_float, c_int, c_longlong, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwlsq_hip.so")

WLSQ_OK = 0
WLSQ_EINVAL = 1
WLSQ_EHIP = 2
WLSQ_ENOMEM = 3
WLSQ_EUNSUPPORTED = 4
LAYOUT_NATURAL = 0
LAYOUT_TAPMAJOR = 1

EXPORTS = [
    "wlsq_last_error", "wlsq_version", "wlsq_set_device", "extrapolate", "wlsq_plain_device",
    "extrapolate_ktable", "wlsq_ktable_device", "wlsq_expand_table_device", "wlsq_batch_device",
    "wlsq_plan_create", "wlsq_plan_info", "wlsq_plan_launch", "wlsq_plan_destroy",
    "wlsq_graph_create", "wlsq_graph_launch", "wlsq_graph_destroy", "wlsq_batch_points",
    "c64fft1dforw", "c64fft1dback", "wlsq_trace_spectra_device", "wlsq_ktable_device_noalloc",
    "wlsq_batch_workspace_bytes", "wlsq_batch_device_ws", "wlsq_probe_fma", "wlsq_batch_round_ws",
    "wlsq_probe_exact_sum",
]


class WlsqError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"libwlsq_hip status {status}: {message}")
        self.status = status


class Tuning(Structure):
    _fields_ = [("shots_per_group", c_int), ("omegas_per_item", c_int), ("block", c_int),
                ("x_per_thread", c_int), ("lds_budget", c_int), ("tables_conjugate", c_int),
                ("image_mode", c_int), ("tables_symmetric", c_int), ("tap_loads", c_int),
                ("strip_gathers", c_int), ("double_rows", c_int), ("depth_steps", c_int),
                ("image_planes", c_int), ("rounding", c_int), ("strips_per_thread", c_int),
                ("exact_image", c_int)]


class Shot(Structure):
    _fields_ = [("f_off", c_longlong), ("b_off", c_longlong), ("img_off", c_longlong),
                ("slow_off", c_longlong), ("ld", c_int), ("img_ld", c_int), ("slow_ld", c_int),
                ("nx", c_int), ("nz", c_int), ("nf", c_int), ("omega_off", c_int),
                ("first_step", c_int)]


_c64p = np.ctypeslib.ndpointer(dtype=np.complex64, flags=("C", "A"))
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags=("C", "A"))
_lib = None


def lib():
    """The loaded library with argtypes set; raises if it cannot be loaded."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m imaging2d_wlsq_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    ci, cf, vp = c_int, c_float, c_void_p
    L.wlsq_last_error.restype = ctypes.c_char_p
    L.wlsq_last_error.argtypes = []
    L.wlsq_version.restype = ci
    L.wlsq_set_device.restype = ci
    L.wlsq_set_device.argtypes = [ci]
    # drop-in (host arrays): argument order of cpp_src/interface_cuda.py:30-38
    L.extrapolate.restype = ci
    L.extrapolate.argtypes = [ci] * 7 + [_c64p, _c64p, _c64p, _c64p, _f32p]
    L.extrapolate_ktable.restype = ci
    L.extrapolate_ktable.argtypes = [ci] * 7 + [_f32p, _f32p, cf, cf, ci, _c64p, _c64p, _c64p,
                                                _c64p, _f32p]
    # device pointers
    L.wlsq_plain_device.restype = ci
    L.wlsq_plain_device.argtypes = [ci] * 7 + [vp] * 5 + [ci, ci, vp]
    L.wlsq_ktable_device.restype = ci
    L.wlsq_ktable_device.argtypes = [ci] * 7 + [vp, ci, vp, cf, cf, ci, vp, vp, vp, vp, vp, ci,
                                                POINTER(Tuning), vp]
    L.wlsq_ktable_device_noalloc.restype = ci
    L.wlsq_ktable_device_noalloc.argtypes = L.wlsq_ktable_device.argtypes
    L.wlsq_expand_table_device.restype = ci
    L.wlsq_expand_table_device.argtypes = [ci] * 4 + [vp, ci, vp, cf, cf, ci, vp, vp, ci, vp]
    L.wlsq_batch_device.restype = ci
    L.wlsq_batch_device.argtypes = [POINTER(Shot), ci, ci, vp, vp, cf, cf, ci, vp, vp, vp, vp, vp,
                                    ci, POINTER(Tuning), vp]
    L.wlsq_batch_workspace_bytes.restype = c_longlong
    L.wlsq_batch_workspace_bytes.argtypes = [POINTER(Shot), ci, ci, ci, POINTER(Tuning)]
    L.wlsq_batch_device_ws.restype = ci
    L.wlsq_batch_device_ws.argtypes = [POINTER(Shot), ci, ci, vp, vp, cf, cf, ci, vp, vp, vp, vp,
                                       vp, ci, POINTER(Tuning), vp, c_longlong, vp]
    L.wlsq_batch_round_ws.restype = ci
    L.wlsq_batch_round_ws.argtypes = [POINTER(Shot), ci, ci, ci, vp, vp, cf, cf, ci, vp, vp, vp,
                                      vp, vp, POINTER(Tuning), vp, c_longlong, vp]
    L.wlsq_plan_create.restype = ci
    L.wlsq_plan_create.argtypes = [POINTER(vp), POINTER(Shot), ci, ci, POINTER(Tuning)]
    L.wlsq_plan_info.restype = ci
    L.wlsq_plan_info.argtypes = [vp, POINTER(c_longlong), POINTER(ci), POINTER(ci), POINTER(ci),
                                 POINTER(ci), POINTER(ci)]
    L.wlsq_plan_launch.restype = ci
    L.wlsq_plan_launch.argtypes = [vp, vp, vp, cf, cf, ci, vp, vp, vp, vp, vp, ci, vp]
    L.wlsq_plan_destroy.restype = ci
    L.wlsq_plan_destroy.argtypes = [vp]
    L.wlsq_graph_create.restype = ci
    L.wlsq_graph_create.argtypes = [POINTER(vp), vp, vp, vp, cf, cf, ci, vp, vp, vp, vp, vp, ci,
                                    c_longlong]
    L.wlsq_graph_launch.restype = ci
    L.wlsq_graph_launch.argtypes = [vp, vp]
    L.wlsq_graph_destroy.restype = ci
    L.wlsq_graph_destroy.argtypes = [vp]
    L.wlsq_probe_fma.restype = ci
    L.wlsq_probe_fma.argtypes = [ci, ci, vp, vp]
    L.wlsq_probe_exact_sum.restype = ci
    L.wlsq_probe_exact_sum.argtypes = [vp, c_longlong, ci, vp, vp]
    L.wlsq_batch_points.restype = c_longlong
    L.wlsq_batch_points.argtypes = [POINTER(Shot), ci]
    # transforms: cpp_src/interface.py:36-45 binds the MKL pair the same way
    L.c64fft1dforw.restype = ci
    L.c64fft1dforw.argtypes = [_c64p, ci]
    L.c64fft1dback.restype = ci
    L.c64fft1dback.argtypes = [_c64p, ci]
    L.wlsq_trace_spectra_device.restype = ci
    L.wlsq_trace_spectra_device.argtypes = [ci, ci, vp, ci, vp, ci, vp]
    _lib = L
    return L


def check(status: int) -> None:
    if status != WLSQ_OK:
        msg = lib().wlsq_last_error().decode(errors="replace")
        raise WlsqError(status, msg)
