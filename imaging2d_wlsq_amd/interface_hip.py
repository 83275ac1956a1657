"""ctypes bindings of libwlsq_hip.so in the shape of the reference's
``cpp_src/interface_cuda.py``.

``cpp_src/interface_cuda.py:10-52`` loads three CUDA libraries at import and
binds each one's ``extrapolate`` symbol as ``extrapolate``,
``extrapolate_revOp`` and ``extrapolate_sh_revOp`` with argtypes
``[c_int]*7 + [ndpointer(complex64, flags=("C","A"))]*4 + [ndpointer(float32)]``.
Here one library serves all three names, it is loaded lazily on first use
(``main.py:16-18`` needs every GPU library even for the ``host`` option), and a
non-zero status raises :class:`~imaging2d_wlsq_amd._lib.WlsqError` instead of
failing silently.

Call (same as the reference)::

    extrapolate(ns, nextrap, nz, nt, nf, nx, M,
                w_op_forw, forw_pulse, w_op_back, back_pulse, image)

with ``w_op_*`` (nextrap, nf, nx, 2M+1) complex64, ``*_pulse`` (ns, nt, nx)
complex64 (rows j < nf overwritten with the final wavefields) and ``image``
(ns, nz, nx) float32 (rows l < nextrap written).
"""
from __future__ import annotations

import numpy as np

from ._lib import check, lib


def _c64(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex64 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous numpy complex64 array")
    return a


def _f32(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous numpy float32 array")
    return a


def _check_sizes(ns, nextrap, nz, nt, nf, nx, M, w_op_forw, forw_pulse, w_op_back, back_pulse, image):
    L = 2 * M + 1
    need = {
        "w_op_forw": nextrap * nf * nx * L, "w_op_back": nextrap * nf * nx * L,
        "forw_pulse": ns * nt * nx, "back_pulse": ns * nt * nx, "image": ns * nz * nx,
    }
    got = {"w_op_forw": w_op_forw.size, "w_op_back": w_op_back.size, "forw_pulse": forw_pulse.size,
           "back_pulse": back_pulse.size, "image": image.size}
    for k, n in need.items():
        if got[k] < n:
            raise ValueError(f"{k} has {got[k]} elements, needs {n}")


def extrapolate(ns, nextrap, nz, nt, nf, nx, M, w_op_forw, forw_pulse, w_op_back, back_pulse, image):
    """Drop-in for ``interface_cuda.extrapolate_revOp`` (and the two other
    names): extrapolate-and-image on the GPU, synchronous, in place."""
    _c64(w_op_forw, "w_op_forw"); _c64(forw_pulse, "forw_pulse")
    _c64(w_op_back, "w_op_back"); _c64(back_pulse, "back_pulse"); _f32(image, "image")
    _check_sizes(ns, nextrap, nz, nt, nf, nx, M, w_op_forw, forw_pulse, w_op_back, back_pulse, image)
    check(lib().extrapolate(int(ns), int(nextrap), int(nz), int(nt), int(nf), int(nx), int(M),
                            w_op_forw, forw_pulse, w_op_back, back_pulse, image))


# cpp_src/interface_cuda.py binds the same signature under three names
extrapolate_revOp = extrapolate
extrapolate_sh_revOp = extrapolate


def extrapolate_ktable(ns, nextrap, nz, nt, nf, nx, M, slowness, omega, kappa0, dkappa, nk,
                       table_forw, forw_pulse, table_back, back_pulse, image):
    """Compact-table form on host arrays: the per-point operator is
    ``table[q]`` with q the nearest kappa sample to omega_j*slowness[l, i]."""
    slowness = _f32(np.ascontiguousarray(slowness, np.float32), "slowness")
    omega = _f32(np.ascontiguousarray(omega, np.float32), "omega")
    _c64(table_forw, "table_forw"); _c64(table_back, "table_back")
    _c64(forw_pulse, "forw_pulse"); _c64(back_pulse, "back_pulse"); _f32(image, "image")
    L = 2 * M + 1
    if slowness.size < nextrap * nx or omega.size < nf:
        raise ValueError("slowness needs (nextrap, nx) and omega (nf,) values")
    if table_forw.size < nk * L or table_back.size < nk * L:
        raise ValueError("tables need (nk, 2M+1) values")
    if forw_pulse.size < ns * nt * nx or back_pulse.size < ns * nt * nx or image.size < ns * nz * nx:
        raise ValueError("wavefields need (ns, nt, nx) and image (ns, nz, nx) values")
    check(lib().extrapolate_ktable(int(ns), int(nextrap), int(nz), int(nt), int(nf), int(nx), int(M),
                                   slowness, omega, float(kappa0), float(dkappa), int(nk),
                                   table_forw, forw_pulse, table_back, back_pulse, image))


def c64fft1dforw(data, N):
    """In-place forward FFT of a complex64 array of length N on the GPU
    (``cpp_src/interface.py:36-40``, MKL in the reference)."""
    _c64(data, "data")
    if data.size < N:
        raise ValueError(f"data has {data.size} elements, needs {N}")
    check(lib().c64fft1dforw(data, int(N)))


def c64fft1dback(data, N):
    """In-place backward FFT scaled by 1/N (``cpp_src/interface.py:41-45``)."""
    _c64(data, "data")
    if data.size < N:
        raise ValueError(f"data has {data.size} elements, needs {N}")
    check(lib().c64fft1dback(data, int(N)))
