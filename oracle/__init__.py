"""CPU oracle for the WLSQ extrapolate-and-image path - TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package. The product (``imaging2d_wlsq_amd``)
never imports it and has no CPU fallback.

Two restatements of the reference live here:

* ``libwlsq_oracle.so`` (``wlsq_oracle.c``): plain-C FP32 restatement of
  ``cpp_src/extrapolation_cpu.cpp:11-107`` (OpenMP over omega, tap-ordered
  unconjugated dot, sequential-omega FP32 imaging, zero halos, write-back with
  shot stride nt*nx), plus the compact-table ("k-table") form and the table
  expansion. Bound below with the reference's CPU argument order
  (``cpp_src/interface.py:50-58``).
* :mod:`oracle.fp64`: vectorised FP64 NumPy restatement of
  ``py_src/wavefieldPropagation/extrapolation.py:38-66``.

How each function is pinned (the reference's own known-answer tests, the
invariants and brute force) is listed in ``tests/test_oracle.py``; every
function here has at least one such check, so none is "parity unpinned".
The reference's own CPU library is not built: it needs ``mkl.h``, which this
image does not have.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwlsq_oracle.so")
_lib = None

_c64p = np.ctypeslib.ndpointer(dtype=np.complex64, flags=("C", "A"))
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags=("C", "A"))
_ci = ctypes.c_int
_cf = ctypes.c_float


def build() -> str:
    """Compile the C oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle.build())")
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_extrapolate.restype = _ci
        L.oracle_extrapolate.argtypes = [_ci] * 7 + [_c64p, _c64p, _c64p, _c64p, _f32p]
        L.oracle_extrapolate_ktable.restype = _ci
        L.oracle_extrapolate_ktable.argtypes = [_ci] * 7 + [
            _f32p, _f32p, _cf, _cf, _ci, _c64p, _c64p, _c64p, _c64p, _f32p]
        L.oracle_expand_table.restype = _ci
        L.oracle_expand_table.argtypes = [_ci] * 4 + [_f32p, _f32p, _cf, _cf, _ci, _c64p, _c64p]
        L.oracle_kappa_index.restype = _ci
        L.oracle_kappa_index.argtypes = [_cf, _cf, _cf, _cf, _ci]
        L.oracle_inv_dkappa.restype = _cf
        L.oracle_inv_dkappa.argtypes = [_cf]
        _lib = L
    return _lib


def extrapolate(ns, nextrap, nz, nx, nf, nt, M, w_op_fs_forw, w_op_fs_back, wave_fs, shot_fs, image):
    """Same arguments and in-place semantics as the reference CPU entry
    (``cpp_src/interface.py:50-58`` -> ``extrapolation_cpu.cpp:11``)."""
    rc = lib().oracle_extrapolate(ns, nextrap, nz, nx, nf, nt, M,
                                  w_op_fs_forw, w_op_fs_back, wave_fs, shot_fs, image)
    if rc != 0:
        raise ValueError(f"oracle_extrapolate failed with status {rc}")


def extrapolate_ktable(ns, nextrap, nz, nx, nf, nt, M, slowness, omega, kappa0, dkappa, nk,
                       table_forw, table_back, wave_fs, shot_fs, image):
    rc = lib().oracle_extrapolate_ktable(ns, nextrap, nz, nx, nf, nt, M,
                                         np.ascontiguousarray(slowness, np.float32),
                                         np.ascontiguousarray(omega, np.float32),
                                         float(kappa0), float(dkappa), int(nk),
                                         table_forw, table_back, wave_fs, shot_fs, image)
    if rc != 0:
        raise ValueError(f"oracle_extrapolate_ktable failed with status {rc}")


def expand_table(nextrap, nf, nx, M, slowness, omega, kappa0, dkappa, nk, table):
    L = 2 * M + 1
    out = np.empty((nextrap, nf, nx, L), np.complex64)
    rc = lib().oracle_expand_table(nextrap, nf, nx, M,
                                   np.ascontiguousarray(slowness, np.float32),
                                   np.ascontiguousarray(omega, np.float32),
                                   float(kappa0), float(dkappa), int(nk),
                                   np.ascontiguousarray(table, np.complex64), out)
    if rc != 0:
        raise ValueError(f"oracle_expand_table failed with status {rc}")
    return out


def kappa_index(omega, slowness, kappa0, dkappa, nk):
    L = lib()
    inv = L.oracle_inv_dkappa(float(dkappa))
    return L.oracle_kappa_index(float(omega), float(slowness), float(kappa0), inv, int(nk))
