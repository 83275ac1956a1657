"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/wlsq_hip.h, and rejects invalid arguments before any HIP
call (the reference does no validation at all, SURVEY of
cpp_src/extrapolation_revOp_gpu.cu:78-231)."""
import ctypes
import os
import re

import numpy as np
import pytest

from imaging2d_wlsq_amd import _lib, interface_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "wlsq_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt, re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_header_declares_expected_entries():
    names = _declared()
    for n in ["extrapolate", "extrapolate_ktable", "wlsq_plain_device", "wlsq_ktable_device",
              "wlsq_batch_device", "wlsq_plan_create", "wlsq_graph_create", "wlsq_last_error"]:
        assert n in names, n


def test_library_loads_and_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in _declared():
        assert hasattr(L, n), f"{n} declared in include/wlsq_hip.h but not exported"
    assert _lib.lib().wlsq_version() >= 1
    assert set(_lib.EXPORTS) <= set(_declared())


@pytest.mark.parametrize("bad", [
    dict(nz=1, nextrap=2),      # nz < nextrap
    dict(nt=1, nf=2),           # nt < nf
    dict(M=-1),
    dict(ns=-1),
])
def test_invalid_shapes_raise_before_touching_the_device(bad):
    a = dict(ns=1, nextrap=1, nz=1, nt=2, nf=2, nx=4, M=1)
    a.update(bad)
    L = 3
    op = np.zeros((2, 2, 4, L), np.complex64)
    F = np.zeros((1, 2, 4), np.complex64)
    B = np.zeros_like(F)
    img = np.zeros((1, 2, 4), np.float32)
    with pytest.raises(_lib.WlsqError) as e:
        _lib.check(_lib.lib().extrapolate(a["ns"], a["nextrap"], a["nz"], a["nt"], a["nf"], a["nx"],
                                          a["M"], op, F, op, B, img))
    assert e.value.status == _lib.WLSQ_EINVAL
    assert _lib.lib().wlsq_last_error()


def test_python_mirror_checks_types_and_sizes():
    op = np.zeros((1, 2, 4, 3), np.complex64)
    F = np.zeros((1, 2, 4), np.complex64)
    img = np.zeros((1, 1, 4), np.float32)
    with pytest.raises(TypeError):
        interface_hip.extrapolate(1, 1, 1, 2, 2, 4, 1, op.astype(np.complex128), F, op, F.copy(), img)
    with pytest.raises(ValueError):
        interface_hip.extrapolate(1, 1, 1, 2, 2, 5, 1, op, F, op, F.copy(), img)
    # the three reference names are one entry
    assert interface_hip.extrapolate_revOp is interface_hip.extrapolate
    assert interface_hip.extrapolate_sh_revOp is interface_hip.extrapolate


def test_plan_rejects_bad_shots_without_device():
    from imaging2d_wlsq_amd._lib import Shot, Tuning
    arr = (Shot * 1)()
    arr[0] = Shot(0, 0, 0, 0, 4, 4, 4, 8, 3, 2, 0, 0)  # ld < nx
    h = ctypes.c_void_p()
    tun = Tuning()
    rc = _lib.lib().wlsq_plan_create(ctypes.byref(h), arr, 1, 2, ctypes.byref(tun))
    # either the shot check (EINVAL) or, with no GPU here, the device query fails first
    assert rc in (_lib.WLSQ_EINVAL, _lib.WLSQ_EHIP)


def test_transform_arguments_are_checked_before_the_device():
    """c64fft1dforw/back (cpp_src/mkl_fft.cpp:9-39 drop-ins) and the trace
    spectra entry reject bad sizes and strides without a HIP call."""
    L = _lib.lib()
    with pytest.raises(_lib.WlsqError) as e:
        _lib.check(L.c64fft1dforw(np.zeros(4, np.complex64), -1))
    assert e.value.status == _lib.WLSQ_EINVAL
    assert L.c64fft1dback(np.zeros(1, np.complex64), 0) == _lib.WLSQ_OK  # empty: no-op
    assert L.wlsq_trace_spectra_device(4, 8, None, 8, None, 8, None) == _lib.WLSQ_EINVAL
    assert L.wlsq_trace_spectra_device(4, 8, 1024, 7, 2048, 8, None) == _lib.WLSQ_EINVAL
    assert L.wlsq_trace_spectra_device(0, 8, None, 8, None, 8, None) == _lib.WLSQ_OK


@pytest.mark.parametrize("empty", [dict(ns=0), dict(nextrap=0), dict(nx=0)])
def test_empty_inputs_are_a_no_op_before_the_device(empty):
    """No shots, no depth steps or no aperture: nothing to compute, status OK,
    outputs untouched (the reference loops simply do not run)."""
    a = dict(ns=1, nextrap=1, nz=1, nt=2, nf=2, nx=4, M=1)
    a.update(empty)
    op = np.zeros((1, 2, 4, 3), np.complex64)
    F = np.full((1, 2, 4), 1 + 2j, np.complex64)
    B = F.copy()
    img = np.full((1, 1, 4), 5.0, np.float32)
    L = _lib.lib()
    assert L.extrapolate(a["ns"], a["nextrap"], a["nz"], a["nt"], a["nf"], a["nx"], a["M"],
                         op, F, op, B, img) == _lib.WLSQ_OK
    sl = np.ones((1, 4), np.float32)
    om = np.ones(2, np.float32)
    tab = np.zeros((3, 3), np.complex64)
    assert L.extrapolate_ktable(a["ns"], a["nextrap"], a["nz"], a["nt"], a["nf"], a["nx"], a["M"],
                                sl, om, 0.0, 1.0, 3, tab, F, tab, B, img) == _lib.WLSQ_OK
    assert (F == 1 + 2j).all() and (img == 5.0).all()
