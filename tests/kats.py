"""Known-answer cases of the reference's own unit tests, restated as data.

Each builder returns a dict with the inputs of one reference test case and
the values that test asserts. Sources (all in the reference tree):

* ``unittests/test_cpp_src.py:115-294``          CPU KATs 1-5 (images 968,
  -3690, -1898, {-3690, -1898}, {-1898, 7298}); identical complex KATs in
  ``unittests/test_py_src_extrapolation.py:150-314``
* ``unittests/test_cpp_src_cuda.py:23-211``      GPU revOp KATs 1-5 (images
  and in-place wavefield write-back)
* ``unittests/test_cpp_src_cuda.py:216-384``     GPU v1 KATs (same inputs,
  image only; covered by the revOp cases)
* ``unittests/test_py_src_extrapolation.py:52-148`` float KATs (-1332, -270,
  and the zero-halo case -493)

``image`` maps (l, i) -> value, ``forw``/``back`` map (j, i) -> complex value
of the final wavefield, ``zero_rows`` lists image rows that must be zero
outside the listed pixels. The reference's tolerance is absolute 1e-5
(``test_cpp_src.py:13``); the values are small integers and exact in FP32.
"""
from __future__ import annotations

import numpy as np

SDELTA = 1e-5


def _cpu_case(name, nextrap, nz, freq, cols, back_sign, pulse_back):
    nw, nx, M = 2, 4, 1
    L = 2 * M + 1
    opF = np.zeros((nextrap, nw, nx, L), np.complex64)
    opB = np.zeros((nextrap, nw, nx, L), np.complex64)
    for l in range(nextrap):
        for i in range(nw):
            for c in cols:
                opF[l, i, c, :] = freq[i]
                opB[l, i, c, :] = back_sign * freq[i]
    F = np.array([[1 + 1j, 2 + 2j, 3 + 3j, 4 + 4j], [5 + 5j, 6 + 6j, 7 + 7j, 8 + 8j]], np.complex64)
    return dict(name=name, ns=1, nextrap=nextrap, nz=nz, nf=nw, nt=nw, nx=nx, M=M,
                opF=opF, opB=opB, F=F[None].copy(), B=pulse_back[None].copy())


def cpu_kats():
    """test_cpp_src.py:115-294 (and test_py_src_extrapolation.py:150-314)."""
    pb_conj = np.array([[1 - 1j, 2 - 2j, 3 - 3j, 4 - 4j], [5 + 5j, 6 + 6j, 7 + 7j, 8 + 8j]], np.complex64)
    pb = np.array([[1 + 1j, 2 + 2j, 3 + 3j, 4 + 4j], [5 + 5j, 6 + 6j, 7 + 7j, 8 + 8j]], np.complex64)
    f1 = np.array([1 + 1j, 2 + 0j], np.complex64)
    f2 = np.array([1 + 0j, 2 + 0j], np.complex64)
    cases = []
    c = _cpu_case("cpu_kat1", 1, 1, f1, [0], +1, pb_conj)
    c.update(image={(0, 0): 968.0}, zero_rows=[0])
    cases.append(c)
    c = _cpu_case("cpu_kat2", 1, 1, f2, [2], -1, pb)
    c.update(image={(0, 2): -3690.0}, zero_rows=[0])
    cases.append(c)
    c = _cpu_case("cpu_kat3", 1, 1, f2, [3], -1, pb)
    c.update(image={(0, 3): -1898.0}, zero_rows=[0])
    cases.append(c)
    c = _cpu_case("cpu_kat4", 1, 1, f2, [2, 3], -1, pb)
    c.update(image={(0, 2): -3690.0, (0, 3): -1898.0}, zero_rows=[0])
    cases.append(c)
    c = _cpu_case("cpu_kat5", 2, 2, f2, [3], -1, pb)
    c.update(image={(0, 3): -1898.0, (1, 3): 7298.0}, zero_rows=[0, 1])
    cases.append(c)
    return cases


def _gpu_case(name, nz, nextrap, nf, nx, M, col, frow, fcols, brow, bcols):
    L = 2 * M + 1
    op = np.zeros((nextrap, nf, nx, L), np.complex64)
    op2 = np.zeros_like(op)
    for l in range(nextrap):
        for i in range(nf):
            for v in range(3 + M):
                op2[l, i, col, v] = 1 + 1j
                op[l, i, col, v] = 1 - 1j
    pulse = np.zeros((nf, nx), np.complex64)
    pulse[frow, fcols] = np.array([1, 2, 3, 4], np.complex64)
    pulse2 = np.zeros_like(pulse)
    pulse2[brow, bcols] = np.array([5 + 1j, 6 - 2j, 7, 8], np.complex64)
    return dict(name=name, ns=1, nextrap=nextrap, nz=nz, nf=nf, nt=nf, nx=nx, M=M,
                opF=op, opB=op2, F=pulse[None].copy(), B=pulse2[None].copy())


def gpu_kats():
    """test_cpp_src_cuda.py:23-211 (revOp) == :216-384 (v1, image only)."""
    cases = []
    c = _gpu_case("gpu_kat1", 5, 1, 2, 6, 2, 1, 1, slice(0, 4), 1, slice(0, 4))
    c.update(forw={(1, 1): 10 - 10j}, back={(1, 1): 27 + 25j}, image={(0, 1): 20.0})
    cases.append(c)
    c = _gpu_case("gpu_kat2", 5, 1, 2, 6, 2, 1, 0, slice(0, 4), 1, slice(0, 4))
    c.update(forw={(0, 1): 10 - 10j}, back={(1, 1): 27 + 25j}, image={}, zero_rows=[0, 1, 2, 3, 4])
    cases.append(c)
    c = _gpu_case("gpu_kat3", 5, 1, 2, 6, 2, 2, 1, slice(1, 5), 1, slice(1, 5))
    c.update(forw={(1, 2): 10 - 10j}, back={(1, 2): 27 + 25j}, image={(0, 2): 20.0},
             zero_rows=[0, 1, 2, 3, 4])
    cases.append(c)
    c = _gpu_case("gpu_kat4", 5, 2, 2, 6, 2, 2, 1, slice(1, 5), 1, slice(1, 5))
    c.update(forw={(1, 2): 0 - 20j}, back={(1, 2): 2 + 52j},
             image={(0, 2): 20.0, (1, 2): -1040.0}, zero_rows=[0, 1, 2, 3, 4])
    cases.append(c)
    c = _gpu_case("gpu_kat5", 10, 2, 20, 16, 5, 2, 1, slice(1, 5), 1, slice(1, 5))
    c.update(forw={(1, 2): 0 - 20j}, back={(1, 2): 2 + 52j},
             image={(0, 2): 20.0, (1, 2): -1040.0}, zero_rows=list(range(10)))
    cases.append(c)
    return cases


def all_kats():
    return cpu_kats() + gpu_kats()


def python_float_kats():
    """test_py_src_extrapolation.py:52-148 (float32 operators and fields)."""
    nz, nextrap, nw, nx, M = 1, 1, 2, 4, 1
    L = 3
    freq = np.array([1, 2])
    pf = np.array([[1, 2, 3, 4], [5, 6, 7, 8]], np.float32)
    out = []
    opF = np.zeros((nextrap, nw, nx, L), np.float32)
    opB = np.zeros((nextrap, nw, nx, L), np.float32)
    for i in range(nw):
        opF[0, i, 1, :] = freq[i]
        opB[0, i, 1, :] = -freq[i]
    out.append(dict(name="py_float_check3", nextrap=1, nz=1, nf=2, nx=4, M=1, opF=opF.copy(),
                    opB=opB.copy(), F=pf, B=pf, image={(0, 1): -1332.0}, zero_rows=[0]))
    for i in range(nw):
        opF[0, i, 1, :] = np.array([0, 0, 1])
    out.append(dict(name="py_float_check4", nextrap=1, nz=1, nf=2, nx=4, M=1, opF=opF.copy(),
                    opB=opB.copy(), F=pf, B=pf, image={(0, 1): -270.0}, zero_rows=[0]))
    opF = np.zeros((nextrap, nw, nx, L), np.float32)
    opB = np.zeros((nextrap, nw, nx, L), np.float32)
    for i in range(nw):
        opF[0, i, 0, :] = freq[i]
        opB[0, i, 0, :] = -freq[i]
    out.append(dict(name="py_float_check5_halo", nextrap=1, nz=1, nf=2, nx=4, M=1, opF=opF,
                    opB=opB, F=pf, B=pf, image={(0, 0): -493.0}, zero_rows=[0]))
    return out


def check_case(case, image, forw=None, back=None, delta=SDELTA):
    """Assert the reference test's expectations on (image (nz, nx), final
    fields (nf, nx)). Returns nothing; raises AssertionError on mismatch."""
    for (l, i), v in case.get("image", {}).items():
        assert abs(float(image[l, i]) - v) <= delta, (case["name"], (l, i), float(image[l, i]), v)
    listed = set(case.get("image", {}).keys())
    for l in case.get("zero_rows", []):
        for i in range(image.shape[1]):
            if (l, i) not in listed:
                assert abs(float(image[l, i])) <= delta, (case["name"], "nonzero", (l, i), float(image[l, i]))
    if forw is not None:
        for (j, i), v in case.get("forw", {}).items():
            assert abs(complex(forw[j, i]) - v) <= delta, (case["name"], "forw", (j, i), complex(forw[j, i]), v)
    if back is not None:
        for (j, i), v in case.get("back", {}).items():
            assert abs(complex(back[j, i]) - v) <= delta, (case["name"], "back", (j, i), complex(back[j, i]), v)
