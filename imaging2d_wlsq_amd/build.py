"""Builds libwlsq_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m imaging2d_wlsq_amd.build`` (or ``build()`` from Python).
Objects go to ``imaging2d_wlsq_amd/_build/``; the library to
``imaging2d_wlsq_amd/libwlsq_hip.so``.

Flags:
* ``--offload-arch=gfx950``: MI355X only.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  (v_pk_*_f32) instructions are generated.
* ``-munsafe-fp-atomics``: float atomic adds compile to the hardware
  global/LDS add instructions (device memory from hipMalloc is coarse-grained).
* ``-ffp-contract=fast`` is the HIP default; the kappa lookup disables
  contraction locally with a pragma so its rounding matches the host rule.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libwlsq_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# (source, object, extra flags): the fused kernel is built once per taps
# (25, 31, 51; 0 = any, runtime loop bounds) and gather width, in parallel
SOURCES = [("wlsq_kernels.hip", "wlsq_kernels.o", []), ("wlsq_capi.hip", "wlsq_capi.o", []),
           ("wlsq_fft.hip", "wlsq_fft.o", [])] + [
    ("wlsq_fused.hip", f"wlsq_fused_L{l}_P{p}.o", [f"-DWLSQ_L={l}", f"-DWLSQ_PAIRED={p}"])
    for l in (25, 31, 51, 0) for p in (0, 1)]

# hipFFT (rocFFT) for the trace spectra and the FFT drop-ins
LINK = ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
    "-I", INCLUDE, "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src, o, extra in SOURCES:
        obj = os.path.join(BUILD, o)
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for r, c in zip(ex.map(_run, cmds), cmds):
            if verbose:
                print(" ".join(c))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp, *objs, *LINK])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
