"""Operator preparation: compact WLSQ operator tables and the kappa grid.

Restates the reference's preparation in FP64 NumPy (the preparation is off
the hot path; only its output tables enter the kernels):

* :func:`wavenumbers`      ``problemConfig.py:42-48`` (kx = fftfreq(nx)*nx*dkx,
  N = (nx-1)/2) - keeps the symmetric set n = -N..N. For odd nx that is all
  of kx; for even nx the unpaired Nyquist wavenumber is dropped, which is the
  only set the reference's design accepts (``preparation.py:67,110``: A is
  L x (2N+1)).
* :func:`wlsq_operator`    ``preparation.py:64-97`` - the symmetric operator
  A = pinv(Gamma^H Gamma) Gamma^H, Gamma[n, m] = exp(i kx_n x_m),
  x_m = (m - M) dx (the one ``defWLSQIndices(sym=True)`` selects for every x,
  ``util.py:19-21``, ``main.py:66``).
* :func:`phase_shift`      ``preparation.py:10-44`` - P = exp(-i kz dz)
  (forward, evanescent kz = -i sqrt(kx^2 - k^2)) and exp(+i kz dz) (backward,
  kz = +i sqrt(kx^2 - k^2)).
* :func:`compact_tables`   rows T[q] = A @ P[q] - the rows that
  ``makeTableOfOperators`` (``preparation.py:101-115``) gathers per point.
* :func:`create_kappa`     ``util.py:39-45`` (kappa_0 = 0, dk = dw/cmax/dense).
* :func:`kappa_index`      the lookup ``find_nearest(kappa, w/v)``
  (``util.py:5-8``, ``preparation.py:109``) on a uniform grid, in the FP32
  form the device uses (see ``include/wlsq_hip.h``).
"""
from __future__ import annotations

import math

import numpy as np


def wavenumbers(nx: int, dx: float) -> np.ndarray:
    """Symmetric horizontal wavenumbers n*dkx, n = -N..N, N = (nx-1)//2,
    dkx = 2*pi/(nx*dx)."""
    N = (nx - 1) // 2
    dkx = 2.0 * math.pi / (nx * dx)
    return np.arange(-N, N + 1, dtype=np.float64) * dkx


def wlsq_operator(kx: np.ndarray, M: int, dx: float) -> np.ndarray:
    """(2M+1, len(kx)) complex128 least-squares operator."""
    x = (np.arange(2 * M + 1, dtype=np.float64) - M) * dx
    gamma = np.exp(1j * np.outer(kx, x))          # (nK, L)
    gh = gamma.conj().T                           # (L, nK)
    return np.linalg.pinv(gh @ gamma) @ gh


def phase_shift(kappa: np.ndarray, kx: np.ndarray, dz: float, forward: bool = True) -> np.ndarray:
    """(nk, nK) complex128 phase-shift operators."""
    k2 = kappa[:, None] ** 2
    x2 = kx[None, :] ** 2
    prop = np.abs(kappa)[:, None] >= np.abs(kx)[None, :]
    kz_r = np.sqrt(np.where(prop, k2 - x2, 0.0))
    kz_e = np.sqrt(np.where(prop, 0.0, x2 - k2))
    if forward:
        kz = np.where(prop, kz_r + 0j, -1j * kz_e)
        return np.exp(-1j * kz * dz)
    kz = np.where(prop, kz_r + 0j, 1j * kz_e)
    return np.exp(1j * kz * dz)


def compact_tables(kappa: np.ndarray, nx: int, dx: float, dz: float, M: int,
                   check_conjugate: bool = True):
    """Forward and backward (nk, 2M+1) complex64 tables.

    For the symmetric kx set and symmetric taps the backward table is the
    complex conjugate of the forward one (A @ conj(P) = conj(A @ P) because
    the imaginary parts of A cancel against P(kx) = P(-kx)); it is computed
    both ways in FP64, checked, and stored as the exact conjugate so the
    fused kernel can read one table.
    """
    kx = wavenumbers(nx, dx)
    A = wlsq_operator(kx, M, dx)
    tf = phase_shift(kappa, kx, dz, True) @ A.T
    if check_conjugate:
        tb = phase_shift(kappa, kx, dz, False) @ A.T
        err = np.max(np.abs(tb - np.conj(tf))) / max(1e-300, np.max(np.abs(tf)))
        if err > 1e-9:
            raise ArithmeticError(f"backward table is not conj(forward): {err:g}")
    tf32 = tf.astype(np.complex64)
    return tf32, np.conj(tf32)


def create_kappa(cmax: float, cmin: float, dw: float, wmax: float, dense: float):
    """(kappa0, dkappa, nk) of util.py:39-45: kappa_i = i*dw/cmax/dense,
    nk = int(wmax/cmin / dk) + ceil(cmax/cmin)."""
    dk = dw / cmax / dense
    nk = int((wmax / cmin) / dk) + math.ceil(cmax / cmin)
    return 0.0, dk, nk


def kappa_span(omega_min: float, omega_max: float, vmin: float, vmax: float, nk: int):
    """(kappa0, dkappa) of nk uniform samples over [w_min/v_max, w_max/v_min]."""
    k0 = omega_min / vmax
    k1 = omega_max / vmin
    return k0, (k1 - k0) / (nk - 1)


def kappa_values(kappa0: float, dkappa: float, nk: int) -> np.ndarray:
    return kappa0 + np.arange(nk, dtype=np.float64) * dkappa


def kappa_index(omega, slowness, kappa0, dkappa, nk):
    """FP32 lookup, one rounding per operation (the device's rule):
    t = w*s; t -= kappa0; u = t*inv_dk; clamp to [0, nk-1]; q = ceil(u-0.5),
    inv_dk = float32(1/float64(float32(dkappa)))."""
    om = np.asarray(omega, np.float32)
    s = np.asarray(slowness, np.float32)
    inv = np.float32(1.0 / float(np.float32(dkappa)))
    t = np.multiply(om, s, dtype=np.float32)
    t = np.subtract(t, np.float32(kappa0), dtype=np.float32)
    u = np.multiply(t, inv, dtype=np.float32)
    u = np.where(u > np.float32(0), u, np.float32(0)).astype(np.float32)
    u = np.minimum(u, np.float32(nk - 1))
    return np.ceil(np.subtract(u, np.float32(0.5), dtype=np.float32)).astype(np.int64)


def expand_table(table, slowness, omega, kappa0, dkappa, nk, nextrap):
    """Host expansion to the reference layout (nextrap, nf, nx, L) (small
    problems only; ``wlsq_expand_table_device`` does it on the GPU)."""
    s = np.asarray(slowness, np.float32)[:nextrap]
    q = kappa_index(np.asarray(omega, np.float32)[None, :, None], s[:, None, :], kappa0, dkappa, nk)
    return np.ascontiguousarray(np.asarray(table)[q])


def ricker_spectrum(f, f0: float):
    """Amplitude spectrum of a Ricker wavelet of peak frequency f0 (Hz):
    (2/sqrt(pi)) f^2/f0^3 exp(-f^2/f0^2)."""
    f = np.asarray(f, np.float64)
    return (2.0 / math.sqrt(math.pi)) * f ** 2 / f0 ** 3 * np.exp(-(f ** 2) / f0 ** 2)
