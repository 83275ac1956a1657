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
* :func:`wlsq_operators`   ``preparation.py:64-97`` - all 2M+1 operators of
  ``createWLSQoperators``: g < M left-edge operators (Gamma columns
  m < M-g zero: no taps past the left edge), g = M the symmetric one,
  g > M right-edge operators (columns m >= 3M+1-g zero).
* :func:`wlsq_indices`     ``util.py:16-37`` (``defWLSQIndices``) - which
  operator each x uses: M everywhere (sym=True, what ``main.py:66`` asks),
  or 0..M-1 on the first M columns and M+1..2M on the last M, the middle
  [extent, nx-extent) reset to M, optionally flipped.
* :func:`phase_shift`      ``preparation.py:10-44`` - P = exp(-i kz dz)
  (forward, evanescent kz = -i sqrt(kx^2 - k^2)) and exp(+i kz dz) (backward,
  kz = +i sqrt(kx^2 - k^2)).
* :func:`compact_tables`   rows T[q] = A @ P[q] - the rows that
  ``makeTableOfOperators`` (``preparation.py:101-115``) gathers per point;
  :func:`compact_tables_all` the same for every operator g (T[g][q] =
  A_g @ P[q]), and :func:`expand_table` with ``gidx`` the per-point gather
  ``opWLSQ[gIdx[i]] @ psOp[q]`` of ``preparation.py:110``.
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


def gamma_matrix(kx: np.ndarray, M: int, dx: float, g: int) -> np.ndarray:
    """(nK, L) complex128 Gamma[n, m] = exp(i kx_n x_m) of operator g
    (``preparation.py:71-91``), with the columns outside the operator's
    support zero: m < M-g for g < M, m >= 3M+1-g for g > M."""
    L = 2 * M + 1
    x = (np.arange(L, dtype=np.float64) - M) * dx
    gamma = np.exp(1j * np.outer(kx, x))
    m = np.arange(L)
    if g < M:
        gamma[:, m < M - g] = 0.0
    elif g > M:
        gamma[:, m >= 3 * M + 1 - g] = 0.0
    return gamma


def wlsq_operators(kx: np.ndarray, M: int, dx: float) -> np.ndarray:
    """(2M+1, 2M+1, len(kx)) complex128 operators of ``createWLSQoperators``
    (``preparation.py:64-97``): op[g] = pinv(Gamma_g^H Gamma_g) Gamma_g^H."""
    L = 2 * M + 1
    ops = np.empty((L, L, len(kx)), np.complex128)
    for g in range(L):
        gm = gamma_matrix(kx, M, dx, g)
        gh = gm.conj().T
        ops[g] = np.linalg.pinv(gh @ gm) @ gh
    return ops


def wlsq_indices(nx: int, M: int, flip: bool = False, sym: bool = True, extent: int = 5) -> np.ndarray:
    """Operator index per x (``defWLSQIndices``, ``util.py:16-37``)."""
    g = np.full(nx, M, np.int16)
    if sym:
        return g
    d1, d2 = 0, M + 1
    for i in range(nx):
        if i < M:
            g[i] = d1
            d1 += 1
        if i >= nx - M:
            g[i] = d2
            d2 += 1
    g[extent:nx - extent] = M
    return np.flip(g).copy() if flip else g


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
    fused kernel can read one table. The symmetric operator's taps are
    symmetric, T[q][k] = T[q][2M-k] (x_{2M-k} = -x_k and the kx set is
    symmetric); the FP64 rows are averaged with their mirror before rounding,
    so the stored taps are symmetric bit for bit and the kernels may pair them.
    """
    kx = wavenumbers(nx, dx)
    A = wlsq_operator(kx, M, dx)
    tf = phase_shift(kappa, kx, dz, True) @ A.T
    tf = 0.5 * (tf + tf[:, ::-1])
    if check_conjugate:
        tb = phase_shift(kappa, kx, dz, False) @ A.T
        err = np.max(np.abs(tb - np.conj(tf))) / max(1e-300, np.max(np.abs(tf)))
        if err > 1e-9:
            raise ArithmeticError(f"backward table is not conj(forward): {err:g}")
    tf32 = tf.astype(np.complex64)
    return tf32, np.conj(tf32)


def compact_tables_all(kappa: np.ndarray, nx: int, dx: float, dz: float, M: int):
    """Forward and backward (2M+1, nk, 2M+1) complex64 tables of every
    operator g: T[g][q] = A_g @ P[q] (FP64, rounded once)."""
    kx = wavenumbers(nx, dx)
    ops = wlsq_operators(kx, M, dx)
    pf = phase_shift(kappa, kx, dz, True)
    pb = phase_shift(kappa, kx, dz, False)
    tf = np.einsum("qn,gmn->gqm", pf, ops)
    tb = np.einsum("qn,gmn->gqm", pb, ops)
    return tf.astype(np.complex64), tb.astype(np.complex64)


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


def expand_table(table, slowness, omega, kappa0, dkappa, nk, nextrap, gidx=None):
    """Host expansion to the reference layout (nextrap, nf, nx, L) (small
    problems only; ``wlsq_expand_table_device`` does it on the GPU). With
    ``gidx`` (one operator index per x) ``table`` holds every operator,
    (2M+1, nk, L), and point (l, j, i) takes table[gidx[i]][q(l, j, i)]
    (``preparation.py:110``)."""
    s = np.asarray(slowness, np.float32)[:nextrap]
    q = kappa_index(np.asarray(omega, np.float32)[None, :, None], s[:, None, :], kappa0, dkappa, nk)
    t = np.asarray(table)
    if gidx is None:
        return np.ascontiguousarray(t[q])
    g = np.asarray(gidx, np.int64)[None, None, :]
    return np.ascontiguousarray(t[g, q])


def ricker_spectrum(f, f0: float):
    """Amplitude spectrum of a Ricker wavelet of peak frequency f0 (Hz):
    (2/sqrt(pi)) f^2/f0^3 exp(-f^2/f0^2)."""
    f = np.asarray(f, np.float64)
    return (2.0 / math.sqrt(math.pi)) * f ** 2 / f0 ** 3 * np.exp(-(f ** 2) / f0 ** 2)
