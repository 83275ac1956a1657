"""FP64 NumPy restatement of the reference hot path - TEST INFRASTRUCTURE ONLY.

Restates ``py_src/wavefieldPropagation/extrapolation.py:38-66`` (one shot):

* zero-padded fields ``old[:, M:nx+M] = wave`` (``:48-51``),
* ``new[j, i+M] = dot(W[l, j, i, :], old[j, i:i+L])`` for ``j < nf`` - an
  unconjugated dot (``np.dot`` at ``:60-61``),
* ``old = new`` (``:62-63``),
* ``image[l, :] = Re(sum_j new[j] * conj(new_b[j]))`` (``:64`` with
  ``imaging`` at ``:22-34``),

vectorised over (j, i) and evaluated in complex128. It also returns the final
wavefields, which the native libraries write back in place
(``cpp_src/extrapolation_cpu.cpp:91-98``).

The per-point operator is given either as the reference-layout table
``(nextrap, nf, nx, L)`` or as a compact table ``(nk, L)`` plus the kappa
lookup, restated here in NumPy float32 with the same separately rounded
operations as the C oracle.
"""
from __future__ import annotations

import numpy as np


def inv_dkappa_f32(dkappa: float) -> np.float32:
    return np.float32(1.0 / float(np.float32(dkappa)))


def kappa_index_f32(omega, slowness, kappa0, dkappa, nk):
    """Vectorised q = nearest kappa sample to omega*slowness, ties -> lower
    index (``util.py:5-8``), with one float32 rounding per operation:
    t = omega*s; t = t - kappa0; u = t*inv_dk; clamp to [0, nk-1]; ceil(u-0.5)."""
    om = np.asarray(omega, np.float32)
    s = np.asarray(slowness, np.float32)
    t = np.multiply(om, s, dtype=np.float32)
    t = np.subtract(t, np.float32(kappa0), dtype=np.float32)
    u = np.multiply(t, inv_dkappa_f32(dkappa), dtype=np.float32)
    u = np.where(u > np.float32(0.0), u, np.float32(0.0)).astype(np.float32)
    u = np.minimum(u, np.float32(nk - 1))
    v = np.subtract(u, np.float32(0.5), dtype=np.float32)
    return np.ceil(v).astype(np.int64)


def _step(W, old, M, nx):
    """new[:, M+i] = sum_k W[:, i, k] * old[:, i+k], complex128."""
    L = 2 * M + 1
    new = np.zeros_like(old)
    acc = np.zeros((old.shape[0], nx), np.complex128)
    for k in range(L):
        acc += W[:, :, k] * old[:, k:k + nx]
    new[:, M:M + nx] = acc
    return new


def extrapolate(nextrap, nz, nx, nf, M, w_op_forw=None, w_op_back=None, wave_fs=None,
                shot_fs=None, *, ktable=None):
    """One shot. Returns (image (nz, nx) float64, final forward field (nf, nx),
    final backward field (nf, nx)), all complex128/float64.

    ``ktable`` = dict(slowness=(>=nextrap, nx), omega=(nf,), kappa0, dkappa,
    nk, forw=(nk, L), back=(nk, L)) replaces the two reference-layout tables.
    """
    L = 2 * M + 1
    dim_x = nx + 2 * M
    F = np.zeros((nf, dim_x), np.complex128)
    B = np.zeros((nf, dim_x), np.complex128)
    F[:, M:M + nx] = np.asarray(wave_fs)[:nf]
    B[:, M:M + nx] = np.asarray(shot_fs)[:nf]
    image = np.zeros((nz, nx), np.float64)
    for l in range(nextrap):
        if ktable is None:
            Wf = np.asarray(w_op_forw[l], np.complex128).reshape(nf, nx, L)
            Wb = np.asarray(w_op_back[l], np.complex128).reshape(nf, nx, L)
        else:
            q = kappa_index_f32(np.asarray(ktable["omega"], np.float32)[:, None],
                                np.asarray(ktable["slowness"], np.float32)[l][None, :],
                                ktable["kappa0"], ktable["dkappa"], ktable["nk"])
            Wf = np.asarray(ktable["forw"], np.complex128)[q]
            Wb = np.asarray(ktable["back"], np.complex128)[q]
        F = _step(Wf, F, M, nx)
        B = _step(Wb, B, M, nx)
        image[l] = np.real(np.sum(F[:, M:M + nx] * np.conj(B[:, M:M + nx]), axis=0))
    return image, F[:, M:M + nx], B[:, M:M + nx]


def extrapolate_shots(ns, nextrap, nz, nx, nf, M, w_op_forw=None, w_op_back=None,
                      wave_fs=None, shot_fs=None, *, ktable=None):
    """All shots; wave_fs/shot_fs are (ns, nt, nx). Returns (image (ns, nz, nx),
    F (ns, nf, nx), B (ns, nf, nx))."""
    imgs, Fs, Bs = [], [], []
    for s in range(ns):
        im, F, B = extrapolate(nextrap, nz, nx, nf, M, w_op_forw, w_op_back,
                               wave_fs[s], shot_fs[s], ktable=ktable)
        imgs.append(im)
        Fs.append(F)
        Bs.append(B)
    return np.stack(imgs), np.stack(Fs), np.stack(Bs)


def brute_force(nextrap, nz, nx, nf, M, w_op_forw, w_op_back, wave_fs, shot_fs):
    """Pure-Python loops, for tiny shapes only: the definition written out
    element by element (extrapolation.py:57-64), used to check the vectorised
    form and the C oracle."""
    L = 2 * M + 1
    F = [[0j] * (nx + 2 * M) for _ in range(nf)]
    B = [[0j] * (nx + 2 * M) for _ in range(nf)]
    for j in range(nf):
        for i in range(nx):
            F[j][i + M] = complex(wave_fs[j][i])
            B[j][i + M] = complex(shot_fs[j][i])
    image = [[0.0] * nx for _ in range(nz)]
    for l in range(nextrap):
        nF = [[0j] * (nx + 2 * M) for _ in range(nf)]
        nB = [[0j] * (nx + 2 * M) for _ in range(nf)]
        for j in range(nf):
            for i in range(nx):
                sf = 0j
                sb = 0j
                for k in range(L):
                    sf += complex(w_op_forw[l][j][i][k]) * F[j][i + k]
                    sb += complex(w_op_back[l][j][i][k]) * B[j][i + k]
                nF[j][i + M] = sf
                nB[j][i + M] = sb
        F, B = nF, nB
        for i in range(nx):
            acc = 0j
            for j in range(nf):
                acc += F[j][i + M] * B[j][i + M].conjugate()
            image[l][i] = acc.real
    return (np.array(image), np.array([r[M:M + nx] for r in F]),
            np.array([r[M:M + nx] for r in B]))
