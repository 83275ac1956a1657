"""Seeded synthetic problems for the benchmark and parity configurations.

Each configuration is a velocity model (smoothed), a frequency list, a
compact WLSQ table pair on a uniform kappa grid, and source / receiver
wavefields:
* source:   Ricker spectrum R(f) * exp(-i w r / v0) / sqrt(r),
            r = sqrt(dz^2 + (x - xs)^2) (the frequency form of the
            reference's makeRickerWavelet, py_src/operatorsPreparation/util.py:47-58)
* receiver: complex standard normal (re, im ~ N(0, 1/2)) times R(f).

``small`` problems build fields with NumPy on the host; the large ones build
them on the GPU with torch (seeded) so nothing of that size crosses PCIe.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
from scipy import ndimage

from . import prep


@dataclass
class Model:
    nz: int
    nx: int
    dx: float
    dz: float
    velocity: np.ndarray  # (nz, nx) float64, smoothed

    @property
    def slowness(self) -> np.ndarray:
        return (1.0 / self.velocity).astype(np.float32)


@dataclass
class Tables:
    M: int
    nk: int
    kappa0: float
    dkappa: float
    forw: np.ndarray  # (nk, 2M+1) complex64
    back: np.ndarray

    @property
    def L(self) -> int:
        return 2 * self.M + 1


@dataclass
class ShotGeom:
    x0: int          # window origin (model column)
    nx: int          # aperture
    nz: int          # depth steps
    xs: float        # source position within the window (cells)
    freqs: np.ndarray  # Hz, float64
    seed: int

    @property
    def omega(self) -> np.ndarray:
        return (2.0 * math.pi * self.freqs).astype(np.float32)

    @property
    def nf(self) -> int:
        return len(self.freqs)


@dataclass
class Problem:
    name: str
    model: Model
    tables: Tables
    shots: List[ShotGeom]
    f0: float = 15.0
    meta: dict = field(default_factory=dict)

    @property
    def points(self) -> int:
        return int(sum(s.nx * s.nz * s.nf for s in self.shots))


def _anomaly(nz, nx, dx, dz, xc, zc, amp, sigma):
    z = np.arange(nz)[:, None] * dz
    x = np.arange(nx)[None, :] * dx
    return amp * np.exp(-((x - xc) ** 2 + (z - zc) ** 2) / (2.0 * sigma ** 2))


def gradient_model(rng, nz, nx, dx, dz, v_top, v_bottom, n_anom, amp=400.0, sigma=300.0,
                   smooth=3.0, vmin=None, vmax=None) -> Model:
    z = np.arange(nz)[:, None] * dz
    v = v_top + (v_bottom - v_top) * z / max(1.0, (nz - 1) * dz) + np.zeros((1, nx))
    for _ in range(n_anom):
        xc = rng.uniform(0.2, 0.8) * nx * dx
        zc = rng.uniform(0.2, 0.8) * nz * dz
        v = v + _anomaly(nz, nx, dx, dz, xc, zc, amp, sigma)
    if vmin is not None or vmax is not None:
        v = np.clip(v, vmin, vmax)
    v = ndimage.gaussian_filter(v, sigma=smooth)
    return Model(nz, nx, dx, dz, v)


def layered_salt_model(rng, nz, nx, dx, dz, smooth=3.0) -> Model:
    """Layers 1500..5000 m/s with dipping interfaces plus a 4500 m/s body."""
    z = np.arange(nz)[:, None] * dz
    x = np.arange(nx)[None, :] * dx
    nlay = 8
    v = np.full((nz, nx), 1500.0)
    depths = np.sort(rng.uniform(0.05, 0.95, nlay)) * nz * dz
    dips = rng.uniform(-0.08, 0.08, nlay)
    vel = np.linspace(1700.0, 5000.0, nlay)
    for d, dip, vv in zip(depths, dips, vel):
        v = np.where(z >= d + dip * (x - 0.5 * nx * dx), vv, v)
    xc, zc = rng.uniform(0.3, 0.7) * nx * dx, rng.uniform(0.35, 0.6) * nz * dz
    ax, az = 0.12 * nx * dx, 0.12 * nz * dz
    body = ((x - xc) / ax) ** 2 + ((z - zc) / az) ** 2 <= 1.0
    v = np.where(body, 4500.0, v)
    v = ndimage.gaussian_filter(v, sigma=smooth)
    return Model(nz, nx, dx, dz, v)


def make_tables(model: Model, M: int, nk: int, fmin: float, fmax: float) -> Tables:
    vmin, vmax = float(model.velocity.min()), float(model.velocity.max())
    k0, dk = prep.kappa_span(2 * math.pi * fmin, 2 * math.pi * fmax, vmin, vmax, nk)
    kap = prep.kappa_values(k0, dk, nk)
    tf, tb = prep.compact_tables(kap, model.nx, model.dx, model.dz, M)
    return Tables(M, nk, float(np.float32(k0)), float(np.float32(dk)), tf, tb)


def _freqs(nf, fmin, fmax):
    if nf == 1:
        return np.array([fmin], np.float64)
    return fmin + (fmax - fmin) * np.arange(nf, dtype=np.float64) / (nf - 1)


# ------------------------------------------------------------------ configs
def config_c1(seed: int = 1) -> Problem:
    """128 x 256 at 10 m, v = 1500 + z + Gaussian (+400, 300 m), nf 64 over
    2-40 Hz, 1 shot at the centre, 25 taps x 1024 kappa samples."""
    rng = np.random.default_rng(seed)
    nz, nx, d = 128, 256, 10.0
    z = np.arange(nz)[:, None] * d
    v = 1500.0 + 1.0 * z + np.zeros((1, nx))
    xc = rng.uniform(0.2, 0.8) * nx * d
    zc = rng.uniform(0.2, 0.8) * nz * d
    v = v + _anomaly(nz, nx, d, d, xc, zc, 400.0, 300.0)
    model = Model(nz, nx, d, d, ndimage.gaussian_filter(v, sigma=3.0))
    tables = make_tables(model, 12, 1024, 2.0, 40.0)
    shots = [ShotGeom(0, nx, nz, nx / 2.0, _freqs(64, 2.0, 40.0), seed * 1000)]
    return Problem("c1_cpu_plain_128x256_nf64_1shot", model, tables, shots)


def config_c2(seed: int = 2, nshots: int = 16) -> Problem:
    """512 x 1024 at 10 m, 1500->4500 m/s gradient + 3 anomalies, nf 256 over
    2-50 Hz, 16 shots evenly spaced, 31 taps x 2048 kappa samples."""
    rng = np.random.default_rng(seed)
    nz, nx, d = 512, 1024, 10.0
    model = gradient_model(rng, nz, nx, d, d, 1500.0, 4500.0, 3, vmin=1500.0, vmax=4500.0)
    tables = make_tables(model, 15, 2048, 2.0, 50.0)
    freqs = _freqs(256, 2.0, 50.0)
    shots = [ShotGeom(0, nx, nz, (k + 0.5) * nx / nshots, freqs, seed * 1000 + k)
             for k in range(nshots)]
    return Problem("c2_plain_512x1024_nf256_16shots", model, tables, shots)


def config_c3(seed: int = 3, nshots: int = 128) -> Problem:
    """800 x 2048 model; 128 shots with windows nx_s ~ U[256, 2048] and
    nf_s ~ U[200, 400] over 2-60 Hz; 31 taps x 4096 kappa samples."""
    rng = np.random.default_rng(seed)
    nz, nx, d = 800, 2048, 10.0
    model = gradient_model(rng, nz, nx, d, d, 1500.0, 4500.0, 3, vmin=1500.0, vmax=4500.0)
    tables = make_tables(model, 15, 4096, 2.0, 60.0)
    shots = []
    for k in range(nshots):
        nxs = int(rng.integers(256, 2049))
        x0 = int(rng.integers(0, nx - nxs + 1))
        nfs = int(rng.integers(200, 401))
        shots.append(ShotGeom(x0, nxs, nz, nxs / 2.0, _freqs(nfs, 2.0, 60.0), seed * 1000 + k))
    return Problem("c3_batched_800x2048_128shots_uneven", model, tables, shots)


def config_c5(seed: int = 5, nshots: int = 1000) -> Problem:
    """2000 x 8192 at 6.25 m, layered 1500-5000 m/s + salt body, nf 1024 over
    2-80 Hz, full-aperture shots, 51 taps x 8192 kappa samples."""
    rng = np.random.default_rng(seed)
    nz, nx, d = 2000, 8192, 6.25
    model = layered_salt_model(rng, nz, nx, d, d)
    tables = make_tables(model, 25, 8192, 2.0, 80.0)
    freqs = _freqs(1024, 2.0, 80.0)
    shots = [ShotGeom(0, nx, nz, (k + 0.5) * nx / nshots, freqs, seed * 1000 + k)
             for k in range(nshots)]
    return Problem("c5_survey_2000x8192_nf1024", model, tables, shots)


# ------------------------------------------------------------------ wavefields
def source_field(p: Problem, shot: ShotGeom) -> np.ndarray:
    """(nf, nx) complex64 source wavefield of one shot (host)."""
    m = p.model
    x = np.arange(shot.nx, dtype=np.float64)
    r = np.sqrt(m.dz ** 2 + ((x - shot.xs) * m.dx) ** 2)
    xs_col = min(m.nx - 1, shot.x0 + int(shot.xs))
    v0 = float(m.velocity[0, xs_col])
    w = 2 * math.pi * shot.freqs[:, None]
    amp = prep.ricker_spectrum(shot.freqs, p.f0)[:, None]
    return (amp * np.exp(-1j * w * r[None, :] / v0) / np.sqrt(r)[None, :]).astype(np.complex64)


def receiver_field(p: Problem, shot: ShotGeom) -> np.ndarray:
    rng = np.random.default_rng(shot.seed)
    g = rng.standard_normal((shot.nf, shot.nx)) + 1j * rng.standard_normal((shot.nf, shot.nx))
    amp = prep.ricker_spectrum(shot.freqs, p.f0)[:, None]
    return (np.sqrt(0.5) * g * amp).astype(np.complex64)


def host_fields(p: Problem, shots: Optional[List[ShotGeom]] = None, nt: Optional[int] = None):
    """(ns, nt, nx) source and receiver arrays for same-shaped shots."""
    shots = shots if shots is not None else p.shots
    nf, nx = shots[0].nf, shots[0].nx
    nt = nt or nf
    F = np.zeros((len(shots), nt, nx), np.complex64)
    B = np.zeros_like(F)
    for i, s in enumerate(shots):
        F[i, :nf] = source_field(p, s)
        B[i, :nf] = receiver_field(p, s)
    return F, B


def device_fields(p: Problem, shots: List[ShotGeom], device="cuda", out_f=None, out_b=None,
                  offsets=None):
    """Fill source/receiver fields of `shots` directly on the GPU (torch,
    seeded per shot). Returns (F, B) flat complex64 tensors and the element
    offsets of each shot (rows of nx_s, row stride nx_s)."""
    import torch

    if offsets is None:
        offsets = []
        o = 0
        for s in shots:
            offsets.append(o)
            o += s.nf * s.nx
        total = o
    else:
        total = max(off + s.nf * s.nx for off, s in zip(offsets, shots))
    F = out_f if out_f is not None else torch.empty(total, dtype=torch.complex64, device=device)
    B = out_b if out_b is not None else torch.empty(total, dtype=torch.complex64, device=device)
    m = p.model
    for off, s in zip(offsets, shots):
        x = torch.arange(s.nx, dtype=torch.float64, device=device)
        r = torch.sqrt(m.dz ** 2 + ((x - s.xs) * m.dx) ** 2)
        xs_col = min(m.nx - 1, s.x0 + int(s.xs))
        v0 = float(m.velocity[0, xs_col])
        fr = torch.as_tensor(s.freqs, dtype=torch.float64, device=device)[:, None]
        amp = (2.0 / math.sqrt(math.pi)) * fr ** 2 / p.f0 ** 3 * torch.exp(-(fr ** 2) / p.f0 ** 2)
        ph = -2 * math.pi * fr * r[None, :] / v0
        src = amp * torch.polar(torch.ones_like(ph), ph) / torch.sqrt(r)[None, :]
        F[off:off + s.nf * s.nx] = src.to(torch.complex64).reshape(-1)
        gen = torch.Generator(device=device)
        gen.manual_seed(int(s.seed))
        g = torch.randn(s.nf, s.nx, 2, generator=gen, device=device, dtype=torch.float32)
        rec = torch.view_as_complex(g.contiguous()) * math.sqrt(0.5) * amp.to(torch.float32)
        B[off:off + s.nf * s.nx] = rec.reshape(-1)
    return F, B, offsets
