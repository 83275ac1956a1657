"""Migration driver: the reference's ``main.py`` with a ``device_hip`` option.

    python -m imaging2d_wlsq_amd.driver velmod.csv problemSetup.txt shotsDir/ outDir device_hip [asym]

Same arguments as ``main.py:30-34``; an optional sixth argument ``asym``
selects the edge operators near the model's sides (``defWLSQIndices`` with
sym=False, extent 5, ``util.py:16-37``) instead of the symmetric operator
everywhere that ``main.py:66`` asks for (``device_hip_expanded`` only). The
steps restate ``main.py``:

1. read the velocity CSV, smooth the slowness with a Gaussian of sigma 2 and
   save ``smoothed_velmod.csv`` (``main.py:39-43``);
2. parse the ``key_value`` setup file (``problemConfig.py:15-54``);
3. WLSQ design (symmetric operator for every x, ``main.py:65-66``), kappa
   grid (``util.py:39-45``), phase-shift operators (``preparation.py:10-44``)
   -> compact operator tables (``prep.py``);
4. shots ``seis<ix>_<nx>.csv`` in ``shotsDir``, sorted by ix (the reference
   keeps glob order, ``readShotFiles.py:5-14``);
5. Ricker source per shot (``util.py:47-58`` with the call of ``main.py:120``)
   and FFT along time of source and receiver (``main.py:121-123``);
6. extrapolation and imaging on the GPU:
   * ``device_hip``: device-resident - trace spectra by hipFFT on the GPU,
     compact tables, operators gathered on the device (``wlsq_ktable_device``);
   * ``device_hip_expanded``: the reference-shaped call - host FFTs, tables
     expanded to (nextrap, nw, nx, L) on the host as ``makeTableOfOperators``
     does (``preparation.py:101-115``) and passed to the drop-in ``extrapolate``;
7. sum the per-shot images and save ``final_image.csv`` (``main.py:184-190``).

The ``host`` options of the reference are its CPU library, which this
package does not provide.
"""
from __future__ import annotations

import glob
import math
import os
import re
import sys
import time
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np
from scipy import ndimage
import scipy.fftpack as spfft

from . import interface_hip, prep


@dataclass
class ProblemConfig:
    """Grids of a ``key_value`` setup file (``problemConfig.py:6-54``)."""
    nz: int
    nx: int
    nt: int
    nw: int
    zmin: float
    zmax: float
    zinit: float
    zfinal: float
    xmin: float
    xmax: float
    tmin: float
    tmax: float
    M: int

    @classmethod
    def parse(cls, text: str, nz: int, nx: int) -> "ProblemConfig":
        vals = {}
        for line in text.splitlines():
            line = line.strip()
            if not line or line.startswith("key"):
                continue
            k, v = line.split("_", 1)
            vals[k] = float(v)
        nt, nw = int(vals["nt"]), int(vals["nw"])
        if nt < nw:
            raise ValueError("nt must be greater or equal to nw")
        return cls(nz, nx, nt, nw, vals["zmin"], vals["zmax"], vals["zinit"], vals["zfinal"],
                   vals["xmin"], vals["xmax"], vals["tmin"], vals["tmax"], int(vals["M"]))

    @property
    def dz(self):
        return (self.zmax - self.zmin) / self.nz

    @property
    def dx(self):
        return (self.xmax - self.xmin) / self.nx

    @property
    def dt(self):
        return (self.tmax - self.tmin) / self.nt

    @property
    def nextrap(self):
        return math.ceil((self.zfinal - self.zinit) / self.dz)

    @property
    def dw(self):
        return 2 * math.pi / (self.tmax - self.tmin)

    @property
    def w(self):
        return np.fft.fftfreq(self.nt) * self.nt * self.dw

    @property
    def wmax(self):
        return (0.5 * self.nt - 1) * self.dw


def smooth_velocity(velocity: np.ndarray, sigma: float = 2.0) -> np.ndarray:
    """1 / gaussian_filter(1 / v) (``main.py:40-42``)."""
    return 1.0 / ndimage.gaussian_filter(1.0 / velocity, sigma=sigma)


def ricker_shot(isx_m: float, isz: float, nt: int, nx: int, tj, xi, init_depth: float, v: float,
                fmax: float = 30.0) -> np.ndarray:
    """(nt, nx) float32 source of ``util.makeRickerWavelet`` (``util.py:47-58``)
    for one source position: sum over traces of ricker(t - r/v) / sqrt(r)."""
    tj = np.asarray(tj, np.float64)
    xi = np.asarray(xi, np.float64)
    r = np.sqrt((init_depth - isz) ** 2 + (xi - isx_m) ** 2)
    tt = (tj[:, None] - r[None, :] / v).astype(np.float32).astype(np.float64)
    term = math.pi ** 2 * fmax ** 2 * tt * tt
    return ((1 - 2 * term) * np.exp(-term) / np.sqrt(r)[None, :]).astype(np.float32)


def find_shots(shots_dir: str) -> List[Tuple[int, str]]:
    """(source index, file) of ``seis<ix>_<nx>.csv`` files, sorted by index."""
    out = []
    for f in glob.glob(os.path.join(shots_dir, "*.csv")):
        m = re.search(r"seis(.+?)_", os.path.basename(f))
        if m:
            out.append((int(m.group(1)), f))
    return sorted(out)


def migrate(velocity: np.ndarray, setup_text: str, shot_ix: List[int], shots: List[np.ndarray],
            option: str = "device_hip", log=print, sym: bool = True, extent: int = 5):
    """Smoothed velocity and final image (nz, nx) float32 of a shot-profile
    migration; ``shots`` are (nt, nx) float32 receiver records. ``sym=False``
    uses the edge operators of ``createWLSQoperators`` on the first and last M
    columns (``defWLSQIndices(nx, M, False, False, extent)``); their per-x
    choice needs the expanded tables of ``device_hip_expanded``."""
    if not sym and option != "device_hip_expanded":
        raise ValueError("edge operators (sym=False) are chosen per x: use device_hip_expanded, "
                         "whose expanded tables carry a per-point operator")
    t0 = time.time()
    vel = smooth_velocity(np.asarray(velocity, np.float64))
    nz, nx = vel.shape
    cfg = ProblemConfig.parse(setup_text, nz, nx)
    L = 2 * cfg.M + 1
    kap0, dk, nk = prep.create_kappa(float(vel.max()), float(vel.min()), cfg.dw, cfg.wmax, 0.25)
    kappa = prep.kappa_values(kap0, dk, nk)
    tf, tb = prep.compact_tables(kappa, nx, cfg.dx, cfg.dz, cfg.M)
    slowness = (1.0 / vel).astype(np.float32)
    omega = cfg.w[:cfg.nw].astype(np.float32)
    log(f"config: nz={nz} nx={nx} nt={cfg.nt} nw={cfg.nw} nextrap={cfg.nextrap} M={cfg.M} "
        f"nk={nk} ({time.time() - t0:.1f} s)")

    ns = len(shots)
    v0 = float(vel[0, 0])  # main.py:111
    tj = [j * cfg.dt for j in range(cfg.nt)]
    xi = [cfg.xmin + (i + 1) * cfg.dx for i in range(nx)]
    sources = [ricker_shot(cfg.xmin + shot_ix[s] * cfg.dx, cfg.zinit, cfg.nt, nx, tj, xi, cfg.dz, v0)
               for s in range(ns)]
    t1 = time.time()
    if option == "device_hip":
        image = _migrate_on_device(cfg, nz, nx, slowness, omega, kap0, dk, nk, tf, tb, sources,
                                   [np.asarray(r, np.float32) for r in shots])
    elif option == "device_hip_expanded":
        # the reference-shaped call: host FFTs (main.py:119-123), expanded tables
        F = np.zeros((ns, cfg.nt, nx), np.complex64)
        B = np.zeros((ns, cfg.nt, nx), np.complex64)
        for s in range(ns):
            F[s] = spfft.fft(sources[s], axis=0)
            B[s] = spfft.fft(np.asarray(shots[s], np.float32), axis=0)
        image = np.zeros((ns, nz, nx), np.float32)
        if sym:
            opF = prep.expand_table(tf, slowness, omega, kap0, dk, nk, cfg.nextrap)
            opB = prep.expand_table(tb, slowness, omega, kap0, dk, nk, cfg.nextrap)
        else:  # main.py:65-66 with sym=False, makeTableOfOperators (preparation.py:101-115)
            tfa, tba = prep.compact_tables_all(kappa, nx, cfg.dx, cfg.dz, cfg.M)
            g = prep.wlsq_indices(nx, cfg.M, flip=False, sym=False, extent=extent)
            opF = prep.expand_table(tfa, slowness, omega, kap0, dk, nk, cfg.nextrap, gidx=g)
            opB = prep.expand_table(tba, slowness, omega, kap0, dk, nk, cfg.nextrap, gidx=g)
        interface_hip.extrapolate(ns, cfg.nextrap, nz, cfg.nt, cfg.nw, nx, cfg.M, opF, F, opB, B, image)
    else:
        raise ValueError(f"unknown option {option!r} (device_hip, device_hip_expanded)")
    log(f"extrapolation and imaging: {ns} shots in {time.time() - t1:.2f} s")
    final = np.zeros((nz, nx), np.float32)
    for s in range(ns):
        final += image[s]
    return vel, final


def _migrate_on_device(cfg: ProblemConfig, nz, nx, slowness, omega, kap0, dk, nk, tf, tb,
                       sources, records) -> np.ndarray:
    """Device-resident ``device_hip`` run: real traces are uploaded, their
    spectra computed on the GPU (``main.py:119-123``), the compact tables
    gathered on the device; returns the per-shot images (ns, nz, nx)."""
    import torch
    from . import autotune, device
    dev = torch.device("cuda")
    ns = len(sources)
    F = torch.empty((ns, cfg.nt, nx), dtype=torch.complex64, device=dev)
    B = torch.empty_like(F)
    for s in range(ns):
        device.trace_spectra_device(torch.from_numpy(np.ascontiguousarray(sources[s])).to(dev), out=F[s])
        device.trace_spectra_device(torch.from_numpy(np.ascontiguousarray(records[s])).to(dev), out=B[s])
    tf_d = torch.from_numpy(np.ascontiguousarray(tf)).to(dev)
    tb_d = torch.from_numpy(np.ascontiguousarray(tb)).to(dev)
    tuning = dict(tables_conjugate=int(device.tables_conjugate(tf_d, tb_d)),
                  tables_symmetric=int(autotune.tables_symmetric(tf_d) and autotune.tables_symmetric(tb_d)))
    image = torch.zeros((ns, nz, nx), dtype=torch.float32, device=dev)
    device.ktable_device(ns, cfg.nextrap, nz, cfg.nt, cfg.nw, nx, cfg.M,
                         torch.from_numpy(np.ascontiguousarray(slowness)).to(dev),
                         torch.from_numpy(np.ascontiguousarray(omega)).to(dev), kap0, dk, nk,
                         tf_d, F, tb_d, B, image, tuning=tuning)
    return image.cpu().numpy()


def main(argv: Optional[List[str]] = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) not in (5, 6) or (len(argv) == 6 and argv[5] != "asym"):
        print(__doc__)
        return 2
    file_vel, file_cfg, shots_dir, out_dir, option = argv[:5]
    sym = len(argv) == 5
    velocity = np.genfromtxt(file_vel, delimiter=",")
    found = find_shots(shots_dir)
    shots = [np.genfromtxt(f, delimiter=",", dtype=np.float32) for _, f in found]
    vel, final = migrate(velocity, open(file_cfg).read(), [ix for ix, _ in found], shots, option,
                         sym=sym)
    os.makedirs(out_dir, exist_ok=True)
    np.savetxt(os.path.join(out_dir, "smoothed_velmod.csv"), vel, delimiter=",")
    np.savetxt(os.path.join(out_dir, "final_image.csv"), final, delimiter=",")
    return 0


if __name__ == "__main__":
    sys.exit(main())
