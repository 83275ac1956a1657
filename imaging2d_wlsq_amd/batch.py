"""Batched extrapolate-and-image of many uneven shots in one launch.

A :class:`Survey` keeps one velocity model (slowness), its frequency lists and
the compact operator tables resident on the GPU. :meth:`Survey.batch` lays
out the source/receiver wavefields of a set of shots (each with its own
aperture window, depth count and frequency list) in two flat pools and
returns a :class:`Batch`: the shot descriptors, the fused kernel's plan
(optionally autotuned) and, on request, a hipGraph of [clear image + launch]
for replay. All shots add into one stacked image of the model's shape
(the per-shot sum of ``main.py:184-186``, done on the device).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import autotune, device, synth


def default_tuning(max_nx: int, shots_share: bool, conj: bool, sym: bool, nshots: int = 1) -> dict:
    """Launch shape without timing: the table properties, with the shape left
    to the library's planner, whose defaults come from an autotuner sweep on
    MI355X over apertures 256-8192 and 1-16 shots per group (2 x per thread
    and 2-4 shots for apertures up to 1024 with shared groups, 4 x per thread
    otherwise, one shot per group and 1024-thread workgroups above 2048)."""
    del max_nx, shots_share, nshots
    return dict(tables_conjugate=int(conj), tables_symmetric=int(sym))


class Survey:
    """One model + tables resident on the device."""

    def __init__(self, problem: synth.Problem, dev="cuda"):
        self.problem = problem
        self.dev = torch.device(dev)
        m, t = problem.model, problem.tables
        self.slowness = torch.from_numpy(m.slowness).to(self.dev)
        self.tf = torch.from_numpy(t.forw).to(self.dev)
        self.tb = torch.from_numpy(t.back).to(self.dev)
        self.conj = device.tables_conjugate(self.tf, self.tb)
        self.sym = autotune.tables_symmetric(self.tf) and autotune.tables_symmetric(self.tb)

    @property
    def tables(self):
        return self.problem.tables

    def batch(self, shots: Sequence[synth.ShotGeom], tune: bool = False, tuning: Optional[dict] = None,
              graph: bool = False, write_back: bool = False, reps: int = 2) -> "Batch":
        return Batch(self, list(shots), tune=tune, tuning=tuning, graph=graph,
                     write_back=write_back, reps=reps)


@dataclass
class Batch:
    survey: Survey
    shots: List[synth.ShotGeom]
    tune: bool = False
    tuning: Optional[dict] = None
    graph: bool = False
    write_back: bool = False
    reps: int = 2
    specs: List[device.ShotSpec] = field(default_factory=list)

    def __post_init__(self):
        sv = self.survey
        p = sv.problem
        m = p.model
        dev = sv.dev
        # frequency pool: shots with identical frequency lists share one entry
        om_list, om_index = [], {}
        omega_off = []
        o = 0
        for s in self.shots:
            key = s.freqs.tobytes()
            if key not in om_index:
                om_index[key] = o
                om_list.append(s.omega)
                o += s.nf
            omega_off.append(om_index[key])
        self.omega = torch.from_numpy(np.concatenate(om_list).astype(np.float32)).to(dev)
        self.F, self.B, offs = synth.device_fields(p, self.shots, device=dev)
        self.image = torch.zeros((m.nz, m.nx), dtype=torch.float32, device=dev)
        self.specs = [device.ShotSpec(f_off=off, b_off=off, img_off=s.x0, slow_off=s.x0, ld=s.nx,
                                      img_ld=m.nx, slow_ld=m.nx, nx=s.nx, nz=s.nz, nf=s.nf,
                                      omega_off=oo)
                      for s, off, oo in zip(self.shots, offs, omega_off)]
        self.points = sum(s.nx * s.nz * s.nf for s in self.shots)
        t = p.tables
        self._args = (sv.slowness, self.omega, t.kappa0, t.dkappa, t.nk, sv.tf, sv.tb)
        if self.tuning is None:
            if self.tune:
                self.tuning = autotune.tune(self.specs, t.M, *self._args, self.F, self.B, self.image,
                                            sv.conj, sv.sym, reps=self.reps)
            else:
                share = len({(s.nx, s.x0, s.nf, s.nz) for s in self.shots}) < len(self.shots)
                self.tuning = default_tuning(max(s.nx for s in self.shots), share, sv.conj, sv.sym,
                                             len(self.shots))
        self.plan = device.Plan(self.specs, t.M, self.tuning)
        self._graph = None
        if self.graph:
            self._graph = self.plan.graph(*self._args, self.F, self.B, self.image,
                                          write_back=self.write_back, zero_count=self.image.numel())

    def run(self, stream=None):
        """Stacked image of all shots (replays the graph when captured)."""
        if self._graph is not None:
            self._graph.launch(stream)
        else:
            self.image.zero_()
            self.plan.launch(*self._args, self.F, self.B, self.image, write_back=self.write_back,
                             stream=stream)
        return self.image

    def close(self):
        if self._graph is not None:
            self._graph.close()
            self._graph = None
        self.plan.close()
