"""Serving loop: continuous admission of migration jobs of mixed size.

Jobs arrive by a Poisson process. Each job names one of the velocity models
kept resident on the GPU (slowness + compact tables), an aperture window, a
depth count, a frequency list and 1-32 shots. The loop

* admits an arrived job as soon as a slot (a HIP stream) is free and its HBM
  footprint (wavefields + stacked image + work list) fits the remaining
  budget, first come first served;
* builds the job's wavefields on its stream and queues ONE fused launch for
  all its shots (shots of a job share window and frequencies, so they are
  grouped and share coefficient loads). The launch's work list is built on
  the host and written, with the padded table copies, into the slot's
  workspace by kernels on the slot's stream (``wlsq_batch_device_ws``): a job
  makes no device allocation, frees nothing and never waits for the device;
* polls completion events, frees the job's memory and slot, and records the
  latency arrival -> image ready.

Jobs admitted while others run overlap on the device (their launches share
the CUs through the hardware queues), so the GPU stays busy across job
boundaries: continuous batching at the granularity of workgroup tasks.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import batch, device, synth


@dataclass
class Job:
    job_id: int
    model: int
    x0: int
    nx: int
    nz: int
    nf: int
    nshots: int
    arrival: float
    seed: int
    fmin: float = 2.0
    fmax: float = 50.0
    rounding: int = 0  # 1: the CPU path's rounding (wlsq_tuning.rounding), bitwise fields

    @property
    def points(self) -> int:
        return self.nx * self.nz * self.nf * self.nshots

    def shots(self) -> List[synth.ShotGeom]:
        freqs = self.fmin + (self.fmax - self.fmin) * np.arange(self.nf) / max(1, self.nf - 1)
        return [synth.ShotGeom(self.x0, self.nx, self.nz, (k + 0.5) * self.nx / self.nshots, freqs,
                               self.seed * 64 + k) for k in range(self.nshots)]

    def footprint(self, M: int = 15, nk: int = 4096) -> int:
        """Device bytes the job holds while it runs: wavefields, image, and its
        launch workspace (group and task descriptors at one shot per group and
        one frequency per task, the most there can be, plus the two row-padded
        table copies of (2M+2) x nk entries)."""
        fields = 2 * self.nshots * self.nf * self.nx * 8
        image = self.nz * self.nx * 4
        workspace = self.nshots * (144 + 16 * self.nf) + 2 * (2 * M + 2) * nk * 8 + 1024
        return fields + image + workspace


def make_jobs(n: int, rate: float, n_models: int, model_nz: int, model_nx: int, seed: int = 4,
              nx_range=(512, 4096), nz_range=(500, 1500), nf_range=(128, 512),
              shots_range=(1, 32)) -> List[Job]:
    """Seeded job stream: Poisson arrivals at `rate` jobs/s, sizes uniform in
    the given ranges (clipped to the model)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nx = int(rng.integers(nx_range[0], min(nx_range[1], model_nx) + 1))
        nz = int(rng.integers(nz_range[0], min(nz_range[1], model_nz) + 1))
        nf = int(rng.integers(nf_range[0], nf_range[1] + 1))
        ns = int(rng.integers(shots_range[0], shots_range[1] + 1))
        x0 = int(rng.integers(0, model_nx - nx + 1))
        jobs.append(Job(i, int(rng.integers(0, n_models)), x0, nx, nz, nf, ns, t, seed * 100000 + i))
    return jobs


def serving_models(n_models: int = 4, nz: int = 1500, nx: int = 4096, seed: int = 4,
                   fmax: float = 50.0, nk: int = 4096, M: int = 15) -> List[synth.Problem]:
    """Resident velocity models of the serving configuration."""
    probs = []
    for k in range(n_models):
        rng = np.random.default_rng(seed * 10 + k)
        m = synth.gradient_model(rng, nz, nx, 10.0, 10.0, 1500.0, 4500.0, 3, vmin=1500.0, vmax=4500.0)
        t = synth.make_tables(m, M, nk, 2.0, fmax)
        probs.append(synth.Problem(f"serving_model_{k}", m, t, []))
    return probs


@dataclass
class _Active:
    job: Job
    slot: int
    handle: object
    t_admit: float


@dataclass
class ServeStats:
    latencies: List[float] = field(default_factory=list)
    queue_waits: List[float] = field(default_factory=list)
    order: List[int] = field(default_factory=list)
    points: int = 0
    wall: float = 0.0
    max_active: int = 0
    peak_bytes: int = 0
    images: Dict[int, np.ndarray] = field(default_factory=dict)

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"jobs": len(self.latencies), "points": self.points, "wall_s": self.wall,
                "points_per_s": self.points / self.wall if self.wall > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_queue_wait_s": float(np.mean(self.queue_waits)) if self.queue_waits else 0.0,
                "max_concurrent_jobs": self.max_active, "peak_job_bytes": self.peak_bytes}


class GpuLauncher:
    """Runs a job on the GPU: wavefields built on the job's stream, one fused
    launch for all its shots with its descriptors in the slot's workspace, a
    completion event."""

    class Handle:
        def __init__(self, event, bufs):
            self.event, self.bufs = event, bufs

        def done(self) -> bool:
            return self.event.query()

        def image(self) -> np.ndarray:
            return self.bufs[3].cpu().numpy()

    def __init__(self, models: Sequence[synth.Problem], slots: int, dev="cuda"):
        self.surveys = [batch.Survey(p, dev) for p in models]
        self.dev = torch.device(dev)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(slots)]
        # one grow-only launch workspace per slot: a slot's next job is queued
        # only after its previous one finished
        self._ws: List[Optional[torch.Tensor]] = [None] * slots

    @property
    def slots(self) -> int:
        return len(self.streams)

    def launch(self, job: Job, slot: int) -> "GpuLauncher.Handle":
        stream = self.streams[slot]
        sv = self.surveys[job.model]
        p = sv.problem
        t = p.tables
        shots = job.shots()
        with torch.cuda.stream(stream):
            F, B, offs = synth.device_fields(p, shots, device=self.dev)
            omega = torch.from_numpy(shots[0].omega).to(self.dev)
            image = torch.zeros((job.nz, job.nx), dtype=torch.float32, device=self.dev)
        specs = [device.ShotSpec(f_off=o, b_off=o, img_off=0, slow_off=job.x0, ld=job.nx,
                                 img_ld=job.nx, slow_ld=p.model.nx, nx=job.nx, nz=job.nz,
                                 nf=job.nf, omega_off=0) for o in offs]
        tun = batch.default_tuning(job.nx, job.nshots > 1, sv.conj, sv.sym, job.nshots)
        if job.rounding:
            tun["rounding"] = 1
        need = device.batch_workspace_bytes(specs, t.M, t.nk, tun)
        ws = self._ws[slot]
        if ws is None or ws.numel() < need:
            with torch.cuda.stream(stream):
                ws = torch.empty(max(need, 4 << 20), dtype=torch.uint8, device=self.dev)
            self._ws[slot] = ws
        device.batch_launch_ws(specs, t.M, sv.slowness, omega, t.kappa0, t.dkappa, t.nk, sv.tf,
                               sv.tb, F, B, image, ws, tuning=tun, stream=stream)
        ev = torch.cuda.Event()
        ev.record(stream)
        return GpuLauncher.Handle(ev, (F, B, omega, image, ws))

    def sync(self):
        torch.cuda.synchronize()

    def close(self):
        torch.cuda.synchronize()
        self._ws = [None] * len(self._ws)


class Server:
    """Admission and completion loop over a launcher (GPU, or a fake one in
    CPU tests)."""

    def __init__(self, launcher, hbm_budget: int = 32 << 30):
        self.launcher = launcher
        self.hbm_budget = int(hbm_budget)

    def run(self, jobs: Sequence[Job], keep_images: bool = False, time_scale: float = 1.0,
            log=None, log_every: int = 1, clock=time.perf_counter,
            idle=lambda: time.sleep(2e-4)) -> ServeStats:
        """Serve `jobs` (arrival times in seconds, divided by `time_scale`);
        `log` gets a line every `log_every` completed jobs."""
        st = ServeStats()
        for j in jobs:
            if j.footprint() > self.hbm_budget:
                raise ValueError(f"job {j.job_id} needs {j.footprint()} bytes, above the budget")
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        waiting: deque = deque()
        active: List[_Active] = []
        free_slots = list(range(self.launcher.slots))
        used = 0
        self.launcher.sync()
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            progressed = False
            while pending and pending[0].arrival / time_scale <= now:
                waiting.append(pending.popleft())
                progressed = True
            for a in list(active):
                if a.handle.done():
                    t_done = clock() - t0
                    st.latencies.append(t_done - a.job.arrival / time_scale)
                    st.order.append(a.job.job_id)
                    st.points += a.job.points
                    if keep_images:
                        st.images[a.job.job_id] = a.handle.image()
                    used -= a.job.footprint()
                    free_slots.append(a.slot)
                    active.remove(a)
                    progressed = True
                    if log and len(st.latencies) % max(1, log_every) == 0:
                        log(f"job {a.job.job_id} done ({len(st.latencies)} of {len(jobs)}): "
                            f"{a.job.points:.3g} pts, latency {st.latencies[-1]:.3f} s")
            while waiting and free_slots:
                job = waiting[0]
                need = job.footprint()
                if used + need > self.hbm_budget:
                    break  # first come first served: wait for memory to free up
                waiting.popleft()
                slot = free_slots.pop(0)
                t_adm = clock() - t0
                st.queue_waits.append(t_adm - job.arrival / time_scale)
                h = self.launcher.launch(job, slot)
                active.append(_Active(job, slot, h, t_adm))
                used += need
                st.peak_bytes = max(st.peak_bytes, used)
                st.max_active = max(st.max_active, len(active))
                progressed = True
            if not progressed:
                idle()
        st.wall = clock() - t0
        return st
