"""Serving loops: continuous admission of migration jobs of mixed size.

Jobs arrive by a Poisson process. Each job names one of the velocity models
kept resident on the GPU (slowness + compact tables), an aperture window, a
depth count, a frequency list and 1-32 shots.

:class:`RoundServer` (continuous batching, the default): the running jobs of a
model advance together in rounds. A round is ONE fused launch that moves
every running job's shots by a number of depth steps from where each one is
(``wlsq_batch_round_ws``); the wavefields stay in HBM between rounds, so a
job admitted while others run joins the very next round of its cohort (its
model's operator tables and its kernel shape) and shares its launches, and a
finished job leaves the batch at its round's end.
Jobs are admitted first come first served as soon as their wavefields and
image fit the device pools (the HBM budget).

:class:`Server` with :class:`GpuLauncher` (one launch per job): the loop

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
boundaries.
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
                ws = torch.empty(max(need, 16 << 20), dtype=torch.uint8, device=self.dev)
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


# --------------------------------------------------------------- continuous batching
def job_shape(job: Job) -> tuple:
    """(x per thread, shots per group) the library's planner picks for this
    job launched alone (2 x per thread for groups of shots up to 1024 wide,
    else 4; two shots per group up to 2048, else one; the CPU path's rounding
    is compiled for 4 x and one shot). Jobs of one model and shape share
    rounds, so a job runs the same kernel shape alone and in the mix."""
    if job.rounding:
        return (4, 1)
    if job.nshots >= 2 and job.nx <= 1024:
        return (2, 4 if job.nshots >= 8 else 2)
    return (4, 2 if job.nshots >= 2 and job.nx <= 2048 else 1)


class RangePool:
    """First-fit allocator of element ranges [0, n) (offsets aligned to
    `align`); freed ranges merge with their neighbours."""

    def __init__(self, n: int, align: int = 64):
        self.n = int(n)
        self.align = int(align)
        self.free: List[List[int]] = [[0, self.n]] if self.n > 0 else []
        self.used = 0

    def _round(self, k: int) -> int:
        return -(-int(k) // self.align) * self.align

    def alloc(self, k: int) -> Optional[int]:
        k = self._round(max(1, k))
        for i, (off, size) in enumerate(self.free):
            if size >= k:
                if size == k:
                    self.free.pop(i)
                else:
                    self.free[i] = [off + k, size - k]
                self.used += k
                return off
        return None

    def release(self, off: int, k: int) -> None:
        k = self._round(max(1, k))
        self.used -= k
        i = 0
        while i < len(self.free) and self.free[i][0] < off:
            i += 1
        self.free.insert(i, [off, k])
        # merge with the right, then the left neighbour
        if i + 1 < len(self.free) and self.free[i][0] + self.free[i][1] == self.free[i + 1][0]:
            self.free[i][1] += self.free[i + 1][1]
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1][1] += self.free[i][1]
            self.free.pop(i)


@dataclass
class _RoundJob:
    job: Job
    key: tuple
    state: object          # the engine's per-job data (pool ranges, shot descriptors)
    t_admit: float
    z: int = 0             # next depth step
    done_round: int = -1   # round after which the job is complete
    joined: bool = False   # takes part in rounds (its inputs are ready and a slot was free)


@dataclass
class _Round:
    rid: int
    handle: object
    jobs: List[_RoundJob]
    steps: int
    points: int


class RoundServer:
    """Continuous batching over an engine (GPU, or a fake one in CPU tests).

    Per cohort (model and kernel shape) at most `depth` rounds are queued on
    the device; each round advances every running job of the cohort by
    `steps` depth steps, chosen so a round holds about `round_points` points
    (between `min_steps` and `max_steps`, a multiple of `min_steps`). Arrived
    jobs are admitted first come first served while fewer than `max_active`
    + `stage` are admitted and the engine can place their wavefields and
    image; an admitted job joins the rounds once its inputs are ready and
    fewer than `max_active` jobs run. With `stage` > 0 the inputs of the next
    jobs are generated while `max_active` jobs run, so a job that leaves is
    replaced at once instead of after its successor's input generation."""

    def __init__(self, engine, round_points: float = 3e9, min_steps: int = 16,
                 max_steps: int = 512, depth: int = 2, max_active: int = 8, lanes: int = 1,
                 stage: int = 0):
        self.engine = engine
        self.stage = max(0, int(stage))
        # jobs of one cohort are spread over `lanes` round sequences (the lane
        # with the least work left takes a new job); the lanes' rounds are
        # independent, so one lane's round can fill the tail of another's
        self.lanes = max(1, int(lanes))
        self.round_points = float(round_points)
        self.min_steps = int(min_steps)
        self.max_steps = int(max_steps)
        self.depth = int(depth)
        self.max_active = int(max_active)
        self.rounds_launched = 0
        # every round: (cohort, [(job id, first step)], steps, points)
        self.round_log: List[tuple] = []

    def _steps(self, jobs: List[_RoundJob]) -> int:
        per_step = sum(r.job.nx * r.job.nf * r.job.nshots for r in jobs)
        d = int(self.round_points // max(1, per_step))
        d = max(self.min_steps, min(self.max_steps, d))
        return max(self.min_steps, d // self.min_steps * self.min_steps)

    def run(self, jobs: Sequence[Job], keep_images: bool = False, time_scale: float = 1.0,
            log=None, log_every: int = 1, clock=time.perf_counter,
            idle=lambda: time.sleep(2e-4)) -> ServeStats:
        """Serve `jobs` (arrival times in seconds, divided by `time_scale`);
        `log` gets a line every `log_every` completed jobs."""
        eng = self.engine
        st = ServeStats()
        for j in jobs:
            if not eng.fits(j):
                raise ValueError(f"job {j.job_id} does not fit the device pools")
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        waiting: deque = deque()
        cohorts: Dict[tuple, List[_RoundJob]] = {}
        inflight: Dict[tuple, deque] = {}
        staged: deque = deque()   # admitted, not yet running (admission order)
        n_admitted = 0            # running and staged jobs
        n_running = 0
        eng.sync()
        t0 = clock()
        while pending or waiting or n_admitted:
            now = clock() - t0
            progressed = False
            while pending and pending[0].arrival / time_scale <= now:
                waiting.append(pending.popleft())
                progressed = True
            # completed rounds: their finished jobs leave
            for key, q in inflight.items():
                while q and eng.done(q[0].handle):
                    rd = q.popleft()
                    progressed = True
                    for r in rd.jobs:
                        if r.done_round != rd.rid:
                            continue
                        t_done = clock() - t0
                        st.latencies.append(t_done - r.job.arrival / time_scale)
                        st.order.append(r.job.job_id)
                        st.points += r.job.points
                        if keep_images:
                            st.images[r.job.job_id] = eng.image(r.state)
                        eng.release(r.state)
                        cohorts[key].remove(r)
                        n_admitted -= 1
                        n_running -= 1
                        if log and len(st.latencies) % max(1, log_every) == 0:
                            log(f"job {r.job.job_id} done ({len(st.latencies)} of {len(jobs)}): "
                                f"{r.job.points:.3g} pts, latency {st.latencies[-1]:.3f} s")
            # admission: first come first served, while the pools place the job
            while waiting and n_admitted < self.max_active + self.stage:
                job = waiting[0]
                state = eng.admit(job)
                if state is None:
                    break  # wait for memory to free up
                waiting.popleft()
                t_adm = clock() - t0
                st.queue_waits.append(t_adm - job.arrival / time_scale)
                base = eng.cohort(job)
                left = lambda k: sum(r.job.nx * r.job.nf * r.job.nshots * (r.job.nz - r.z)  # noqa: E731
                                     for r in cohorts.get(k, []))
                key = min((base + (ln,) for ln in range(self.lanes)), key=left)
                rj = _RoundJob(job, key, state, t_adm)
                cohorts.setdefault(key, []).append(rj)
                inflight.setdefault(key, deque())
                staged.append(rj)
                n_admitted += 1
                st.peak_bytes = max(st.peak_bytes, eng.bytes_in_use())
                progressed = True
            # admitted jobs start running, in admission order among those whose
            # inputs are ready, while a slot is free
            for r in list(staged):
                if n_running >= self.max_active:
                    break
                if eng.ready(r.state):
                    staged.remove(r)
                    r.joined = True
                    n_running += 1
                    st.max_active = max(st.max_active, n_running)
                    progressed = True
            # rounds: every cohort keeps up to `depth` rounds queued
            for key, members in cohorts.items():
                q = inflight[key]
                while len(q) < self.depth:
                    run = [r for r in members if r.z < r.job.nz and r.joined]
                    if not run:
                        break
                    steps = self._steps(run)
                    rid = self.rounds_launched
                    self.rounds_launched += 1
                    pts = sum(r.job.nx * r.job.nf * r.job.nshots * (min(r.job.nz, r.z + steps) - r.z)
                              for r in run)
                    h = eng.launch_round(key, [(r.state, r.z) for r in run], steps)
                    self.round_log.append((key, [(r.job.job_id, r.z) for r in run], steps, pts))
                    for r in run:
                        r.z = min(r.job.nz, r.z + steps)
                        if r.z >= r.job.nz:
                            r.done_round = rid
                    q.append(_Round(rid, h, run, steps, pts))
                    progressed = True
            if not progressed:
                idle()
        st.wall = clock() - t0
        return st


class _PoolJob:
    """A job's place in the engine's pools and its shot descriptors."""

    def __init__(self, job, f_off, f_len, i_off, i_len, o_off, specs, nbytes):
        self.job = job
        self.f_off, self.f_len = f_off, f_len
        self.i_off, self.i_len = i_off, i_len
        self.o_off = o_off
        self.specs = specs
        self.nbytes = nbytes


class GpuRoundEngine:
    """Device side of :class:`RoundServer`: the resident models, three device
    pools (wavefields, images, frequency lists) carved out of the HBM budget,
    one HIP stream and two launch workspaces per cohort (table family, kernel
    shape, rounding). Models whose operator tables and kappa grids are equal
    bit for bit form one table family: their slowness models sit side by side
    in one device array, so jobs on any of them share rounds. A job's
    wavefields are generated in its pool range on its cohort's stream; a
    round is one ``wlsq_batch_round_ws`` launch."""

    def __init__(self, models: Sequence[synth.Problem], hbm_budget: int = 32 << 30, dev="cuda",
                 image_share: float = 0.1, omega_slots: int = 1 << 20,
                 one_shot_tuning: Optional[dict] = None):
        self.surveys = [batch.Survey(p, dev) for p in models]
        self.dev = torch.device(dev)
        self.M = models[0].tables.M
        # extra tuning of the cohorts with one shot per group (fast path):
        # by default two strips per thread with the whole LDS, so a task holds
        # two frequency rows of apertures up to about 5000 (measured on the
        # 2000-job stream: 6-7 % above one strip per thread)
        self.one_shot_tuning = dict(strips_per_thread=2, lds_budget=160 << 10) \
            if one_shot_tuning is None else dict(one_shot_tuning)
        # table families: the first model of each family holds its tables
        self.family: List[int] = []
        for i, p in enumerate(models):
            t = p.tables
            fam = i
            for k in range(i):
                u = models[k].tables
                if (self.family[k] == k and (t.M, t.nk, t.kappa0, t.dkappa) == (u.M, u.nk, u.kappa0, u.dkappa)
                        and np.array_equal(t.forw.view(np.uint32), u.forw.view(np.uint32))
                        and np.array_equal(t.back.view(np.uint32), u.back.view(np.uint32))):
                    fam = k
                    break
            self.family.append(fam)
        # the slowness models of a family in one array: model i at slow_base[i]
        self.slow_base: List[int] = [0] * len(models)
        self.family_slowness: Dict[int, torch.Tensor] = {}
        for fam in sorted(set(self.family)):
            members = [i for i, f in enumerate(self.family) if f == fam]
            o = 0
            for i in members:
                self.slow_base[i] = o
                o += models[i].model.slowness.size
            self.family_slowness[fam] = torch.cat(
                [self.surveys[i].slowness.reshape(-1) for i in members]).contiguous()
        img_bytes = int(hbm_budget * image_share)
        self.fpool = RangePool((int(hbm_budget) - img_bytes) // 8)
        self.ipool = RangePool(img_bytes // 4)
        self.opool = RangePool(omega_slots, align=16)
        self.fields = torch.empty(self.fpool.n, dtype=torch.complex64, device=self.dev)
        self.images = torch.zeros(self.ipool.n, dtype=torch.float32, device=self.dev)
        self.omegas = torch.zeros(self.opool.n, dtype=torch.float32, device=self.dev)
        self.streams: Dict[tuple, torch.cuda.Stream] = {}
        self.ws: Dict[tuple, List[Optional[torch.Tensor]]] = {}
        self.flip: Dict[tuple, int] = {}
        # inputs of admitted jobs are generated here; a job joins rounds once
        # they are ready, so running jobs never wait for a new job's inputs
        self.gen_stream = torch.cuda.Stream(device=self.dev)

    def cohort(self, job: Job) -> tuple:
        return (self.family[job.model],) + job_shape(job) + (job.rounding,)

    def fits(self, job: Job) -> bool:
        return (2 * job.nshots * job.nf * job.nx <= self.fpool.n and job.nz * job.nx <= self.ipool.n
                and job.nf <= self.opool.n)

    def bytes_in_use(self) -> int:
        return 8 * self.fpool.used + 4 * self.ipool.used

    def _stream(self, key):
        if key not in self.streams:
            self.streams[key] = torch.cuda.Stream(device=self.dev)
            self.ws[key] = [None, None]
            self.flip[key] = 0
        return self.streams[key]

    def admit(self, job: Job):
        n = job.nshots * job.nf * job.nx
        f_off = self.fpool.alloc(2 * n)
        if f_off is None:
            return None
        i_off = self.ipool.alloc(job.nz * job.nx)
        if i_off is None:
            self.fpool.release(f_off, 2 * n)
            return None
        o_off = self.opool.alloc(job.nf)
        if o_off is None:
            self.fpool.release(f_off, 2 * n)
            self.ipool.release(i_off, job.nz * job.nx)
            return None
        p = self.surveys[job.model].problem
        shots = job.shots()
        self._stream(self.cohort(job))
        with torch.cuda.stream(self.gen_stream):
            # F of the job's shots in the first half of its range, B in the second
            _, _, offs = synth.device_fields(p, shots, device=self.dev,
                                             out_f=self.fields[f_off:f_off + n],
                                             out_b=self.fields[f_off + n:f_off + 2 * n])
            self.images[i_off:i_off + job.nz * job.nx].zero_()
            self.omegas[o_off:o_off + job.nf].copy_(torch.from_numpy(shots[0].omega))
        ready = torch.cuda.Event()
        ready.record(self.gen_stream)
        specs = [device.ShotSpec(f_off=f_off + o, b_off=f_off + n + o, img_off=i_off,
                                 slow_off=self.slow_base[job.model] + job.x0, ld=job.nx,
                                 img_ld=job.nx, slow_ld=p.model.nx,
                                 nx=job.nx, nz=job.nz, nf=job.nf, omega_off=o_off) for o in offs]
        st = _PoolJob(job, f_off, 2 * n, i_off, job.nz * job.nx, o_off, specs,
                      8 * 2 * n + 4 * job.nz * job.nx)
        st.ready = ready
        return st

    def ready(self, state) -> bool:
        """The job's inputs are on the device (it may join a round)."""
        return state.ready.query()

    def warm(self, lanes: int = 1, shapes=((2, 2), (2, 4), (4, 2), (4, 1))) -> None:
        """One one-step round of a tiny shot per table family, kernel shape
        and lane, so each kernel variant is loaded and each lane's stream and
        workspaces exist before the first job (a server's start-up, not part
        of its service)."""
        for fam in sorted(set(self.family)):
            p = self.surveys[fam].problem
            for R, S in shapes:
                f_off, i_off, o_off = self.fpool.alloc(128), self.ipool.alloc(64), self.opool.alloc(1)
                spec = device.ShotSpec(f_off=f_off, b_off=f_off + 64, img_off=i_off,
                                       slow_off=self.slow_base[fam], ld=64, img_ld=64,
                                       slow_ld=p.model.nx, nx=64, nz=1, nf=1, omega_off=o_off)
                self.fields[f_off:f_off + 128].zero_()
                self.omegas[o_off:o_off + 1].fill_(10.0)
                torch.cuda.synchronize()
                for lane in range(max(1, lanes)):
                    for _ in range(2):  # both workspaces of the lane
                        self._launch((fam, R, S, 0, lane), [spec], 1)
                torch.cuda.synchronize()
                self.fpool.release(f_off, 128)
                self.ipool.release(i_off, 64)
                self.opool.release(o_off, 1)

    def _tuning(self, key) -> dict:
        fam, R, S, rounding = key[:4]
        sv = self.surveys[fam]
        t = dict(tables_conjugate=int(sv.conj), tables_symmetric=int(sv.sym), x_per_thread=R,
                 shots_per_group=S)
        if rounding:
            t["rounding"] = 1
        elif S == 1:
            t.update(self.one_shot_tuning)
        return t

    def _ws_bytes(self, specs, nk) -> int:
        """Upper bound of wlsq_batch_workspace_bytes: a group per shot (144
        bytes), a task per (shot, frequency) (16 bytes), two row-padded
        table copies, each part 256-byte aligned."""
        al = lambda b: (b + 255) // 256 * 256  # noqa: E731
        L = 2 * self.M + 1
        return (al(144 * len(specs)) + al(16 * sum(s.nf for s in specs)) +
                2 * al(8 * (L + 1) * nk) + 256)

    def launch_round(self, key, members, steps):
        specs = []
        for state, z in members:
            for sp in state.specs:
                sp.first_step = z
                specs.append(sp)
        return self._launch(key, specs, steps)

    def _launch(self, key, specs, steps):
        fam = key[0]
        sv = self.surveys[fam]
        t = sv.problem.tables
        stream = self._stream(key)
        k = self.flip[key]
        self.flip[key] = k ^ 1
        need = self._ws_bytes(specs, t.nk)
        ws = self.ws[key][k]
        if ws is None or ws.numel() < need:
            with torch.cuda.stream(stream):
                ws = torch.empty(max(need, 4 << 20), dtype=torch.uint8, device=self.dev)
            self.ws[key][k] = ws
        device.batch_round_ws(specs, steps, t.M, self.family_slowness[fam], self.omegas, t.kappa0,
                              t.dkappa, t.nk,
                              sv.tf, sv.tb, self.fields, self.fields, self.images, ws,
                              tuning=self._tuning(key), stream=stream, check_extents=False)
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev

    def done(self, handle) -> bool:
        return handle.query()

    def image(self, state) -> np.ndarray:
        j = state.job
        return self.images[state.i_off:state.i_off + j.nz * j.nx].view(j.nz, j.nx).cpu().numpy()

    def release(self, state) -> None:
        self.fpool.release(state.f_off, state.f_len)
        self.ipool.release(state.i_off, state.i_len)
        self.opool.release(state.o_off, state.job.nf)

    def sync(self):
        torch.cuda.synchronize()

    def close(self):
        torch.cuda.synchronize()
        self.ws = {k: [None, None] for k in self.ws}
