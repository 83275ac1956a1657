"""Serving loop logic on the CPU: a fake launcher with a simulated clock
stands in for the GPU, so admission by HBM footprint and free slots,
first-come-first-served order, latency accounting and failure handling run
without a device."""
import numpy as np
import pytest

from imaging2d_wlsq_amd import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def idle(self):
        self.t += 1e-3


class FakeLauncher:
    """Each slot runs one job at `rate` points/s; records what was running."""

    def __init__(self, clock, slots=3, rate=1e9, budget=None):
        self.clock = clock
        self.slots = slots
        self.rate = rate
        self.running = {}
        self.max_running = 0
        self.max_bytes = 0
        self.launched = []

    class H:
        def __init__(self, outer, job, t_end):
            self.outer, self.job, self.t_end = outer, job, t_end

        def done(self):
            if self.outer.clock() >= self.t_end:
                self.outer.running.pop(self.job.job_id, None)
                return True
            return False

        def image(self):
            return np.full((2, 2), float(self.job.job_id))

    def launch(self, job, slot):
        assert slot not in [s for s, _ in self.running.values()], "slot reused while busy"
        self.running[job.job_id] = (slot, job)
        self.launched.append(job.job_id)
        self.max_running = max(self.max_running, len(self.running))
        self.max_bytes = max(self.max_bytes, sum(j.footprint() for _, j in self.running.values()))
        return FakeLauncher.H(self, job, self.clock() + job.points / self.rate)

    def sync(self):
        pass


def _jobs(n=40, rate=50.0, seed=7, small=True):
    kw = dict(nx_range=(64, 256), nz_range=(10, 40), nf_range=(4, 16), shots_range=(1, 4)) if small else {}
    return serve.make_jobs(n, rate, 2, 100, 512, seed=seed, **kw)


def test_job_stream_is_seeded_and_within_ranges():
    a = serve.make_jobs(200, 10.0, 4, 1500, 4096, seed=4)
    b = serve.make_jobs(200, 10.0, 4, 1500, 4096, seed=4)
    assert [(j.nx, j.nz, j.nf, j.nshots, j.x0, j.model) for j in a] == \
        [(j.nx, j.nz, j.nf, j.nshots, j.x0, j.model) for j in b]
    assert all(512 <= j.nx <= 4096 and 500 <= j.nz <= 1500 and 128 <= j.nf <= 512 and
               1 <= j.nshots <= 32 and 0 <= j.x0 <= 4096 - j.nx and 0 <= j.model < 4 for j in a)
    arr = np.array([j.arrival for j in a])
    assert (np.diff(arr) > 0).all()
    assert 0.6 < 200 / arr[-1] / 10.0 < 1.5  # Poisson rate ~ 10 jobs/s
    mean_pts = np.mean([j.points for j in a])
    assert 0.6e10 < mean_pts < 2.0e10  # ~1.2e10 points per job on average


@pytest.mark.parametrize("slots,budget_jobs", [(1, 100), (3, 100), (4, 2)])
def test_admission_respects_slots_and_budget(slots, budget_jobs):
    jobs = _jobs()
    clock = FakeClock()
    budget = budget_jobs * max(j.footprint() for j in jobs)
    L = FakeLauncher(clock, slots=slots)
    st = serve.Server(L, hbm_budget=budget).run(jobs, clock=clock, idle=clock.idle, keep_images=True)
    assert sorted(st.order) == [j.job_id for j in jobs]
    assert L.launched == sorted(L.launched)            # first come, first served
    assert L.max_running <= slots and st.max_active <= slots
    assert L.max_bytes <= budget and st.peak_bytes <= budget
    svc = {j.job_id: j.points / L.rate for j in jobs}
    for jid, lat in zip(st.order, st.latencies):
        assert lat >= svc[jid] - 1e-9                   # latency covers the service time
    assert all(st.images[j.job_id][0, 0] == j.job_id for j in jobs)
    s = st.summary()
    assert s["jobs"] == len(jobs) and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_more_slots_lower_latency_under_load():
    jobs = _jobs(rate=200.0)
    res = {}
    for slots in (1, 4):
        clock = FakeClock()
        st = serve.Server(FakeLauncher(clock, slots=slots), hbm_budget=1 << 40).run(
            jobs, clock=clock, idle=clock.idle)
        res[slots] = st.summary()["p50_latency_s"]
    assert res[4] < res[1]


def test_oversized_job_is_rejected_up_front():
    jobs = _jobs(n=5)
    clock = FakeClock()
    with pytest.raises(ValueError):
        serve.Server(FakeLauncher(clock), hbm_budget=1000).run(jobs, clock=clock, idle=clock.idle)


def test_footprint_counts_fields_image_and_work_list():
    j = serve.Job(0, 0, 0, nx=1000, nz=500, nf=100, nshots=3, arrival=0.0, seed=1)
    fields_image = 2 * 3 * 100 * 1000 * 8 + 500 * 1000 * 4
    tables = 2 * 32 * 4096 * 8       # the launch's two padded table copies (M 15, nk 4096)
    descriptors = 3 * 144 + 3 * 100 * 16
    assert j.footprint() >= fields_image + tables + descriptors
    assert j.footprint(M=25, nk=8192) >= fields_image + 2 * 52 * 8192 * 8
    assert j.points == 1000 * 500 * 100 * 3


# ------------------------------------------------------------ continuous batching
class FakeRoundEngine:
    """One device that runs queued rounds back to back at `rate` points/s,
    with `capacity` bytes of pools; records the depth range every round gave
    every job."""

    def __init__(self, clock, rate=1e9, capacity=1 << 40, gen_time=0.0):
        self.clock = clock
        self.rate = rate
        self.capacity = capacity
        # inputs of admitted jobs are generated one after another, gen_time each
        self.gen_time = gen_time
        self.t_gen = 0.0
        self.held = 0        # admitted and not yet released
        self.max_held = 0
        self.used = 0
        self.t_free = 0.0
        self.ranges = {}     # job id -> [(first step, end step), ...]
        self.rounds = []     # [(cohort, [(job id, first step)], steps)]
        self.admitted = []

    class S:
        def __init__(self, job, nbytes, t_ready=0.0):
            self.job, self.nbytes, self.t_ready = job, nbytes, t_ready

    class H:
        def __init__(self, t_end):
            self.t_end = t_end

    def nbytes(self, job):
        return 2 * job.nshots * job.nf * job.nx * 8 + job.nz * job.nx * 4

    def cohort(self, job):
        return (job.model,) + serve.job_shape(job) + (job.rounding,)

    def fits(self, job):
        return self.nbytes(job) <= self.capacity

    def admit(self, job):
        if self.used + self.nbytes(job) > self.capacity:
            return None
        self.used += self.nbytes(job)
        self.admitted.append(job.job_id)
        self.held += 1
        self.max_held = max(self.max_held, self.held)
        self.t_gen = max(self.clock(), self.t_gen) + self.gen_time
        return FakeRoundEngine.S(job, self.nbytes(job), self.t_gen)

    def launch_round(self, key, members, steps):
        pts = 0
        for s, z in members:
            j = s.job
            assert all(key[:len(self.cohort(m.job))] == self.cohort(m.job) for m, _ in members)
            prev = self.ranges.setdefault(j.job_id, [])
            assert (prev[-1][1] if prev else 0) == z, "a round must continue where the last ended"
            end = min(j.nz, z + steps)
            assert end > z
            prev.append((z, end))
            pts += j.nx * j.nf * j.nshots * (end - z)
        self.rounds.append((key, [(s.job.job_id, z) for s, z in members], steps))
        start = max(self.clock(), self.t_free)
        self.t_free = start + pts / self.rate
        return FakeRoundEngine.H(self.t_free)

    def done(self, h):
        return self.clock() >= h.t_end

    def ready(self, s):
        return self.clock() >= s.t_ready

    def image(self, s):
        return np.full((2, 2), float(s.job.job_id))

    def release(self, s):
        self.used -= s.nbytes
        self.held -= 1

    def bytes_in_use(self):
        return self.used

    def sync(self):
        pass


@pytest.mark.parametrize("max_active,cap_jobs,lanes", [(24, 1000, 1), (3, 1000, 1), (24, 2, 1),
                                                       (24, 1000, 3)])
def test_round_server_runs_every_step_of_every_job(max_active, cap_jobs, lanes):
    """Each job's rounds tile its depth range [0, nz) in order; admission is
    first come first served within the pools and the active-job cap; every job
    completes with its own image."""
    jobs = _jobs(n=60, rate=100.0)
    clock = FakeClock()
    cap = cap_jobs * max(2 * j.nshots * j.nf * j.nx * 8 + j.nz * j.nx * 4 for j in jobs)
    eng = FakeRoundEngine(clock, capacity=cap)
    srv = serve.RoundServer(eng, round_points=2e6, min_steps=4, max_steps=32, max_active=max_active,
                            lanes=lanes)
    st = srv.run(jobs, clock=clock, idle=clock.idle, keep_images=True)
    assert sorted(st.order) == [j.job_id for j in jobs]
    assert eng.admitted == sorted(eng.admitted)              # first come, first served
    assert st.max_active <= max_active and st.peak_bytes <= cap and eng.used == 0
    for j in jobs:
        r = eng.ranges[j.job_id]
        assert r[0][0] == 0 and r[-1][1] == j.nz
        assert all(a[1] == b[0] for a, b in zip(r, r[1:]))
        assert st.images[j.job_id][0, 0] == j.job_id
    for _, members, steps in eng.rounds:
        assert 4 <= steps <= 32 and steps % 4 == 0
    used = {k[-1] for k, _, _ in eng.rounds}
    assert used <= set(range(lanes)) and (len(used) > 1) == (lanes > 1)
    assert st.points == sum(j.points for j in jobs)


def test_jobs_join_running_rounds():
    """Continuous batching: a job that arrives while others run joins the next
    round of its cohort beside them (a round holds jobs at different depths),
    and a short job that arrives behind a long one finishes first."""
    clock = FakeClock()
    long = serve.Job(0, 0, 0, nx=256, nz=400, nf=16, nshots=1, arrival=0.0, seed=1)
    short = serve.Job(1, 0, 0, nx=256, nz=20, nf=16, nshots=1, arrival=0.01, seed=2)
    eng = FakeRoundEngine(clock, rate=1e8)
    srv = serve.RoundServer(eng, round_points=256 * 16 * 8, min_steps=8, max_steps=8)
    st = srv.run([long, short], clock=clock, idle=clock.idle)
    assert st.order == [1, 0]
    mixed = [m for _, m, _ in eng.rounds if len(m) == 2]
    assert mixed and all(dict(m)[0] > dict(m)[1] for m in mixed)


def test_staged_jobs_hide_input_generation():
    """With `stage` > 0 the next jobs are admitted while `max_active` run and
    their inputs are generated then, so a job that leaves is replaced at once:
    the same jobs finish sooner, never more than `max_active` run, at most
    `max_active` + `stage` are admitted, and every job's rounds still tile its
    depth range."""
    jobs = _jobs(n=30, rate=100.0)
    for j in jobs:
        j.arrival = 0.0
    walls, held = {}, {}
    for stage in (0, 2):
        clock = FakeClock()
        mean_run = np.mean([j.points for j in jobs]) / 1e9
        eng = FakeRoundEngine(clock, rate=1e9, gen_time=0.5 * mean_run)
        srv = serve.RoundServer(eng, round_points=2e6, min_steps=4, max_steps=32, max_active=3,
                                stage=stage)
        st = srv.run(jobs, clock=clock, idle=clock.idle)
        assert sorted(st.order) == [j.job_id for j in jobs] and eng.used == 0
        assert st.max_active == 3
        for j in jobs:
            r = eng.ranges[j.job_id]
            assert r[0][0] == 0 and r[-1][1] == j.nz
            assert all(a[1] == b[0] for a, b in zip(r, r[1:]))
        walls[stage], held[stage] = st.wall, eng.max_held
    assert held[0] == 3 and held[2] == 5
    assert walls[2] < 0.95 * walls[0], walls


def test_round_steps_follow_the_round_size():
    eng = FakeRoundEngine(FakeClock())
    srv = serve.RoundServer(eng, round_points=1e6, min_steps=16, max_steps=128)
    mk = lambda nx, nf, ns: serve._RoundJob(serve.Job(0, 0, 0, nx, 100, nf, ns, 0.0, 1), (), None, 0.0)  # noqa: E731
    assert srv._steps([mk(100, 10, 1)]) == 128        # 1000 points per step: capped
    assert srv._steps([mk(1000, 100, 1)]) == 16       # 1e5 per step: at least min_steps
    assert srv._steps([mk(1000, 25, 1)]) == 32        # 2.5e4 per step: 40 steps, cut to 16k


def test_job_shape_matches_the_planner_defaults():
    J = lambda nx, ns, r=0: serve.Job(0, 0, 0, nx, 10, 10, ns, 0.0, 1, rounding=r)  # noqa: E731
    assert serve.job_shape(J(1024, 2)) == (2, 2) and serve.job_shape(J(1024, 8)) == (2, 4)
    assert serve.job_shape(J(1025, 2)) == (4, 2) and serve.job_shape(J(2048, 32)) == (4, 2)
    assert serve.job_shape(J(2049, 2)) == (4, 1) and serve.job_shape(J(512, 1)) == (4, 1)
    assert serve.job_shape(J(512, 8, r=1)) == (4, 1)


def test_range_pool_allocates_disjoint_ranges_and_merges():
    rng = np.random.default_rng(3)
    pool = serve.RangePool(10000, align=16)
    live = {}
    occupied = np.zeros(10000, bool)
    for step in range(2000):
        if live and (rng.random() < 0.45 or len(live) > 40):
            off = list(live)[int(rng.integers(len(live)))]
            k = live.pop(off)
            occupied[off:off + pool._round(k)] = False
            pool.release(off, k)
        else:
            k = int(rng.integers(1, 700))
            off = pool.alloc(k)
            if off is None:
                # first fit failed: no free range holds it
                assert all(size < pool._round(k) for _, size in pool.free)
                continue
            assert off % 16 == 0 and not occupied[off:off + pool._round(k)].any()
            occupied[off:off + pool._round(k)] = True
            live[off] = k
        assert pool.used == int(occupied.sum())
    for off, k in list(live.items()):
        pool.release(off, k)
    assert pool.free == [[0, 10000]] and pool.used == 0
