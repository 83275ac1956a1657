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
