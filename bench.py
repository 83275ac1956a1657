"""Benchmark: extrapolated points/s of the WLSQ extrapolate-and-image path.

Workload (one GPU): 512 x 1024 model at 10 m, 16 shots, nf = 256 over
2-50 Hz, 31-tap operators on 2048 kappa samples (complex64, FP32 math).
One step = all 16 shots through every depth step, both wavefields plus the
imaging condition, i.e. 16 * 512 * 1024 * 256 = 2.147e9 points.

* value ("batched"): the fused compact-table kernel - one launch for all 16
  shots, operators gathered on the device by the kappa lookup, wavefield rows
  resident in LDS for all depth steps, image accumulated on the device.
* plain: the drop-in path on the same inputs - expanded tap-major tables
  (2 x 33.3 GB) resident in HBM, one launch sequence per shot (one launch per
  depth step per shot), fixed-order image reduction.
* cpu_baseline: the CPU oracle (C/OpenMP restatement of the reference CPU
  path, compact-table mode) on a bounded sample on this host, the 128 x 256
  case with the reference's expanded tables, and the error of the GPU and CPU
  paths against an FP64 re-run on stated shot subsets.
* c3_batched, c5_survey, serving: the uneven 128-shot batch (hipGraph), one
  GPU's 125-shot share of the 2000 x 8192 survey, and the 2000-job serving
  stream (p50/p99 latency).

Inputs are resident in HBM before the timed region. Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "extrapolated points/s (nx·nz·nf·shots) batched vs plain; p50/p99 job latency"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# vector L1 of 256 CUs at 64 B per clock and 2.4 GHz: the ceiling for the
# compact-table coefficient gathers, which HBM does not see (the tables stay in L2)
L1_PEAK_TBS = 256 * 64 * 2.4e9 / 1e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_point(L):
    return 16 * L + 4


def plain_bytes_per_point(L):
    return 16 * L + 32


def read_profile_traffic(name):
    """HBM bytes per launch of a kernel from a committed rocprofv3 summary
    (profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _fp64_subset(p, shot, tunings, log, label):
    """One shot through every depth step on the GPU (each tuning), the CPU
    oracle and the FP64 re-run on the same inputs: image and final-field
    relative L2 of each pair."""
    import numpy as np
    import torch

    import oracle
    from oracle import fp64
    from imaging2d_wlsq_amd import batch

    m, t = p.model, p.tables
    sv = batch.Survey(p)
    res = {}
    Fi = Bi = None
    dev = {}
    for name, tun in tunings.items():
        base = dict(tables_conjugate=int(sv.conj), tables_symmetric=int(sv.sym))
        b = sv.batch([shot], tuning=dict(base, **tun), write_back=True)
        if Fi is None:
            Fi = b.F.cpu().numpy().reshape(shot.nf, shot.nx)
            Bi = b.B.cpu().numpy().reshape(shot.nf, shot.nx)
        img = b.run().clone()
        torch.cuda.synchronize()
        dev[name] = (img.cpu().numpy()[:shot.nz, shot.x0:shot.x0 + shot.nx],
                     b.F.cpu().numpy().reshape(shot.nf, shot.nx), b.B.cpu().numpy().reshape(shot.nf, shot.nx))
        b.close()
    sl = np.ascontiguousarray(m.slowness[:shot.nz, shot.x0:shot.x0 + shot.nx])
    Fo, Bo = Fi[None].copy(), Bi[None].copy()
    io = np.zeros((1, shot.nz, shot.nx), np.float32)
    oracle.extrapolate_ktable(1, shot.nz, shot.nz, shot.nx, shot.nf, shot.nf, t.M, sl, shot.omega,
                              t.kappa0, t.dkappa, t.nk, t.forw, t.back, Fo, Bo, io)
    t0 = time.perf_counter()
    kt = dict(slowness=sl, omega=shot.omega, kappa0=t.kappa0, dkappa=t.dkappa, nk=t.nk,
              forw=t.forw, back=t.back)
    i64, F64, B64 = fp64.extrapolate(shot.nz, shot.nz, shot.nx, shot.nf, t.M, wave_fs=Fi,
                                     shot_fs=Bi, ktable=kt)
    t64 = time.perf_counter() - t0

    def rel(a, b):
        a = np.asarray(a, np.complex128)
        b = np.asarray(b, np.complex128)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    fo = np.concatenate([Fo[0].ravel(), Bo[0].ravel()])
    f64 = np.concatenate([F64.ravel(), B64.ravel()])
    res["cpu_oracle_vs_fp64"] = {"image": rel(io[0], i64), "fields": rel(fo, f64)}
    for name, (img, Fd, Bd) in dev.items():
        fd = np.concatenate([Fd.ravel(), Bd.ravel()])
        res[f"gpu_{name}_vs_fp64"] = {"image": rel(img, i64), "fields": rel(fd, f64)}
        res[f"gpu_{name}_vs_cpu_oracle"] = {"image": rel(img, io[0]), "fields": rel(fd, fo),
                                            "fields_bitwise": bool(np.array_equal(
                                                fd.view(np.uint64), fo.view(np.uint64)))}
    res["fp64_seconds"] = t64
    log(f"{label}: " + ", ".join(f"{k} {v['image']:.2e}" for k, v in res.items() if isinstance(v, dict)))
    return res


def cpu_baseline(p, log, nshots_sample=16):
    """The CPU leg: the oracle (port of the reference CPU path) timed on this
    host, and the accuracy of both paths against the FP64 re-run.

    * value: compact-table mode on `nshots_sample` shots of the workload (all
      16: about 7 s on 16 cores), all depth steps, every frequency;
    * c1_expanded: the 128 x 256 single-shot case with the expanded
      (nextrap, nf, nx, L) table, as the reference CPU path reads it
      (cpp_src/extrapolation_cpu.cpp:39-89);
    * fp64_errors: image and final-field relative L2 of the GPU (fast path,
      and the CPU path's rounding where stated), the CPU oracle and an FP64
      re-run (py_src/wavefieldPropagation/extrapolation.py:38-66) on shot
      subsets of the workload (one shot, 16 of its 256 frequencies, all 512
      steps) and of the survey (one full-aperture shot, 2 frequencies, all
      2000 steps)."""
    import numpy as np

    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle
    from imaging2d_wlsq_amd import prep, synth

    oracle.lib()
    t, m = p.tables, p.model
    shots = p.shots[:nshots_sample]
    F, B = synth.host_fields(p, shots)
    s = shots[0]
    img = np.zeros((len(shots), m.nz, m.nx), np.float32)
    t0 = time.perf_counter()
    oracle.extrapolate_ktable(len(shots), m.nz, m.nz, m.nx, s.nf, s.nf, t.M, m.slowness, s.omega,
                              t.kappa0, t.dkappa, t.nk, t.forw, t.back, F, B, img)
    dt = time.perf_counter() - t0
    pts = len(shots) * m.nz * m.nx * s.nf
    cpu = _cpu_model()
    out = {"value": pts / dt, "unit": "points/s", "cores": cores, "kind": "port",
           "sample": f"{len(shots)} of {len(p.shots)} shots, all {m.nz} depth steps, nf {s.nf}, "
                     f"nx {m.nx} ({pts:.3g} points, {dt:.1f} s, compact-table mode, {cpu})"}
    log(f"cpu (compact tables, 512x1024): {out['value']:.3e} points/s on {cores} cores")

    c1 = synth.config_c1()
    s1 = c1.shots[0]
    t1, m1 = c1.tables, c1.model
    Wf = prep.expand_table(t1.forw, m1.slowness, s1.omega, t1.kappa0, t1.dkappa, t1.nk, m1.nz)
    Wb = prep.expand_table(t1.back, m1.slowness, s1.omega, t1.kappa0, t1.dkappa, t1.nk, m1.nz)
    F1, B1 = synth.host_fields(c1, c1.shots)
    im1 = np.zeros((1, m1.nz, m1.nx), np.float32)
    reps = []
    for _ in range(3):
        Fa, Ba = F1.copy(), B1.copy()
        t0 = time.perf_counter()
        oracle.extrapolate(1, m1.nz, m1.nz, m1.nx, s1.nf, s1.nf, t1.M, Wf, Wb, Fa, Ba, im1)
        reps.append(time.perf_counter() - t0)
    p1 = m1.nz * m1.nx * s1.nf
    out["c1_expanded"] = {"value": p1 / min(reps), "unit": "points/s", "cores": cores,
                          "seconds": min(reps), "points": p1,
                          "table_bytes_per_direction": int(Wf.nbytes),
                          "what": "128 x 256, nf 64, 1 shot, 25 taps: expanded (nextrap, nf, nx, L) "
                                  "tables as the reference CPU path reads them; best of 3"}
    log(f"cpu (expanded tables, 128x256): {out['c1_expanded']['value']:.3e} points/s")
    del Wf, Wb

    errs = {}
    s2 = p.shots[3]
    sub = synth.ShotGeom(0, m.nx, m.nz, s2.xs, s2.freqs[::16], s2.seed)
    errs["c2_shot3_16freqs_512steps"] = _fp64_subset(
        p, sub, {"fast": {}, "cpu_rounding": dict(rounding=1)}, log, "fp64 errors, 512x1024 subset")
    p5 = synth.config_c5(nshots=4)
    s5 = p5.shots[2]
    sub5 = synth.ShotGeom(0, p5.model.nx, p5.model.nz, s5.xs, s5.freqs[[128, 768]], s5.seed)
    errs["c5_shot_2freqs_2000steps"] = _fp64_subset(
        p5, sub5, {"fast": {}, "cpu_rounding": dict(rounding=1)}, log, "fp64 errors, survey subset")
    out["fp64_errors"] = errs
    return out


def onchip_coefficients(M, L, conj, sym, shots_per_group, points, kernel_ms):
    """The other ceiling of the compact-table kernel: coefficient bytes gathered
    from L1/L2 per point (M + 1 taps per x with tap-symmetric tables, else L;
    one table when they are conjugate; shared by the shots of a group)."""
    taps = (M + 1) if sym else L
    per_point = taps * 8 * (1 if conj else 2) / max(1, shots_per_group)
    tbs = per_point * points / (kernel_ms / 1e3) / 1e12
    return {"what": "coefficient gathers (vector L1/L2)", "bytes_per_point": per_point,
            "achieved": tbs, "peak": L1_PEAK_TBS, "unit": "TB/s", "frac": tbs / L1_PEAK_TBS}


PROBE_BLOCKS, PROBE_ITERS = 8192, 8192


def probe_fma(log):
    """The FMA probe: 8192 workgroups x 4 waves, each wave 8 x 8192 FP32 FMAs
    in independent chains; HIP-event time of the second of two launches."""
    import torch

    from imaging2d_wlsq_amd._lib import check, lib

    out = torch.empty(PROBE_BLOCKS * 256, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(2):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        check(lib().wlsq_probe_fma(PROBE_BLOCKS, PROBE_ITERS, out.data_ptr(), stream.cuda_stream))
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    waves = PROBE_BLOCKS * 4
    fma = waves * 8 * PROBE_ITERS
    rate = fma / (ms[-1] / 1e3)
    log(f"fma probe: {ms[-1]:.3f} ms, {rate:.3e} wave64 FMA instructions/s "
        f"({rate * 64 * 2 / 1e12:.1f} TFLOP/s)")
    return {"ms": ms[-1], "wave_fma_instructions": fma, "wave_instructions_per_s": rate,
            "tflops": rate * 64 * 2 / 1e12}


def bench_c3(args, log):
    """128 shots of uneven aperture / frequency count on an 800 x 2048 model in
    one launch, replayed as a hipGraph; launch shape autotuned on the first
    --c3-tune-shots shots (all 128 by default)."""
    import torch

    from imaging2d_wlsq_amd import batch, synth

    t0 = time.perf_counter()
    p = synth.config_c3()
    sv = batch.Survey(p)
    tuning = None
    if args.c3_tuning:
        tuning = dict(json.loads(args.c3_tuning), tables_conjugate=int(sv.conj),
                      tables_symmetric=int(sv.sym))
    elif not args.no_tune:
        t_tune = time.perf_counter()
        probe = sv.batch(p.shots[:args.c3_tune_shots], tune=True, reps=1)
        tuning = probe.tuning
        log(f"c3: tuned on {min(args.c3_tune_shots, len(p.shots))} shots in "
            f"{time.perf_counter() - t_tune:.1f} s: {tuning}")
        probe.close()
        del probe
    b = sv.batch(p.shots, tuning=tuning, graph=True)
    setup = time.perf_counter() - t0
    b.run()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.c3_steps)]
    t1 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        b.run()
        e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t1) / args.c3_steps
    ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
    res = {"workload": p.name, "shots": len(p.shots), "points_per_step": b.points,
           "value": b.points / (ms / 1e3), "unit": "points/s", "ms_per_step": ms,
           "wall_ms_per_step": wall * 1e3, "tuning": b.tuning, "plan": b.plan.info(),
           "graph": True, "setup_s": setup,
           "achieved_tflops": flops_per_point(p.tables.L) * b.points / (ms / 1e3) / 1e12}
    log(f"c3: {ms:.1f} ms/step, {res['value']:.3e} points/s, {res['achieved_tflops']:.1f} TFLOP/s")
    b.close()
    return res


def bench_c5(args, log):
    """Survey-scale shots (2000 x 8192 at 6.25 m, nf 1024 over 2-80 Hz, 51 taps
    x 8192 kappa samples): --c5-shots full-aperture shots in one launch."""
    import torch

    from imaging2d_wlsq_amd import batch, synth

    t0 = time.perf_counter()
    p = synth.config_c5()
    sv = batch.Survey(p)
    shots = p.shots[:: max(1, len(p.shots) // args.c5_shots)][:args.c5_shots]
    warm = sv.batch(shots[:1])  # the same kernel on one shot (first-launch costs)
    warm.run()
    torch.cuda.synchronize()
    warm.close()
    del warm
    b = sv.batch(shots)
    setup = time.perf_counter() - t0
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    pts_s = b.points / (ms / 1e3)
    res = {"workload": p.name, "shots_measured": len(shots), "shots_in_survey": len(p.shots),
           "points_per_launch": b.points, "value": pts_s, "unit": "points/s", "ms_per_launch": ms,
           "tuning": b.tuning, "plan": b.plan.info(), "setup_s": setup,
           "achieved_tflops": flops_per_point(p.tables.L) * b.points / (ms / 1e3) / 1e12,
           "survey_points": p.points, "survey_estimated_s": p.points / pts_s}
    log(f"c5: {len(shots)} shots {ms:.0f} ms, {pts_s:.3e} points/s, {res['achieved_tflops']:.1f} TFLOP/s, "
        f"full survey ~{res['survey_estimated_s']:.0f} s")
    b.close()
    return res


def bench_serving(args, log):
    """Continuous admission of jobs of mixed size against 4 resident models.

    --serve-mode rounds (default): continuous batching, the running jobs of a
    model advanced together by round launches that later jobs join;
    streams: one launch per job on one of --serve-slots streams.
    A saturation pass (the first jobs all arrive at once) measures the loop's
    service rate; the timed run then draws Poisson arrivals at
    --serve-load times that rate."""
    import numpy as np

    from imaging2d_wlsq_amd import serve

    t0 = time.perf_counter()
    models = serve.serving_models()
    budget = 32 << 30
    rounds = args.serve_mode == "rounds"
    if rounds:
        extra = {"strips_per_thread": args.serve_strips, "lds_budget": args.serve_lds_budget}
        launcher = serve.GpuRoundEngine(models, hbm_budget=budget, one_shot_tuning=extra)
        launcher.warm(lanes=args.serve_lanes)
        mk = lambda: serve.RoundServer(launcher, round_points=args.serve_round_points,  # noqa: E731
                                       max_active=args.serve_max_active, lanes=args.serve_lanes,
                                       min_steps=args.serve_min_steps, stage=args.serve_stage)
    else:
        launcher = serve.GpuLauncher(models, args.serve_slots)
        mk = lambda: serve.Server(launcher, hbm_budget=budget)  # noqa: E731
    setup = time.perf_counter() - t0
    nzm, nxm = models[0].model.nz, models[0].model.nx
    calib = serve.make_jobs(min(48, args.serve_jobs), 1.0, len(models), nzm, nxm, seed=4)
    for j in calib:
        j.arrival = 0.0
    cst = mk().run(calib)
    cap_pts = cst.points / cst.wall
    log(f"serving saturation pass: {len(calib)} jobs at once, {cap_pts:.3e} points/s, "
        f"p50 {cst.summary()['p50_latency_s']:.3f} s")
    probe = serve.make_jobs(2000, 1.0, len(models), nzm, nxm, seed=4)
    mean_pts = float(np.mean([j.points for j in probe]))
    rate = args.serve_load * cap_pts / mean_pts
    jobs = serve.make_jobs(args.serve_jobs, rate, len(models), nzm, nxm, seed=4)
    srv = mk()
    st = srv.run(jobs, log=log if args.serve_jobs >= 500 else None, log_every=100)
    launcher.close()
    s = st.summary()
    if rounds:
        n_in = [len(m) for _, m, _, _ in srv.round_log]
        s.update({"mode": "continuous batching: rounds of one fused launch per model and kernel "
                          "shape, jobs joining between rounds",
                  "rounds": srv.rounds_launched, "mean_jobs_per_round": float(np.mean(n_in)),
                  "rounds_with_jobs_at_different_depths":
                      sum(len({z for _, z in m}) > 1 for _, m, _, _ in srv.round_log),
                  "round_points": args.serve_round_points, "max_active": args.serve_max_active,
                  "staged": args.serve_stage,
                  "lanes": args.serve_lanes, "min_steps": args.serve_min_steps,
                  "one_shot_tuning": launcher.one_shot_tuning})
    else:
        s.update({"mode": "one launch per job on its own stream", "slots": args.serve_slots})
    s.update({"arrival_rate_jobs_per_s": rate, "offered_load": args.serve_load,
              "saturated_points_per_s": cap_pts,
              "hbm_budget_bytes": budget, "models": len(models), "model_shape": [nzm, nxm],
              "setup_s": setup,
              "note": f"first {args.serve_jobs} jobs of the seeded 2000-job stream; arrival rate = "
                      f"{args.serve_load} x the saturated service rate of the first "
                      f"{len(calib)} jobs"})
    log(f"serving: {s['jobs']} jobs, {s['points_per_s']:.3e} points/s (saturated {cap_pts:.3e}), "
        f"p50 {s['p50_latency_s']:.3f} s, p99 {s['p99_latency_s']:.3f} s, "
        f"max concurrent {s['max_concurrent_jobs']}")
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--plain-shots", type=int, default=0,
                    help="shots in the plain leg (0: all; fewer bounds the launch count for profiling)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--shots-per-group", type=int, default=0)
    ap.add_argument("--omegas-per-item", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--x-per-thread", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true",
                    help="uneven-batch leg: use the planner's default shape instead of the autotuner")
    ap.add_argument("--tune-c2", action="store_true",
                    help="main leg: autotune the launch shape (default: the planner's shape, which "
                         "the committed profile covers)")
    ap.add_argument("--exact-steps", type=int, default=5,
                    help="timed launches of the main workload with exact image sums (0: skip)")
    ap.add_argument("--c3-steps", type=int, default=3, help="timed steps of the uneven batch (0: skip)")
    ap.add_argument("--c3-tune-shots", type=int, default=128,
                    help="shots of the uneven batch the launch shape is tuned on")
    ap.add_argument("--c3-tuning", type=str, default="",
                    help="launch shape of the uneven batch as JSON (no tuning run), e.g. the "
                         "tuner's choice, for profiler passes")
    ap.add_argument("--serve-jobs", type=int, default=2000,
                    help="jobs of the serving run (0: skip; the configuration is 2000)")
    ap.add_argument("--serve-slots", type=int, default=8, help="streams of --serve-mode streams")
    ap.add_argument("--serve-mode", choices=["rounds", "streams"], default="rounds")
    ap.add_argument("--serve-round-points", type=float, default=3e9,
                    help="points per round of --serve-mode rounds")
    ap.add_argument("--serve-max-active", type=int, default=8,
                    help="running jobs at most (--serve-mode rounds)")
    ap.add_argument("--serve-stage", type=int, default=2,
                    help="jobs admitted beyond --serve-max-active whose inputs are generated ahead "
                         "(--serve-mode rounds)")
    ap.add_argument("--serve-lanes", type=int, default=4,
                    help="round sequences per cohort (--serve-mode rounds)")
    ap.add_argument("--serve-strips", type=int, default=2, choices=[1, 2],
                    help="strips per thread of the one-shot cohorts (--serve-mode rounds)")
    ap.add_argument("--serve-lds-budget", type=int, default=160 << 10,
                    help="LDS budget of the one-shot cohorts (--serve-mode rounds)")
    ap.add_argument("--serve-min-steps", type=int, default=64,
                    help="depth steps per round at least (--serve-mode rounds)")
    ap.add_argument("--c5-shots", type=int, default=125,
                    help="shots of the 2000 x 8192 survey measured in one launch (0: skip; 125 is "
                         "one GPU's share of the 1000-shot survey over 8 GPUs)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="arrival rate as a multiple of the measured saturated service rate")
    ap.add_argument("--tune-verbose", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device copy (calibrates HBM byte counters)")
    ap.add_argument("--probe-fma", action="store_true",
                    help="run the FMA probe of known instruction count (calibrates the VALU "
                         "counters; its time gives the reachable FMA issue rate)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "shot across GPUs with one image reduction at the end, which is not built here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from imaging2d_wlsq_amd import autotune, device, synth
    from imaging2d_wlsq_amd._lib import LAYOUT_TAPMAJOR, lib

    lib()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("building workload ...")
    p = synth.config_c2()
    t, m = p.tables, p.model
    L = t.L
    ns, nz, nx = len(p.shots), m.nz, m.nx
    nf = p.shots[0].nf
    points = p.points
    slow = torch.from_numpy(m.slowness).to(dev)
    omega = torch.from_numpy(p.shots[0].omega).to(dev)
    tf = torch.from_numpy(t.forw).to(dev)
    tb = torch.from_numpy(t.back).to(dev)
    conj = device.tables_conjugate(tf, tb)
    sym = autotune.tables_symmetric(tf) and autotune.tables_symmetric(tb)
    F0, B0, offs = synth.device_fields(p, p.shots, device=dev)
    F0 = F0.view(ns, nf, nx)
    B0 = B0.view(ns, nf, nx)
    stream = torch.cuda.current_stream()
    if args.calib_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)  # 1 GiB read + 1 GiB written
        torch.cuda.synchronize()
        del a, b

    probe = probe_fma(log) if args.probe_fma else None

    # ---------------------------------------------------------------- batched
    specs = [device.ShotSpec(f_off=s * nf * nx, b_off=s * nf * nx, img_off=0, slow_off=0, ld=nx,
                             img_ld=nx, slow_ld=nx, nx=nx, nz=nz, nf=nf, omega_off=0)
             for s in range(ns)]
    F = F0.clone()
    B = B0.clone()
    image = torch.zeros((nz, nx), dtype=torch.float32, device=dev)
    forced = any([args.shots_per_group, args.omegas_per_item, args.block, args.x_per_thread])
    if not args.tune_c2 or forced:
        tuning = dict(tables_conjugate=int(conj), tables_symmetric=int(sym),
                      shots_per_group=args.shots_per_group, omegas_per_item=args.omegas_per_item,
                      block=args.block, x_per_thread=args.x_per_thread)
        tune_s = 0.0
    else:
        t_tune = time.perf_counter()
        tuning = autotune.tune(specs, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tf, tb, F, B,
                               image, conj, sym, reps=2, stream=stream, verbose=args.tune_verbose)
        tune_s = time.perf_counter() - t_tune
    log("tuning", tuning, f"({tune_s:.1f} s)")
    plan = device.Plan(specs, t.M, tuning)
    info = plan.info()
    log("plan", info)

    def fused_step():
        image.zero_()
        plan.launch(slow, omega, t.kappa0, t.dkappa, t.nk, tf, tb, F, B, image, write_back=False,
                    stream=stream)

    for _ in range(args.warmup):
        fused_step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        image.zero_()
        e0.record(stream)
        plan.launch(slow, omega, t.kappa0, t.dkappa, t.nk, tf, tb, F, B, image, write_back=False,
                    stream=stream)
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [e0.elapsed_time(e1) for e0, e1 in evs]
    k_avg = sum(kernel_ms) / len(kernel_ms)
    ms_per_step = wall * 1e3 / args.steps
    value = points / (ms_per_step / 1e3)
    achieved_tflops = flops_per_point(L) * points / (k_avg / 1e3) / 1e12
    img_fused = image.clone()
    log(f"fused: {ms_per_step:.2f} ms/step, kernel {k_avg:.2f} ms, {value:.3e} points/s, "
        f"{achieved_tflops:.1f} TFLOP/s")

    # the same launches with exact image sums (bitwise-repeatable images): their time and image
    exact = None
    if args.exact_steps > 0 and conj:
        xplan = device.Plan(specs, t.M, dict(tuning, exact_image=1))
        ximgs = []
        xms = []
        for k in range(args.exact_steps + 1):
            image.zero_()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            xplan.launch(slow, omega, t.kappa0, t.dkappa, t.nk, tf, tb, F, B, image, write_back=False,
                         stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            if k > 0:
                xms.append(e0.elapsed_time(e1))
            ximgs.append(image.clone())
        xplan.close()
        x_avg = sum(xms) / len(xms)
        rep = all(torch.equal(ximgs[0].view(torch.int32), x.view(torch.int32)) for x in ximgs[1:])
        xrel = float((ximgs[0].double() - img_fused.double()).norm() / img_fused.double().norm())
        exact = {"what": "exact_image = 1: task sums added as 64-bit integers on the 2^-149 grid, "
                         "rounded once (fill, fused kernel, finish kernel per launch)",
                 "ms_per_launch": x_avg, "vs_float_atomics": x_avg / k_avg,
                 "launches_bitwise_equal": rep, "launches": len(ximgs),
                 "image_rel_l2_vs_float_atomics": xrel}
        log(f"exact image sums: {x_avg:.2f} ms per launch ({x_avg / k_avg:.3f} x), "
            f"{len(ximgs)} launches bitwise equal: {rep}, rel L2 vs float atomics {xrel:.2e}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (complex64)",
        "data": "synthetic: seeded 512x1024 gradient+anomaly model, Ricker source, seeded receiver spectra",
        "config": {"workload": "c2_512x1024_nf256_16shots_L31_nk2048", "nz": nz, "nx": nx,
                   "nf": nf, "shots": ns, "taps": L, "nk": t.nk, "points_per_step": points,
                   "path": "batched fused compact-table kernel (one launch, all shots)",
                   "plan": info, "tuning": tuning, "tables_conjugate": conj,
                   "tables_symmetric": sym},
        "roofline": {"bound": "mfma", "pipe": "fp32 VALU (no MFMA: per-point gather-convolution)",
                     "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP32_PEAK_TFLOPS,
                     "traffic": read_profile_traffic("fused_kernel_c2_512x1024_nf256_16shots.txt"),
                     "kernel": "fused_kernel (L=31)", "kernel_ms": k_avg,
                     "flops_per_point": flops_per_point(L),
                     "onchip": onchip_coefficients(t.M, L, conj, sym, info["shots_per_group"],
                                                   points, k_avg)},
    }

    if probe is not None:
        out["probe_fma"] = probe
    if exact is not None:
        out["exact_image"] = exact

    # ---------------------------------------------------------------- plain
    if not args.no_plain:
        log("expanding tables (2 x %.1f GB) ..." % (nz * nf * nx * L * 8 / 1e9))
        tapF = torch.empty(nz * nf * L * nx, dtype=torch.complex64, device=dev)
        tapB = torch.empty_like(tapF)
        device.expand_table_device(nz, nf, nx, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tf, tapF,
                                   layout=LAYOUT_TAPMAJOR, stream=stream)
        device.expand_table_device(nz, nf, nx, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tb, tapB,
                                   layout=LAYOUT_TAPMAJOR, stream=stream)
        nsp = min(ns, args.plain_shots) if args.plain_shots > 0 else ns
        P0, Q0 = F0[:nsp].contiguous(), B0[:nsp].contiguous()
        Fp = P0.clone()
        Bp = Q0.clone()
        pimg = torch.zeros((nsp, nz, nx), dtype=torch.float32, device=dev)

        def plain_step():
            Fp.copy_(P0)
            Bp.copy_(Q0)
            device.plain_device(nsp, nz, nz, nf, nf, nx, t.M, tapF, Fp, tapB, Bp, pimg,
                                layout=LAYOUT_TAPMAJOR, shots_per_launch=1, stream=stream)

        plain_step()  # warm-up (and scratch allocation)
        torch.cuda.synchronize()
        pe = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.plain_steps)]
        for e0, e1 in pe:
            Fp.copy_(P0)
            Bp.copy_(Q0)
            e0.record(stream)
            device.plain_device(nsp, nz, nz, nf, nf, nx, t.M, tapF, Fp, tapB, Bp, pimg,
                                layout=LAYOUT_TAPMAJOR, shots_per_launch=1, stream=stream)
            e1.record(stream)
        torch.cuda.synchronize()
        p_ms = sum(e0.elapsed_time(e1) for e0, e1 in pe) / len(pe)
        p_points = nsp * nz * nf * nx
        p_value = p_points / (p_ms / 1e3)
        per_launch_ms = p_ms / (nsp * nz)
        gbs = plain_bytes_per_point(L) * (nf * nx) / (per_launch_ms / 1e3) / 1e9
        if nsp == ns:
            stacked = pimg.sum(0)
            err = float(torch.linalg.norm(stacked - img_fused) / torch.linalg.norm(stacked))
        else:
            err = None
        log(f"plain: {p_ms:.1f} ms/step, {p_value:.3e} points/s, {gbs:.0f} GB/s per step launch, "
            f"fused-vs-plain image rel L2 {err}")
        out["plain"] = {"value": p_value, "unit": "points/s", "ms_per_step": p_ms,
                        "path": "drop-in plain path, expanded tap-major tables in HBM, one launch "
                                "sequence per shot", "launches_per_step": nsp * (nz + 3), "shots": nsp}
        out["speedup_vs_plain"] = value / p_value
        out["image_rel_l2_batched_vs_plain"] = err
        out["roofline_plain"] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                                 "traffic": read_profile_traffic("plain_step_kernel_c2_512x1024_nf256_16shots.txt"),
                                 "kernel": "plain_step_kernel<1,31>",
                                 "bytes_per_point": plain_bytes_per_point(L),
                                 "note": "per-launch time = step time / (shots*depths); includes "
                                         "the small copy/reduce launches"}
        del tapF, tapB
        torch.cuda.empty_cache()

    if args.c3_steps > 0:
        out["c3_batched"] = bench_c3(args, log)
        torch.cuda.empty_cache()
    if args.c5_shots > 0:
        out["c5_survey"] = bench_c5(args, log)
        torch.cuda.empty_cache()
    if args.serve_jobs > 0:
        out["serving"] = bench_serving(args, log)
        torch.cuda.empty_cache()

    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
