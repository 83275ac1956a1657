"""Autotuner for the fused kernel's launch shape.

For a batch of shots it times candidate tunings - shots per group (coefficient
reuse), x per thread (register blocking), threads per workgroup, frequencies
per task (LDS rows per workgroup), the tap-pairing variant when the tables
allow it, then depth steps per launch for the best of them - with HIP events
on the caller's stream, and returns the fastest.
Results are cached in memory and, optionally, in a JSON file keyed by the
workload signature (apertures, depths, frequency counts, taps, device name).

Candidates that the planner rejects (LDS, thread limits) are skipped; every
candidate computes the same operator, so the choice only affects speed (the
tap-paired variant changes rounding order; the tests bound it).
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Dict, Iterable, List, Optional, Sequence

import torch

from . import device
from ._lib import WlsqError

_MEM: Dict[str, dict] = {}


def signature(shots: Sequence[device.ShotSpec], M: int, conj: bool, sym: bool) -> str:
    h = hashlib.sha1()
    for s in shots:
        h.update(f"{s.nx},{s.nz},{s.nf},{s.omega_off},{s.slow_off},{int(s.img_off)};".encode())
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{name}|M{M}|c{int(conj)}|s{int(sym)}|n{len(shots)}|{h.hexdigest()[:16]}"


def candidates(conj: bool, sym: bool, max_nx: int) -> List[dict]:
    out = []
    syms = [1, 0] if sym else [0]
    for sy in syms:
        for S in (1, 2, 4):
            for R in (2, 4, 8):
                if R == 8 and S != 1:
                    continue
                strips = (max_nx + R - 1) // R
                wide = R == 4 and S == 1  # 1024-thread workgroups, up to 2 strips per thread
                if strips > (1024 if R == 8 else 2048 if wide else 512):
                    continue
                if R == 2 and S == 1:
                    continue
                blocks = ([256, 512, 1024] if wide else [256, 512]) if R == 4 else [0]
                for block in blocks:
                    if block and block * (2 if wide and block > 512 else 1) < strips:
                        continue
                    for W in (0, 1):
                        for tl in ((1, 2) if S <= 2 and R <= 4 else (1,)):
                            out.append(dict(shots_per_group=S, x_per_thread=R, block=block,
                                            omegas_per_item=W, tables_conjugate=int(conj),
                                            tables_symmetric=sy, tap_loads=tl))
                    # apertures that one strip per thread covers: also two
                    # strips per thread (twice the rows per task, one copy of
                    # them) at LDS budgets near the device's
                    if wide and block == 1024 and strips <= 1024:
                        for budget in (120 << 10, 160 << 10):
                            out.append(dict(shots_per_group=S, x_per_thread=R, block=block,
                                            omegas_per_item=0, tables_conjugate=int(conj),
                                            tables_symmetric=sy, tap_loads=2,
                                            strips_per_thread=2, lds_budget=budget))
    return out


def longest_run(shots: Sequence[device.ShotSpec]) -> int:
    """Most consecutive shots that can share a group (same window, depth
    count, frequency list and strides, as the planner requires): candidates
    with more shots per group would run groups of fewer shots."""
    key = lambda s: (s.nx, s.nz, s.nf, s.omega_off, s.slow_off, s.slow_ld, s.ld, s.img_ld,  # noqa: E731
                     s.first_step)
    best = run = 1
    for a, b in zip(shots, shots[1:]):
        run = run + 1 if key(a) == key(b) else 1
        best = max(best, run)
    return best


def _time_plan(plan: device.Plan, args, reps: int, stream) -> float:
    slowness, omega, kappa0, dkappa, nk, tf, tb, F, B, image = args
    plan.launch(slowness, omega, kappa0, dkappa, nk, tf, tb, F, B, image, stream=stream)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        plan.launch(slowness, omega, kappa0, dkappa, nk, tf, tb, F, B, image, stream=stream)
        e1.record(stream)
    torch.cuda.synchronize()
    return min(e0.elapsed_time(e1) for e0, e1 in ev)


def tune(shots: Sequence[device.ShotSpec], M: int, slowness, omega, kappa0, dkappa, nk, tf, tb,
         F, B, image, conj: bool, sym: bool, reps: int = 3, cache_file: Optional[str] = None,
         cands: Optional[Iterable[dict]] = None, stream=None, verbose=False) -> dict:
    """Best tuning dict for this batch (``image`` is used as scratch and
    left holding garbage; F/B are not written back)."""
    cands_given = cands
    key = signature(shots, M, conj, sym)
    if key in _MEM:
        return dict(_MEM[key])
    if cache_file and os.path.exists(cache_file):
        try:
            disk = json.load(open(cache_file))
            if key in disk:
                _MEM[key] = disk[key]["tuning"]
                return dict(_MEM[key])
        except (OSError, ValueError):
            pass
    stream = stream or torch.cuda.current_stream()
    max_nx = max(s.nx for s in shots)
    results = []
    args = (slowness, omega, kappa0, dkappa, nk, tf, tb, F, B, image)
    cands = list(cands) if cands is not None else [c for c in candidates(conj, sym, max_nx)
                                                   if c["shots_per_group"] <= longest_run(shots)]
    for c in cands:
        try:
            plan = device.Plan(shots, M, c)
        except WlsqError:
            continue
        try:
            ms = _time_plan(plan, args, reps, stream)
        except WlsqError:
            plan.close()
            continue
        info = plan.info()
        plan.close()
        results.append((ms, c, info))
        if verbose:
            print(f"  tune {c} -> {ms:.3f} ms {info}", flush=True)
    if not results:
        raise RuntimeError("no launch shape fits this batch")
    results.sort(key=lambda r: r[0])
    best = results[0][1]
    # depth steps per launch: all of them, or the range split in 2 or 4
    # launches (each restarts from the written-back fields)
    max_nz = max(s.nz for s in shots)
    if cands_given is None and max_nz >= 64:
        for parts in (2, 4):
            c = dict(best, depth_steps=(max_nz + parts - 1) // parts)
            try:
                plan = device.Plan(shots, M, c)
            except WlsqError:
                continue
            try:
                ms = _time_plan(plan, args, reps, stream)
            except WlsqError:
                plan.close()
                continue
            info = plan.info()
            plan.close()
            if verbose:
                print(f"  tune {c} -> {ms:.3f} ms {info}", flush=True)
            if ms < results[0][0]:
                results.insert(0, (ms, c, info))
        best = results[0][1]
    _MEM[key] = best
    if cache_file:
        try:
            disk = json.load(open(cache_file)) if os.path.exists(cache_file) else {}
        except (OSError, ValueError):
            disk = {}
        disk[key] = {"tuning": best, "ms": results[0][0]}
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return dict(best)


def tables_symmetric(table: torch.Tensor) -> bool:
    """True when table[q][k] == table[q][L-1-k] bit for bit for every row."""
    v = torch.view_as_real(table).contiguous().view(torch.int32)
    return bool(torch.equal(v, torch.flip(v, dims=[1])))
