"""Device-pointer entries of libwlsq_hip.so for data resident in HBM.

PyTorch is only the allocator and stream provider here: tensors are passed as
raw pointers, and every computation runs in the HIP library. Complex data is
``torch.complex64`` (8-byte interleaved, the layout of ``wlsq_c32``).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int, c_longlong, c_void_p
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from ._lib import LAYOUT_NATURAL, LAYOUT_TAPMAJOR, Shot, Tuning, check, lib


def _ptr(t: torch.Tensor, dtype, name: str) -> int:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must be on the GPU")
    if not t.is_contiguous():
        raise TypeError(f"{name} must be contiguous")
    return t.data_ptr()


def _need(t: torch.Tensor, n: int, name: str) -> None:
    """The C ABI takes raw pointers and no buffer sizes: a launch whose
    offsets reach past a tensor would read or write outside it on the device,
    so every extent is checked here before the call."""
    if n > t.numel():
        raise ValueError(f"{name} has {t.numel()} elements, the call reads or writes {n}")


def _extent(off: int, rows: int, ld: int, nx: int) -> int:
    """Elements from the start of a buffer to the end of `rows` rows of `nx`
    values at stride `ld` from element `off` (0 when there are none)."""
    if rows <= 0 or nx <= 0:
        return 0
    return off + (rows - 1) * ld + nx


def _check_uniform(ns, nextrap, nz, nt, nf, nx, forw, back, image):
    """Fields (ns, nt, nx), rows j < nf used; image (ns, nz, nx), rows l < nextrap."""
    if ns <= 0 or nx <= 0:
        return
    _need(forw, _extent((ns - 1) * nt * nx, nf, nx, nx), "forw")
    _need(back, _extent((ns - 1) * nt * nx, nf, nx, nx), "back")
    _need(image, _extent((ns - 1) * nz * nx, nextrap, nx, nx), "image")


def _stream(stream) -> int:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream
    return int(stream)


def _tuning(t: Optional[dict]) -> Tuning:
    tun = Tuning()
    for k, v in (t or {}).items():
        setattr(tun, k, int(v))
    return tun


def plain_device(ns, nextrap, nz, nt, nf, nx, M, op_forw, forw, op_back, back, image,
                 layout=LAYOUT_TAPMAJOR, shots_per_launch=1, stream=None):
    """Plain path on device tensors: one launch per depth step streaming the
    expanded (nextrap, nf, [L, nx] | [nx, L]) tables; fields (ns, nt, nx) in
    place; image (ns, nz, nx) rows < nextrap written."""
    c = torch.complex64
    args = (_ptr(op_forw, c, "op_forw"), _ptr(forw, c, "forw"), _ptr(op_back, c, "op_back"),
            _ptr(back, c, "back"), _ptr(image, torch.float32, "image"))
    if min(ns, nextrap, nf, nx) > 0 and M >= 0:
        _check_uniform(ns, nextrap, nz, nt, nf, nx, forw, back, image)
        _need(op_forw, nextrap * nf * (2 * M + 1) * nx, "op_forw")
        _need(op_back, nextrap * nf * (2 * M + 1) * nx, "op_back")
    check(lib().wlsq_plain_device(ns, nextrap, nz, nt, nf, nx, M, _ptr(op_forw, c, "op_forw"),
                                  _ptr(forw, c, "forw"), _ptr(op_back, c, "op_back"),
                                  _ptr(back, c, "back"), _ptr(image, torch.float32, "image"),
                                  layout, shots_per_launch, _stream(stream)))


def ktable_device(ns, nextrap, nz, nt, nf, nx, M, slowness, omega, kappa0, dkappa, nk,
                  table_forw, forw, table_back, back, image, write_back=True, tuning=None,
                  stream=None, slow_ld=None, no_alloc=False):
    """Fused compact-table path on device tensors. image rows < nextrap of
    every shot are cleared and then accumulated. ``no_alloc`` calls the
    allocation-free entry (library-owned 16 MiB scratch per GPU; calls that
    overlap on one GPU race)."""
    c = torch.complex64
    tun = _tuning(tuning)
    ld = int(slow_ld if slow_ld is not None else nx)
    for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                        (table_forw, c, "table_forw"), (forw, c, "forw"), (table_back, c, "table_back"),
                        (back, c, "back"), (image, torch.float32, "image")):
        _ptr(t, dt, name)
    if min(ns, nextrap, nx) > 0 and M >= 0:
        _need(image, _extent((ns - 1) * nz * nx, nextrap, nx, nx), "image")
        if nf > 0:
            _check_uniform(ns, nextrap, nz, nt, nf, nx, forw, back, image)
            _need(slowness, _extent(0, nextrap, ld, nx), "slowness")
            _need(omega, nf, "omega")
            _need(table_forw, int(nk) * (2 * M + 1), "table_forw")
            _need(table_back, int(nk) * (2 * M + 1), "table_back")
    fn = lib().wlsq_ktable_device_noalloc if no_alloc else lib().wlsq_ktable_device
    check(fn(ns, nextrap, nz, nt, nf, nx, M,
                                   _ptr(slowness, torch.float32, "slowness"),
                                   ld, _ptr(omega, torch.float32, "omega"), float(kappa0), float(dkappa),
                                   int(nk), _ptr(table_forw, c, "table_forw"), _ptr(forw, c, "forw"),
                                   _ptr(table_back, c, "table_back"), _ptr(back, c, "back"),
                                   _ptr(image, torch.float32, "image"), int(bool(write_back)),
                                   byref(tun), _stream(stream)))


def expand_table_device(nextrap, nf, nx, M, slowness, omega, kappa0, dkappa, nk, table, out,
                        layout=LAYOUT_TAPMAJOR, stream=None, slow_ld=None):
    c = torch.complex64
    ld = int(slow_ld if slow_ld is not None else nx)
    for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                        (table, c, "table"), (out, c, "out")):
        _ptr(t, dt, name)
    if min(nextrap, nf, nx) > 0 and M >= 0:
        _need(slowness, _extent(0, nextrap, ld, nx), "slowness")
        _need(omega, nf, "omega")
        _need(table, int(nk) * (2 * M + 1), "table")
        _need(out, nextrap * nf * nx * (2 * M + 1), "out")
    check(lib().wlsq_expand_table_device(nextrap, nf, nx, M, _ptr(slowness, torch.float32, "slowness"),
                                         int(slow_ld if slow_ld is not None else nx),
                                         _ptr(omega, torch.float32, "omega"), float(kappa0),
                                         float(dkappa), int(nk), _ptr(table, c, "table"),
                                         _ptr(out, c, "out"), int(layout), _stream(stream)))


def trace_spectra_device(traces: torch.Tensor, out: Optional[torch.Tensor] = None,
                         stream=None) -> torch.Tensor:
    """(nt, nx) complex64 spectra along time of (nt, nx) float32 traces on the
    GPU: ``scipy.fftpack.fft(traces, axis=0)`` of ``main.py:119-123``. ``out``
    may be a row-strided view (ld >= nx) into a larger buffer."""
    if traces.dim() != 2:
        raise ValueError("traces must be (nt, nx)")
    nt, nx = traces.shape
    tp = _ptr(traces, torch.float32, "traces")
    if out is None:
        out = torch.empty((nt, nx), dtype=torch.complex64, device=traces.device)
    if out.dim() != 2 or out.shape[0] < nt or out.shape[1] < nx or out.stride(1) != 1:
        raise ValueError("out must be (>= nt, >= nx) with unit column stride")
    if out.dtype != torch.complex64 or out.device.type != "cuda":
        raise TypeError("out must be complex64 on the GPU")
    check(lib().wlsq_trace_spectra_device(nt, nx, tp, nx, out.data_ptr(), int(out.stride(0)),
                                          _stream(stream)))
    return out


def exact_sum(terms: torch.Tensor, stream=None) -> torch.Tensor:
    """(m,) float32: the float nearest the exact sum over rows of an (n, m)
    float32 device tensor, through the fused kernel's exact image
    accumulation (wlsq_probe_exact_sum; a measurement and test aid)."""
    if terms.dim() != 2:
        raise ValueError("terms must be (n, m)")
    n, m = terms.shape
    out = torch.empty(m, dtype=torch.float32, device=terms.device)
    check(lib().wlsq_probe_exact_sum(_ptr(terms, torch.float32, "terms"), int(n), int(m),
                                     out.data_ptr(), _stream(stream)))
    return out


def tables_conjugate(table_forw: torch.Tensor, table_back: torch.Tensor) -> bool:
    """True when table_back == conj(table_forw) bit for bit (then the fused
    kernel reads one table)."""
    a = torch.view_as_real(table_forw).view(torch.int32)
    b = torch.view_as_real(table_back).view(torch.int32)
    neg = torch.view_as_real(table_forw.conj().resolve_conj()).view(torch.int32)
    return bool(torch.equal(a[..., 0], b[..., 0]) and torch.equal(neg[..., 1], b[..., 1]))


@dataclass
class ShotSpec:
    """One shot of a batch (element offsets into the pools)."""
    f_off: int
    b_off: int
    img_off: int
    slow_off: int
    ld: int
    img_ld: int
    slow_ld: int
    nx: int
    nz: int
    nf: int
    omega_off: int
    first_step: int = 0  # rounds of a continuous batch: the next depth step


def check_shot_extents(shots: Sequence[ShotSpec], M: int, nk: int, slowness, omega, table_forw,
                       table_back, forw, back, image) -> None:
    """Every shot's rows lie inside the tensors handed to the batched launch
    (raises ValueError before any device call)."""
    L = 2 * M + 1
    for i, s in enumerate(shots):
        if s.nx <= 0 or s.nz <= 0 or s.nf <= 0:
            continue
        for t, n, name in ((forw, _extent(s.f_off, s.nf, s.ld, s.nx), "forw"),
                           (back, _extent(s.b_off, s.nf, s.ld, s.nx), "back"),
                           (image, _extent(s.img_off, s.nz, s.img_ld, s.nx), "image"),
                           (slowness, _extent(s.slow_off, s.nz, s.slow_ld, s.nx), "slowness"),
                           (omega, s.omega_off + s.nf, "omega")):
            if n > t.numel():
                raise ValueError(f"shot {i}: {name} has {t.numel()} elements, the shot reaches {n}")
    if any(s.nx > 0 and s.nz > 0 and s.nf > 0 for s in shots):
        _need(table_forw, int(nk) * L, "table_forw")
        _need(table_back, int(nk) * L, "table_back")


def _shots_array(shots: Sequence[ShotSpec]):
    arr = (Shot * max(1, len(shots)))()
    for i, s in enumerate(shots):
        arr[i] = Shot(s.f_off, s.b_off, s.img_off, s.slow_off, s.ld, s.img_ld, s.slow_ld, s.nx,
                      s.nz, s.nf, s.omega_off, s.first_step)
    return arr


def batch_workspace_bytes(shots: Sequence[ShotSpec], M: int, nk: int,
                          tuning: Optional[dict] = None) -> int:
    """Device bytes :func:`batch_launch_ws` needs for this batch."""
    tun = _tuning(tuning)
    n = int(lib().wlsq_batch_workspace_bytes(_shots_array(shots), len(shots), M, int(nk),
                                             byref(tun)))
    if n < 0:
        raise ValueError(f"invalid batch: {lib().wlsq_last_error().decode(errors='replace')}")
    return n


def batch_launch_ws(shots: Sequence[ShotSpec], M: int, slowness, omega, kappa0, dkappa, nk,
                    table_forw, table_back, forw, back, image, workspace: torch.Tensor,
                    write_back=False, tuning: Optional[dict] = None, stream=None) -> None:
    """Queue one fused launch of `shots` with its descriptors and table copies
    in `workspace` (a uint8 device tensor of at least
    :func:`batch_workspace_bytes`): no device allocation, no wait. Adds into
    `image`."""
    c = torch.complex64
    for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                        (table_forw, c, "table_forw"), (table_back, c, "table_back"),
                        (forw, c, "forw"), (back, c, "back"), (image, torch.float32, "image"),
                        (workspace, torch.uint8, "workspace")):
        _ptr(t, dt, name)
    check_shot_extents(shots, M, nk, slowness, omega, table_forw, table_back, forw, back, image)
    tun = _tuning(tuning)
    check(lib().wlsq_batch_device_ws(_shots_array(shots), len(shots), M, slowness.data_ptr(),
                                     omega.data_ptr(), float(kappa0), float(dkappa), int(nk),
                                     table_forw.data_ptr(), table_back.data_ptr(), forw.data_ptr(),
                                     back.data_ptr(), image.data_ptr(), int(bool(write_back)),
                                     byref(tun), workspace.data_ptr(), int(workspace.numel()),
                                     _stream(stream)))


def batch_round_ws(shots: Sequence[ShotSpec], steps: int, M: int, slowness, omega, kappa0, dkappa,
                   nk, table_forw, table_back, forw, back, image, workspace: torch.Tensor,
                   tuning: Optional[dict] = None, stream=None, check_extents: bool = True) -> None:
    """Queue one round of a continuous batch: every shot runs `steps` depth
    steps from its ``first_step`` (cut at its nz) on the wavefields its rows
    hold, writes them back and adds its image rows. Descriptors and table
    copies go into `workspace` (at least :func:`batch_workspace_bytes` of the
    same shots); no device allocation, no wait. ``check_extents=False`` skips
    the per-shot extent checks for callers that placed every shot inside the
    tensors themselves (the serving loop's pools)."""
    c = torch.complex64
    for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                        (table_forw, c, "table_forw"), (table_back, c, "table_back"),
                        (forw, c, "forw"), (back, c, "back"), (image, torch.float32, "image"),
                        (workspace, torch.uint8, "workspace")):
        _ptr(t, dt, name)
    if check_extents:
        check_shot_extents(shots, M, nk, slowness, omega, table_forw, table_back, forw, back, image)
    tun = _tuning(tuning)
    check(lib().wlsq_batch_round_ws(_shots_array(shots), len(shots), int(steps), M,
                                    slowness.data_ptr(), omega.data_ptr(), float(kappa0),
                                    float(dkappa), int(nk), table_forw.data_ptr(),
                                    table_back.data_ptr(), forw.data_ptr(), back.data_ptr(),
                                    image.data_ptr(), byref(tun), workspace.data_ptr(),
                                    int(workspace.numel()), _stream(stream)))


class Plan:
    """Work list of a batch of uneven shots for the single fused launch."""

    def __init__(self, shots: Sequence[ShotSpec], M: int, tuning: Optional[dict] = None):
        self._h = c_void_p()
        self.shots = list(shots)
        self.M = M
        arr = _shots_array(self.shots)
        self._tun = _tuning(tuning)
        check(lib().wlsq_plan_create(byref(self._h), arr, len(self.shots), M, byref(self._tun)))

    def info(self) -> dict:
        p = c_longlong()
        n, b, lds, r, s = c_int(), c_int(), c_int(), c_int(), c_int()
        check(lib().wlsq_plan_info(self._h, byref(p), byref(n), byref(b), byref(lds), byref(r), byref(s)))
        return dict(points=p.value, items=n.value, block=b.value, lds_bytes=lds.value,
                    x_per_thread=r.value, shots_per_group=s.value)

    def launch(self, slowness, omega, kappa0, dkappa, nk, table_forw, table_back, forw, back, image,
               write_back=False, stream=None):
        c = torch.complex64
        for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                            (table_forw, c, "table_forw"), (table_back, c, "table_back"),
                            (forw, c, "forw"), (back, c, "back"), (image, torch.float32, "image")):
            _ptr(t, dt, name)
        check_shot_extents(self.shots, self.M, nk, slowness, omega, table_forw, table_back, forw,
                           back, image)
        check(lib().wlsq_plan_launch(self._h, _ptr(slowness, torch.float32, "slowness"),
                                     _ptr(omega, torch.float32, "omega"), float(kappa0), float(dkappa),
                                     int(nk), _ptr(table_forw, c, "table_forw"),
                                     _ptr(table_back, c, "table_back"), _ptr(forw, c, "forw"),
                                     _ptr(back, c, "back"), _ptr(image, torch.float32, "image"),
                                     int(bool(write_back)), _stream(stream)))

    def graph(self, slowness, omega, kappa0, dkappa, nk, table_forw, table_back, forw, back, image,
              write_back=False, zero_count=0) -> "Graph":
        return Graph(self, slowness, omega, kappa0, dkappa, nk, table_forw, table_back, forw, back,
                     image, write_back, zero_count)

    def close(self):
        if self._h:
            lib().wlsq_plan_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """hipGraph of [image clear] + one plan launch; pointers are fixed."""

    def __init__(self, plan: Plan, slowness, omega, kappa0, dkappa, nk, table_forw, table_back,
                 forw, back, image, write_back=False, zero_count=0):
        c = torch.complex64
        self._h = c_void_p()
        for t, dt, name in ((slowness, torch.float32, "slowness"), (omega, torch.float32, "omega"),
                            (table_forw, c, "table_forw"), (table_back, c, "table_back"),
                            (forw, c, "forw"), (back, c, "back"), (image, torch.float32, "image")):
            _ptr(t, dt, name)
        check_shot_extents(plan.shots, plan.M, nk, slowness, omega, table_forw, table_back, forw,
                           back, image)
        if zero_count > image.numel():
            raise ValueError(f"zero_count {zero_count} above the image's {image.numel()} elements")
        self._keep = (plan, slowness, omega, table_forw, table_back, forw, back, image)
        check(lib().wlsq_graph_create(byref(self._h), plan._h,
                                      _ptr(slowness, torch.float32, "slowness"),
                                      _ptr(omega, torch.float32, "omega"), float(kappa0),
                                      float(dkappa), int(nk), _ptr(table_forw, c, "table_forw"),
                                      _ptr(table_back, c, "table_back"), _ptr(forw, c, "forw"),
                                      _ptr(back, c, "back"), _ptr(image, torch.float32, "image"),
                                      int(bool(write_back)), int(zero_count)))

    def launch(self, stream=None):
        check(lib().wlsq_graph_launch(self._h, _stream(stream)))

    def close(self):
        if self._h:
            lib().wlsq_graph_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_points(shots: Sequence[ShotSpec]) -> int:
    return int(lib().wlsq_batch_points(_shots_array(shots), len(shots)))
