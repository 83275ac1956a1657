"""GPU parity at the benchmark configurations' real sizes: whole shots of the
512 x 1024 x 256-frequency workload, of the uneven 800 x 2048 batch and of
the 2000 x 8192 survey model (every depth step), through the batched device
path, against the CPU oracle on the same device-generated inputs. Images are
compared at relative L2 <= 1e-5, and so are the final wavefields where the
depth count is the workload's (512, 800); the survey shot's 2000-step fields
are judged against FP64 (see that test). The fused path is also compared with
the plain drop-in path on the device's expanded tables."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    oracle.build()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def crel(a, b):
    return rel(np.concatenate([np.real(a).ravel(), np.imag(a).ravel()]),
               np.concatenate([np.real(b).ravel(), np.imag(b).ravel()]))


def _run_and_check(p, shots, tuning=None, field_rtol=RTOL):
    """Run `shots` as one batch with write-back; compare the stacked image and
    every shot's final fields with the oracle on the same inputs."""
    import torch
    from imaging2d_wlsq_amd import batch
    sv = batch.Survey(p)
    b = sv.batch(shots, tuning=tuning, write_back=True)
    print("batch", b.tuning, b.plan.info(), flush=True)
    F0, B0 = b.F.cpu().numpy(), b.B.cpu().numpy()  # inputs, before the run overwrites them
    img = b.run().clone()
    torch.cuda.synchronize()
    Fd, Bd = b.F.cpu().numpy(), b.B.cpu().numpy()
    m, t = p.model, p.tables
    ref = np.zeros((m.nz, m.nx))
    outs = []  # per shot: oracle F, B, device F, B (final), inputs F, B
    for spec, s in zip(b.specs, shots):
        n = s.nf * s.nx
        F = F0[spec.f_off:spec.f_off + n].reshape(1, s.nf, s.nx).copy()
        B = B0[spec.b_off:spec.b_off + n].reshape(1, s.nf, s.nx).copy()
        io = np.zeros((1, s.nz, s.nx), np.float32)
        sl = np.ascontiguousarray(m.slowness[:s.nz, s.x0:s.x0 + s.nx])
        oracle.extrapolate_ktable(1, s.nz, s.nz, s.nx, s.nf, s.nf, t.M, sl, s.omega, t.kappa0,
                                  t.dkappa, t.nk, t.forw, t.back, F, B, io)
        ref[:s.nz, s.x0:s.x0 + s.nx] += io[0]
        ef = crel(Fd[spec.f_off:spec.f_off + n].reshape(s.nf, s.nx), F[0])
        eb = crel(Bd[spec.b_off:spec.b_off + n].reshape(s.nf, s.nx), B[0])
        print(f"  shot x0={s.x0} nx={s.nx} nz={s.nz} nf={s.nf}: final fields rel L2 {ef:.2e} {eb:.2e}",
              flush=True)
        outs.append((F[0], B[0], Fd[spec.f_off:spec.f_off + n].reshape(s.nf, s.nx),
                     Bd[spec.b_off:spec.b_off + n].reshape(s.nf, s.nx),
                     F0[spec.f_off:spec.f_off + n].reshape(s.nf, s.nx),
                     B0[spec.b_off:spec.b_off + n].reshape(s.nf, s.nx)))
        if field_rtol is not None:
            assert ef <= field_rtol and eb <= field_rtol
    err = rel(img.cpu().numpy(), ref)
    print(f"{p.name}: {len(shots)} shots, image rel L2 vs oracle {err:.2e}", flush=True)
    assert err <= RTOL
    b.close()
    return outs


def test_two_shots_of_the_512x1024_workload():
    from imaging2d_wlsq_amd import synth
    p = synth.config_c2()
    _run_and_check(p, [p.shots[3], p.shots[12]])


def test_widest_shots_of_the_uneven_batch():
    from imaging2d_wlsq_amd import synth
    p = synth.config_c3(nshots=128)
    shots = sorted(p.shots, key=lambda s: -s.nx * s.nf)[:2]
    _run_and_check(p, shots)


@pytest.mark.parametrize("mode", ["reference_rounding", "paired_taps", "taps_in_order"])
def test_survey_shot_every_depth_step(mode):
    """One full-aperture shot of the 2000 x 8192 model, all 2000 depth steps,
    two frequencies (11.8 and 60.6 Hz), the default survey shape (two strips
    per thread, 1024 threads), against the FP32 CPU oracle and an FP64 re-run.

    * reference_rounding (rounding=1): every complex product rounded as the
      CPU path rounds it (cpp_src/extrapolation_cpu.cpp:52-53,74-79,
      cpp_src/mkl_linalg.cpp:10), so the final fields are bit for bit the
      oracle's and the image is within 1e-5 of it (the image sums the same
      two terms per point);
    * paired_taps / taps_in_order (rounding=0, the fast path with and without
      tap pairing): fused multiply-adds round differently from the oracle's
      separate products, and over 2000 FP32 steps that difference grows to
      about 6e-5 (image) - above 1e-5 but closer to FP64 than the FP32 oracle
      itself. Asserted: at most 1e-4 from the oracle, and no farther from
      FP64 than twice the oracle's own distance."""
    import torch
    from imaging2d_wlsq_amd import batch, synth
    from oracle import fp64
    p = synth.config_c5(nshots=4)
    s0 = p.shots[2]
    freqs = s0.freqs[[128, 768]]
    shot = synth.ShotGeom(0, p.model.nx, p.model.nz, s0.xs, freqs, s0.seed)
    m, t = p.model, p.tables
    sv = batch.Survey(p)
    tuning = dict(tables_conjugate=1, tables_symmetric=int(mode == "paired_taps"),
                  rounding=int(mode == "reference_rounding"))
    b = sv.batch([shot], tuning=tuning, write_back=True)
    print("batch", b.tuning, b.plan.info(), flush=True)
    Fi, Bi = b.F.cpu().numpy().reshape(2, m.nx), b.B.cpu().numpy().reshape(2, m.nx)
    img = b.run().clone().cpu().numpy()
    Fd, Bd = b.F.cpu().numpy().reshape(2, m.nx), b.B.cpu().numpy().reshape(2, m.nx)
    b.close()
    Fo, Bo = Fi[None].copy(), Bi[None].copy()
    io = np.zeros((1, m.nz, m.nx), np.float32)
    oracle.extrapolate_ktable(1, m.nz, m.nz, m.nx, 2, 2, t.M, m.slowness, shot.omega, t.kappa0,
                              t.dkappa, t.nk, t.forw, t.back, Fo, Bo, io)
    kt = dict(slowness=m.slowness, omega=shot.omega, kappa0=t.kappa0, dkappa=t.dkappa, nk=t.nk,
              forw=t.forw, back=t.back)
    i64, F64, B64 = fp64.extrapolate(m.nz, m.nz, m.nx, 2, t.M, wave_fs=Fi, shot_fs=Bi, ktable=kt)
    both = lambda F, B: np.concatenate([F.ravel(), B.ravel()])
    e_img = (rel(img, io[0]), rel(img, i64), rel(io[0], i64))
    e_fld = (crel(both(Fd, Bd), both(Fo[0], Bo[0])), crel(both(Fd, Bd), both(F64, B64)),
             crel(both(Fo[0], Bo[0]), both(F64, B64)))
    bitwise_f = np.array_equal(Fd.view(np.uint32), Fo[0].view(np.uint32)) and \
        np.array_equal(Bd.view(np.uint32), Bo[0].view(np.uint32))
    bitwise_i = np.array_equal(img.view(np.uint32), io[0].view(np.uint32))
    print(f"  2000 steps, {mode}: image rel L2 device-oracle {e_img[0]:.2e}, "
          f"device-FP64 {e_img[1]:.2e}, oracle-FP64 {e_img[2]:.2e}; final fields "
          f"{e_fld[0]:.2e} {e_fld[1]:.2e} {e_fld[2]:.2e}; fields bitwise {bitwise_f}, "
          f"image bitwise {bitwise_i}", flush=True)
    if mode == "reference_rounding":
        assert bitwise_f, "final fields differ from the oracle's"
        assert e_img[0] <= RTOL
        return
    assert e_img[1] <= 2 * e_img[2] + 1e-6
    assert e_fld[1] <= 2 * e_fld[2] + 1e-6
    assert e_img[0] <= 1e-4 and e_fld[0] <= 1e-3, (
        f"fast path over 2000 steps: image {e_img[0]:.2e} and fields {e_fld[0]:.2e} from the FP32 "
        f"oracle (measured before at about 6e-5; the 1e-5 bound holds with rounding=1)")


def test_fused_equals_plain_path_at_full_size():
    """Two shots of the 512 x 1024 workload through the fused kernel and through
    the drop-in plain kernels on the device-expanded (512, 256, 31, 1024) tables
    (33 GB each): the same operator, so the images agree to rounding."""
    import torch
    from imaging2d_wlsq_amd import device, synth
    from imaging2d_wlsq_amd._lib import LAYOUT_TAPMAJOR
    p = synth.config_c2()
    shots = [p.shots[5], p.shots[10]]
    m, t = p.model, p.tables
    nz, nx, nf, ns = m.nz, m.nx, shots[0].nf, len(shots)
    dev = torch.device("cuda")
    slow = torch.from_numpy(m.slowness).to(dev)
    omega = torch.from_numpy(shots[0].omega).to(dev)
    tf = torch.from_numpy(t.forw).to(dev)
    tb = torch.from_numpy(t.back).to(dev)
    F, B, _ = synth.device_fields(p, shots, device=dev)
    F = F.view(ns, nf, nx)
    B = B.view(ns, nf, nx)
    img_f = torch.zeros((ns, nz, nx), device=dev)
    conj = device.tables_conjugate(tf, tb)
    device.ktable_device(ns, nz, nz, nf, nf, nx, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tf,
                         F.clone(), tb, B.clone(), img_f, tuning=dict(tables_conjugate=int(conj)))
    tapF = torch.empty(nz * nf * t.L * nx, dtype=torch.complex64, device=dev)
    tapB = torch.empty_like(tapF)
    device.expand_table_device(nz, nf, nx, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tf, tapF,
                               layout=LAYOUT_TAPMAJOR)
    device.expand_table_device(nz, nf, nx, t.M, slow, omega, t.kappa0, t.dkappa, t.nk, tb, tapB,
                               layout=LAYOUT_TAPMAJOR)
    img_p = torch.zeros((ns, nz, nx), device=dev)
    Fp, Bp = F.clone(), B.clone()
    device.plain_device(ns, nz, nz, nf, nf, nx, t.M, tapF, Fp, tapB, Bp, img_p, layout=LAYOUT_TAPMAJOR)
    torch.cuda.synchronize()
    err = float(torch.linalg.norm(img_f - img_p) / torch.linalg.norm(img_p))
    print(f"fused vs plain at 512 x 1024 x 256: image rel L2 {err:.2e}", flush=True)
    assert err <= RTOL
    del tapF, tapB
    torch.cuda.empty_cache()
