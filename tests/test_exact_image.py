"""GPU tests of the exact image accumulation (wlsq_tuning.exact_image): the
device's exact sums against the Python checker (correct rounding, ties,
subnormals, cancellation, infinities, NaN), and fused launches whose images are
bit for bit the same across repeated launches, graph replays and depth-split
launches, within rounding of the float-atomic image."""
import numpy as np
import pytest

from exact_sum import exact_sum_f32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32)) or \
        (np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]))


def test_exact_sums_match_the_checker():
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(3)
    n = 600
    cols = [rng.standard_normal(n).astype(np.float32) * np.float32(2.0) ** rng.integers(-60, 60, n).astype(np.float32)
            for _ in range(24)]
    cols += [rng.standard_normal(n).astype(np.float32) * np.float32(2.0) ** rng.integers(-149, -120, n).astype(np.float32)
             for _ in range(4)]                                          # subnormal and tiny
    c = np.zeros(n, np.float32)
    c[:3] = [1e30, 1.0, -1e30]
    cols.append(c)                                                       # cancellation
    for base in (16777216.0, 16777218.0):                                # ties to even
        c = np.zeros(n, np.float32)
        c[:2] = [base, 1.0]
        cols.append(c)
    c = np.full(n, 3.0e38, np.float32)
    cols.append(c)                                                       # overflow
    for special in ([np.inf], [np.nan], [np.inf, -np.inf], [-np.inf]):
        c = rng.standard_normal(n).astype(np.float32)
        c[:len(special)] = special
        cols.append(c)
    terms = np.stack(cols, axis=1).astype(np.float32)                  # (n, m)
    got = device.exact_sum(torch.from_numpy(terms).cuda()).cpu().numpy()
    want = np.array([exact_sum_f32(terms[:, j]) for j in range(terms.shape[1])], np.float32)
    bad = [j for j in range(terms.shape[1]) if not _same(got[j:j + 1], want[j:j + 1])]
    assert not bad, [(j, got[j], want[j]) for j in bad]
    again = device.exact_sum(torch.from_numpy(terms[::-1].copy()).cuda()).cpu().numpy()
    assert _same(again, got)                                            # order of the terms


@pytest.mark.parametrize("shots_per_group", [1, 4])
def test_exact_images_repeat_bit_for_bit(shots_per_group):
    import torch
    from imaging2d_wlsq_amd import batch, synth
    from test_hip_batch_serve import _small_problem, _uneven_shots, rel
    p = _small_problem(seed=23, nz=40, nx=700)
    base = _uneven_shots(p, 3, 61)
    shots = [synth.ShotGeom(s.x0, s.nx, s.nz, s.xs + 0.5 * k, s.freqs, s.seed + k)
             for s in base for k in range(4)]
    sv = batch.Survey(p)
    tun = dict(tables_conjugate=int(sv.conj), tables_symmetric=int(sv.sym),
               shots_per_group=shots_per_group, x_per_thread=4 if shots_per_group == 1 else 2)
    ref_b = sv.batch(shots, tuning=tun, write_back=True)
    ref = ref_b.run().clone()
    torch.cuda.synchronize()
    ref_F = ref_b.F.clone()
    ref_b.close()
    imgs = []
    for variant in (dict(), dict(depth_steps=7)):
        for graph in (False, True):
            # no write-back: every run starts from the same fields
            b = sv.batch(shots, tuning=dict(tun, exact_image=1, **variant), graph=graph)
            for rep in range(2):
                img = b.run().clone()
                torch.cuda.synchronize()
                imgs.append(img.cpu().numpy())
            print(variant, "graph" if graph else "eager", b.plan.info(), flush=True)
            b.close()
    b = sv.batch(shots, tuning=dict(tun, exact_image=1), write_back=True)
    b.run()
    torch.cuda.synchronize()
    assert torch.equal(b.F, ref_F)                                     # the fields do not change
    b.close()
    for k, img in enumerate(imgs):
        assert np.array_equal(img.view(np.uint32), imgs[0].view(np.uint32)), k
    assert rel(imgs[0], ref.cpu().numpy()) <= 1e-6


def test_exact_image_refusals():
    import torch
    from imaging2d_wlsq_amd import batch, device
    from imaging2d_wlsq_amd._lib import WLSQ_EINVAL, WLSQ_EUNSUPPORTED, WlsqError
    from test_hip_batch_serve import _small_problem, _uneven_shots
    p = _small_problem(seed=24, nz=20, nx=300)
    shots = _uneven_shots(p, 2, 62)
    sv = batch.Survey(p)
    t = p.tables
    b = sv.batch(shots, write_back=True)
    tun = dict(b.tuning, exact_image=1)
    ws = torch.empty(1 << 22, dtype=torch.uint8, device="cuda")
    with pytest.raises(WlsqError) as e:
        device.batch_launch_ws(b.specs, t.M, sv.slowness, b.omega, t.kappa0, t.dkappa, t.nk, sv.tf,
                               sv.tb, b.F, b.B, b.image, ws, tuning=tun)
    assert e.value.status == WLSQ_EINVAL
    with pytest.raises(WlsqError) as e:
        device.Plan(b.specs, t.M, dict(tun, image_mode=2))
    assert e.value.status == WLSQ_EUNSUPPORTED
    with pytest.raises(WlsqError) as e:
        device.Plan(b.specs, t.M, dict(tun, tables_conjugate=0))
    assert e.value.status == WLSQ_EUNSUPPORTED
    b.close()
