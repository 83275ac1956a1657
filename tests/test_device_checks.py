"""The device-pointer entries take raw pointers and no sizes, so the Python
wrappers check every extent a call would touch before it reaches the library
(a shot spec reaching past a tensor would otherwise read or write outside it
on the GPU). These checks need no GPU."""
import pytest
import torch

from imaging2d_wlsq_amd import device


def _spec(**kw):
    base = dict(f_off=0, b_off=0, img_off=0, slow_off=0, ld=100, img_ld=100, slow_ld=100, nx=100,
                nz=10, nf=4, omega_off=0)
    base.update(kw)
    return device.ShotSpec(**base)


def _tensors(n_fields=400, n_img=1000, n_slow=1000, n_omega=4, nk=16, M=2):
    L = 2 * M + 1
    c = torch.complex64
    return dict(slowness=torch.zeros(n_slow), omega=torch.zeros(n_omega),
                table_forw=torch.zeros(nk * L, dtype=c), table_back=torch.zeros(nk * L, dtype=c),
                forw=torch.zeros(n_fields, dtype=c), back=torch.zeros(n_fields, dtype=c),
                image=torch.zeros(n_img))


def test_extent_arithmetic():
    assert device._extent(0, 4, 100, 100) == 400
    assert device._extent(7, 3, 50, 20) == 7 + 2 * 50 + 20
    assert device._extent(5, 0, 50, 20) == 0


def test_shots_inside_their_tensors_pass():
    t = _tensors()
    device.check_shot_extents([_spec(), _spec(f_off=0, nx=60, ld=100)], 2, 16, **t)


@pytest.mark.parametrize("bad,name", [
    (dict(f_off=1), "forw"), (dict(b_off=301), "back"), (dict(ld=101), "forw"),
    (dict(img_off=1), "image"), (dict(img_ld=101), "image"), (dict(slow_off=1), "slowness"),
    (dict(slow_ld=101), "slowness"), (dict(omega_off=1), "omega"), (dict(nf=5), "forw"),
    (dict(nz=11), "image")])
def test_a_shot_past_a_tensor_raises(bad, name):
    t = _tensors()
    with pytest.raises(ValueError, match=name):
        device.check_shot_extents([_spec(), _spec(**bad)], 2, 16, **t)


def test_tables_smaller_than_nk_rows_raise():
    t = _tensors(nk=16)
    with pytest.raises(ValueError, match="table_forw"):
        device.check_shot_extents([_spec()], 2, 17, **t)


def test_empty_shots_need_no_room():
    t = _tensors(n_fields=0, n_img=0, n_slow=0, n_omega=0, nk=0)
    device.check_shot_extents([_spec(nx=0), _spec(nz=0), _spec(nf=0)], 2, 16, **t)


def test_uniform_layout_checks():
    c = torch.complex64
    ns, nextrap, nz, nt, nf, nx = 3, 5, 6, 8, 7, 10
    F = torch.zeros(ns * nt * nx, dtype=c)
    img = torch.zeros(ns * nz * nx)
    device._check_uniform(ns, nextrap, nz, nt, nf, nx, F, F, img)
    # the last shot's rows j < nf end at (ns-1)*nt*nx + nf*nx: shorter tensors fail
    short = torch.zeros((ns - 1) * nt * nx + nf * nx - 1, dtype=c)
    with pytest.raises(ValueError, match="forw"):
        device._check_uniform(ns, nextrap, nz, nt, nf, nx, short, F, img)
    with pytest.raises(ValueError, match="image"):
        device._check_uniform(ns, nextrap, nz, nt, nf, nx, F, F, img[:-(nz - nextrap) * nx - 1])
