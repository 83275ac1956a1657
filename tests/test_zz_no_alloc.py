"""The allocation-free entry wlsq_ktable_device_noalloc, called through the C
ABI: the same results as wlsq_ktable_device (final fields bit for bit, images
within the rounding of the atomic order) on two calls in a row, the second
smaller, with no change in the device's free memory; sizes above its scratch
maximum are refused before anything is queued. Apart from the scratch array
the library loads with its code on the first call on a device (16 MiB, kept
until the process ends), the calls take no device memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _case(seed, ns, nextrap, nf, nx, M, nk, conj):
    import torch
    rng = np.random.default_rng(seed)
    L = 2 * M + 1
    nz, nt = nextrap + 3, nf + 2
    slow = (1.0 / rng.uniform(1500, 4000, (nextrap, nx))).astype(np.float32)
    omega = np.linspace(12.0, 250.0, nf).astype(np.float32)
    k0 = np.float32(omega.min() * slow.min())
    dk = np.float32((omega.max() * slow.max() - k0) / (nk - 1))
    tf = (0.1 * (rng.standard_normal((nk, L)) + 1j * rng.standard_normal((nk, L)))).astype(np.complex64)
    tb = np.conj(tf) if conj else \
        (0.1 * (rng.standard_normal((nk, L)) + 1j * rng.standard_normal((nk, L)))).astype(np.complex64)
    F = (rng.standard_normal((ns, nt, nx)) + 1j * rng.standard_normal((ns, nt, nx))).astype(np.complex64)
    B = (rng.standard_normal((ns, nt, nx)) + 1j * rng.standard_normal((ns, nt, nx))).astype(np.complex64)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return dict(ns=ns, nextrap=nextrap, nz=nz, nt=nt, nf=nf, nx=nx, M=M, nk=nk, k0=float(k0),
                dk=float(dk), slow=d(slow), omega=d(omega), tf=d(tf), tb=d(tb), F=d(F), B=d(B),
                tuning=dict(tables_conjugate=int(conj)))


def _run(c, no_alloc):
    import torch
    from imaging2d_wlsq_amd import device
    F, B = c["F"].clone(), c["B"].clone()
    img = torch.full((c["ns"], c["nz"], c["nx"]), 3.0, device="cuda")
    device.ktable_device(c["ns"], c["nextrap"], c["nz"], c["nt"], c["nf"], c["nx"], c["M"],
                         c["slow"], c["omega"], c["k0"], c["dk"], c["nk"], c["tf"], F, c["tb"], B,
                         img, tuning=c["tuning"], no_alloc=no_alloc)
    torch.cuda.synchronize()
    return F.cpu().numpy(), B.cpu().numpy(), img.cpu().numpy()


def test_no_alloc_entry_matches_the_allocating_one():
    import torch
    cases = [_case(1, ns=3, nextrap=40, nf=12, nx=700, M=15, nk=1500, conj=True),
             _case(2, ns=2, nextrap=25, nf=7, nx=300, M=12, nk=600, conj=False),
             _case(3, ns=1, nextrap=9, nf=3, nx=5000, M=25, nk=800, conj=True)]
    refs = [_run(c, no_alloc=False) for c in cases]
    # the first call on the device loads the library's scratch array (held
    # until the process ends); after that the calls take no device memory
    _run(cases[-1], no_alloc=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for c, (Fr, Br, ir) in zip(cases, refs):
        Fn, Bn, inn = _run(c, no_alloc=True)
        print(f"case ns={c['ns']} nextrap={c['nextrap']} nf={c['nf']} nx={c['nx']} M={c['M']} "
              f"nk={c['nk']}", flush=True)
        np.testing.assert_array_equal(Fn.view(np.uint32), Fr.view(np.uint32))
        np.testing.assert_array_equal(Bn.view(np.uint32), Br.view(np.uint32))
        nz = c["nextrap"]
        err = float(np.linalg.norm(inn[:, :nz] - ir[:, :nz]) / np.linalg.norm(ir[:, :nz]))
        print(f"  image rel L2 vs the allocating entry {err:.2e}", flush=True)
        assert err <= 1e-6
        assert (inn[:, nz:] == 3.0).all()       # rows >= nextrap untouched
    free1 = torch.cuda.mem_get_info()[0]
    print("device free memory before/after the allocation-free calls", free0, free1, flush=True)
    assert abs(free1 - free0) <= (2 << 20)


def test_no_alloc_entry_refuses_sizes_above_its_scratch():
    import torch
    from imaging2d_wlsq_amd import device
    from imaging2d_wlsq_amd._lib import WLSQ_EINVAL, WLSQ_ENOMEM, WlsqError
    c = _case(4, ns=1, nextrap=4, nf=2, nx=64, M=25, nk=21000, conj=False)  # 2 x 8.7 MB of table copies
    F, B = c["F"].clone(), c["B"].clone()
    img = torch.zeros((1, c["nz"], c["nx"]), device="cuda")
    with pytest.raises(WlsqError) as e:
        device.ktable_device(1, c["nextrap"], c["nz"], c["nt"], c["nf"], c["nx"], c["M"], c["slow"],
                             c["omega"], c["k0"], c["dk"], c["nk"], c["tf"], F, c["tb"], B, img,
                             tuning=c["tuning"], no_alloc=True)
    assert e.value.status == WLSQ_ENOMEM and "maximum" in str(e.value)
    with pytest.raises(WlsqError) as e:
        device.ktable_device(1, c["nextrap"], c["nz"], c["nt"], c["nf"], c["nx"], c["M"], c["slow"],
                             c["omega"], c["k0"], c["dk"], 100, c["tf"], F, c["tb"], B, img,
                             tuning=dict(c["tuning"], depth_steps=2), no_alloc=True)
    assert e.value.status == WLSQ_EINVAL
    torch.cuda.synchronize()
    assert torch.equal(F, c["F"]) and not img.any()   # nothing was queued
