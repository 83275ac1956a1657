"""GPU tests of the transforms: the drop-ins for the reference's MKL FFT
wrappers, restating unittests/test_cpp_src.py:77-111 (lengths 10, 33, 155,
values uniform in [-2, 2], each output within 1e-5 of numpy), and the trace
spectra of main.py:119-123 against numpy / scipy.fftpack on the same traces.
numpy's FFT is the oracle here: a library routine, evaluated in float64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SDELTA = 1e-5  # unittests/test_cpp_src.py:13


@pytest.fixture(scope="module", autouse=True)
def _ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _cvec(rng, n):
    a = rng.uniform(-2, 2, n).astype(np.complex64)
    a.imag = rng.uniform(-2, 2, n).astype(np.float32)
    return a


@pytest.mark.parametrize("N", [10, 33, 155])
def test_forward_matches_numpy_elementwise(N):
    from imaging2d_wlsq_amd import interface_hip
    rng = np.random.default_rng(N)
    a = _cvec(rng, N)
    expect = np.fft.fft(a.astype(np.complex128))
    interface_hip.c64fft1dforw(a, N)
    np.testing.assert_allclose(a.real, expect.real, rtol=0, atol=SDELTA)
    np.testing.assert_allclose(a.imag, expect.imag, rtol=0, atol=SDELTA)


@pytest.mark.parametrize("N", [10, 33, 155])
def test_backward_matches_numpy_elementwise(N):
    from imaging2d_wlsq_amd import interface_hip
    rng = np.random.default_rng(100 + N)
    a = _cvec(rng, N)
    expect = np.fft.ifft(a.astype(np.complex128))
    interface_hip.c64fft1dback(a, N)
    np.testing.assert_allclose(a.real, expect.real, rtol=0, atol=SDELTA)
    np.testing.assert_allclose(a.imag, expect.imag, rtol=0, atol=SDELTA)


@pytest.mark.parametrize("N", [1, 2, 800, 1024, 4097])
def test_round_trip_and_large_lengths(N):
    from imaging2d_wlsq_amd import interface_hip
    rng = np.random.default_rng(7 + N)
    a = _cvec(rng, N)
    f = a.copy()
    interface_hip.c64fft1dforw(f, N)
    ref = np.fft.fft(a.astype(np.complex128))
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < 1e-6
    interface_hip.c64fft1dback(f, N)
    assert np.linalg.norm(f - a) / np.linalg.norm(a) < 1e-6


@pytest.mark.parametrize("nt,nx", [(800, 151), (7, 3), (8, 5), (1, 4), (2, 9), (1000, 1024)])
def test_trace_spectra_match_numpy(nt, nx):
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(nt * 31 + nx)
    traces = rng.standard_normal((nt, nx)).astype(np.float32)
    print("trace spectra", nt, nx, flush=True)
    out = device.trace_spectra_device(torch.from_numpy(traces).cuda())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = np.fft.fft(traces.astype(np.float64), axis=0)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-6
    np.testing.assert_allclose(got, ref, rtol=0, atol=SDELTA * max(1.0, np.abs(ref).max() / 10))
    if nt > 1:  # the reference's own call on float32 records
        import scipy.fftpack as spfft
        sp = spfft.fft(traces, axis=0)
        assert np.linalg.norm(got - sp) / np.linalg.norm(sp) < 2e-6


def test_trace_spectra_into_strided_rows_of_a_pulse_buffer():
    """Write into the (nt, nx) rows of one shot of an (ns, nt, ld) buffer with
    ld > nx; the columns beyond nx and the other shots stay untouched."""
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(5)
    ns, nt, nx, ld = 3, 64, 50, 64
    traces = rng.standard_normal((nt, nx)).astype(np.float32)
    buf = torch.full((ns, nt, ld), 7 + 7j, dtype=torch.complex64, device="cuda")
    view = buf[1, :, :nx]
    device.trace_spectra_device(torch.from_numpy(traces).cuda(), out=view)
    torch.cuda.synchronize()
    b = buf.cpu().numpy()
    ref = np.fft.fft(traces.astype(np.float64), axis=0)
    assert np.linalg.norm(b[1, :, :nx] - ref) / np.linalg.norm(ref) < 1e-6
    assert (b[1, :, nx:] == 7 + 7j).all() and (b[0] == 7 + 7j).all() and (b[2] == 7 + 7j).all()


def test_transform_errors_are_reported():
    import torch
    from imaging2d_wlsq_amd import _lib, interface_hip
    with pytest.raises(_lib.WlsqError):
        _lib.check(_lib.lib().c64fft1dforw(np.zeros(4, np.complex64), -1))
    t = torch.zeros((4, 8), device="cuda")
    o = torch.zeros((4, 8), dtype=torch.complex64, device="cuda")
    with pytest.raises(_lib.WlsqError):  # output row stride below nx
        _lib.check(_lib.lib().wlsq_trace_spectra_device(4, 8, t.data_ptr(), 8, o.data_ptr(), 4, None))
    with pytest.raises(ValueError):
        interface_hip.c64fft1dforw(np.zeros(3, np.complex64), 4)
