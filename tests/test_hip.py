"""GPU parity tests: the HIP library against the pinned CPU oracle.

Tolerances: the reference's KATs are exact small integers and are checked to
the reference's absolute 1e-5. Random and synthetic cases compare relative L2
errors of the image and the final wavefields with the FP32 CPU oracle at
<= 1e-5; the FP64 oracle error of both paths is
checked on the C1 configuration.
"""
import math

import numpy as np
import pytest

import kats
import oracle
from oracle import fp64

pytestmark = pytest.mark.gpu

RTOL = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _ready():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    oracle.build()
    from imaging2d_wlsq_amd import _lib
    _lib.lib()
    print(f"GPU: {torch.cuda.get_device_name(0)}", flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def crel(a, b):
    return rel(np.concatenate([np.real(a).ravel(), np.imag(a).ravel()]),
               np.concatenate([np.real(b).ravel(), np.imag(b).ravel()]))


def _oracle_run(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B):
    Fo, Bo = F.copy(), B.copy()
    img = np.zeros((ns, nz, nx), np.float32)
    oracle.extrapolate(ns, nextrap, nz, nx, nf, nt, M, opF, opB, Fo, Bo, img)
    return img, Fo, Bo


def _hip_dropin(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B, fill=0.0):
    from imaging2d_wlsq_amd import interface_hip
    Fh, Bh = F.copy(), B.copy()
    img = np.full((ns, nz, nx), fill, np.float32)
    interface_hip.extrapolate(ns, nextrap, nz, nt, nf, nx, M, opF, Fh, opB, Bh, img)
    return img, Fh, Bh


@pytest.mark.parametrize("case", kats.all_kats(), ids=lambda c: c["name"])
def test_drop_in_reference_kats(case):
    print("case", case["name"], flush=True)
    img, F, B = _hip_dropin(case["ns"], case["nextrap"], case["nz"], case["nt"], case["nf"],
                            case["nx"], case["M"], case["opF"], case["opB"], case["F"], case["B"])
    kats.check_case(case, img[0], F[0, :case["nf"]], B[0, :case["nf"]])


def _rand(rng, shape, scale=1.0):
    return (scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)


SHAPES = [  # ns, nextrap, nz, nt, nf, nx, M
    (1, 3, 4, 5, 4, 7, 2),
    (3, 5, 5, 6, 6, 130, 3),
    (5, 4, 6, 9, 9, 64, 0),
    (2, 6, 6, 3, 1, 200, 12),
    (4, 7, 9, 17, 17, 97, 15),
    (1, 2, 2, 2, 2, 1, 1),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_drop_in_matches_oracle_random(shape):
    ns, nextrap, nz, nt, nf, nx, M = shape
    print("shape", shape, flush=True)
    rng = np.random.default_rng(sum(shape))
    L = 2 * M + 1
    opF = _rand(rng, (nextrap, nf, nx, L), 0.7 / math.sqrt(L))
    opB = _rand(rng, (nextrap, nf, nx, L), 0.7 / math.sqrt(L))
    F = _rand(rng, (ns, nt, nx))
    B = _rand(rng, (ns, nt, nx))
    ri, rF, rB = _oracle_run(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B)
    hi, hF, hB = _hip_dropin(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B, fill=3.5)
    assert rel(hi[:, :nextrap], ri[:, :nextrap]) <= RTOL
    assert (hi[:, nextrap:] == 3.5).all()           # rows >= nextrap untouched
    assert crel(hF, rF) <= RTOL and crel(hB, rB) <= RTOL
    np.testing.assert_array_equal(hF[:, nf:], F[:, nf:])  # rows >= nf untouched
    # deterministic: fixed-order reduction, bitwise repeatable
    hi2, hF2, _ = _hip_dropin(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B, fill=3.5)
    np.testing.assert_array_equal(hi2, hi)
    np.testing.assert_array_equal(hF2, hF)


def test_plain_device_shot_batching_and_layouts_are_bitwise_equal():
    import torch
    from imaging2d_wlsq_amd import device, _lib
    rng = np.random.default_rng(17)
    ns, nextrap, nz, nt, nf, nx, M = 5, 6, 7, 10, 9, 150, 4
    L = 2 * M + 1
    opF = _rand(rng, (nextrap, nf, nx, L), 0.3)
    opB = _rand(rng, (nextrap, nf, nx, L), 0.3)
    F = _rand(rng, (ns, nt, nx))
    B = _rand(rng, (ns, nt, nx))
    ri, rF, rB = _oracle_run(ns, nextrap, nz, nt, nf, nx, M, opF, opB, F, B)
    results = []
    tapF = np.ascontiguousarray(opF.transpose(0, 1, 3, 2))
    tapB = np.ascontiguousarray(opB.transpose(0, 1, 3, 2))
    for layout, (tf, tb) in [(_lib.LAYOUT_NATURAL, (opF, opB)), (_lib.LAYOUT_TAPMAJOR, (tapF, tapB))]:
        for spl in (1, 2, 4):
            print("layout", layout, "shots_per_launch", spl, flush=True)
            dF = torch.from_numpy(F).cuda()
            dB = torch.from_numpy(B).cuda()
            img = torch.zeros((ns, nz, nx), dtype=torch.float32, device="cuda")
            device.plain_device(ns, nextrap, nz, nt, nf, nx, M, torch.from_numpy(tf).cuda(), dF,
                                torch.from_numpy(tb).cuda(), dB, img, layout=layout,
                                shots_per_launch=spl)
            torch.cuda.synchronize()
            results.append((img.cpu().numpy(), dF.cpu().numpy(), dB.cpu().numpy()))
    for r in results[1:]:
        for a, b in zip(r, results[0]):
            np.testing.assert_array_equal(a, b)
    hi, hF, hB = results[0]
    assert rel(hi[:, :nextrap], ri[:, :nextrap]) <= RTOL
    assert crel(hF, rF) <= RTOL


def _ktable_problem(rng, nx, M, nf, nextrap, nk=300, conj=True):
    L = 2 * M + 1
    slow = (1.0 / rng.uniform(1500, 4000, (nextrap, nx))).astype(np.float32)
    omega = np.linspace(2 * np.pi * 2, 2 * np.pi * 45, nf).astype(np.float32)
    k0 = float(omega.min() * slow.min())
    dk = float((omega.max() * slow.max() - k0) / (nk - 1))
    tf = _rand(rng, (nk, L), 0.6 / math.sqrt(L))
    tb = np.conj(tf).astype(np.complex64) if conj else _rand(rng, (nk, L), 0.6 / math.sqrt(L))
    return slow, omega, np.float32(k0), np.float32(dk), tf, tb


KT_SHAPES = [  # ns, nextrap, nf, nx, M, conj
    (1, 4, 3, 1, 0, True),
    (2, 5, 5, 7, 3, False),
    (3, 6, 33, 100, 12, True),
    (5, 3, 7, 513, 15, False),
    (4, 8, 9, 1024, 15, True),
    (1, 4, 4, 300, 25, True),
    (2, 3, 2, 2100, 15, True),
    (1, 3, 2, 4100, 15, True),
    (1, 2, 2, 8192, 25, True),
    (1, 2, 3, 40, 21, False),
]


@pytest.mark.parametrize("shape", KT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_ktable_fused_matches_oracle(shape):
    from imaging2d_wlsq_amd import interface_hip
    ns, nextrap, nf, nx, M, conj = shape
    print("ktable shape", shape, flush=True)
    rng = np.random.default_rng(1000 + sum(shape[:5]))
    slow, omega, k0, dk, tf, tb = _ktable_problem(rng, nx, M, nf, nextrap, conj=conj)
    nt, nz = nf + 1, nextrap + 2
    F = _rand(rng, (ns, nt, nx))
    B = _rand(rng, (ns, nt, nx))
    Fo, Bo = F.copy(), B.copy()
    io = np.zeros((ns, nz, nx), np.float32)
    oracle.extrapolate_ktable(ns, nextrap, nz, nx, nf, nt, M, slow, omega, k0, dk, 300, tf, tb, Fo, Bo, io)
    Fh, Bh = F.copy(), B.copy()
    ih = np.full((ns, nz, nx), -2.0, np.float32)
    interface_hip.extrapolate_ktable(ns, nextrap, nz, nt, nf, nx, M, slow, omega, k0, dk, 300,
                                     tf, Fh, tb, Bh, ih)
    assert rel(ih[:, :nextrap], io[:, :nextrap]) <= RTOL
    assert (ih[:, nextrap:] == -2.0).all()
    assert crel(Fh, Fo) <= RTOL and crel(Bh, Bo) <= RTOL
    np.testing.assert_array_equal(Fh[:, nf:], F[:, nf:])


@pytest.mark.parametrize("tuning", [
    dict(shots_per_group=1), dict(shots_per_group=2), dict(shots_per_group=4),
    dict(x_per_thread=8), dict(omegas_per_item=1), dict(omegas_per_item=3, block=512),
    dict(tables_conjugate=1), dict(image_mode=1), dict(image_mode=2),
    dict(image_mode=2, shots_per_group=4), dict(x_per_thread=2, shots_per_group=4),
    dict(tap_loads=2), dict(tap_loads=2, shots_per_group=1), dict(tap_loads=1, shots_per_group=1),
    dict(tap_loads=2, x_per_thread=8),
    dict(x_per_thread=4, shots_per_group=1, block=1024),
    dict(x_per_thread=4, shots_per_group=1, block=1024, tap_loads=2, omegas_per_item=2),
    dict(x_per_thread=4, shots_per_group=1, block=640),
    dict(double_rows=2), dict(double_rows=1, shots_per_group=1),
    dict(double_rows=2, x_per_thread=4, shots_per_group=1, block=1024),
    dict(x_per_thread=2, shots_per_group=2, block=1024),
    dict(depth_steps=2), dict(depth_steps=4, double_rows=2), dict(depth_steps=1, shots_per_group=1),
    dict(image_mode=3), dict(image_mode=3, omegas_per_item=1),
    dict(image_mode=3, omegas_per_item=1, shots_per_group=4),
    dict(image_mode=3, omegas_per_item=1, x_per_thread=4, shots_per_group=1, block=1024),
    dict(image_mode=3, omegas_per_item=1, x_per_thread=8),
    dict(image_mode=3, omegas_per_item=1, x_per_thread=2, shots_per_group=2),
    dict(image_mode=3, omegas_per_item=1, double_rows=2, shots_per_group=1),
    dict(image_mode=3, omegas_per_item=1, depth_steps=2, shots_per_group=1),
], ids=lambda t: ",".join(f"{k}={v}" for k, v in t.items()))
def test_ktable_device_tunings_agree(tuning):
    """Every launch shape computes the same fields bit for bit (taps are summed
    in the same order); images agree to rounding of the atomic order."""
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(77)
    ns, nextrap, nf, nx, M = 5, 6, 7, 333, 15
    slow, omega, k0, dk, tf, tb = _ktable_problem(rng, nx, M, nf, nextrap, conj=True)
    nt, nz = nf, nextrap
    F = _rand(rng, (ns, nt, nx))
    B = _rand(rng, (ns, nt, nx))
    Fo, Bo = F.copy(), B.copy()
    io = np.zeros((ns, nz, nx), np.float32)
    oracle.extrapolate_ktable(ns, nextrap, nz, nx, nf, nt, M, slow, omega, k0, dk, 300, tf, tb, Fo, Bo, io)
    outs = []
    for t in (dict(), tuning):
        print("tuning", t, flush=True)
        dF, dB = torch.from_numpy(F).cuda(), torch.from_numpy(B).cuda()
        img = torch.full((ns, nz, nx), 9.0, device="cuda")
        device.ktable_device(ns, nextrap, nz, nt, nf, nx, M, torch.from_numpy(slow).cuda(),
                             torch.from_numpy(omega).cuda(), k0, dk, 300, torch.from_numpy(tf).cuda(),
                             dF, torch.from_numpy(tb).cuda(), dB, img, tuning=t)
        torch.cuda.synchronize()
        outs.append((img.cpu().numpy(), dF.cpu().numpy(), dB.cpu().numpy()))
    np.testing.assert_array_equal(outs[1][1], outs[0][1])
    np.testing.assert_array_equal(outs[1][2], outs[0][2])
    assert rel(outs[1][0], outs[0][0]) <= 1e-6
    assert rel(outs[1][0], io) <= RTOL


@pytest.mark.parametrize("tuning", [
    dict(shots_per_group=1), dict(shots_per_group=2), dict(shots_per_group=4),
    dict(x_per_thread=8), dict(omegas_per_item=3, block=512), dict(tables_conjugate=1),
    dict(x_per_thread=2, shots_per_group=2), dict(x_per_thread=2, shots_per_group=4),
    dict(shots_per_group=1, tap_loads=2), dict(shots_per_group=2, tap_loads=2),
    dict(x_per_thread=8, tap_loads=2), dict(x_per_thread=2, shots_per_group=2, tap_loads=2),
], ids=lambda t: ",".join(f"{k}={v}" for k, v in t.items()))
@pytest.mark.parametrize("M", [15, 4, 0, 25, 1])
def test_tap_paired_variant_matches_oracle(tuning, M):
    """Tap-symmetric tables (the WLSQ design gives them): the paired-tap
    kernel computes the same operator in a different rounding order."""
    import torch
    from imaging2d_wlsq_amd import autotune, device
    rng = np.random.default_rng(400 + M)
    ns, nextrap, nf, nx = 5, 6, 7, 301
    slow, omega, k0, dk, tf, _ = _ktable_problem(rng, nx, M, nf, nextrap, conj=True)
    tf = np.ascontiguousarray((tf + tf[:, ::-1]) * 0.5).astype(np.complex64)
    tb = np.conj(tf).astype(np.complex64)
    assert autotune.tables_symmetric(torch.from_numpy(tf).cuda())
    F = _rand(rng, (ns, nf, nx))
    B = _rand(rng, (ns, nf, nx))
    Fo, Bo = F.copy(), B.copy()
    io = np.zeros((ns, nextrap, nx), np.float32)
    oracle.extrapolate_ktable(ns, nextrap, nextrap, nx, nf, nf, M, slow, omega, k0, dk, 300, tf, tb, Fo, Bo, io)
    outs = []
    for sym in (0, 1):
        t = dict(tuning, tables_symmetric=sym)
        print("tuning", t, "M", M, flush=True)
        dF, dB = torch.from_numpy(F).cuda(), torch.from_numpy(B).cuda()
        img = torch.zeros((ns, nextrap, nx), device="cuda")
        device.ktable_device(ns, nextrap, nextrap, nf, nf, nx, M, torch.from_numpy(slow).cuda(),
                             torch.from_numpy(omega).cuda(), k0, dk, 300, torch.from_numpy(tf).cuda(),
                             dF, torch.from_numpy(tb).cuda(), dB, img, tuning=t)
        torch.cuda.synchronize()
        outs.append((img.cpu().numpy(), dF.cpu().numpy(), dB.cpu().numpy()))
    for o in outs:
        assert rel(o[0], io) <= RTOL
        assert crel(o[1], Fo) <= RTOL and crel(o[2], Bo) <= RTOL
    assert crel(outs[1][1], outs[0][1]) <= 1e-6


def _blocky_slowness(rng, nextrap, nx, noisy_from):
    """Rows of constant runs (8-40 x long: at most one kappa change over any
    8 consecutive x) up to x = noisy_from, then a random value per x."""
    slow = np.empty((nextrap, nx), np.float32)
    for l in range(nextrap):
        x = 0
        while x < noisy_from:
            n = int(rng.integers(8, 41))
            slow[l, x:x + n] = 1.0 / rng.uniform(1500, 4000)
            x += n
        slow[l, noisy_from:] = 1.0 / rng.uniform(1500, 4000, nx - noisy_from)
    return slow


@pytest.mark.parametrize("tuning", [
    dict(), dict(x_per_thread=2, shots_per_group=4), dict(x_per_thread=2, shots_per_group=2, tap_loads=2),
    dict(x_per_thread=4, shots_per_group=1, tap_loads=2), dict(x_per_thread=4, shots_per_group=1, tap_loads=1),
    dict(x_per_thread=4, shots_per_group=2), dict(x_per_thread=8, tap_loads=2),
    dict(x_per_thread=8, tap_loads=1), dict(x_per_thread=8, tables_symmetric=0),
    dict(x_per_thread=4, shots_per_group=1, block=1024),
    dict(x_per_thread=4, shots_per_group=1, block=1024, tables_symmetric=0, tap_loads=2),
    dict(x_per_thread=4, shots_per_group=1, block=1024, image_mode=3, omegas_per_item=1),
], ids=lambda t: ",".join(f"{k}={v}" for k, v in t.items()) or "auto")
@pytest.mark.parametrize("pattern", ["constant", "blocky", "blocky_then_noise"])
@pytest.mark.parametrize("conj", [True, False])
def test_strip_gathers_are_bitwise_the_per_x_gathers(tuning, pattern, conj):
    """One or two coefficient gathers per strip (wavefronts whose kappa rows
    change at most once per strip) use the same coefficients as one gather
    per x: the fields are bitwise equal, and both match the oracle."""
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(900 + len(pattern) + 7 * len(tuning) + conj)
    ns, nextrap, nf, nx, M = 4, 5, 6, 1000, 15
    if (tuning.get("x_per_thread") == 8 or tuning.get("block") == 1024) and pattern == "blocky":
        nx = 5000  # two strips per thread
    _, omega, k0, dk, tf, tb = _ktable_problem(rng, nx, M, nf, nextrap, conj=conj)
    tf = np.ascontiguousarray((tf + tf[:, ::-1]) * 0.5).astype(np.complex64)
    tb = np.conj(tf).astype(np.complex64) if conj else np.ascontiguousarray(
        (tb + tb[:, ::-1]) * 0.5).astype(np.complex64)
    if pattern == "constant":
        slow = np.full((nextrap, nx), 1.0 / 2500.0, np.float32)
    else:
        slow = _blocky_slowness(rng, nextrap, nx, nx if pattern == "blocky" else nx // 2)
    F = _rand(rng, (ns, nf, nx))
    B = _rand(rng, (ns, nf, nx))
    Fo, Bo = F.copy(), B.copy()
    io = np.zeros((ns, nextrap, nx), np.float32)
    oracle.extrapolate_ktable(ns, nextrap, nextrap, nx, nf, nf, M, slow, omega, k0, dk, 300, tf, tb, Fo, Bo, io)
    outs = []
    for sg in (2, 0):
        t = dict(tuning, strip_gathers=sg, tables_conjugate=int(conj))
        print("tuning", t, pattern, "nx", nx, flush=True)
        dF, dB = torch.from_numpy(F).cuda(), torch.from_numpy(B).cuda()
        img = torch.zeros((ns, nextrap, nx), device="cuda")
        device.ktable_device(ns, nextrap, nextrap, nf, nf, nx, M, torch.from_numpy(slow).cuda(),
                             torch.from_numpy(omega).cuda(), k0, dk, 300, torch.from_numpy(tf).cuda(),
                             dF, torch.from_numpy(tb).cuda(), dB, img, tuning=t)
        torch.cuda.synchronize()
        outs.append((img.cpu().numpy(), dF.cpu().numpy(), dB.cpu().numpy()))
    np.testing.assert_array_equal(outs[1][1], outs[0][1])
    np.testing.assert_array_equal(outs[1][2], outs[0][2])
    assert rel(outs[1][0], outs[0][0]) <= 1e-6
    assert rel(outs[1][0], io) <= RTOL
    assert crel(outs[1][1], Fo) <= RTOL and crel(outs[1][2], Bo) <= RTOL


@pytest.mark.parametrize("layout", [0, 1])
def test_device_expansion_is_bitwise_the_oracle_expansion(layout):
    import torch
    from imaging2d_wlsq_amd import device, prep
    rng = np.random.default_rng(5 + layout)
    nextrap, nf, nx, M, nk = 6, 11, 77, 5, 300
    slow, omega, k0, dk, tf, _ = _ktable_problem(rng, nx, M, nf, nextrap)
    ref = oracle.expand_table(nextrap, nf, nx, M, slow, omega, k0, dk, nk, tf)
    np.testing.assert_array_equal(prep.expand_table(tf, slow, omega, k0, dk, nk, nextrap), ref)
    L = 2 * M + 1
    out = torch.empty(nextrap * nf * nx * L, dtype=torch.complex64, device="cuda")
    device.expand_table_device(nextrap, nf, nx, M, torch.from_numpy(slow).cuda(),
                               torch.from_numpy(omega).cuda(), k0, dk, nk,
                               torch.from_numpy(tf).cuda(), out, layout=layout)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(nextrap, nf, nx, L) if layout == 0 else \
        out.cpu().numpy().reshape(nextrap, nf, L, nx).transpose(0, 1, 3, 2)
    np.testing.assert_array_equal(got, ref)


def test_ktable_and_expanded_paths_give_identical_fields():
    """k-table (fused, on-device lookup) == plain path on the expanded table:
    the same q everywhere, taps summed in the same order."""
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(31)
    ns, nextrap, nf, nx, M, nk = 3, 5, 10, 260, 15, 300
    slow, omega, k0, dk, tf, tb = _ktable_problem(rng, nx, M, nf, nextrap)
    L = 2 * M + 1
    F = _rand(rng, (ns, nf, nx))
    B = _rand(rng, (ns, nf, nx))
    dslow, domg = torch.from_numpy(slow).cuda(), torch.from_numpy(omega).cuda()
    tapF = torch.empty(nextrap * nf * L * nx, dtype=torch.complex64, device="cuda")
    tapB = torch.empty_like(tapF)
    device.expand_table_device(nextrap, nf, nx, M, dslow, domg, k0, dk, nk, torch.from_numpy(tf).cuda(), tapF)
    device.expand_table_device(nextrap, nf, nx, M, dslow, domg, k0, dk, nk, torch.from_numpy(tb).cuda(), tapB)
    pF, pB = torch.from_numpy(F).cuda(), torch.from_numpy(B).cuda()
    pimg = torch.zeros((ns, nextrap, nx), device="cuda")
    device.plain_device(ns, nextrap, nextrap, nf, nf, nx, M, tapF, pF, tapB, pB, pimg)
    kF, kB = torch.from_numpy(F).cuda(), torch.from_numpy(B).cuda()
    kimg = torch.zeros((ns, nextrap, nx), device="cuda")
    device.ktable_device(ns, nextrap, nextrap, nf, nf, nx, M, dslow, domg, k0, dk, nk,
                         torch.from_numpy(tf).cuda(), kF, torch.from_numpy(tb).cuda(), kB, kimg)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(kF.cpu().numpy(), pF.cpu().numpy())
    np.testing.assert_array_equal(kB.cpu().numpy(), pB.cpu().numpy())
    assert rel(kimg.cpu().numpy(), pimg.cpu().numpy()) <= 1e-6


def test_c1_configuration_against_both_oracles():
    """C1 in full: 128 x 256, nf 64, L 25, one shot. Drop-in (expanded table)
    and k-table paths vs the FP32 CPU oracle (<= 1e-5) and the FP64 oracle."""
    from imaging2d_wlsq_amd import interface_hip, prep, synth
    p = synth.config_c1()
    t, m, s = p.tables, p.model, p.shots[0]
    nz, nx, nf, M = m.nz, m.nx, s.nf, t.M
    slow, omega = m.slowness, s.omega
    opF = prep.expand_table(t.forw, slow, omega, t.kappa0, t.dkappa, t.nk, nz)
    opB = prep.expand_table(t.back, slow, omega, t.kappa0, t.dkappa, t.nk, nz)
    F, B = synth.host_fields(p)
    ri, rF, _ = _oracle_run(1, nz, nz, nf, nf, nx, M, opF, opB, F, B)
    hi, hF, _ = _hip_dropin(1, nz, nz, nf, nf, nx, M, opF, opB, F, B)
    e_drop = rel(hi, ri)
    Fk, Bk = F.copy(), B.copy()
    ki = np.zeros((1, nz, nx), np.float32)
    interface_hip.extrapolate_ktable(1, nz, nz, nf, nf, nx, M, slow, omega, t.kappa0, t.dkappa, t.nk,
                                     t.forw, Fk, t.back, Bk, ki)
    e_kt = rel(ki, ri)
    d64, _, _ = fp64.extrapolate(nz, nz, nx, nf, M, wave_fs=F[0], shot_fs=B[0],
                                 ktable=dict(slowness=slow, omega=omega, kappa0=t.kappa0,
                                             dkappa=t.dkappa, nk=t.nk, forw=t.forw, back=t.back))
    print(f"C1 rel L2: drop-in vs CPU {e_drop:.2e}, k-table vs CPU {e_kt:.2e}, "
          f"CPU vs FP64 {rel(ri[0], d64):.2e}, drop-in vs FP64 {rel(hi[0], d64):.2e}, "
          f"k-table vs FP64 {rel(ki[0], d64):.2e}", flush=True)
    assert e_drop <= RTOL and e_kt <= RTOL
    assert crel(hF, rF) <= RTOL
    assert rel(hi[0], d64) <= 1e-4 and rel(ri[0], d64) <= 1e-4


def _batch_setup(rng, conj=True):
    """A small uneven batch: 6 shots, windows of a 40 x 300 model, own
    frequency lists, stacked image."""
    nzm, nxm, M, nk = 12, 300, 7, 300
    slow = (1.0 / rng.uniform(1500, 4000, (nzm, nxm))).astype(np.float32)
    L = 2 * M + 1
    tf = _rand(rng, (nk, L), 0.6 / math.sqrt(L))
    tb = np.conj(tf).astype(np.complex64) if conj else _rand(rng, (nk, L), 0.6 / math.sqrt(L))
    shots = []
    omegas = []
    o_off = 0
    f_off = 0
    for k in range(6):
        nx = int(rng.integers(5, 300))
        x0 = int(rng.integers(0, nxm - nx + 1))
        nf = int(rng.integers(1, 9))
        nz = int(rng.integers(1, nzm + 1))
        om = np.sort(rng.uniform(2 * np.pi * 2, 2 * np.pi * 40, nf)).astype(np.float32)
        shots.append(dict(nx=nx, x0=x0, nf=nf, nz=nz, o_off=o_off, f_off=f_off))
        omegas.append(om)
        o_off += nf
        f_off += nf * nx
    omega = np.concatenate(omegas)
    k0 = float(omega.min() * slow.min())
    dk = float((omega.max() * slow.max() - k0) / (nk - 1))
    F = _rand(rng, (f_off,))
    B = _rand(rng, (f_off,))
    return dict(nzm=nzm, nxm=nxm, M=M, nk=nk, slow=slow, tf=tf, tb=tb, shots=shots, omega=omega,
                k0=np.float32(k0), dk=np.float32(dk), F=F, B=B)


def _batch_oracle(b):
    img = np.zeros((b["nzm"], b["nxm"]), np.float64)
    F_out, B_out = b["F"].copy(), b["B"].copy()
    for s in b["shots"]:
        nx, nf, nz, x0 = s["nx"], s["nf"], s["nz"], s["x0"]
        Fs = b["F"][s["f_off"]:s["f_off"] + nf * nx].reshape(1, nf, nx).copy()
        Bs = b["B"][s["f_off"]:s["f_off"] + nf * nx].reshape(1, nf, nx).copy()
        io = np.zeros((1, nz, nx), np.float32)
        sl = np.ascontiguousarray(b["slow"][:nz, x0:x0 + nx])
        om = b["omega"][s["o_off"]:s["o_off"] + nf]
        oracle.extrapolate_ktable(1, nz, nz, nx, nf, nf, b["M"], sl, om, b["k0"], b["dk"], b["nk"],
                                  b["tf"], b["tb"], Fs, Bs, io)
        img[:nz, x0:x0 + nx] += io[0]
        F_out[s["f_off"]:s["f_off"] + nf * nx] = Fs.ravel()
        B_out[s["f_off"]:s["f_off"] + nf * nx] = Bs.ravel()
    return img, F_out, B_out


def _batch_specs(b):
    from imaging2d_wlsq_amd.device import ShotSpec
    return [ShotSpec(f_off=s["f_off"], b_off=s["f_off"], img_off=s["x0"], slow_off=s["x0"],
                     ld=s["nx"], img_ld=b["nxm"], slow_ld=b["nxm"], nx=s["nx"], nz=s["nz"],
                     nf=s["nf"], omega_off=s["o_off"]) for s in b["shots"]]


@pytest.mark.parametrize("conj", [True, False])
def test_batched_uneven_shots_match_per_shot_oracle(conj):
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(90 + conj)
    b = _batch_setup(rng, conj)
    ref_img, ref_F, ref_B = _batch_oracle(b)
    plan = device.Plan(_batch_specs(b), b["M"], dict(tables_conjugate=int(conj)))
    print("plan", plan.info(), flush=True)
    dF, dB = torch.from_numpy(b["F"]).cuda(), torch.from_numpy(b["B"]).cuda()
    img = torch.zeros((b["nzm"], b["nxm"]), device="cuda")
    plan.launch(torch.from_numpy(b["slow"]).cuda(), torch.from_numpy(b["omega"]).cuda(), b["k0"],
                b["dk"], b["nk"], torch.from_numpy(b["tf"]).cuda(), torch.from_numpy(b["tb"]).cuda(),
                dF, dB, img, write_back=True)
    torch.cuda.synchronize()
    assert rel(img.cpu().numpy(), ref_img) <= RTOL
    assert crel(dF.cpu().numpy(), ref_F) <= RTOL and crel(dB.cpu().numpy(), ref_B) <= RTOL


def test_graph_replays_match_eager():
    """Captured [clear image + batched launch], replayed several times with
    eager work and synchronisation between replays; every replay equals the
    eager launch."""
    import torch
    from imaging2d_wlsq_amd import device
    rng = np.random.default_rng(123)
    b = _batch_setup(rng, True)
    plan = device.Plan(_batch_specs(b), b["M"], dict(tables_conjugate=1))
    args = [torch.from_numpy(b["slow"]).cuda(), torch.from_numpy(b["omega"]).cuda(), b["k0"], b["dk"],
            b["nk"], torch.from_numpy(b["tf"]).cuda(), torch.from_numpy(b["tb"]).cuda()]
    dF, dB = torch.from_numpy(b["F"]).cuda(), torch.from_numpy(b["B"]).cuda()
    eager = torch.zeros((b["nzm"], b["nxm"]), device="cuda")
    plan.launch(*args, dF, dB, eager, write_back=False)
    torch.cuda.synchronize()
    img = torch.full((b["nzm"], b["nxm"]), 5.0, device="cuda")
    g = plan.graph(*args, dF, dB, img, write_back=False, zero_count=img.numel())
    for rep in range(4):
        g.launch()
        torch.cuda.synchronize()
        got = img.cpu().numpy()
        assert rel(got, eager.cpu().numpy()) <= 1e-6, rep
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        img.fill_(float(rep))
        torch.cuda.synchronize()
        del junk
    g.close()
    plan.close()


def test_errors_are_reported_not_silent():
    from imaging2d_wlsq_amd import _lib, device
    import torch
    rng = np.random.default_rng(1)
    b = _batch_setup(rng)
    specs = _batch_specs(b)
    specs[0].nx = 5000
    specs[0].ld = 5000
    specs[0].img_ld = 5000
    specs[0].slow_ld = 5000
    with pytest.raises(_lib.WlsqError):  # 1250 strips of 2 shots > 512 threads
        device.Plan(specs, b["M"], dict(x_per_thread=4, shots_per_group=2))
    with pytest.raises(_lib.WlsqError):  # two strips per thread of 4 x need > 512 threads
        device.Plan(specs, b["M"], dict(x_per_thread=4, shots_per_group=1, block=512))
    device.Plan(specs, b["M"], dict(x_per_thread=4, shots_per_group=1)).close()  # 1024 threads
    for bad in (-1, 17):
        with pytest.raises(_lib.WlsqError):
            device.Plan(specs, b["M"], dict(image_mode=3, image_planes=bad))
    with pytest.raises(_lib.WlsqError):
        device.Plan(specs, b["M"], dict(image_mode=5))
    with pytest.raises(_lib.WlsqError):
        device.Plan(specs, b["M"], dict(rounding=2))
    with pytest.raises(_lib.WlsqError):  # the CPU path's rounding needs one shot per group
        device.Plan(specs, b["M"], dict(rounding=1, shots_per_group=2))
    with pytest.raises(_lib.WlsqError):
        device.Plan(specs, b["M"], dict(x_per_thread=8, shots_per_group=2))
    for k in ("nx", "ld", "img_ld", "slow_ld"):
        setattr(specs[0], k, 9000)
    with pytest.raises(_lib.WlsqError):  # 2250 strips of 4 x > 2 x 1024 threads
        device.Plan(specs, b["M"], dict(x_per_thread=4, shots_per_group=1))
    with pytest.raises(_lib.WlsqError):  # 9000-wide rows twice do not fit 160 KB of LDS
        device.Plan(specs, b["M"], dict(x_per_thread=8, shots_per_group=1, double_rows=1))
