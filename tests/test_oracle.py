"""Pins the CPU oracle before anything is checked against it (CPU only).

Every oracle function is checked here against something that does not
restate it:
* the reference's own known-answer tests (tests/kats.py cites them),
* brute force (pure-Python loops of the definition) on tiny random shapes,
* invariants the mathematics fixes: identity and shift operators, linearity,
  forward/backward symmetry of the image, zero operator, untouched rows,
  shot additivity,
* the kappa lookup against an FP64 argmin over the explicit kappa grid
  (py_src/operatorsPreparation/util.py:5-8, util.py:39-45), with ties.
Planted mistakes (a conjugated dot, a dropped halo zero, a transposed table)
are shown to fail these checks, on the CPU.
"""
import numpy as np
import pytest

import oracle
from oracle import fp64
import kats


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def _run_c(case):
    ns, nt, nx = case["ns"], case["nt"], case["nx"]
    F = np.ascontiguousarray(case["F"].astype(np.complex64))
    B = np.ascontiguousarray(case["B"].astype(np.complex64))
    img = np.zeros((ns, case["nz"], nx), np.float32)
    oracle.extrapolate(ns, case["nextrap"], case["nz"], nx, case["nf"], nt, case["M"],
                       np.ascontiguousarray(case["opF"], np.complex64),
                       np.ascontiguousarray(case["opB"], np.complex64), F, B, img)
    return img, F, B


@pytest.mark.parametrize("case", kats.all_kats(), ids=lambda c: c["name"])
def test_reference_kats_c_oracle(case):
    img, F, B = _run_c(case)
    kats.check_case(case, img[0], F[0, :case["nf"]], B[0, :case["nf"]])


@pytest.mark.parametrize("case", kats.all_kats(), ids=lambda c: c["name"])
def test_reference_kats_fp64_oracle(case):
    img, F, B = fp64.extrapolate(case["nextrap"], case["nz"], case["nx"], case["nf"], case["M"],
                                 case["opF"], case["opB"], case["F"][0], case["B"][0])
    kats.check_case(case, img, F, B)


@pytest.mark.parametrize("case", kats.python_float_kats(), ids=lambda c: c["name"])
def test_reference_float_kats(case):
    img, _, _ = fp64.extrapolate(case["nextrap"], case["nz"], case["nx"], case["nf"], case["M"],
                                 case["opF"], case["opB"], case["F"], case["B"])
    kats.check_case(case, img)
    c = dict(case, ns=1, nt=case["nf"], F=case["F"][None].astype(np.complex64),
             B=case["B"][None].astype(np.complex64),
             opF=case["opF"].astype(np.complex64), opB=case["opB"].astype(np.complex64))
    img_c, _, _ = _run_c(c)
    kats.check_case(case, img_c[0])


def _rand_case(rng, ns, nextrap, nz, nf, nt, nx, M, scale=0.3):
    L = 2 * M + 1
    def c(shape):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    return dict(name="rand", ns=ns, nextrap=nextrap, nz=nz, nf=nf, nt=nt, nx=nx, M=M,
                opF=(scale * c((nextrap, nf, nx, L))).astype(np.complex64),
                opB=(scale * c((nextrap, nf, nx, L))).astype(np.complex64),
                F=c((ns, nt, nx)), B=c((ns, nt, nx)))


def test_brute_force_agrees_with_both_oracles():
    rng = np.random.default_rng(11)
    for (nextrap, nf, nx, M) in [(3, 3, 7, 2), (2, 4, 5, 0), (4, 2, 3, 3), (1, 1, 1, 1)]:
        case = _rand_case(rng, 1, nextrap, nextrap + 1, nf, nf + 1, nx, M)
        bi, bF, bB = fp64.brute_force(nextrap, nextrap + 1, nx, nf, M, case["opF"], case["opB"],
                                      case["F"][0], case["B"][0])
        vi, vF, vB = fp64.extrapolate(nextrap, nextrap + 1, nx, nf, M, case["opF"], case["opB"],
                                      case["F"][0], case["B"][0])
        np.testing.assert_allclose(vi, bi, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(vF, bF, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(vB, bB, rtol=1e-12, atol=1e-12)
        ci, cF, cB = _run_c(case)
        np.testing.assert_allclose(ci[0], bi, rtol=1e-5, atol=1e-5 * np.abs(bi).max() + 1e-6)
        np.testing.assert_allclose(cF[0, :nf], bF, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(cB[0, :nf], bB, rtol=1e-5, atol=1e-5)
        # rows j >= nf of the caller's buffers are not touched
        np.testing.assert_array_equal(cF[0, nf:], case["F"][0, nf:])


def _planted_conjugated_dot(case):
    """A plausible mistake: conjugating the operator in the dot."""
    return fp64.extrapolate(case["nextrap"], case["nz"], case["nx"], case["nf"], case["M"],
                            np.conj(case["opF"]), np.conj(case["opB"]), case["F"][0], case["B"][0])


def _planted_transposed_table(case):
    """A plausible mistake: reading the table tap-major as if natural."""
    nextrap, nf, nx, L = case["opF"].shape
    t = lambda op: np.ascontiguousarray(op).reshape(nextrap, nf, L, nx).transpose(0, 1, 3, 2)
    return fp64.extrapolate(case["nextrap"], case["nz"], case["nx"], case["nf"], case["M"],
                            t(case["opF"]), t(case["opB"]), case["F"][0], case["B"][0])


def _planted_periodic_halo(case):
    """A plausible mistake: wrapping around instead of zero halos."""
    M, nx, nf = case["M"], case["nx"], case["nf"]
    L = 2 * M + 1
    F = case["F"][0][:nf].astype(np.complex128)
    B = case["B"][0][:nf].astype(np.complex128)
    img = np.zeros((case["nz"], nx))
    for l in range(case["nextrap"]):
        nF = np.zeros_like(F)
        nB = np.zeros_like(B)
        for k in range(L):
            nF += case["opF"][l][:, :, k] * np.roll(F, M - k, axis=1)
            nB += case["opB"][l][:, :, k] * np.roll(B, M - k, axis=1)
        F, B = nF, nB
        img[l] = np.real(np.sum(F * np.conj(B), axis=0))
    return img, F, B


@pytest.mark.parametrize("planted", [_planted_conjugated_dot, _planted_transposed_table,
                                     _planted_periodic_halo])
def test_planted_mistakes_fail_the_kats(planted):
    failed = 0
    for case in kats.all_kats():
        img, F, B = planted(case)
        try:
            kats.check_case(case, img, F, B)
        except AssertionError:
            failed += 1
    rng = np.random.default_rng(5)
    case = _rand_case(rng, 1, 2, 3, 3, 3, 6, 2)
    bi, _, _ = fp64.brute_force(2, 3, 6, 3, 2, case["opF"], case["opB"], case["F"][0], case["B"][0])
    pi, _, _ = planted(case)
    assert failed > 0 or not np.allclose(pi, bi)


def test_identity_and_shift_operators():
    rng = np.random.default_rng(3)
    case = _rand_case(rng, 2, 3, 4, 3, 5, 9, 2)
    M = case["M"]
    for tap, shift in [(M, 0), (M + 1, 1), (M - 2, -2)]:
        op = np.zeros_like(case["opF"])
        op[..., tap] = 1.0
        c = dict(case, opF=op, opB=op.copy())
        img, F, B = _run_c(c)
        F0 = case["F"][:, :3]
        exp = np.zeros_like(F0)
        n = case["nextrap"] * shift
        nx = case["nx"]
        for i in range(nx):
            if 0 <= i + n < nx:
                exp[:, :, i] = F0[:, :, i + n]
        np.testing.assert_array_equal(F[:, :3], exp)
        if shift == 0:
            B0 = case["B"][:, :3]
            for l in range(3):
                np.testing.assert_allclose(img[:, l], np.real(np.sum(F0 * np.conj(B0), axis=1)),
                                           rtol=1e-6, atol=1e-5)


def test_linearity_symmetry_zero_and_untouched_rows():
    rng = np.random.default_rng(4)
    case = _rand_case(rng, 1, 3, 6, 4, 4, 10, 3)
    img, F, B = _run_c(case)
    # image is linear in each field: scale F by 2 (exact in binary FP)
    c2 = dict(case, F=(2 * case["F"]).astype(np.complex64))
    img2, F2, _ = _run_c(c2)
    np.testing.assert_array_equal(img2, 2 * img)
    np.testing.assert_array_equal(F2, 2 * F)
    # swapping the (operator, field) pairs leaves Re(F conj B) unchanged
    cs = dict(case, opF=case["opB"], opB=case["opF"], F=case["B"], B=case["F"])
    imgs, _, _ = _run_c(cs)
    np.testing.assert_allclose(imgs, img, rtol=1e-6, atol=1e-5 * np.abs(img).max())
    # zero operator -> zero fields and image
    cz = dict(case, opF=np.zeros_like(case["opF"]), opB=np.zeros_like(case["opB"]))
    imgz, Fz, _ = _run_c(cz)
    assert not imgz.any() and not Fz[:, :4].any()
    # rows l >= nextrap are never written
    ns, nz, nx = 1, case["nz"], case["nx"]
    sentinel = np.full((ns, nz, nx), 7.25, np.float32)
    Fc = case["F"].copy(); Bc = case["B"].copy()
    oracle.extrapolate(ns, 3, nz, nx, 4, 4, 3, case["opF"], case["opB"], Fc, Bc, sentinel)
    assert (sentinel[:, 3:] == 7.25).all()
    np.testing.assert_array_equal(sentinel[:, :3], img[:, :3])


def test_shot_additivity():
    """Each shot of a multi-shot call equals the single-shot call (bitwise)."""
    rng = np.random.default_rng(6)
    case = _rand_case(rng, 3, 4, 5, 3, 6, 12, 2)
    img, F, B = _run_c(case)
    for s in range(3):
        c1 = dict(case, ns=1, F=case["F"][s:s + 1], B=case["B"][s:s + 1])
        i1, F1, B1 = _run_c(c1)
        np.testing.assert_array_equal(i1[0], img[s])
        np.testing.assert_array_equal(F1[0], F[s])
        np.testing.assert_array_equal(B1[0], B[s])


def test_c_oracle_against_fp64_random():
    rng = np.random.default_rng(7)
    case = _rand_case(rng, 2, 6, 8, 5, 7, 33, 4, scale=0.25)
    img, F, B = _run_c(case)
    ref, rF, rB = fp64.extrapolate_shots(2, 6, 8, 33, 5, 4, case["opF"], case["opB"], case["F"], case["B"])
    rel = np.linalg.norm(img - ref) / np.linalg.norm(ref)
    assert rel < 1e-5, rel
    assert np.linalg.norm(F[:, :5] - rF) / np.linalg.norm(rF) < 1e-5


def test_argument_validation():
    case = kats.cpu_kats()[0]
    F = case["F"].copy(); B = case["B"].copy()
    img = np.zeros((1, 1, 4), np.float32)
    with pytest.raises(ValueError):   # nz < nextrap
        oracle.extrapolate(1, 2, 1, 4, 2, 2, 1, case["opF"], case["opB"], F, B, img)
    with pytest.raises(ValueError):   # nt < nf
        oracle.extrapolate(1, 1, 1, 4, 2, 1, 1, case["opF"], case["opB"], F, B, img)


# ---------------------------------------------------------------- kappa lookup

def test_kappa_index_matches_fp64_argmin_and_numpy_f32():
    """q = argmin |kappa - omega*s| over kappa_q = q*dk (createKappa starts at 0,
    util.py:39-45); numpy argmin returns the first minimum (util.py:5-8)."""
    rng = np.random.default_rng(8)
    nk = 777
    dk = np.float32(2.7e-4)
    kappa = np.arange(nk, dtype=np.float64) * float(dk)
    om = rng.uniform(2 * np.pi * 2, 2 * np.pi * 50, 4000).astype(np.float32)
    s = (1.0 / rng.uniform(1500, 4500, 4000)).astype(np.float32)
    q_np = fp64.kappa_index_f32(om, s, 0.0, dk, nk)
    q_c = np.array([oracle.kappa_index(o, v, 0.0, dk, nk) for o, v in zip(om, s)])
    np.testing.assert_array_equal(q_np, q_c)
    target = np.float64(om) * np.float64(s)
    q_ref = np.array([np.argmin(np.abs(kappa - t)) for t in target])
    u = target / float(dk)
    away = np.abs(u - np.floor(u) - 0.5) > 1e-3
    assert away.mean() > 0.99
    np.testing.assert_array_equal(q_np[away], q_ref[away])
    # out-of-range targets clamp to the ends, as argmin does
    assert oracle.kappa_index(1e9, 1.0, 0.0, dk, nk) == nk - 1
    assert oracle.kappa_index(-5.0, 1.0, 0.0, dk, nk) == 0


def test_kappa_index_ties_go_to_lower_index():
    # omega*s lands exactly half-way between kappa_2 and kappa_3 (dk = 1)
    assert oracle.kappa_index(2.5, 1.0, 0.0, 1.0, 10) == 2
    assert int(fp64.kappa_index_f32(2.5, 1.0, 0.0, 1.0, 10)) == 2
    assert np.argmin(np.abs(np.arange(10.0) - 2.5)) == 2
    assert oracle.kappa_index(2.5000002, 1.0, 0.0, 1.0, 10) == 3
    assert oracle.kappa_index(7.0, 0.5, 0.5, 0.25, 100) == 12  # (3.5-0.5)/0.25


def _ktable_case(rng, ns=2, nextrap=5, nf=6, nx=17, M=3, nk=64):
    L = 2 * M + 1
    slow = (1.0 / rng.uniform(1500, 3000, (nextrap + 1, nx))).astype(np.float32)
    omega = np.linspace(2 * np.pi * 3, 2 * np.pi * 40, nf).astype(np.float32)
    dk = np.float32((omega.max() * slow.max()) / (nk - 3))
    tf = (0.3 * (rng.standard_normal((nk, L)) + 1j * rng.standard_normal((nk, L)))).astype(np.complex64)
    tb = np.conj(tf).astype(np.complex64)
    F = (rng.standard_normal((ns, nf + 2, nx)) + 1j * rng.standard_normal((ns, nf + 2, nx))).astype(np.complex64)
    B = (rng.standard_normal((ns, nf + 2, nx)) + 1j * rng.standard_normal((ns, nf + 2, nx))).astype(np.complex64)
    return dict(ns=ns, nextrap=nextrap, nz=nextrap + 1, nf=nf, nt=nf + 2, nx=nx, M=M, nk=nk,
                slow=slow, omega=omega, dk=dk, tf=tf, tb=tb, F=F, B=B)


def test_ktable_equals_expanded_table_bitwise():
    rng = np.random.default_rng(9)
    k = _ktable_case(rng)
    ef = oracle.expand_table(k["nextrap"], k["nf"], k["nx"], k["M"], k["slow"], k["omega"], 0.0,
                             k["dk"], k["nk"], k["tf"])
    eb = oracle.expand_table(k["nextrap"], k["nf"], k["nx"], k["M"], k["slow"], k["omega"], 0.0,
                             k["dk"], k["nk"], k["tb"])
    # the expansion is the argmin gather, checked against numpy fancy indexing
    q = fp64.kappa_index_f32(k["omega"][None, :, None], k["slow"][:k["nextrap"], None, :], 0.0,
                             k["dk"], k["nk"])
    np.testing.assert_array_equal(ef, k["tf"][q])
    F1, B1 = k["F"].copy(), k["B"].copy()
    img1 = np.zeros((k["ns"], k["nz"], k["nx"]), np.float32)
    oracle.extrapolate(k["ns"], k["nextrap"], k["nz"], k["nx"], k["nf"], k["nt"], k["M"], ef, eb, F1, B1, img1)
    F2, B2 = k["F"].copy(), k["B"].copy()
    img2 = np.zeros_like(img1)
    oracle.extrapolate_ktable(k["ns"], k["nextrap"], k["nz"], k["nx"], k["nf"], k["nt"], k["M"],
                              k["slow"], k["omega"], 0.0, k["dk"], k["nk"], k["tf"], k["tb"], F2, B2, img2)
    np.testing.assert_array_equal(img1, img2)
    np.testing.assert_array_equal(F1, F2)
    np.testing.assert_array_equal(B1, B2)
    # and the FP64 oracle's own k-table gather agrees with the C oracle
    ref, _, _ = fp64.extrapolate_shots(k["ns"], k["nextrap"], k["nz"], k["nx"], k["nf"], k["M"],
                                       wave_fs=k["F"], shot_fs=k["B"],
                                       ktable=dict(slowness=k["slow"], omega=k["omega"], kappa0=0.0,
                                                   dkappa=k["dk"], nk=k["nk"], forw=k["tf"], back=k["tb"]))
    assert np.linalg.norm(img2 - ref) / np.linalg.norm(ref) < 1e-5
