"""GPU parity of the Sigma_c part: MWM, spline, convolution, Pade/QP."""
import os

import numpy as np
import pytest

from oracle import chain
from pygw_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _herm_w(pb, nom, rng):
    x = rng.normal(size=(pb, pb, nom)) + 1j * rng.normal(size=(pb, pb, nom))
    w = (x + np.conj(x.transpose(1, 0, 2))) / 2
    return np.asfortranarray(w)


@pytest.mark.parametrize("nk,nb1,nb2,pb,nom", [(2, 3, 7, 70, 3), (3, 2, 33, 129, 2)])
def test_mwm_vs_oracle(nk, nb1, nb2, pb, nom):
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(nk * 100 + pb)
    M = rng.normal(size=(nk, nb1 * nb2, pb)) + 1j * rng.normal(size=(nk, nb1 * nb2, pb))
    W = _herm_w(pb, nom, rng)
    print(f"case mwm nk={nk} rows={nb1*nb2} pb={pb}", flush=True)
    Wt = torch.as_tensor(W.reshape(-1, order="F"), device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    got = P.calc_mwm(M, Wt, 0.125).cpu().numpy()
    for k in range(nk):
        want = chain.calc_mwm(M[k], W, 0.125)
        assert _rel(got[k], want) < 1e-13


def test_convolution_vs_reference_fixture():
    from pygw_amd import polarization as P
    g = np.load(os.path.join(GOLD, "conv.npz"))
    mwm = g["mwm"]  # [nom, nb1, nb2]
    print("case conv fixture", flush=True)
    sc = P.compute_selfc_inside(g["enk"][None, :], mwm[None], g["omega"], g["womeg"]).cpu().numpy()[0]
    np.testing.assert_allclose(sc, g["sc"], rtol=1e-12, atol=1e-14)


def test_spline_convolution_vs_oracle():
    """Spline branch (precise mesh) against scipy CubicSpline + the
    convolution formula (oracle.chain.selfc_spline)."""
    from pygw_amd import polarization as P
    rng = np.random.default_rng(8)
    fm = mesh.FrequencyMesh(16, 0.02, 20.0, iop_multiple=5)
    om, wo, omp, wop = fm.omega, fm.womeg, fm.omega_precise, fm.womeg_precise
    nk, nb1, nb2 = 2, 2, 6
    a = rng.uniform(0.3, 1.5, (nk, nb1, nb2))
    mwm = np.empty((nk, om.size, nb1, nb2), dtype=complex)
    for k in range(nk):
        mwm[k] = (1.0 / (om[:, None, None] ** 2 + a[k][None] ** 2)) * (1 + 0.1j * rng.normal(size=(1, nb1, nb2)))
    enk = rng.uniform(-0.6, 0.6, (nk, nb2))
    print("case spline conv", flush=True)
    got = P.compute_selfc_inside(enk, mwm, om, wo, omp, wop).cpu().numpy()
    for k in range(nk):
        want = chain.selfc_spline(enk[k], mwm[k], om, omp, wop)
        np.testing.assert_allclose(got[k], want, rtol=1e-11, atol=1e-13)


def test_spline_convolution_real_mwm_and_many_rows():
    """The spline + convolution where the imaginary part is not splined (mwm
    real: PyGW's sum|Im|/n <= 1e-10 rule) and with more rows per k than a
    block of the convolution has waves (nb1 = 11 > 8), against the oracle."""
    from pygw_amd import polarization as P
    rng = np.random.default_rng(9)
    fm = mesh.FrequencyMesh(12, 0.02, 20.0, iop_multiple=5)
    om, wo, omp, wop = fm.omega, fm.womeg, fm.omega_precise, fm.womeg_precise
    nk, nb1, nb2 = 2, 11, 5
    a = rng.uniform(0.3, 1.5, (nk, nb1, nb2))
    mwm = (1.0 / (om[None, :, None, None] ** 2 + a[:, None] ** 2)).astype(complex)
    mwm[1] *= 1 + 0.2j
    enk = rng.uniform(-0.6, 0.6, (nk, nb2))
    print("case spline conv real / 11 rows", flush=True)
    got = P.compute_selfc_inside(enk, mwm, om, wo, omp, wop).cpu().numpy()
    for k in range(nk):
        want = chain.selfc_spline(enk[k], mwm[k], om, omp, wop)
        np.testing.assert_allclose(got[k], want, rtol=1e-11, atol=1e-13)


def test_pade_qp_vs_reference_fixture():
    from pygw_amd import polarization as P
    g = np.load(os.path.join(GOLD, "pade.npz"))
    z, y = g["z"], g["y"]
    om = z.imag
    ene = np.array([0.1, -0.3])
    print("case pade", flush=True)
    out = P.pade_quasiparticles(np.vstack([y, y]), om, ene, sigx=np.array([0.2, 0.0]), vxc=np.array([0.05, 0.1]),
                                npade=z.size).cpu().numpy()
    for s, e in enumerate(ene):
        yc, dyc = chain.analytic_continuation(om, y, e, npts=z.size)
        np.testing.assert_allclose(out[s, 0] + 1j * out[s, 1], yc, rtol=1e-9)
        np.testing.assert_allclose(out[s, 2] + 1j * out[s, 3], dyc, rtol=1e-7)
        zz = 1 / (1 - dyc.real)
        zz = 1.0 if (zz > 1 or zz < 0) else zz
        want = e + zz * (yc.real + [0.2, 0.0][s] - [0.05, 0.1][s])
        np.testing.assert_allclose(out[s, 5], want, rtol=1e-9)


def test_mwm_mixed_vs_fp64():
    """Split-precision MWM against the FP64 oracle: relative error per
    frequency <= 1e-6 (M rounded to complex64 on both sides)."""
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(77)
    nk, nb1, nb2, pb, nom = 2, 4, 9, 133, 3
    M = (rng.normal(size=(nk, nb1 * nb2, pb)) + 1j * rng.normal(size=(nk, nb1 * nb2, pb))).astype(np.complex64)
    W = _herm_w(pb, nom, rng)
    print("case mwm mixed", flush=True)
    Wt = torch.as_tensor(W.reshape(-1, order="F"), device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    got = P.calc_mwm(M, Wt, 0.5, precision="mixed").cpu().numpy()
    for k in range(nk):
        want = chain.calc_mwm(M[k].astype(complex), W, 0.5)
        for w in range(nom):
            err = _rel(got[k, w], want[w])
            print(f"  k={k} w={w} rel={err:.2e}", flush=True)
            assert err < 1e-6


@pytest.mark.parametrize("shape", [(3, 5, 30, 129, 3), (1, 8, 17, 260, 2)])
def test_mwm_mixed_slabs_and_ragged(shape):
    """Split-precision MWM with rows and PB off the 128 tile and the
    frequencies walked in slabs (a workspace one frequency short): same
    result as one slab, both within 1e-6 of the FP64 oracle."""
    import ctypes as C
    import torch
    from pygw_amd import _lib
    from pygw_amd import polarization as P
    nk, nb1, nb2, pb, nom = shape
    rng = np.random.default_rng(91 + pb)
    M = (rng.normal(size=(nk, nb1 * nb2, pb)) + 1j * rng.normal(size=(nk, nb1 * nb2, pb))).astype(np.complex64)
    W = _herm_w(pb, nom, rng)
    print(f"case mwm mixed slabs {shape}", flush=True)
    Wt = torch.as_tensor(W.reshape(-1, order="F"), device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    full = P.calc_mwm(M, Wt, 0.25, precision="mixed").cpu().numpy()
    sz = C.c_size_t(0)
    _lib.call("pygw_mwm_workspace", nk, nb1 * nb2, pb, nom, 1, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    Md = torch.as_tensor(M, device="cuda")
    out = torch.empty((nk, nom, nb1 * nb2), dtype=torch.complex128, device="cuda")
    _lib.call("pygw_mwm", C.c_void_p(Md.data_ptr()), nk, nb1 * nb2, pb, C.c_void_p(Wt.data_ptr()), nom, 0.25, 1,
              C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()), sz.value - 1,
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    np.testing.assert_array_equal(out.cpu().numpy(), full)
    for k in range(nk):
        want = chain.calc_mwm(M[k].astype(complex), W, 0.25)
        for w in range(nom):
            assert _rel(full[k, w], want[w]) < 1e-6


def test_spline_convolution_vs_reference_fixture():
    """The spline branch against PyGW's compiled few_fr_convolutions
    (tests/golden/conv_spline.npz).  At E = 0 (band 0) the atan argument
    lies on its branch cut; the kernel divides as C does and takes the
    -pi/2 side, the amdflang build of PyGW the +pi/2 side: there the two
    differ by exactly mwm(w_n) (checked on the oracle in test_oracle_pins).
    The last frequency is left out: there w_n = Omega_max, so for the E = 0
    band the argument is i (1 + rounding), at the logarithmic singularity of
    atan, where no two implementations agree."""
    from pygw_amd import polarization as P
    g = np.load(os.path.join(GOLD, "conv_spline.npz"))
    print("case spline conv fixture", flush=True)
    got = P.compute_selfc_inside(g["enk"][None, :], g["mwm"][None], g["omega"], g["womeg"], g["omega_p"],
                                 g["womeg_p"]).cpu().numpy()[0]
    cut = (g["omega_p"][-1] / g["omega"] > 1.0 + 1e-9)[None, :]
    want = g["sc"] - np.where(cut, g["mwm"][:, :, 0].T, 0.0)
    np.testing.assert_allclose(got[:, :-1], want[:, :-1], rtol=1e-11, atol=1e-13)


def test_padecof_padeg_vs_reference_fixture():
    """The standalone Pade pair against PyGW's compiled padecof/padeg
    (tests/golden/pade.npz: a five-pole function on 10 imaginary nodes,
    continued to 9 real points), and a batch of functions at once."""
    from pygw_amd import polarization as P
    g = np.load(os.path.join(GOLD, "pade.npz"))
    print("case padecof/padeg fixture", flush=True)
    # the reciprocal differences amplify the rounding of each complex
    # division (the kernel's and the Fortran's differ in the last bit): the
    # bar is that of pygw_pade_qp against the same fixture
    coef = P.padecof(g["y"], g["z"]).cpu().numpy()
    np.testing.assert_allclose(coef, g["apar"], rtol=1e-9, atol=1e-12)
    gx = P.padeg(g["x"], g["z"], coef).cpu().numpy()
    np.testing.assert_allclose(gx, g["g"], rtol=1e-9, atol=1e-12)
    # three functions in one call, each equal to its own single call
    ys = np.stack([g["y"], 2.0 * g["y"], g["y"] + 0.1j])
    cs = P.padecof(ys, g["z"]).cpu().numpy()
    gs = P.padeg(g["x"], g["z"], cs).cpu().numpy()
    for i in range(3):
        c1 = P.padecof(ys[i], g["z"]).cpu().numpy()
        assert np.array_equal(cs[i], c1)
        assert np.array_equal(gs[i], P.padeg(g["x"], g["z"], c1).cpu().numpy())
