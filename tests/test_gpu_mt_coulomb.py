"""GPU parity of the muffin-tin Coulomb block and its q = 0 term
(mt_coulomb.hip through pygw_amd.coulomb) against the oracle
(oracle/coulomb.py): the Ewald structure constants, MT_Coulomb, the radial
overlaps of Calc_Olap_PW_Product (spherical Bessel functions by Miller's
recurrence on the device, scipy's in the oracle) and Zero_Momentum_Coulomb.
At PyGW's own size of the structure constants (lam_max = 4 (lmbmax + 1) =
16) and a cell of 8 atoms the Ewald sum is checked through its independence
of the splitting parameter eta."""
import numpy as np
import pytest

from oracle import coulomb as O
from pygw_amd import coulomb as C
from pygw_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(got, want):
    return np.abs(np.asarray(got) - want).max() / np.abs(want).max()


def _ew_gpu(s, pq, lam_max, eta=None, rcut=7.0, gcut=14.0):
    eta = s.eta if eta is None else eta
    return C.ewald_summation(pq, lam_max, s.pos, s.br2, s.rbas, s.alat, s.vol, s.ortho, rcut * eta, gcut / eta,
                             eta).cpu().numpy()


@pytest.mark.parametrize("ortho", [False, True])
@pytest.mark.parametrize("q0", [False, True])
@pytest.mark.parametrize("lam_max", [6, 16])
def test_ewald_matches_oracle(ortho, q0, lam_max):
    s = synth.mt_system(ortho=ortho)
    pq = np.zeros(3) if q0 else s.pq
    print(f"ewald ortho={ortho} q0={q0} lam_max={lam_max}", flush=True)
    got = _ew_gpu(s, pq, lam_max)
    want = O.ewald_summation(pq, lam_max, s.pos, s.br2, s.rbas, s.alat, s.vol, s.ortho, 7 * s.eta, 14 / s.eta, s.eta)
    assert got.shape == want.shape
    # per l: the constants fall by orders of magnitude with l
    for l in range(lam_max + 1):
        blk = slice(l * l, (l + 1) ** 2)
        assert np.abs(got[blk] - want[blk]).max() <= TOL * max(np.abs(want[blk]).max(), 1e-300)


def _cell8(seed=3):
    """A cell of 8 atoms (4 types, mult 2 each) in a monoclinic lattice."""
    rng = np.random.default_rng(seed)
    s = synth.mt_system()
    rbas = np.array([[7.9, 0.0, 0.0], [1.3, 8.4, 0.0], [0.0, 0.0, 10.6]])
    s.rbas, s.br2, s.vol = rbas, 2 * np.pi * np.linalg.inv(rbas), abs(np.linalg.det(rbas))
    s.alat = np.array([7.9, 8.5, 10.6])
    s.pos = rng.uniform(0, 1, (3, 8))
    return s


def test_ewald_full_size_is_independent_of_eta():
    """PyGW's lam_max 16 over 36 atom pairs: the sum at q != 0 agrees between
    two splittings eta, l by l."""
    s = _cell8()
    print("ewald 8 atoms lam_max 16", flush=True)
    a, b = _ew_gpu(s, s.pq, 16, eta=1.9), _ew_gpu(s, s.pq, 16, eta=2.7)
    assert a.shape == (289, 36)
    for l in range(17):
        blk = slice(l * l, (l + 1) ** 2)
        assert np.abs(a[blk] - b[blk]).max() <= 1e-9 * np.abs(a[blk]).max()


def _mt_args(s, vq):
    sgm = O.ewald_summation(vq, s.lam_max, s.pos, s.br2, s.rbas, s.alat, s.vol, s.ortho, 7 * s.eta, 14 / s.eta,
                            s.eta)
    return sgm, (s.loctmatsize, s.big_l_f, s.nmix, s.im_start, s.tilg, s.rrint, s.djmm, s.rtlij, sgm, s.mult)


@pytest.mark.parametrize("q0", [False, True])
@pytest.mark.parametrize("aligned", [False, True])
def test_mt_coulomb_matches_oracle(q0, aligned):
    s = synth.mt_system(identity_rotations=aligned)
    vq = np.zeros(3) if q0 else s.pq
    sgm, args = _mt_args(s, vq)
    print(f"mt_coulomb q0={q0} aligned={aligned}", flush=True)
    total = 0
    for iat in range(2):
        want = O.mt_coulomb(vq, iat, *args)
        got = C.mt_coulomb(vq, iat, *args, s.lam_max).cpu().numpy()
        assert _rel(got, want) < TOL
        total = total + want
    got = C.mt_coulomb(vq, None, *args, s.lam_max).cpu().numpy()
    assert _rel(got, total) < TOL
    if q0:
        l0 = np.array([l == 0 for iat in range(2) for _ in range(s.mult[iat]) for l in s.big_l[iat]
                       for _ in range(2 * l + 1)])
        assert np.all(got[np.ix_(l0, l0)] == 0)
    if aligned:
        assert np.abs(got - got.conj().T).max() < 1e-10 * np.abs(got).max()


def test_mt_coulomb_with_device_structure_constants():
    """The chain as PyGW runs it: sgm from the device Ewald sum fed to the
    device MT_Coulomb, summed over the atom types."""
    s = synth.mt_system()
    sgm, args = _mt_args(s, s.pq)
    want = sum(O.mt_coulomb(s.pq, iat, *args) for iat in range(2))
    sg = C.ewald_summation(s.pq, s.lam_max, s.pos, s.br2, s.rbas, s.alat, s.vol, s.ortho, 7 * s.eta, 14 / s.eta,
                           s.eta)
    got = C.mt_coulomb(s.pq, None, *args[:8], sg, s.mult, s.lam_max).cpu().numpy()
    assert _rel(got, want) < 1e-11


@pytest.mark.parametrize("iat", [0, 1])
def test_calc_olap_pw_product_matches_oracle(iat):
    """|q + G| from 1e-6 (the power-series branch) to 30 (k r up to 66)."""
    s = synth.mt_system()
    r0, dh, npt = s.meshes[iat]
    rx = O.radial_mesh(r0, dh, npt)
    gq = np.r_[1e-6, 3e-3, np.geomspace(0.02, 30.0, 57)]
    lmax = int(max(max(b) for b in s.big_l))
    want = O.calc_olap_pw_product(gq, lmax, rx, dh, s.rel, s.big_l[iat], s.ul[iat], s.us[iat])
    got = C.calc_olap_pw_product(gq, lmax, rx, dh, s.rel, s.big_l[iat], s.ul[iat], s.us[iat]).cpu().numpy()
    for irm in range(len(s.big_l[iat])):
        assert np.abs(got[irm] - want[irm]).max() <= TOL * np.abs(want[irm]).max()


@pytest.mark.parametrize("npt", [781, 640])
@pytest.mark.parametrize("L", [0, 1, 4, 9])
def test_radial_olap_bessel_of_high_order(L, npt):
    """A radial function r^(L+1) e^{-r} against j_L for L up to 9: the
    relative error of each overlap, the oracle's scipy j_L as reference."""
    r0, dh = 1e-4, np.log(2.4e4) / (npt - 1)
    rx = O.radial_mesh(r0, dh, npt)
    u = (rx ** (L + 1) * np.exp(-rx))[None, :]
    gq = np.geomspace(1e-4, 40.0, 64)
    want = O.calc_olap_pw_product(gq, L, rx, dh, False, [L], u, 0 * u)
    got = C.calc_olap_pw_product(gq, L, rx, dh, False, [L], u, 0 * u).cpu().numpy()
    big = np.abs(want) > 1e-250
    assert np.all(np.abs(got - want)[big] <= 1e-11 * np.abs(want)[big] + 1e-15 * np.abs(want).max())


def test_zero_momentum_matches_oracle():
    s = synth.mt_system()
    mb = s.loctmatsize + 11
    args = (s.glen, s.gindex, s.pos, s.vol, mb, s.rel, s.meshes, s.big_l, s.ul, s.us, s.atm, s.iprod)
    want = O.zero_momentum_coulomb(*args)
    got = C.zero_momentum_coulomb(*args).cpu().numpy()
    assert _rel(got, want) < TOL
    nz = np.zeros(mb, bool)
    nz[list(s.iprod.values())] = True
    assert np.all(got[~nz, :] == 0) and np.all(got[:, ~nz] == 0)
