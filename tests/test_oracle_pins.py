"""Pins of the oracle functions that the reference's compiled weights do not
cover: the chi0 contraction of calc_eps (closed forms and invariants), the
spline branch of the Sigma_c convolution (a fixture of PyGW's compiled
few_fr_convolutions), the k - q table (PyGW's compiled k_m_q_id) and the
tetrahedra of the k-mesh (tiling invariants of KQmesh.tetra)."""
import itertools
import os

import numpy as np
import pytest

from oracle import chain, ref
from pygw_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cplx(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


# --- chi0 (calc_eps, pygw.py:3161-3200) -------------------------------------

def test_chi0_rank_one_closed_form():
    """One k-point, one pair: eps[:, :, w] = -2 w(w) m m^H exactly (the
    conjugate on the second index: a transposed or unconjugated operand
    gives conj(m) m^T or m m^T instead)."""
    rng = np.random.default_rng(1)
    m = _cplx(rng, 1, 7)
    w = -rng.uniform(0.1, 2.0, (1, 5))
    eps = chain.calc_eps_accumulate([(0, m)], lambda ik: w)
    for iw in range(5):
        want = -2.0 * w[0, iw] * np.outer(m[0], np.conj(m[0]))
        np.testing.assert_allclose(eps[:, :, iw], want, rtol=1e-15, atol=1e-15)
    assert not np.allclose(eps[:, :, 0], -2.0 * w[0, 0] * np.outer(np.conj(m[0]), m[0]))


def test_chi0_linear_in_weights_and_k_additive():
    rng = np.random.default_rng(2)
    M = _cplx(rng, 3, 11, 6)
    w1, w2 = -rng.uniform(0, 1, (3, 11, 4)), -rng.uniform(0, 1, (3, 11, 4))
    a, b = 0.7, -1.9
    f = lambda w: chain.calc_eps_accumulate(((k, M[k]) for k in range(3)), lambda k: w[k])  # noqa: E731
    np.testing.assert_allclose(f(a * w1 + b * w2), a * f(w1) + b * f(w2), rtol=1e-13, atol=1e-13)
    # the k-sum is a plain sum: each k alone, added, equals all k at once
    parts = sum(chain.calc_eps_accumulate([(k, M[k])], lambda _k, k=k: w1[k]) for k in range(3))
    np.testing.assert_allclose(f(w1), parts, rtol=1e-13, atol=1e-13)


def test_chi0_hermitian_psd_and_trace():
    """w <= 0 with coef -2: every frequency's eps is Hermitian positive
    semidefinite, and tr eps(w) = -2 sum_{k,p} w_kp(w) |M_kp|^2."""
    rng = np.random.default_rng(3)
    nk, npair, pb, nom = 4, 9, 13, 3
    M = _cplx(rng, nk, npair, pb)
    w = -rng.uniform(0, 1, (nk, npair, nom))
    eps = chain.calc_eps_accumulate(((k, M[k]) for k in range(nk)), lambda k: w[k])
    for iw in range(nom):
        e = eps[:, :, iw]
        np.testing.assert_allclose(e, e.conj().T, atol=1e-13)
        assert np.linalg.eigvalsh(e).min() > -1e-12 * np.abs(e).max()
        tr = -2.0 * np.sum(w[:, :, iw] * (np.abs(M) ** 2).sum(axis=2))
        assert abs(np.trace(e) - tr) < 1e-12 * abs(tr)
    # positive weights give a negative semidefinite eps (the sign is the coef's)
    e = chain.calc_eps_accumulate(((k, M[k]) for k in range(nk)), lambda k: -w[k])[:, :, 0]
    assert np.linalg.eigvalsh(e).max() < 1e-12 * np.abs(e).max()


# --- Sigma_c convolution, spline branch (mcommon.py:126-144, fnc.f90:161-176)

def test_spline_convolution_matches_reference_fixture():
    """The oracle's spline (scipy CubicSpline, as PyGW calls it) and its
    restatement of few_fr_convolutions against PyGW's compiled routine
    (fixture from tests/golden/make_golden.py).

    Band 0 sits at E = 0, where Omega_max/(E - i w) lands on the branch cut
    of atan with a signed-zero real part: the oracle (numpy) and the GPU
    kernel divide as C does (Smith's algorithm, real part -0, atan real part
    -pi/2); PyGW's Fortran built with amdflang takes +pi/2.  The two differ
    there by exactly mwm(w_n) of that band (pi times mwm/pi) and agree to
    rounding everywhere else."""
    g = np.load(os.path.join(GOLD, "conv_spline.npz"))
    nb1, nom = g["sc"].shape
    assert g["enk"][0] == 0.0 and np.all(g["enk"][1:] != 0.0)
    for ie1 in range(nb1):
        for ie2 in range(g["enk"].size):
            mp = chain.spline_precise(g["mwm"][:, ie1, ie2], g["omega"], g["omega_p"])
            np.testing.assert_allclose(mp, g["mwm_p"][ie1, ie2], rtol=1e-14, atol=1e-16)
        for iom in range(nom):
            s = chain.few_fr_convolutions(g["enk"], g["mwm_p"][ie1], g["mwm"][iom, ie1, :], g["omega_p"],
                                          g["omega"][iom], g["womeg_p"])
            # on the cut only where |Omega_max/w| > 1 (not at w = Omega_max)
            on_cut = g["omega_p"][-1] / g["omega"][iom] > 1.0 + 1e-9
            jump = g["mwm"][iom, ie1, 0] if on_cut else 0.0
            assert abs(s + jump - g["sc"][ie1, iom]) <= 1e-13 * abs(g["sc"][ie1, iom])
    got = chain.selfc_spline(g["enk"], g["mwm"], g["omega"], g["omega_p"], g["womeg_p"])
    cut = (g["omega_p"][-1] / g["omega"] > 1.0 + 1e-9)[None, :]
    np.testing.assert_allclose(got + np.where(cut, g["mwm"][:, :, 0].T, 0.0), g["sc"], rtol=1e-13, atol=1e-15)


@pytest.mark.skipif(not ref.available("libref_fnc.so"), reason="PyGW's Fortran not built here")
def test_few_fr_convolutions_live_reference():
    rng = np.random.default_rng(7)
    fm = mesh.FrequencyMesh(10, 0.03, 10.0, iop_multiple=4)
    nb2 = 5
    mp = _cplx(rng, nb2, fm.omega_precise.size) / (1 + fm.omega_precise[None, :] ** 2)
    mi = _cplx(rng, nb2)
    enk = rng.uniform(-1, 1, nb2)
    for iom in (0, 4, 9):
        a = chain.few_fr_convolutions(enk, mp, mi, fm.omega_precise, fm.omega[iom], fm.womeg_precise)
        b = ref.few_fr_convolutions(enk, mp, mi, fm.omega_precise, fm.omega[iom], fm.womeg_precise)
        assert abs(a - b) <= 1e-13 * abs(b)


# --- k - q table (for_kpts.f90:21-45) and tetrahedra (kqmesh.py:188-291) ----

DIVS = [(4, 4, 4), (4, 2, 3), (3, 5, 2), (8, 8, 8), (1, 3, 2)]


@pytest.mark.skipif(not ref.available("libref_kpts.so"), reason="PyGW's Fortran not built here")
@pytest.mark.parametrize("div", DIVS)
def test_kqid_equals_compiled_k_m_q_id(div):
    assert np.array_equal(mesh.k_minus_q(div), ref.k_m_q_id(div))
    for k in itertools.product(*(range(d) for d in div)):
        assert int(mesh.k2indx(np.array(k), np.array(div))) == ref.k2indx(k, div)


@pytest.mark.parametrize("div", DIVS[:3])
def test_kqid_group_structure(div):
    """k - q is a group operation on the mesh: kqid[., 0] is the identity,
    each column a permutation, and (k - q1) - q2 = k - (q1 + q2)."""
    kq = mesh.k_minus_q(div)
    nk = kq.shape[0]
    assert np.array_equal(kq[:, 0], np.arange(nk))
    assert all(np.array_equal(np.sort(kq[:, q]), np.arange(nk)) for q in range(nk))
    neg = kq[0]                                   # index of -q
    for q1 in range(nk):
        for q2 in range(0, nk, 3):
            s = kq[q1, neg[q2]]                   # q1 + q2 = q1 - (-q2)
            assert np.array_equal(kq[kq[:, q1], q2], kq[:, s])


def _tet_volume6(c):
    return abs(np.linalg.det(np.array([c[1] - c[0], c[2] - c[0], c[3] - c[0]], dtype=float)))


@pytest.mark.parametrize("mnd", [0, 1, 2, 3])
def test_cube_tetrahedra_tile_the_cube(mnd):
    """The six tetrahedra of diagonal mnd each have volume 1/6, all contain
    both ends of that diagonal, and every point of the cube lies in exactly
    one of them (sampled)."""
    tet = mesh.cube_tetrahedra(mnd)
    p0 = np.array([[0, 0, 0, 0, 1, 1, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1], [0, 1, 0, 1, 0, 1, 0, 1]])
    a, b = p0[:, mnd], p0[:, 7 - mnd]
    for t in tet:
        assert _tet_volume6(t) == pytest.approx(1.0)
        assert any(np.array_equal(c, a) for c in t) and any(np.array_equal(c, b) for c in t)
    rng = np.random.default_rng(mnd)
    pts = rng.uniform(0, 1, (4000, 3))
    count = np.zeros(len(pts), dtype=int)
    for t in tet:
        T = np.array([t[1] - t[0], t[2] - t[0], t[3] - t[0]], dtype=float).T
        lam = np.linalg.solve(T, (pts - t[0]).T).T
        inside = (lam >= 0).all(1) & (lam.sum(1) <= 1)
        count += inside
    assert np.all(count == 1)


def test_main_diagonal_is_the_shortest():
    """A sheared reciprocal lattice makes one cube diagonal the shortest;
    main_diagonal picks it, and the tetrahedra share it."""
    p0 = np.array([[0, 0, 0, 0, 1, 1, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1], [0, 1, 0, 1, 0, 1, 0, 1]])
    div = np.array([4, 4, 4])
    for target in range(4):
        d = (p0[:, target] - p0[:, 7 - target]) / div
        # a lattice that shrinks the direction of the target diagonal
        u = d / np.linalg.norm(d)
        gbas = np.eye(3) - 0.5 * np.outer(u, u)
        lens = [np.sum((((p0[:, i] - p0[:, 7 - i]) / div) @ gbas) ** 2) for i in range(4)]
        assert mesh.main_diagonal(div, gbas) == int(np.argmin(lens)) == target


@pytest.mark.parametrize("div", DIVS[:3])
def test_tetc_covers_the_mesh(div):
    """tetc: 6 tetrahedra per cube, each k a corner of 24 tetrahedra, the
    corners of a tetrahedron are its cube's corners (offsets in {0,1}^3
    modulo the mesh), and the cube volumes add up to the whole zone."""
    tetc = mesh.tetra_corners(div)
    nk = int(np.prod(div))
    assert tetc.shape == (4, 6 * nk)
    assert np.array_equal(np.bincount(tetc.ravel(), minlength=nk), np.full(nk, 24))
    pts = np.array(list(itertools.product(*(range(d) for d in div))))
    for it in range(6 * nk):
        org = pts[it // 6]
        off = (pts[tetc[:, it]] - org) % np.array(div)
        off = np.where(np.array(div) == 1, 0, off)
        assert off.max() <= 1
        assert tetc[0, it] == it // 6                 # every tetrahedron starts at its cube's origin
