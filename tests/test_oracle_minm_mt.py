"""The oracle of the muffin-tin and core matrix elements (oracle/minm_mt.py:
the Gaunt table, calc_minm_MT, calc_minc, calc_minc2) against PyGW's compiled
gaunt.f90 and for_vxcnn.f90 (oracle/_ref, where built) and against an
independent form of the same integrals: every angular integral
int Y* Y Y dOmega by Gauss-Legendre x trapezoid quadrature of scipy's Y_lm,
and the radial channels (u, udot, local orbitals) summed densely, without
the table, the triangle and m selection rules or the reference's loop order.
(calc_minm_MT itself calls BLAS zgemm, which this image lacks, so it is pinned
by the second form only.)"""
import numpy as np
import pytest

from oracle import minm_mt as O
from oracle import ref
from pygw_amd import synth

need_gaunt = pytest.mark.skipif(not ref.available("libref_gaunt.so"), reason="oracle/_ref not built")
need_vx = pytest.mark.skipif(not ref.available("libref_vxcnn.so"), reason="oracle/_ref not built")


def _quad(lmax):
    """Y_lm (l <= lmax) on a product rule exact for the degrees used here."""
    from scipy.special import sph_harm_y
    x, w = np.polynomial.legendre.leggauss(2 * lmax + 2)
    nph = 4 * lmax + 4
    th, ph = np.meshgrid(np.arccos(x), np.linspace(0, 2 * np.pi, nph, endpoint=False), indexing="ij")
    wt = (w[:, None] * np.full(nph, 2 * np.pi / nph)[None, :]).ravel()
    Y = np.array([sph_harm_y(l, m, th, ph).ravel() for l in range(lmax + 1) for m in range(-l, l + 1)])
    return Y, wt


def _channels(s, iat, lmax):
    """Radial channels of an atom type: (kind, l, m, ilo, s3r index) in a dense list."""
    nc = int(s.ncore[iat])
    out = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            out.append((0, l, m, 0, nc + l))
            out.append((1, l, m, 0, nc + l + lmax + 1))
            lo0 = nc + 2 * (lmax + 1) + int(np.sum(s.nLO_at[:l, iat]))
            for i in range(int(s.nLO_at[l, iat]) if l < s.nLO_at.shape[0] else 0):
                out.append((2, l, m, i, lo0 + i))
    return out


def _coef(a, b, g, ch, idf):
    """[bands, channels] coefficients (alfa / beta / gama) of one atom."""
    cols = []
    for kind, l, m, i, _ in ch:
        lm = l * l + l + m
        cols.append(a[:, lm, idf] if kind == 0 else b[:, lm, idf] if kind == 1 else g[:, i, lm, idf])
    return np.array(cols).T


def _atoms(s):
    idf = 0
    for iat in range(len(s.mult)):
        for _ in range(s.mult[iat]):
            yield iat, idf
            idf += 1


def minm_mt_by_quadrature(s, nst, nend, mst, mend):
    lmax = s.lmax
    lb_max = int(s.big_l.max())
    Y, wt = _quad(max(lmax, lb_max))
    nlm = (lmax + 1) ** 2
    idx = lambda l, m: l * l + l + m  # noqa: E731
    out = np.zeros((nend - nst, mend - mst, s.loctmatsize), dtype=complex)
    im = 0
    for iat, idf in _atoms(s):
        ch = _channels(s, iat, lmax)
        A = _coef(s.alfa[nst:nend], s.beta[nst:nend], s.gama[nst:nend], ch, idf)
        P = _coef(s.alfp[mst:mend], s.betp[mst:mend], s.gamp[mst:mend], ch, idf)
        phs = np.exp(-2j * np.pi * np.dot(s.ql, s.vpos[:, idf]))
        for irm in range(s.nmix[iat]):
            lb = int(s.big_l[irm, iat])
            for mb in range(-lb, lb + 1):
                # G[X1, X2] = int Y*_l1m1 Y_l2m2 Y_lbmb
                G = np.array([[np.sum(wt * np.conj(Y[idx(c1[1], c1[2])]) * Y[idx(c2[1], c2[2])] * Y[idx(lb, mb)])
                               for c2 in ch] for c1 in ch])
                R = s.s3r[irm][np.ix_([c[4] for c in ch], [c[4] for c in ch])][:, :, iat]
                out[:, :, im] = phs * A @ (G * R) @ P.T
                im += 1
    assert nlm > 0
    return out


def minc_by_quadrature(s, nst, nend, cst, cend, second):
    lmax = s.lmax
    Y, wt = _quad(max(lmax, int(s.big_l.max()), max(max(c) for c in s.core_l)))
    idx = lambda l, m: l * l + l + m  # noqa: E731
    off, im = {}, 0
    for iat, idf in _atoms(s):
        for irm in range(s.nmix[iat]):
            off[(irm, idf)] = im
            im += 2 * int(s.big_l[irm, iat]) + 1
    nb = nend - nst
    out = np.zeros(((cend - cst), nb, s.loctmatsize) if second else (nb, cend - cst, s.loctmatsize), dtype=complex)
    for icg in range(cst, cend):
        iat, idf, ic, lc, mc = [int(x) for x in s.corind[icg]]
        ch = _channels(s, iat, lmax)
        if second:
            A = _coef(s.alfp[nst:nend], s.betp[nst:nend], s.gamp[nst:nend], ch, idf)
            phs = np.exp(2j * np.pi * np.dot(s.kl, s.vpos[:, idf]))
        else:
            A = _coef(s.alfa[nst:nend], s.beta[nst:nend], s.gama[nst:nend], ch, idf)
            phs = np.exp(-2j * np.pi * np.dot(s.kl, s.vpos[:, idf]))
        for irm in range(s.nmix[iat]):
            lb = int(s.big_l[irm, iat])
            for mb in range(-lb, lb + 1):
                if second:   # int Y*_lc,mc Y_l1m1 Y_lbmb
                    g = np.array([np.sum(wt * np.conj(Y[idx(lc, mc)]) * Y[idx(c[1], c[2])] * Y[idx(lb, mb)])
                                  for c in ch])
                else:        # int Y*_l1m1 Y_lc,mc Y_lbmb
                    g = np.array([np.sum(wt * np.conj(Y[idx(c[1], c[2])]) * Y[idx(lc, mc)] * Y[idx(lb, mb)])
                                  for c in ch])
                r = np.array([s.s3r[irm, ic, c[4], iat] for c in ch])
                v = phs * A @ (g * r)
                j = off[(irm, idf)] + lb + mb
                if second:
                    out[icg - cst, :, j] = v
                else:
                    out[:, icg - cst, j] = v
    return out


@need_gaunt
def test_gaunt_table_matches_reference():
    """cmp_all_Gaunt of the oracle and of the package against the compiled
    table (maxj 8), and getcgcoef's lookup (both m2 signs, both l orders)."""
    from pygw_amd.coulomb import cmp_all_gaunt
    want = ref.cmp_all_gaunt(8)
    assert np.abs(O.cmp_all_gaunt(8) - want).max() < 1e-15
    assert np.abs(cmp_all_gaunt(8) - want).max() < 1e-15
    rng = np.random.default_rng(1)
    for _ in range(300):
        l1, l2 = rng.integers(0, 7, 2)
        l3 = int(rng.integers(abs(l1 - l2), l1 + l2 + 1))
        m1, m2 = int(rng.integers(-l1, l1 + 1)), int(rng.integers(-l2, l2 + 1))
        assert abs(O.getcgcoef(l1, l2, l3, m1, m2, want) - ref.getcgcoef(l1, l2, l3, m1, m2, want)) < 1e-15


def test_gaunt_table_is_the_angular_integral():
    """getcgcoef(l1, l2, l3, m1, m2) = int Y*_l3,m1+m2 Y_l1m1 Y_l2m2 by quadrature."""
    cg = O.cmp_all_gaunt(5)
    Y, wt = _quad(10)
    idx = lambda l, m: l * l + l + m  # noqa: E731
    for l1 in range(4):
        for l2 in range(4):
            for l3 in range(abs(l1 - l2), l1 + l2 + 1):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(-l2, l2 + 1):
                        if abs(m1 + m2) > l3:
                            continue
                        q = np.sum(wt * np.conj(Y[idx(l3, m1 + m2)]) * Y[idx(l1, m1)] * Y[idx(l2, m2)])
                        assert abs(O.getcgcoef(l1, l2, l3, m1, m2, cg) - q) < 1e-12


@pytest.mark.parametrize("los", [True, False])
def test_minm_mt_is_the_angular_integral(los):
    s = synth.minm_mt_system(los=los)
    got = O.calc_minm_mt(s.ql, 1, 6, 2, 9, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, s.s3r, s.vpos, s.mult,
                         s.nmix, s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
    want = minm_mt_by_quadrature(s, 1, 6, 2, 9)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("second", [False, True])
def test_minc_is_the_angular_integral(second):
    s = synth.minm_mt_system()
    if second:
        got = O.calc_minc2(s.kl, 1, 5, 2, 8, s.alfp, s.betp, s.gamp, s.s3r, s.corind, s.vpos, s.mult, s.nmix,
                           s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
    else:
        got = O.calc_minc(s.kl, 1, 6, 1, 5, s.alfa, s.beta, s.gama, s.s3r, s.corind, s.vpos, s.mult, s.nmix,
                          s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
    want = minc_by_quadrature(s, 2, 8, 1, 5, True) if second else minc_by_quadrature(s, 1, 6, 1, 5, False)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@need_vx
@pytest.mark.parametrize("los", [True, False])
def test_minc_matches_reference(los):
    s = synth.minm_mt_system(los=los)
    ncg = len(s.corind)
    args = (s.s3r, s.corind, s.vpos, s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
    want = ref.calc_minc(s.kl, 1, 6, 0, ncg, s.alfa, s.beta, s.gama, *args)
    got = O.calc_minc(s.kl, 1, 6, 0, ncg, s.alfa, s.beta, s.gama, *args)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    want = ref.calc_minc2(s.kl, 1, ncg - 1, 2, 8, s.alfp, s.betp, s.gamp, *args)
    got = O.calc_minc2(s.kl, 1, ncg - 1, 2, 8, s.alfp, s.betp, s.gamp, *args)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
