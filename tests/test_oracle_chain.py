"""Pins of the numpy chain oracle (CPU only): golden fixtures from PyGW's
Fortran (convolution, Pade), closed forms and linear-algebra identities."""
import os

import numpy as np

from oracle import chain
from pygw_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_convolution_matches_reference_fixture():
    g = np.load(os.path.join(GOLD, "conv.npz"))
    sc = chain.all_fr_convolutions(g["enk"], g["mwm"], g["omega"], g["womeg"])
    np.testing.assert_allclose(sc, g["sc"], rtol=1e-13, atol=1e-15)
    sv = chain.all_fr_convolutions_vec(g["enk"], g["mwm"], g["omega"], g["womeg"])
    np.testing.assert_allclose(sv, g["sc"], rtol=1e-12, atol=1e-14)


def test_pade_matches_reference_fixture():
    g = np.load(os.path.join(GOLD, "pade.npz"))
    a = chain.padecof(g["y"], g["z"])
    np.testing.assert_allclose(a, g["apar"], rtol=1e-10)
    np.testing.assert_allclose(chain.padeg(g["x"], g["z"], g["apar"]), g["g"], rtol=1e-10)


def test_pade_continues_rational_function():
    """A two-pole function sampled at i*w_n is continued to the real axis."""
    z = 1j * np.geomspace(0.01, 8.0, 20)
    f = lambda x: 0.4 / (x - 0.3 + 0.1j) + 0.6 / (x + 0.7 - 0.05j)  # noqa: E731
    a = chain.padecof(f(z), z)
    x = np.linspace(-1.5, 1.5, 13)
    np.testing.assert_allclose(chain.padeg(x, z, a), f(x), rtol=1e-8)


def test_convolution_closed_form():
    """mwm(W) = 1/(W^2+a^2): (1/pi) int_0^inf mwm(W) xi/(xi^2+W^2) dW
    = xi / (2 a s (a+s)), s = sqrt(xi^2) with Re s > 0, xi = E - i w."""
    a = 0.7
    # Gauss-Legendre in t on [0, atan(W_max/c)], W = c tan t: a converged
    # rule on [0, W_max] (PyGW's own tan mesh weights extend past W_max,
    # mcommon.py:28-30, which this check must not depend on)
    c, wmax = 0.5, 400.0
    t, wt = np.polynomial.legendre.leggauss(3000)
    tm = np.arctan(wmax / c)
    t = 0.5 * tm * (t + 1)
    om = c * np.tan(t)
    wom = 0.5 * tm * wt * c / np.cos(t) ** 2
    om = np.append(om, wmax)       # fr_convolution2 takes W_max = omega(nom)
    wom = np.append(wom, 0.0)
    enk = np.array([-0.4, 0.3, 1.1])
    wn = om[om > 0.05][::400]   # below ~0.05 the Lorentzian at W=|E| is narrower than the mesh
    mwm_coarse = 1.0 / (wn ** 2 + a * a)
    for ie, e in enumerate(enk):
        for iw, w in enumerate(wn):
            got = chain.fr_convolution2(e, 1.0 / (om ** 2 + a * a), mwm_coarse[iw], om, w, wom)
            xi = complex(e, -w)
            s = xi if xi.real > 0 else -xi
            want = xi / (2 * a * s * (a + s))
            assert abs(got - want) < 2e-5 * abs(want), (e, w, got, want)


def test_eps_inverse_identities():
    rng = np.random.default_rng(3)
    pb, nom = 12, 3
    x = rng.normal(size=(40, pb)) + 1j * rng.normal(size=(40, pb))
    eps = np.zeros((pb, pb, nom), dtype=complex)
    for i in range(nom):
        w = rng.uniform(0.1, 1.0, 40)
        eps[:, :, i] = (x.T * w) @ np.conj(x)
    inv = chain.eps_inverse(eps)
    for i in range(nom):
        a = (eps[:, :, i] + eps[:, :, i].conj().T) / 2 + np.eye(pb)
        np.testing.assert_allclose(a @ (inv[:, :, i] + np.eye(pb)), np.eye(pb), atol=1e-12)
        ev = np.linalg.eigvalsh(inv[:, :, i] + np.eye(pb))
        assert ev.min() > 0 and ev.max() <= 1 + 1e-12


def test_mwm_matches_definition():
    rng = np.random.default_rng(5)
    r, pb, nom = 7, 9, 2
    m = rng.normal(size=(r, pb)) + 1j * rng.normal(size=(r, pb))
    w = rng.normal(size=(pb, pb, nom)) + 1j * rng.normal(size=(pb, pb, nom))
    got = chain.calc_mwm(m, w, 0.25)
    want = 0.25 * np.einsum("ri,ijw,rj->wr", np.conj(m), w, m)
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_tan_mesh_solves_its_equations():
    from scipy import optimize
    x0, L, nw = 0.02, 20.0, 32
    om, dom = mesh.give2tan_mesh(x0, L, nw)

    def fun(x):
        d, w = x
        return [L - w / np.tan(d), x0 - w * np.tan(np.pi / (2 * nw) - d / nw)]
    tnw = np.tan(np.pi / (2 * nw))
    d0 = nw / 2.0 * (tnw - np.sqrt(tnw ** 2 - 4 * x0 / L / nw))
    sol = optimize.root(fun, [d0, L * d0])
    d, w = sol.x
    xt = np.linspace(0.0, 1.0, 2 * nw) * (np.pi - 2 * d) - np.pi / 2 + d
    np.testing.assert_allclose(om, w * np.tan(xt), rtol=1e-10, atol=1e-14)
    assert abs(om[-1] - L) < 1e-9 * L


def _herm_pd(rng, n, nom, scale=0.3):
    x = rng.normal(size=(nom, n, 2 * n)) + 1j * rng.normal(size=(nom, n, 2 * n))
    e = np.einsum("wik,wjk->ijw", x, np.conj(x)) * scale / n
    return e


def test_wings_are_the_extra_column_of_chi0():
    """The q = 0 wing equals the extra column of the chi0 contraction whose
    M carries conj(pm) as one more product-basis function."""
    rng = np.random.default_rng(11)
    nk, npair, pb, nom = 3, 7, 9, 4
    M = rng.normal(size=(nk, npair, pb)) + 1j * rng.normal(size=(nk, npair, pb))
    pm = rng.normal(size=(nk, npair)) + 1j * rng.normal(size=(nk, npair))
    kcw = -np.abs(rng.normal(size=(nk, npair, nom)))
    w1 = chain.calc_eps_wings(((k, M[k]) for k in range(nk)), lambda k: kcw[k], lambda k: pm[k])
    Mx = np.concatenate([M, np.conj(pm)[:, :, None]], axis=2)
    ex = chain.calc_eps_accumulate(((k, Mx[k]) for k in range(nk)), lambda k: kcw[k])
    np.testing.assert_allclose(w1, ex[:pb, pb, :], rtol=0, atol=1e-12 * np.abs(w1).max())


def test_head_update_is_the_bordered_inverse_for_real_wings():
    """With real wings the q = 0 update of pygw.py:3246-3263 is the block
    inverse of the bordered matrix [[1 + head, epsw2^T], [epsw1, E]]
    (E = hermitized eps + 1): head + 1, both wings, the body and emac."""
    rng = np.random.default_rng(12)
    pb, nom = 10, 3
    eps = _herm_pd(rng, pb, nom)
    head = 0.5 + rng.normal(size=nom) + 0.2j * rng.normal(size=nom)
    w1 = 0.3 * rng.normal(size=(pb, nom)) + 0j
    out, h, e1, e2, emac = chain.eps_inverse_head(eps, head, w1)
    for iom in range(nom):
        E = (eps[:, :, iom] + np.conj(eps[:, :, iom].T)) / 2 + np.eye(pb)
        B = np.zeros((pb + 1, pb + 1), dtype=complex)
        B[0, 0] = 1 + head[iom]
        B[0, 1:] = np.conj(w1[:, iom])
        B[1:, 0] = w1[:, iom]
        B[1:, 1:] = E
        Bi = np.linalg.inv(B)
        np.testing.assert_allclose(h[iom] + 1, Bi[0, 0], rtol=1e-12)
        np.testing.assert_allclose(e1[:, iom], Bi[1:, 0], rtol=0, atol=1e-12)
        np.testing.assert_allclose(e2[:, iom], Bi[0, 1:], rtol=0, atol=1e-12)
        np.testing.assert_allclose(out[:, :, iom] + np.eye(pb), Bi[1:, 1:], rtol=0, atol=1e-12)
        np.testing.assert_allclose(emac[iom, 0], 1 / Bi[0, 0], rtol=1e-12)
    # complex wings: the reference's row wing conj(A epsw2) is not the block
    # inverse's (A^T epsw2), so the update departs from it (kept as PyGW has it)
    w1c = w1 + 0.3j * rng.normal(size=(pb, nom))
    _, _, _, e2c, _ = chain.eps_inverse_head(eps, head, w1c)
    E = (eps[:, :, 0] + np.conj(eps[:, :, 0].T)) / 2 + np.eye(pb)
    B = np.block([[np.array([[1 + head[0]]]), np.conj(w1c[:, :1]).T], [w1c[:, :1], E]])
    assert np.abs(e2c[:, 0] - np.linalg.inv(B)[0, 1:]).max() > 1e-3


def test_mwm_head_terms_are_the_bordered_product():
    """With coefs2 = coefs1^2 and wkq = 1 the q = 0 MWM terms complete
    conj(Mf) Wf Mf over the bordered basis: Mf = [coefs1, M],
    Wf = [[head, epsw2^T], [epsw1, W]] (pygw.py:3359-3364)."""
    rng = np.random.default_rng(13)
    nb1, ncg, nb2v, pb, nom = 3, 2, 4, 6, 2
    nb2 = ncg + nb2v
    minm = rng.normal(size=(nb1, nb2, pb)) + 1j * rng.normal(size=(nb1, nb2, pb))
    W = rng.normal(size=(pb, pb, nom)) + 1j * rng.normal(size=(pb, pb, nom))
    head = rng.normal(size=nom) + 1j * rng.normal(size=nom)
    e1 = rng.normal(size=(pb, nom)) + 1j * rng.normal(size=(pb, nom))
    e2 = rng.normal(size=(pb, nom)) + 1j * rng.normal(size=(pb, nom))
    c1 = 0.7
    mwm = chain.calc_mwm(minm.reshape(nb1 * nb2, pb), W, 1.0).reshape(nom, nb1, nb2)
    chain.calc_mwm_head(minm, mwm, head, e1, e2, c1, c1 * c1, ncg)
    for iom in range(nom):
        Wf = np.zeros((pb + 1, pb + 1), dtype=complex)
        Wf[0, 0] = head[iom]
        Wf[0, 1:] = e2[:, iom]
        Wf[1:, 0] = e1[:, iom]
        Wf[1:, 1:] = W[:, :, iom]
        for ib1 in range(nb1):
            for ib2 in range(nb2):
                mf = np.concatenate([[c1 if ib2 == ib1 + ncg else 0.0], minm[ib1, ib2]])
                want = np.conj(mf) @ Wf @ mf
                np.testing.assert_allclose(mwm[iom, ib1, ib2], want, rtol=1e-12)
