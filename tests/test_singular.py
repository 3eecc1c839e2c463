"""The q = 0 Coulomb singularity terms (coulomb.singular_coefficients,
pygw.py:2456-2474) and <chi|1> (coulomb.fwi0, pygw.py:2216-2228), host
arithmetic: the integrals intf1, intf2 against quadrature of
Vol/(2 pi)^3 int d^3Q e^{-alpha Q^2}/Q^n, and the lattice sums converging to
them (singc2 ~ N_k^(-1/3) h-scaling of the excluded cell, singc1 ~ N_k^(-2/3))
on a simple cubic cell, which a wrong normalisation, a kept Q = 0 term or a
wrong alpha breaks."""
import numpy as np
from scipy import integrate

from pygw_amd import coulomb as C


def _gqlen(a, n, gmax=6):
    """|q + G| of a simple cubic cell (side a) on an n^3 q-mesh, G = 0 first at q = 0."""
    b = 2 * np.pi / a
    r = np.arange(-gmax, gmax + 1)
    G = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    G = G[np.argsort((G ** 2).sum(1), kind="stable")]
    out = []
    for q in np.stack(np.meshgrid(*[np.arange(n) / n] * 3, indexing="ij"), -1).reshape(-1, 3):
        q = np.where(q > 0.5, q - 1, q)
        out.append(np.linalg.norm((G + q) * b, axis=1))
    return out


def test_integrals_are_the_continuum_limits():
    vol = 270.0
    alpha = (vol / (6 * np.pi ** 2)) ** (1 / 3)
    for n in (1, 2):
        f = lambda Q: 4 * np.pi * Q ** 2 * np.exp(-alpha * Q ** 2) / Q ** n  # noqa: E731
        want = vol / (2 * np.pi) ** 3 * integrate.quad(f, 0, np.inf, epsabs=1e-14)[0]
        got = vol / (4 * np.pi ** 2) * (1 / alpha if n == 1 else np.sqrt(np.pi / alpha))
        assert abs(got - want) < 1e-10 * want
    c1, c2 = C.singular_coefficients([np.array([0.0])], vol, (1, 1, 1))
    assert np.isclose(c1, vol / (4 * np.pi ** 2 * alpha)) and np.isclose(c2, vol * np.sqrt(np.pi / alpha) / (4 * np.pi ** 2))


def test_lattice_sums_converge_to_the_integrals():
    a = 6.0
    vals = [(n, *C.singular_coefficients(_gqlen(a, n), a ** 3, (n, n, n))) for n in (4, 6, 8)]
    # the excluded cell: singc2 ~ h^-... scales as 1/n (int d^3Q/Q^2 over a cell of side b/n), singc1 as 1/n^2
    s2 = [c2 * n for n, _, c2 in vals]
    s1 = [c1 * n * n for n, c1, _ in vals]
    assert abs(s2[2] - s2[1]) < 0.005 * abs(s2[2]) and abs(s1[2] - s1[1]) < 0.005 * abs(s1[2])
    assert all(c2 > 0 for _, _, c2 in vals)


def test_fwi0_places_the_L0_functions():
    w = C.fwi0(100.0, [1, 2], [[0, 1, 0], [1, 0]], [[0.5, 9.0, 0.25], [9.0, 2.0]])
    f = np.sqrt(4 * np.pi / 100.0)
    want = np.concatenate([[0.5 * f], [0, 0, 0], [0.25 * f], [0, 0, 0], [2.0 * f], [0, 0, 0], [2.0 * f]])
    assert np.allclose(w, want, rtol=0, atol=1e-15)
