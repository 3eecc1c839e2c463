"""The Coulomb v^(1/2) oracle (oracle/coulomb.py) against closed forms: the
2 x 2 Hermitian eigenproblem, matrices of known spectrum, the q = 0 rule of
Coul_setev, and v^(1/2) as a square root of v."""
import numpy as np

from oracle import coulomb as O


def test_eigh_2x2_closed_form():
    a, b, c = 2.0, -0.5, 0.7 - 1.1j
    ev, V = O.eigh_lower(np.array([[a, np.conj(c)], [c, b]]))
    m, d = (a + b) / 2, np.sqrt(((a - b) / 2) ** 2 + abs(c) ** 2)
    assert np.allclose(ev, [m - d, m + d], rtol=0, atol=1e-15)
    # an eigenvector of the upper value: (c*, a' - a) with a' = m + d
    x = np.array([np.conj(c), m + d - a])
    x /= np.linalg.norm(x)
    assert abs(abs(np.vdot(x, V[:, 1])) - 1.0) < 1e-14


def test_setev_known_spectrum_and_order():
    rng = np.random.default_rng(1)
    lam = np.concatenate([10 ** rng.uniform(-3, 2, 40), 10 ** rng.uniform(-12, -9, 6), [-3e-11, -2e-12]])
    A, Q = O.hermitian_with_spectrum(lam, rng)
    for evtol in (1e-8, -1e-10):
        ev, V, barcev, barcvm, immax = O.coul_setev(A, evtol)
        keep = lam > evtol
        assert immax == -1 and len(barcev) == keep.sum() == barcvm.shape[1]
        assert np.all(np.diff(barcev) <= 0)
        assert np.allclose(barcev, np.sort(lam[keep])[::-1], rtol=0, atol=1e-13 * lam.max())
        P = (Q[:, keep] * np.abs(lam[keep])) @ Q[:, keep].conj().T
        assert np.abs(barcvm @ barcvm.conj().T - P).max() < 1e-13 * lam.max()
        # each column is an eigenvector scaled by sqrt|ev|
        for m in (0, len(barcev) // 2, len(barcev) - 1):
            x = barcvm[:, m]
            assert abs(np.vdot(x, x).real - abs(barcev[m])) < 1e-13 * lam.max()
            assert np.abs(A @ x - barcev[m] * x).max() < 1e-12 * lam.max()


def test_setev_q0_immax_rule():
    rng = np.random.default_rng(2)
    lam = np.sort(10 ** rng.uniform(-4, 1, 30))
    A, Q = O.hermitian_with_spectrum(lam, rng)
    j0 = 17
    wi0 = 2.5 * Q[:, j0] + 1e-3 * (rng.normal(size=30) + 1j * rng.normal(size=30))
    _, _, barcev, barcvm, immax = O.coul_setev(A, 1e-8, wi0, 0)
    assert immax == j0 and len(barcev) == 29 and not np.any(np.isclose(barcev, lam[j0], rtol=1e-10))
    # iop_coul_x != 0 keeps it even below evtol
    _, _, barcev, _, immax = O.coul_setev(A, lam[j0] * 1.5, wi0, 1)
    assert immax == j0 and np.isclose(barcev[-1], lam[j0], rtol=1e-10)
    assert len(barcev) == np.sum(lam > lam[j0] * 1.5) + 1


def test_minm_is_v_half():
    """minm minm^H = mmat conj(V) mmat^H when nothing is truncated; the minc
    product over the first n_in rows equals the full product of the
    zero-padded mmat_c."""
    rng = np.random.default_rng(3)
    lam = 10 ** rng.uniform(-2, 1, 24)
    A, _ = O.hermitian_with_spectrum(lam, rng)
    _, _, _, barcvm, _ = O.coul_setev(A, 0.0)
    mm = rng.normal(size=(3, 5, 24)) + 1j * rng.normal(size=(3, 5, 24))
    M = O.minm(mm, barcvm).reshape(15, -1)
    X = mm.reshape(15, 24)
    assert np.abs(M @ M.conj().T - X @ np.conj(A) @ X.conj().T).max() < 1e-12 * np.abs(M).max() ** 2
    mc = rng.normal(size=(7, 10)) + 1j * rng.normal(size=(7, 10))
    pad = np.concatenate([mc, np.zeros((7, 14))], axis=1)
    longer = np.concatenate([mc, rng.normal(size=(7, 4))], axis=1)    # rows of ld 14, n_in = 10
    assert np.abs(O.minm(longer, barcvm, 10) - O.minm(pad, barcvm)).max() < 1e-13
