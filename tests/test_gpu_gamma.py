"""GPU parity of the q = 0 (Gamma point) head and wings against the oracle:
the wing of calc_eps, the inversion's head/wings update, and the MWM terms
(pygw.py:3177-3193, 3202-3263, 3359-3364)."""
import numpy as np
import pytest

from oracle import chain

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_wings_vs_oracle(dtype):
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(21)
    nk, npair, pb, nom = 5, 37, 83, 7
    M = (rng.normal(size=(nk, npair, pb)) + 1j * rng.normal(size=(nk, npair, pb))).astype(dtype)
    pm = rng.normal(size=(nk, npair)) + 1j * rng.normal(size=(nk, npair))
    kcw = -np.abs(rng.normal(size=(nk, nom, npair)))
    print(f"case wings {dtype}", flush=True)
    want = chain.calc_eps_wings(((k, M[k].astype(complex)) for k in range(nk)), lambda k: kcw[k].T,
                                lambda k: pm[k])
    got = P.chi0_wings(M, kcw, pm).cpu().numpy()
    assert _rel(got, want) < 1e-13
    # two k-chunks accumulated into one wing (k0 offset into kcw)
    w = P.chi0_wings(M[:2], kcw, pm[:2])
    P.chi0_wings(M[2:], kcw, pm[2:], epsw1=w, k0=2)
    assert _rel(w.cpu().numpy(), want) < 1e-13
    torch.cuda.synchronize()


def test_eps_inverse_head_vs_oracle():
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(22)
    pb, nom = 70, 4
    x = rng.normal(size=(nom, pb, 2 * pb)) + 1j * rng.normal(size=(nom, pb, 2 * pb))
    eps = np.einsum("wik,wjk->ijw", x, np.conj(x)) * 0.3 / pb
    head = 0.4 + 0.3 * rng.normal(size=nom) + 0.1j * rng.normal(size=nom)
    w1 = 0.2 * (rng.normal(size=(pb, nom)) + 1j * rng.normal(size=(pb, nom)))
    print("case eps inverse head", flush=True)
    want = chain.eps_inverse_head(eps, head, w1)
    e = torch.empty(pb * pb * nom, dtype=torch.complex128, device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    e.copy_(torch.as_tensor(eps))
    got = P.eps_inverse_head(e, head, w1)
    names = ["eps", "head", "epsw1", "epsw2", "emac"]
    for n, g, wv in zip(names, got, want):
        err = _rel(g.cpu().numpy(), wv)
        print(f"  {n} rel={err:.2e}", flush=True)
        assert err < 1e-11, n
    torch.cuda.synchronize()


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_mwm_head_vs_oracle(dtype):
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(23)
    nk, nb1, ncg, nb2v, pb, nom = 3, 4, 2, 5, 66, 3
    nb2 = ncg + nb2v
    Ms = (rng.normal(size=(nk, nb1 * nb2, pb)) + 1j * rng.normal(size=(nk, nb1 * nb2, pb))).astype(dtype)
    W = rng.normal(size=(pb, pb, nom)) + 1j * rng.normal(size=(pb, pb, nom))
    head = rng.normal(size=nom) + 1j * rng.normal(size=nom)
    e1 = rng.normal(size=(pb, nom)) + 1j * rng.normal(size=(pb, nom))
    e2 = rng.normal(size=(pb, nom)) + 1j * rng.normal(size=(pb, nom))
    c1, c2, wkq = 0.37, 0.21, 1.0 / 8
    print(f"case mwm head {dtype}", flush=True)
    Wd = torch.empty(pb * pb * nom, dtype=torch.complex128, device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    Wd.copy_(torch.as_tensor(W))
    prec = "fp64" if dtype == "complex128" else "mixed"
    mwm = P.calc_mwm(torch.as_tensor(Ms, device="cuda"), Wd, wkq, precision=prec)
    base = mwm.clone()
    P.calc_mwm_head(torch.as_tensor(Ms, device="cuda"), mwm, head, e1, e2, c1, c2, ncg, nb1, nb2)
    got = (mwm - base).cpu().numpy()
    for k in range(nk):
        want = np.zeros((nom, nb1, nb2), dtype=complex)
        chain.calc_mwm_head(Ms[k].astype(complex).reshape(nb1, nb2, pb), want, head, e1, e2, c1, c2, ncg)
        assert _rel(got[k].reshape(nom, nb1, nb2), want) < 1e-13
    torch.cuda.synchronize()
