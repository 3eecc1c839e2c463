"""GPU parity of the LAPW coefficients (lapw.hip through pygw_amd.lapw) against
the oracle (oracle/lapw.py, itself pinned to PyGW's compiled lapwcoeff.f90):
gap2_set_lapwcoef with and without local orbitals, rotations and iop = 2, at
k = 0 (a zero |k+G|), at a PyGW-sized lmax 10; the band coefficients
against numpy products; and the chain that calc_minm runs, coefficients ->
alfa/beta/gama at k and k - q -> calc_minm_MT, against the oracle's chain."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import lapw as O
from oracle import minm_mt as OM
from pygw_amd import coulomb as C
from pygw_amd import lapw as L
from pygw_amd import synth

pytestmark = pytest.mark.gpu


def _gpu(s, iop=1, ifrot=True):
    plan = L.LapwPlan(*s.plan_args())
    return plan, L.gap2_set_lapwcoef(plan, s.kl, s.indgk, iop, ifrot, s.nvk, s.gindex)


def _close(got, want, tol=1e-12):
    g = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    assert g.shape == want.shape
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(g - want).max() / scale
    assert err < tol, err


@pytest.mark.parametrize("iop,ifrot,los", [(1, True, True), (2, True, True), (1, False, True), (1, True, False)])
def test_lapw_coef_matches_oracle(iop, ifrot, los):
    s = synth.lapw_system(los=los)
    print(f"lapw coef iop={iop} ifrot={ifrot} los={los}", flush=True)
    _, (a, b, c) = _gpu(s, iop, ifrot)
    wa, wb, wc = O.gap2_set_lapwcoef(*s.coef_args(iop, ifrot))
    _close(a, wa)
    _close(b, wb)
    if los:
        _close(c, wc)
    else:
        assert c is None and wc.shape[1] == 0


def test_lapw_coef_gamma_point():
    s = synth.lapw_system(kl=(0.0, 0.0, 0.0), nvk=30)
    print("lapw coef k = 0", flush=True)
    _, (a, b, c) = _gpu(s)
    wa, wb, wc = O.gap2_set_lapwcoef(*s.coef_args())
    for g, w in ((a, wa), (b, wb), (c, wc)):
        _close(g, w)


def test_lapw_coef_lmax10():
    """lmax 10 (121 (l, m)), four atom types of two atoms, 300 plane waves."""
    s = synth.lapw_system(lmax=10, lomax=3, nvk=300, mult=(2, 2, 2, 2), apw=(True, False, True, False), gmax=6,
                          alat=10.0)
    print("lapw coef lmax 10", flush=True)
    _, (a, b, c) = _gpu(s)
    wa, wb, wc = O.gap2_set_lapwcoef(*s.coef_args())
    for g, w in ((a, wa), (b, wb), (c, wc)):
        _close(g, w)


def test_lapw_coef_reuses_out_and_rejects_bad_rows():
    s = synth.lapw_system()
    plan, first = _gpu(s)
    again = L.gap2_set_lapwcoef(plan, s.kl, s.indgk, 1, True, s.nvk, s.gindex, out=first)
    assert again[0].data_ptr() == first[0].data_ptr()
    _close(again[0], O.gap2_set_lapwcoef(*s.coef_args())[0])
    bad = s.indgk.copy()
    bad[5] = len(s.gindex) + 3       # not a row of gindex: that row reads as NaN, the others are unaffected
    a, _, _ = L.gap2_set_lapwcoef(plan, s.kl, bad, 1, True, s.nvk, s.gindex)
    a = a.cpu().numpy()
    assert np.isnan(a[5]).all() and not np.isnan(np.delete(a, 5, axis=0)).any()


@pytest.mark.parametrize("conj", [False, True])
def test_band_coefficients_match_numpy(conj):
    s = synth.lapw_system(nbmax=70, nvk=150, gmax=5)
    print(f"band coefficients conj={conj}", flush=True)
    _, (a, b, c) = _gpu(s)
    alfa, beta, gama = L.band_coefficients(s.Aeig, a, b, c, conj=conj)
    A = np.conj(s.Aeig) if conj else s.Aeig
    wa, wb, wc = O.band_coefficients(A, a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy())
    for g, w in ((alfa, wa), (beta, wb), (gama, wc)):
        _close(g, w, 1e-13)


def test_lapw_chain_into_minm_mt():
    """gap2_set_lapwcoef at k and k - q, alfa/beta/gama (Aeigq conjugated as
    PyGW forms it), then calc_minm_MT on the GPU, against the same chain in
    the oracle."""
    sk = synth.lapw_system(lmax=3, lomax=3, kl=(0.25, -0.125, 0.375))
    sq = synth.lapw_system(lmax=3, lomax=3, kl=(0.0, 0.125, 0.25))
    base = synth.minm_mt_system(lmax=3, nbmax=sk.Aeig.shape[0], mbmax=sq.Aeig.shape[0])
    # the product basis of the matrix-element system on the LAPW system's atoms
    nat = len(sk.mult)
    ncore = np.asarray(base.ncore)
    nfnc = max(int(ncore[a]) + 2 * 4 + int(sk.nLO_at[:, a].sum()) for a in range(nat))
    rng = np.random.default_rng(4)
    s3r = rng.normal(size=(base.s3r.shape[0], nfnc, nfnc, nat))
    s3r = 0.5 * (s3r + s3r.transpose(0, 2, 1, 3))
    print("lapw -> minm_mt chain", flush=True)
    plan = L.LapwPlan(*sk.plan_args())
    ck = L.gap2_set_lapwcoef(plan, sk.kl, sk.indgk, 1, True, sk.nvk, sk.gindex)
    cq = L.gap2_set_lapwcoef(plan, sq.kl, sq.indgk, 2, True, sq.nvk, sq.gindex)
    alfa, beta, gama = L.band_coefficients(sk.Aeig, *ck)
    alfp, betp, gamp = L.band_coefficients(sq.Aeig, *cq, conj=True)
    s = dataclasses.replace(base, mult=sk.mult, nLO_at=sk.nLO_at, s3r=s3r, vpos=sk.pos)
    mplan = C.MinmMTPlan(s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.lmax)
    got = C.calc_minm_mt(mplan, s.ql, 1, 10, 2, 11, alfa, beta, gama, alfp, betp, gamp, s.s3r, s.vpos,
                         s.cgcoef).cpu().numpy()
    wk = O.band_coefficients(sk.Aeig, *O.gap2_set_lapwcoef(*sk.coef_args(1, True)))
    wq = O.band_coefficients(np.conj(sq.Aeig), *O.gap2_set_lapwcoef(*sq.coef_args(2, True)))
    want = OM.calc_minm_mt(s.ql, 1, 10, 2, 11, *wk, *wq, s.s3r, s.vpos, s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore,
                           s.cgcoef, s.lmax, s.loctmatsize)
    _close(got, want, 1e-12)
