"""GPU parity of the q = 0 head (head.hip through pygw_amd.head) against the
oracle (oracle/head.py, pinned to PyGW's compiled for_q_0.f90 and to the
momentum operator): the valence-valence elements (muffin-tin + interstitial,
both orders averaged) with and without local orbitals and for overlapping
windows (metals), the core-valence elements in the reference's form and the
exact one, the head sums (insulator, metal with the Drude term, SVD basis),
the wing's pm, and calc_head end to end from the LAPW coefficients."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import head as O
from oracle import lapw as OL
from pygw_amd import head as H
from pygw_amd import lapw as L
from pygw_amd import synth

pytestmark = pytest.mark.gpu


def _tables(s):
    return H.HeadTables(s.mult, s.nLO_at, s.vv_tables, s.core_off, s.cv_tables, s.corind)


def _pw(s):
    return H.PlaneWaves(s.ipwint, s.gindex, s.i_g0, s.k2cartes)


def _close(got, want, tol=1e-12):
    g = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    assert g.shape == want.shape, (g.shape, want.shape)
    err = np.abs(g - want).max() / max(np.abs(want).max(), 1e-300)
    assert err < tol, err


def _oracle_elements(s, irk, nst, nend, mst, mend, exact=False):
    vv, cv = O.momentum_elements(s.abc[irk], s.Aeig[irk], s.kl[irk], s.indgk[irk], s.nvk[irk], s.ipwint, s.gindex,
                                 s.i_g0, s.k2cartes, s.nLO_at, s.mult, s.vv_tables, s.corind, s.core_off,
                                 s.cv_tables, nst, nend, mst, mend)
    if exact:
        cv = np.zeros_like(cv)
        a, b, g = s.abc[irk]
        for ie2 in range(mst, mend):
            for iat in sorted(set(int(x) for x in s.corind[:, 0])):
                cv[ie2 - mst] += O.calcmmatcv(iat, ie2, s.corind, s.core_off, a, b, g, *s.cv_tables, s.nLO_at,
                                              exact=True)
    return vv, cv


@pytest.mark.parametrize("los,metal", [(True, False), (False, False), (True, True)])
def test_momentum_elements_match_oracle(los, metal):
    s = synth.head_system(los=los, metal=metal, nomx=4 if metal else 3, numin=2 if metal else 4)
    nst, nend, mst, mend = 0, s.nomx + 1, s.numin, s.nbmaxpol
    print(f"momentum elements los={los} metal={metal}", flush=True)
    vv, cv = H.momentum_elements(_tables(s), _pw(s), s.abc[1], s.Aeig[1], s.kl[1], s.indgk[1], s.nvk[1], nst, nend,
                                 mst, mend)
    wvv, wcv = _oracle_elements(s, 1, nst, nend, mst, mend)
    _close(vv, wvv)
    _close(cv, wcv)


def test_core_elements_exact_form():
    s = synth.head_system()
    print("core elements, exact form", flush=True)
    _, cv = H.momentum_elements(_tables(s), _pw(s), s.abc[0], s.Aeig[0], s.kl[0], s.indgk[0], s.nvk[0], 0, 4, 4, 9,
                                exact=True)
    _close(cv, _oracle_elements(s, 0, 0, 4, 4, 9, exact=True)[1])


def _head_args(s):
    return (s.Ebnd, s.eig_core, s.kcw, s.k_ind, s.weight, s.nomx, s.numin, s.nbmaxpol, s.ncg_p, s.vol)


@pytest.mark.parametrize("case", ["insulator", "metal_drude", "metal_drude_svd"])
def test_head_sums_match_oracle(case):
    metal = case != "insulator"
    s = synth.head_system(metal=metal, nirk=3, nomx=4 if metal else 3, numin=2 if metal else 4)
    rng = np.random.default_rng(3)
    n_e, nv, ncg = s.nbmaxpol - s.numin, s.nomx + 1, len(s.corind)
    vv = rng.normal(size=(3, n_e, nv, 3)) + 1j * rng.normal(size=(3, n_e, nv, 3))
    cv = rng.normal(size=(3, n_e, ncg, 3)) + 1j * rng.normal(size=(3, n_e, ncg, 3))
    kw = dict(Eg=-1.0 if metal else 1.0, kwfer=s.kwfer, iop_drude=1 if metal else 0, eta_head=0.01, omega=s.omega)
    kcw = s.kcw
    if case == "metal_drude_svd":
        dUl = rng.normal(size=(len(s.omega), 4))
        kw["dUl"] = dUl
        kcw = np.asfortranarray(s.kcw[:, :, :, :4])
    args = list(_head_args(s))
    args[2] = kcw
    print(f"head sums {case}", flush=True)
    got = H.head_from_elements(torch.as_tensor(vv, device="cuda"), torch.as_tensor(cv, device="cuda"), *args, **kw)
    want = O.head_from_elements(vv, cv, s.Ebnd, s.eig_core, s.corind, kcw, s.k_ind, s.weight, s.nomx, s.numin,
                                s.nbmaxpol, s.ncg_p, s.vol, **kw)
    _close(got, want)


def test_wing_pm_matches_oracle():
    s = synth.head_system(nirk=3)
    rng = np.random.default_rng(5)
    n_e, nv, ncg = s.nbmaxpol - s.numin, s.nomx + 1, len(s.corind)
    vv = rng.normal(size=(3, n_e, nv, 3)) + 1j * rng.normal(size=(3, n_e, nv, 3))
    cv = rng.normal(size=(3, n_e, ncg, 3)) + 1j * rng.normal(size=(3, n_e, ncg, 3))
    kii = np.array([2, 0, 1, 1, 2])
    got = H.wing_pm(torch.as_tensor(vv, device="cuda"), torch.as_tensor(cv, device="cuda"), s.Ebnd, s.eig_core, kii,
                    s.nomx, s.numin, s.nbmaxpol, s.ncg_p, s.vol).cpu().numpy()
    for k, irk in enumerate(kii):
        want = O.wing_pm(vv[irk], cv[irk], np.concatenate([s.eig_core, s.Ebnd[irk]]), s.ncg_p, s.nomx, s.numin,
                         s.nbmaxpol, s.vol)
        _close(got[k], want)


def test_calc_head_from_lapw_coefficients():
    """calc_head end to end: the LAPW coefficients (ifrot False) at each irk,
    alfa/beta/gama, the momentum elements and the head, against the oracle's
    composition of the same steps."""
    nirk = 2
    ls = [synth.lapw_system(lmax=3, lomax=3, nvk=30, nbmax=9, gmax=4, kl=kl, seed=23)
          for kl in ((0.25, -0.125, 0.375), (0.0, 0.125, -0.25))]
    g = ls[0].gindex
    where = {tuple(v): i for i, v in enumerate(g)}                 # one G list for both k-points
    ls[1] = dataclasses.replace(ls[1], gindex=g, indgk=np.array([where[tuple(ls[1].gindex[i])] for i in ls[1].indgk]))
    for x in ls:
        assert np.abs(x.gindex[x.indgk[:x.nvk]]).max() <= 2      # differences stay inside the +-4 box
    base = synth.head_system(nirk=nirk, nbmax=9, mult=(1, 2))
    gm = 4
    i_g0 = -np.ones((2 * gm + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gm, g[:, 1] + gm, g[:, 2] + gm] = np.arange(len(g))
    rng = np.random.default_rng(2)
    ipw = (rng.normal(size=len(g)) + 1j * rng.normal(size=len(g))) * np.exp(-0.3 * (g ** 2).sum(1))
    neg = i_g0[-g[:, 0] + gm, -g[:, 1] + gm, -g[:, 2] + gm]
    ipwint = 0.5 * (ipw + np.conj(ipw[neg]))
    # radial tables for the LAPW system's local orbitals
    nlo = max(int(ls[0].nLO_at.max()), 1)
    R = lambda *sh: rng.normal(size=sh)  # noqa: E731
    vv_t = tuple(R(2, 3, 2) for _ in range(4)) + tuple(R(2, nlo, 4, 2) for _ in range(4)) + (R(2, nlo, nlo, 4, 2),)
    tabs = H.HeadTables(ls[0].mult, ls[0].nLO_at, vv_t, base.core_off, base.cv_tables, base.corind)
    pw = H.PlaneWaves(ipwint, g, i_g0, ls[0].k2cartes)
    plan = L.LapwPlan(*ls[0].plan_args())
    print("calc_head chain", flush=True)
    head, mcv, mvv, mst = H.calc_head(tabs, pw, plan, [x.kl for x in ls], [x.indgk for x in ls], [x.nvk for x in ls],
                                      [x.Aeig for x in ls], *_head_args(base), gindex=g)
    assert mst == base.numin
    wvv, wcv = [], []
    for x in ls:
        abc = OL.band_coefficients(x.Aeig, *OL.gap2_set_lapwcoef(*x.coef_args(iop=1, ifrot=False)))
        v, c = O.momentum_elements(abc, x.Aeig, x.kl, x.indgk, x.nvk, ipwint, g, i_g0, x.k2cartes, x.nLO_at, x.mult,
                                   vv_t, base.corind, base.core_off, base.cv_tables, 0, base.nomx + 1, base.numin,
                                   base.nbmaxpol)
        wvv.append(v)
        wcv.append(c)
    wvv, wcv = np.array(wvv), np.array(wcv)
    _close(mvv, wvv)
    _close(mcv, wcv)
    want = O.head_from_elements(wvv, wcv, base.Ebnd, base.eig_core, base.corind, base.kcw, base.k_ind, base.weight,
                                base.nomx, base.numin, base.nbmaxpol, base.ncg_p, base.vol)
    _close(head, want)


def test_indices_outside_the_arrays_read_as_nan():
    """A k whose irreducible index is not a row of mmatvv gives a NaN pm row
    (the others unchanged); a k_ind outside kcw's k-points a NaN head."""
    s = synth.head_system(nirk=2)
    rng = np.random.default_rng(6)
    n_e, nv, ncg = s.nbmaxpol - s.numin, s.nomx + 1, len(s.corind)
    vv = torch.as_tensor(rng.normal(size=(2, n_e, nv, 3)) + 0j, device="cuda")
    cv = torch.as_tensor(rng.normal(size=(2, n_e, ncg, 3)) + 0j, device="cuda")
    pm = H.wing_pm(vv, cv, s.Ebnd, s.eig_core, np.array([1, 2, 0]), s.nomx, s.numin, s.nbmaxpol, s.ncg_p,
                   s.vol).cpu().numpy()
    assert np.isnan(pm[1]).all() and not np.isnan(pm[[0, 2]]).any()
    args = list(_head_args(s))
    args[3] = np.array([0, 5])                       # k 5 is not a k-point of kcw (2 of them)
    head = H.head_from_elements(vv, cv, *args).cpu().numpy()
    assert np.isnan(head).all()
