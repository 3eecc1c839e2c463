"""ORACLE (test infrastructure only): numpy FP64 restatement of the
contractions and the self-energy tail of PyGW's polarization path.

  calc_eps_accumulate  chi0 contraction of calc_eps      (pygw.py:3161-3200)
  calc_eps_wings       q = 0 wing of calc_eps            (pygw.py:3177-3193)
  eps_inverse          hermitize, +1, inverse, -1        (pygw.py:3229-3271)
  eps_inverse_head     the same with the q = 0 head/wings update
                                                         (pygw.py:3202-3263)
  calc_mwm             MWM of calc_selfc                 (pygw.py:3329-3366)
  calc_mwm_head        q = 0 head/wing terms of MWM      (pygw.py:3310-3311,3359-3364)
  fr_convolution2 / all_fr_convolutions / few_fr_convolutions
                                                         (fnc.f90:86-176)
  spline_precise       the (w^2+1) cubic spline onto the precise mesh
                                                         (mcommon.py:126-140)
  selfc_spline         spline branch of Compute_selfc_inside
                                                         (mcommon.py:126-144)
  padecof / padeg      Thiele continued fraction         (for_pade.f90:3-51)
  analytic_continuation (iop_ac = 0 branch)              (mcommon.py:315-343)

Array conventions follow the reference: M per k is minm2[(i_o*n_e+i_e), I]
(pygw.py:3170), kcw is [n_o, n_e, N_k, n_om] (Fortran), eps is
[PB, PB, n_om] with eps[:, :, iom] a matrix, mwm is [n_om, nb1, nb2].
"""
from __future__ import annotations

import numpy as np

PI_F = 3.14159265358979  # pi as written in fnc.f90:71,100


def calc_eps_accumulate(minm_k, kcw_k, coef=-2.0):
    """sum_k sum_p coef*kcw[p,k,w] M[k,p,I] conj(M[k,p,J]) -> [PB, PB, n_om].

    minm_k: iterable of (ik, M[N_p, PB]); kcw_k(ik) -> [N_p, n_om] (pair p =
    i_o*n_e + i_e, the ravel order of pygw.py:3182)."""
    eps = None
    for ik, m in minm_k:
        cm = np.conj(m)
        mTc = m.T * coef                  # pygw.py:3173
        w = kcw_k(ik)                     # [N_p, n_om]
        if eps is None:
            eps = np.zeros((m.shape[1], m.shape[1], w.shape[1]), dtype=complex)
        for iom in range(w.shape[1]):
            tmat = mTc * w[:, iom]        # pygw.py:3182
            eps[:, :, iom] += tmat @ cm   # pygw.py:3185
    return eps


def calc_eps_wings(minm_k, kcw_k, pm_k, coef=-2.0):
    """q = 0 wing: epsw1[I, iom] = sum_k dot(tmat, pm) with
    tmat = minm2.T*coef*kcw (pygw.py:3177-3193).  pm_k(ik) -> [N_p]
    momentum matrix elements in the pair order of minm2."""
    epsw1 = None
    for ik, m in minm_k:
        mTc = m.T * coef
        w = kcw_k(ik)
        pm = pm_k(ik)
        if epsw1 is None:
            epsw1 = np.zeros((m.shape[1], w.shape[1]), dtype=complex)
        for iom in range(w.shape[1]):
            tmat = mTc * w[:, iom]
            epsw1[:, iom] += np.dot(tmat, pm)
    return epsw1


def eps_inverse_head(eps, head, epsw1):
    """q = 0 inversion (pygw.py:3202-3263): epsw2 = conj(epsw1); per
    frequency hermitize, +1, inverse A, then
      bw1 = A epsw1, w2b = conj(A epsw2), ww = dot(epsw2, bw1),
      emac = (1 + head - ww, 1 + head), h = 1/(1 + head - ww),
      epsw1 = -h bw1, epsw2 = -h w2b, A += h outer(bw1, w2b), head = h - 1,
    and -1.  Returns (eps, head, epsw1, epsw2, emac)."""
    out = np.empty_like(eps)
    head = np.array(head, dtype=complex)
    epsw1 = np.array(epsw1, dtype=complex)
    epsw2 = np.conj(epsw1)
    nom = eps.shape[2]
    emac = np.zeros((nom, 2), dtype=complex)
    ident = np.eye(eps.shape[0])
    for iom in range(nom):
        a = (eps[:, :, iom] + np.conj(eps[:, :, iom].T)) / 2.0 + ident
        a = np.linalg.inv(a)
        bw1 = np.dot(a, epsw1[:, iom])
        w2b = np.conj(np.dot(a, epsw2[:, iom]))
        ww = np.dot(epsw2[:, iom], bw1)
        emac[iom, 0] = 1.0 + head[iom] - ww
        emac[iom, 1] = 1.0 + head[iom]
        h = 1.0 / (1.0 + head[iom] - ww)
        epsw1[:, iom] = -h * bw1
        epsw2[:, iom] = -h * w2b
        a = a + h * np.multiply.outer(bw1, w2b)
        head[iom] = h - 1.0
        out[:, :, iom] = a - ident
    return out, head, epsw1, epsw2, emac


def eps_inverse(eps):
    """(1 + eps)^-1 - 1 per frequency after hermitizing (pygw.py:3229-3271,
    q != 0: no head/wings)."""
    out = np.empty_like(eps)
    ident = np.eye(eps.shape[0])
    for iom in range(eps.shape[2]):
        a = (eps[:, :, iom] + np.conj(eps[:, :, iom].T)) / 2.0
        a = a + ident
        out[:, :, iom] = np.linalg.inv(a) - ident
    return out


def calc_mwm(minm2, w, wkq):
    """mwm[iom, ib1, ib2] = wkq * sum_IJ conj(M[(b1,b2),I]) W[I,J,iom] M[(b1,b2),J]
    (pygw.py:3353-3358).  minm2: [nb1*nb2, PB] -> [n_om, nb1*nb2]."""
    nom = w.shape[2]
    out = np.empty((nom, minm2.shape[0]), dtype=complex)
    cm = np.conj(minm2)
    for iom in range(nom):
        mw = cm @ np.ascontiguousarray(w[:, :, iom])  # a strided slice would miss BLAS
        out[iom] = wkq * np.sum(mw * minm2, axis=-1)
    return out


def calc_mwm_head(minm, mwm, head, epsw1, epsw2, coefs1, coefs2, ncg_c):
    """q = 0 terms of calc_selfc (pygw.py:3359-3364), added in place to
    mwm[iom, ib1, ib1 + ncg_c] for every ib1 (ie1 - ibgw):
      coefs2 head[iom] + coefs1 dot(minm[ib1, ib1+ncg_c, :], epsw2[:, iom])
                       + coefs1 vdot(minm[ib1, ib1+ncg_c, :], epsw1[:, iom]).
    minm: [nb1, nb2, PB]; coefs1 = singc1 sqrt(4 pi/Vol), coefs2 =
    singc2 4 pi/Vol (pygw.py:3310-3311)."""
    nb1 = minm.shape[0]
    for iom in range(mwm.shape[0]):
        for ib1 in range(nb1):
            ib2 = ib1 + ncg_c
            m = minm[ib1, ib2, :]
            mwm[iom, ib1, ib2] += (coefs2 * head[iom] + coefs1 * np.dot(m, epsw2[:, iom])
                                   + coefs1 * np.vdot(m, epsw1[:, iom]))
    return mwm


def fr_convolution2(enk, mwm, mwm_iom, omega, omega_iom, womeg):
    """fnc.f90:86-119."""
    eps_om = complex(enk, -omega_iom)
    eps_om2 = eps_om * eps_om
    sc0 = np.sum((mwm - mwm_iom) * womeg / (omega ** 2 + eps_om2))
    res = eps_om * sc0 / PI_F
    if abs(mwm_iom) > 1e-16:
        res = res + mwm_iom / PI_F * np.arctan(omega[-1] / eps_om)
    return res


def all_fr_convolutions(enk, mwm, omega, womeg):
    """sc_p[ie1, iom] = sum_ie2 fr_convolution2(...) (fnc.f90:139-159).
    mwm: [n_om, nb1, nb2]."""
    nom, nb1, nb2 = mwm.shape
    sc = np.zeros((nb1, nom), dtype=complex)
    for ie1 in range(nb1):
        for iom in range(nom):
            s = 0.0
            for ie2 in range(nb2):
                s = s + fr_convolution2(enk[ie2], mwm[:, ie1, ie2], mwm[iom, ie1, ie2], omega, omega[iom], womeg)
            sc[ie1, iom] = s
    return sc


def all_fr_convolutions_vec(enk, mwm, omega, womeg):
    """Vectorised all_fr_convolutions (same formula, numpy summation order)."""
    nom, nb1, nb2 = mwm.shape
    z = enk[None, :] - 1j * omega[:, None]                     # [nom(n), nb2]
    den = omega[None, :, None] ** 2 + (z ** 2)[:, None, :]      # [n, l, nb2]
    num = mwm[None, :, :, :] - mwm[:, None, :, :]               # [n, l, nb1, nb2]
    sc0 = np.einsum("nlab,nlb,l->nab", num, 1.0 / den, womeg)
    at = np.arctan(omega[-1] / z)                               # [n, nb2]
    tail = np.where(np.abs(mwm) > 1e-16, mwm / PI_F * at[:, None, :], 0.0)
    res = z[:, None, :] * sc0 / PI_F + tail
    return res.sum(axis=2).T


def few_fr_convolutions(enk, mwm_p, mwm_iom, omega_p, omega_iom, womeg_p):
    """sum_ie2 fr_convolution2(enk[ie2], mwm_p[ie2, :], mwm_iom[ie2], ...)
    (fnc.f90:161-176).  mwm_p: [nb2, n_om_precise]."""
    s = 0.0
    for ie2 in range(len(enk)):
        s = s + fr_convolution2(enk[ie2], mwm_p[ie2], mwm_iom[ie2], omega_p, omega_iom, womeg_p)
    return s


def spline_precise(mwm_col, omega, omega_p):
    """mwm of one (ie1, ie2) on the precise mesh: cubic spline of (w^2+1)*mwm
    with bc ((2,0),(1,0)), extrapolated, divided by (w_p^2+1)
    (mcommon.py:126-140; the imaginary part only when its mean |.| > 1e-10)."""
    from scipy import interpolate

    nom = omega.size
    fa, fp = omega ** 2 + 1.0, omega_p ** 2 + 1.0
    sr = interpolate.CubicSpline(omega, fa * mwm_col.real, bc_type=((2, 0.0), (1, 0.0)), extrapolate=True)
    if np.sum(np.abs(mwm_col.imag)) / nom > 1e-10:
        si = interpolate.CubicSpline(omega, fa * mwm_col.imag, bc_type=((2, 0.0), (1, 0.0)), extrapolate=True)
        return (sr(omega_p) + 1j * si(omega_p)) / fp
    return sr(omega_p) / fp + 0j


def selfc_spline(enk, mwm, omega, omega_p, womeg_p):
    """Spline branch of Compute_selfc_inside (mcommon.py:126-144): cubic
    spline of (w^2+1)*mwm with bc ((2,0),(1,0)) onto the precise mesh, then
    few_fr_convolutions against the coarse-mesh value mwm(w_n)."""
    nom, nb1, nb2 = mwm.shape
    sc = np.zeros((nb1, nom), dtype=complex)
    for ie1 in range(nb1):
        mp = np.array([spline_precise(mwm[:, ie1, ie2], omega, omega_p) for ie2 in range(nb2)])
        for iom in range(nom):
            sc[ie1, iom] = few_fr_convolutions(enk, mp, mwm[iom, ie1, :], omega_p, omega[iom], womeg_p)
    return sc


def padecof(gn, zn):
    """Thiele reciprocal differences (for_pade.f90:3-25)."""
    n = len(gn)
    p = np.zeros((n, n), dtype=complex)
    p[0, :] = gn
    for j in range(1, n):
        for i in range(1, j + 1):
            p[i, j] = (p[i - 1, i - 1] - p[i - 1, j]) / (zn[j] - zn[i - 1]) / p[i - 1, j]
    return np.array([p[j, j] for j in range(n)])


def padeg(x, zn, pt):
    """Continued-fraction evaluation (for_pade.f90:30-51)."""
    n = len(zn)
    out = np.zeros(len(x), dtype=complex)
    for jx, xv in enumerate(x):
        aw = [0j] * (n + 1)
        bw = [0j] * (n + 1)
        aw[0], aw[1] = 0j, pt[0]
        bw[0], bw[1] = 1 + 0j, 1 + 0j
        for i in range(1, n):
            aw[i + 1] = aw[i] + (xv - zn[i - 1]) * pt[i] * aw[i - 1]
            bw[i + 1] = bw[i] + (xv - zn[i - 1]) * pt[i] * bw[i - 1]
        out[jx] = aw[n] / bw[n]
    return out


def analytic_continuation(omega, sc, enk, npts=32):
    """iop_ac = 0 branch of AnalyticContinuation (mcommon.py:330-343):
    Pade through the first min(32, n_om) points; value and central
    difference derivative (+-1e-3) at enk."""
    n = min(npts, len(omega))
    z = omega[:n] * 1j
    y = sc[:n]
    a = padecof(y, z)
    yc = padeg([enk], z, a)[0]
    dyc = (padeg([enk + 1e-3], z, a)[0] - padeg([enk - 1e-3], z, a)[0]) / 2e-3
    return yc, dyc
