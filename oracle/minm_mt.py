"""ORACLE (test infrastructure only): the muffin-tin and core parts of the
product-basis matrix elements, restated from PyGW's Fortran:

* the Gaunt table: ``Gaunt`` (gaunt.f90:105-120), ``cmp_all_Gaunt``
  (gaunt.f90:1-34) and ``getcgcoef`` (gaunt.f90:36-101, for_vxcnn.f90:407-470);
* ``calc_minm_MT`` (for_vxcnn.f90:839-1028, called at pygw.py:2655);
* ``calc_minc`` and ``calc_minc2`` (for_vxcnn.f90:1136-1397, pygw.py:2677-2680).

Arrays keep PyGW's index order (the f2py arrays, Fortran semantics, 0-based
here): alfa[nbmax, (lmax+1)^2, ndf], gama[nbmax, nLOmax, (lmax+1)^2, ndf],
s3r[n_mix, n_fnc, n_fnc, nat], big_l[max_nmix, nat], nLO_at[lomax+1, nat],
corind[ncg, 5]; mmat[n1, n2, loctmatsize].
"""
from __future__ import annotations

from fractions import Fraction
from math import factorial

import numpy as np

PI_F = 3.14159265358979     # the value of pi these Fortran routines carry


def wigner3j(j1, j2, j3, m1, m2, m3):
    """The Wigner 3j symbol by Racah's formula in exact rational arithmetic."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    tri = Fraction(factorial(j1 + j2 - j3) * factorial(j1 - j2 + j3) * factorial(-j1 + j2 + j3),
                   factorial(j1 + j2 + j3 + 1))
    pre = tri * factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) * factorial(j2 - m2) \
        * factorial(j3 + m3) * factorial(j3 - m3)
    s = Fraction(0)
    for t in range(max(0, j2 - j3 - m1, j1 - j3 + m2), min(j1 + j2 - j3, j1 - m1, j2 + m2) + 1):
        s += Fraction((-1) ** t, factorial(t) * factorial(j3 - j2 + t + m1) * factorial(j3 - j1 + t - m2)
                      * factorial(j1 + j2 - j3 - t) * factorial(j1 - t - m1) * factorial(j2 - t + m2))
    sign = -1 if (j1 - j2 - m3) % 2 else 1
    # sqrt of the exact rational prefactor, then the exact sum
    return sign * float(s) * np.sqrt(float(pre))


def gaunt(l1, m1, l2, m2, l3, m3):
    """<Y_l1m1 | Y_l2m2 | Y_l3m3> = int Y*_l1m1 Y_l2m2 Y_l3m3 (gaunt.f90:105-120)."""
    return ((-1) ** (m1 % 2) * np.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * np.pi))
            * wigner3j(l1, l2, l3, 0, 0, 0) * wigner3j(l1, l2, l3, -m1, m2, m3))


def cmp_all_gaunt(maxj):
    """cmp_All_Gaunt (gaunt.f90:1-34): for l2 <= l1 <= maxj, l3 in
    [l1-l2, l1+l2], -l1 <= m1 <= l1, 0 <= m2 <= l2: <Y_l3,m1+m2 | Y_l2m2 | Y_l1m1>
    (0 when l1 + l2 + l3 is odd or |m1 + m2| > l3)."""
    out = []
    for l1 in range(maxj + 1):
        for l2 in range(l1 + 1):
            for l3 in range(l1 - l2, l1 + l2 + 1):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(l2 + 1):
                        m3 = m1 + m2
                        out.append(gaunt(l3, m3, l2, m2, l1, m1) if (l1 + l2 + l3) % 2 == 0 and abs(m3) <= l3
                                   else 0.0)
    return np.array(out)


def getcgcoef(l1, l2, l3, m1, m2, cg):
    """getcgcoef (gaunt.f90:36-101): <Y_l3,m1+m2 | Y_l1m1 Y_l2m2> from the table."""
    par = abs(l1 + l2 - l3) % 2
    if not (abs(m1 + m2) <= l3 and abs(l1 - l2) <= l3 and l1 + l2 >= l3):
        return 0.0
    if l1 < l2:
        j1, mj1, j2, mj2 = l2, m2, l1, m1
    else:
        j1, mj1, j2, mj2 = l1, m1, l2, m2
    fact = 1.0
    if mj2 < 0:
        mj2, mj1, fact = -mj2, -mj1, -2.0 * par + 1.0
    ind1 = (16 * j1 * j1 - 3 * j1 - 3) * (j1 + 2) * (j1 + 1) * j1 // 60
    ind2 = j1 * j2 * (j2 + 1) * (4 * j2 - 1) // 3
    ind3 = j2 * (j2 - 1) * (4 * j2 + 7) // 6
    ind4 = (2 * j1 + 1) * (j2 + 1) * (l3 - j1 + j2)
    ing = ind1 + ind2 + ind3 + ind4 + (j2 + 1) * (mj1 + j1) + mj2 + j2 + 1
    return fact * cg[ing - 1]


def _lo_start(nLO_at, iat, l, ncore, lmax):
    """0-based s3r index of the first local orbital of angular momentum l."""
    return ncore + 2 * (lmax + 1) + int(np.sum(nLO_at[:l, iat]))


def _nlo(nLO_at, l, iat):
    """nLO_at(l+1, iat); l above lomax has no local orbitals (the Fortran
    reads past the array's first extent there, so PyGW runs with lomax >=
    lmax)."""
    return int(nLO_at[l, iat]) if l < nLO_at.shape[0] else 0


def _im_offsets(mult, nmix, big_l):
    """locmixind3: the first product-basis index of (irm, idf)."""
    off, im = {}, 0
    idf = 0
    for iat in range(len(mult)):
        for _ in range(mult[iat]):
            for irm in range(nmix[iat]):
                off[(irm, idf)] = im
                im += 2 * int(big_l[irm, iat]) + 1
            idf += 1
    return off


def calc_minm_mt(ql, nstart, nend, mstart, mend, alfa, beta, gama, alfp, betp, gamp, s3r, pos, mult, nmix, big_l,
                 nLO_at, ncore, cgcoef, lmax, loctmatsize):
    """calc_minm_MT (for_vxcnn.f90:839-1028):
      mmat[ie1, ie2, im] = < u_im Y_lb,mb | psi*_ie2 psi_ie1 > e^{-2 pi i ql.pos}
    with, per product function (irm, lb, mb) of atom idf and per (l1, m1,
    l2, m2 = m1 - mb), the Gaunt coefficient <Y_l1m1 | Y_l2m2 Y_lbmb> times
    s3r of the radial channels (u, udot, local orbitals) of l1 and l2, the
    coefficients alfa/beta/gama of band ie1 at (l1, m1) and alfp/betp/gamp of
    band ie2 at (l2, m2)."""
    n1, n2 = nend - nstart, mend - mstart
    mmat = np.zeros((n1, n2, loctmatsize), dtype=complex)
    big_l = np.asarray(big_l)
    nLO_at = np.asarray(nLO_at)
    pos = np.asarray(pos, dtype=float)
    idf, im = 0, 0
    for iat in range(len(mult)):
        nc = int(ncore[iat])
        for _ in range(mult[iat]):
            phs = np.exp(-np.dot(ql, pos[:, idf]) * 2.0 * PI_F * 1j)
            for irm in range(nmix[iat]):
                lb = int(big_l[irm, iat])
                for mb in range(-lb, lb + 1):
                    # na[ie1, lm2], nb[ie1, lm2], nc[(l2, m2, ilo2), ie1]
                    na = np.zeros((n1, (lmax + 1) ** 2), dtype=complex)
                    nb = np.zeros_like(na)
                    ncl = {}
                    for l1 in range(lmax + 1):
                        la1, lb1 = nc + l1, nc + l1 + lmax + 1
                        lc10 = _lo_start(nLO_at, iat, l1, nc, lmax)
                        for l2 in range(abs(lb - l1), min(lb + l1, lmax) + 1):
                            la2, lb2 = nc + l2, nc + l2 + lmax + 1
                            lc20 = _lo_start(nLO_at, iat, l2, nc, lmax)
                            for m1 in range(-l1, l1 + 1):
                                lm1 = l1 * l1 + l1 + m1
                                m2 = m1 - mb
                                if abs(m2) > l2:
                                    continue
                                lm2 = l2 * l2 + l2 + m2
                                ang = getcgcoef(l2, lb, l1, m2, mb, cgcoef)
                                if abs(ang) < 1e-8:
                                    continue
                                pa = phs * ang
                                a1 = alfa[nstart:nend, lm1, idf]
                                b1 = beta[nstart:nend, lm1, idf]
                                g1 = [gama[nstart:nend, ilo1, lm1, idf] for ilo1 in range(_nlo(nLO_at, l1, iat))]
                                na[:, lm2] += pa * s3r[irm, la1, la2, iat] * a1 + pa * s3r[irm, lb1, la2, iat] * b1
                                nb[:, lm2] += pa * s3r[irm, la1, lb2, iat] * a1 + pa * s3r[irm, lb1, lb2, iat] * b1
                                for ilo1 in range(_nlo(nLO_at, l1, iat)):
                                    na[:, lm2] += pa * s3r[irm, lc10 + ilo1, la2, iat] * g1[ilo1]
                                    nb[:, lm2] += pa * s3r[irm, lc10 + ilo1, lb2, iat] * g1[ilo1]
                                for ilo2 in range(_nlo(nLO_at, l2, iat)):
                                    ccsm = sum(pa * s3r[irm, lc10 + ilo1, lc20 + ilo2, iat] * g1[ilo1]
                                               for ilo1 in range(_nlo(nLO_at, l1, iat)))
                                    key = (l2, m2, ilo2)
                                    ncl[key] = ncl.get(key, 0) + ccsm + pa * s3r[irm, la1, lc20 + ilo2, iat] * a1 \
                                        + pa * s3r[irm, lb1, lc20 + ilo2, iat] * b1
                    res = na @ alfp[mstart:mend, :, idf].T + nb @ betp[mstart:mend, :, idf].T
                    for (l2, m2, ilo2), v in ncl.items():
                        res += np.outer(v, gamp[mstart:mend, ilo2, l2 * l2 + l2 + m2, idf])
                    mmat[:, :, im] = res
                    im += 1
            idf += 1
    return mmat


def calc_minc(kl, nstart, nend, cstart, cend, alfa, beta, gama, s3r, corind, pos, mult, nmix, big_l, nLO_at, ncore,
              cgcoef, lmax, loctmatsize):
    """calc_minc (for_vxcnn.f90:1136-1273): mmat[ie1, ic, im] = < u_im Y_lb,mb |
    psi_ie1 psi*_core > e^{-2 pi i kl.pos} for the core states cstart..cend-1
    (corind rows: atom type, atom, core index, l, m), with m1 = mc + mb and
    the Gaunt coefficient getcgcoef(lc, lb, l1, mc, mb)."""
    n1 = nend - nstart
    mmat = np.zeros((n1, cend - cstart, loctmatsize), dtype=complex)
    off = _im_offsets(mult, nmix, big_l)
    nLO_at = np.asarray(nLO_at)
    for icg in range(cstart, cend):
        iat, idf, ic, l2, m2 = [int(x) for x in corind[icg]]
        phs = np.exp(-np.dot(kl, np.asarray(pos)[:, idf]) * 2.0 * PI_F * 1j)
        nc = int(ncore[iat])
        for irm in range(nmix[iat]):
            lb = int(big_l[irm, iat])
            for mb in range(-lb, lb + 1):
                im = off[(irm, idf)] + lb + mb
                for l1 in range(abs(lb - l2), min(lb + l2, lmax) + 1):
                    la1, lb1 = nc + l1, nc + l1 + lmax + 1
                    lc10 = _lo_start(nLO_at, iat, l1, nc, lmax)
                    m1 = m2 + mb
                    ang = getcgcoef(l2, lb, l1, m2, mb, cgcoef)
                    if abs(m1) > l1 or abs(ang) < 1e-8:
                        continue
                    lm1 = l1 * l1 + l1 + m1
                    pa = phs * ang
                    csm = pa * s3r[irm, ic, la1, iat] * alfa[nstart:nend, lm1, idf] \
                        + pa * s3r[irm, ic, lb1, iat] * beta[nstart:nend, lm1, idf]
                    for ilo1 in range(_nlo(nLO_at, l1, iat)):
                        csm = csm + pa * s3r[irm, ic, lc10 + ilo1, iat] * gama[nstart:nend, ilo1, lm1, idf]
                    mmat[:, icg - cstart, im] += csm
    return mmat


def calc_minc2(kmql, cstart, cend, mstart, mend, alfp, betp, gamp, s3r, corind, pos, mult, nmix, big_l, nLO_at,
               ncore, cgcoef, lmax, loctmatsize):
    """calc_minc2 (for_vxcnn.f90:1275-1397): mmat[ic, ie2, im] = < u_im Y_lb,mb |
    psi_core psi*_ie2 > e^{+2 pi i kmql.pos}, with m1 = mc - mb and the Gaunt
    coefficient getcgcoef(l1, lb, lc, m1, mb)."""
    n2 = mend - mstart
    mmat = np.zeros((cend - cstart, n2, loctmatsize), dtype=complex)
    off = _im_offsets(mult, nmix, big_l)
    nLO_at = np.asarray(nLO_at)
    for icg in range(cstart, cend):
        iat, idf, ic, l2, m2 = [int(x) for x in corind[icg]]
        phs = np.exp(np.dot(kmql, np.asarray(pos)[:, idf]) * 2.0 * PI_F * 1j)
        nc = int(ncore[iat])
        for irm in range(nmix[iat]):
            lb = int(big_l[irm, iat])
            for mb in range(-lb, lb + 1):
                im = off[(irm, idf)] + lb + mb
                for l1 in range(abs(lb - l2), min(lb + l2, lmax) + 1):
                    la1, lb1 = nc + l1, nc + l1 + lmax + 1
                    lc10 = _lo_start(nLO_at, iat, l1, nc, lmax)
                    m1 = m2 - mb
                    ang = getcgcoef(l1, lb, l2, m1, mb, cgcoef)
                    if abs(m1) > l1 or abs(ang) < 1e-8:
                        continue
                    lm1 = l1 * l1 + l1 + m1
                    pa = phs * ang
                    csm = pa * s3r[irm, ic, la1, iat] * alfp[mstart:mend, lm1, idf] \
                        + pa * s3r[irm, ic, lb1, iat] * betp[mstart:mend, lm1, idf]
                    for ilo1 in range(_nlo(nLO_at, l1, iat)):
                        csm = csm + pa * s3r[irm, ic, lc10 + ilo1, iat] * gamp[mstart:mend, ilo1, lm1, idf]
                    mmat[icg - cstart, :, im] += csm
    return mmat
