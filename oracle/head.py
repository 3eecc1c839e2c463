"""ORACLE (test infrastructure only): the momentum matrix elements and the
head of the dielectric matrix at q = 0, restated from PyGW:

* ``calcmmatvv_mt2`` / ``calcmmatvv_mt`` (for_q_0.f90:236-458 / 10-234):
  <psi_n1| -i nabla |psi_n2> inside the spheres from the band coefficients
  and the radial integrals of d/dr - l/r and d/dr + (l+2)/r;
* ``calcmmatvv_is`` (for_q_0.f90:460-507): the interstitial part, sum over
  plane-wave pairs of conj(A) A (k+G2) ipwint(G2-G1);
* ``calcmmatcv`` (for_q_0.f90:510-683): core-valence elements, the average
  of <core|p|psi> and conj(<psi|p|core>);
* ``calc_head`` (pygw.py:2808-3037): the loop over irreducible k building
  mmatvv[irk, ie2-mst, ie1-nst, 3] (the Hermitian average of the two
  orders) and mmatcv[irk, ie2-mst, icg, 3], then the head
  -4 pi fspin kwt/Vol sum |p|^2/3/(E_c - E_v)^2 kcw(omega), its core part,
  the metallic intraband term and the Drude plasmon term;
* ``wing_pm`` (pygw.py:3122-3141): pm[ie12] of calc_eps's q = 0 wing.

Index conventions are 0-based; arrays keep the f2py shapes: alfa/beta
[nbmax, (lmax+1)^2, ndf], gama [nbmax, nLOmax, (lmax+1)^2, ndf], nLO_at
[lomax+1, nat], iul_ul & co [2, lmax, nat], iulol_ul/iulol_udl/iul_ulol/
iudl_ulol [2, nlo, lomx, nat], iulol_ulol [2, nlo, nlo, lomx, nat].  Core
tables are concatenated over atoms in corind's order of core index: iul_ucl
& co [2, ncore_total], iulol_ucl / iucl_ulol [2, nloat, ncore_total] (the
first local-orbital slot of PyGW's per-atom arrays dropped, as calc_head
passes them, pygw.py:2918-2920).  A local-orbital count nLO_at[l] for an l
beyond the table's rows is zero (the Fortran reads past the array there, but
every such term multiplies a gama entry that gap2_set_lapwcoef leaves 0).
"""
from __future__ import annotations

import numpy as np


def _lm(l, m):
    return l * (l + 1) + m


def _nlo(nLO_at, l, iat):
    return int(nLO_at[l, iat]) if 0 <= l < nLO_at.shape[0] else 0


def calcmmatvv_mt2(ca1, cb1, cg1, a2, b2, g2, nLO_at, mult, iul_ul, iul_udl, iudl_ul, iudl_udl, iulol_ul,
                   iulol_udl, iul_ulol, iudl_ulol, iulol_ulol):
    """for_q_0.f90:236-458: mt12[3] from the conjugated coefficients of band
    n1 (ca1/cb1 [ntnt, ndf], cg1 [nLOmax, ntnt, ndf]) and those of band n2."""
    lmax = iul_ul.shape[1]
    pxpy = pxmy = pz = 0j
    idf = 0
    for iat in range(len(mult)):
        for _ in range(mult[iat]):
            for lm in range(lmax * lmax):
                l = int(np.sqrt(lm) + 1e-6)
                m = lm - l * (l + 1)
                den = float((2 * l + 1) * (2 * l + 3))
                l1mm1, l1m, l1m1 = _lm(l + 1, m - 1), _lm(l + 1, m), _lm(l + 1, m + 1)
                flm1 = -np.sqrt((l + m + 1) * (l + m + 2) / den)
                flm3 = np.sqrt((l - m + 1) * (l - m + 2) / den)
                flm5 = np.sqrt((l - m + 1) * (l + m + 1) / den)
                r1 = (iul_ul[0, l, iat], iul_udl[0, l, iat], iudl_ul[0, l, iat], iudl_udl[0, l, iat])
                r2 = (iul_ul[1, l, iat], iul_udl[1, l, iat], iudl_ul[1, l, iat], iudl_udl[1, l, iat])

                def ab(x, y, r):   # <A x u + B x udot | op | A y u + B y udot>
                    (xa, xb), (ya, yb) = x, y
                    return xa * ya * r[0] + xa * yb * r[1] + xb * ya * r[2] + xb * yb * r[3]

                n1 = lambda t: (ca1[t, idf], cb1[t, idf])  # noqa: E731
                n2 = lambda t: (a2[t, idf], b2[t, idf])    # noqa: E731
                xpy1 = ab(n1(l1m1), n2(lm), r1)
                xpy2 = ab(n1(lm), n2(l1mm1), r2)
                xmy1 = ab(n1(l1mm1), n2(lm), r1)
                xmy2 = ab(n1(lm), n2(l1m1), r2)
                z1 = ab(n1(l1m), n2(lm), r1)
                z2 = ab(n1(lm), n2(l1m), r2)
                for jlo in range(_nlo(nLO_at, l, iat)):
                    g = g2[jlo, lm, idf]
                    for (t, acc) in ((l1m1, "xpy1"), (l1mm1, "xmy1"), (l1m, "z1")):
                        v = ca1[t, idf] * g * iul_ulol[0, jlo, l, iat] + cb1[t, idf] * g * iudl_ulol[0, jlo, l, iat]
                        if acc == "xpy1":
                            xpy1 += v
                        elif acc == "xmy1":
                            xmy1 += v
                        else:
                            z1 += v
                for jlo in range(_nlo(nLO_at, l + 1, iat)):
                    a, b = ca1[lm, idf], cb1[lm, idf]
                    xpy2 += a * g2[jlo, l1mm1, idf] * iul_ulol[1, jlo, l, iat] + \
                        b * g2[jlo, l1mm1, idf] * iudl_ulol[1, jlo, l, iat]
                    xmy2 += a * g2[jlo, l1m1, idf] * iul_ulol[1, jlo, l, iat] + \
                        b * g2[jlo, l1m1, idf] * iudl_ulol[1, jlo, l, iat]
                    z2 += a * g2[jlo, l1m, idf] * iul_ulol[1, jlo, l, iat] + \
                        b * g2[jlo, l1m, idf] * iudl_ulol[1, jlo, l, iat]
                for ilo in range(_nlo(nLO_at, l + 1, iat)):
                    a, b = a2[lm, idf], b2[lm, idf]
                    xpy1 += cg1[ilo, l1m1, idf] * a * iulol_ul[0, ilo, l, iat] + \
                        cg1[ilo, l1m1, idf] * b * iulol_udl[0, ilo, l, iat]
                    xmy1 += cg1[ilo, l1mm1, idf] * a * iulol_ul[0, ilo, l, iat] + \
                        cg1[ilo, l1mm1, idf] * b * iulol_udl[0, ilo, l, iat]
                    z1 += cg1[ilo, l1m, idf] * a * iulol_ul[0, ilo, l, iat] + \
                        cg1[ilo, l1m, idf] * b * iulol_udl[0, ilo, l, iat]
                for ilo in range(_nlo(nLO_at, l, iat)):
                    c = cg1[ilo, lm, idf]
                    xpy2 += c * a2[l1mm1, idf] * iulol_ul[1, ilo, l, iat] + c * b2[l1mm1, idf] * iulol_udl[1, ilo, l, iat]
                    xmy2 += c * a2[l1m1, idf] * iulol_ul[1, ilo, l, iat] + c * b2[l1m1, idf] * iulol_udl[1, ilo, l, iat]
                    z2 += c * a2[l1m, idf] * iulol_ul[1, ilo, l, iat] + c * b2[l1m, idf] * iulol_udl[1, ilo, l, iat]
                for jlo in range(_nlo(nLO_at, l, iat)):
                    for ilo in range(_nlo(nLO_at, l + 1, iat)):
                        i1 = iulol_ulol[0, ilo, jlo, l, iat]
                        i2 = iulol_ulol[1, jlo, ilo, l, iat]
                        xpy1 += cg1[ilo, l1m1, idf] * g2[jlo, lm, idf] * i1
                        xmy1 += cg1[ilo, l1mm1, idf] * g2[jlo, lm, idf] * i1
                        z1 += cg1[ilo, l1m, idf] * g2[jlo, lm, idf] * i1
                        xpy2 += cg1[jlo, lm, idf] * g2[ilo, l1mm1, idf] * i2
                        xmy2 += cg1[jlo, lm, idf] * g2[ilo, l1m1, idf] * i2
                        z2 += cg1[jlo, lm, idf] * g2[ilo, l1m, idf] * i2
                pxpy += flm1 * xpy1 + flm3 * xpy2
                pxmy += flm3 * xmy1 + flm1 * xmy2
                pz += flm5 * (z1 + z2)
            idf += 1
    return np.array([-(pxpy + pxmy) * 1j / 2.0, (-pxpy + pxmy) / 2.0, -pz * 1j])


def _zg(gama, n):
    """Band n of gama, or PyGW's (1, 1, 1, 1) placeholder without local orbitals."""
    return gama[n] if n < gama.shape[0] else gama[0]


def calcmmatvv_mt(n1, n2, alfa, beta, gama, nLO_at, mult, *tables):
    """for_q_0.f90:10-234: the same element with the bands taken by index."""
    return calcmmatvv_mt2(np.conj(alfa[n1]), np.conj(beta[n1]), np.conj(_zg(gama, n1)), alfa[n2], beta[n2],
                          _zg(gama, n2), nLO_at, mult, *tables)


def calcmmatvv_is(i, j, kl, Ak, indgk, nvk, ipwint, gindex, i_g0, k2cartes):
    """for_q_0.f90:460-507: sum_G2 A[j,G2] (k+G2) sum_G1 ipwint(G2-G1) conj(A[i,G1])."""
    kcar = k2cartes @ np.asarray(kl, dtype=np.float64)
    G = np.asarray(gindex)[np.asarray(indgk)[:nvk]]
    kpG = kcar[None, :] + G @ k2cartes.T
    n = [(i_g0.shape[c] - 1) // 2 for c in range(3)]
    dG = G[:, None, :] - G[None, :, :]
    idg = i_g0[dG[..., 0] + n[0], dG[..., 1] + n[1], dG[..., 2] + n[2]]
    intmom = (ipwint[idg] * np.conj(Ak[i, :nvk])[None, :]).sum(axis=1)
    return (kpG * (Ak[j, :nvk] * intmom)[:, None]).sum(axis=0)


def calcmmatcv(iat, iv, corind, core_off, alfa, beta, gama, iul_ucl, iudl_ucl, iucl_ul, iucl_udl, iulol_ucl,
               iucl_ulol, nLO_at, exact=False):
    """for_q_0.f90:510-683 for band iv and the core states of atom type iat
    (rows of other types stay 0).  core_off[iat] + ic indexes the
    concatenated core tables.  The valence-to-core x-y term with a local
    orbital at l-1 multiplies iudl_ucl, not iulol_ucl (for_q_0.f90:648), as
    the reference does; exact=True takes iulol_ucl there."""
    ncg = corind.shape[0]
    mcv = np.zeros((ncg, 3), dtype=complex)
    A = lambda t, d: alfa[iv, t, d]  # noqa: E731
    B = lambda t, d: beta[iv, t, d]  # noqa: E731
    Gm = lambda o, t, d: _zg(gama, iv)[o, t, d]  # noqa: E731
    for icg in range(ncg):
        if corind[icg, 0] != iat:
            continue
        idf, ic, l, m = (int(x) for x in corind[icg, 1:5])
        c = core_off[iat] + ic
        d1, d2 = float((2 * l + 1) * (2 * l + 3)), float((2 * l - 1) * (2 * l + 1))
        lm1mm1, lm1m, lm1m1 = _lm(l - 1, m - 1), _lm(l - 1, m), _lm(l - 1, m + 1)
        l1mm1, l1m, l1m1 = _lm(l + 1, m - 1), _lm(l + 1, m), _lm(l + 1, m + 1)
        flm1 = -np.sqrt((l + m + 1) * (l + m + 2) / d1)
        # (products of consecutive integers: never negative; at l = 0 they are
        # 0 over d2 = -1)
        flm2 = np.sqrt((l - m - 1) * (l - m) / d2)
        flm3 = np.sqrt((l - m + 1) * (l - m + 2) / d1)
        flm4 = -np.sqrt((l + m - 1) * (l + m) / d2)
        flm5 = np.sqrt((l - m + 1) * (l + m + 1) / d1)
        flm6 = np.sqrt((l - m) * (l + m) / d2)
        xpy1 = xmy1 = z1 = 0j
        nlm1 = _nlo(nLO_at, l - 1, iat) if l > 0 else 0
        nlp1 = _nlo(nLO_at, l + 1, iat)
        if l > 0 and l + m - 2 >= 0:
            xpy1 = A(lm1mm1, idf) * iucl_ul[0, c] + B(lm1mm1, idf) * iucl_udl[0, c]
            for o in range(nlm1):
                xpy1 += Gm(o, lm1mm1, idf) * iucl_ulol[0, o, c]
        if l > 0 and l - m - 2 >= 0:
            xmy1 = A(lm1m1, idf) * iucl_ul[0, c] + B(lm1m1, idf) * iucl_udl[0, c]
            for o in range(nlm1):
                xmy1 += Gm(o, lm1m1, idf) * iucl_ulol[0, o, c]
        if l > 0 and l + m - 1 >= 0 and l - m - 1 >= 0:
            z1 = A(lm1m, idf) * iucl_ul[0, c] + B(lm1m, idf) * iucl_udl[0, c]
            for o in range(nlm1):
                z1 += Gm(o, lm1m, idf) * iucl_ulol[0, o, c]
        xpy2 = A(l1mm1, idf) * iucl_ul[1, c] + B(l1mm1, idf) * iucl_udl[1, c]
        xmy2 = A(l1m1, idf) * iucl_ul[1, c] + B(l1m1, idf) * iucl_udl[1, c]
        z2 = A(l1m, idf) * iucl_ul[1, c] + B(l1m, idf) * iucl_udl[1, c]
        for o in range(nlp1):
            xpy2 += Gm(o, l1mm1, idf) * iucl_ulol[1, o, c]
            xmy2 += Gm(o, l1m1, idf) * iucl_ulol[1, o, c]
            z2 += Gm(o, l1m, idf) * iucl_ulol[1, o, c]
        pxpy, pxmy, pz = flm4 * xpy1 + flm3 * xpy2, flm2 * xmy1 + flm1 * xmy2, flm6 * z1 + flm5 * z2
        mxcv, mycv, mzcv = -(pxpy + pxmy) * 1j / 2.0, -(pxpy - pxmy) / 2.0, -1j * pz
        cA = lambda t: np.conj(A(t, idf))  # noqa: E731
        cB = lambda t: np.conj(B(t, idf))  # noqa: E731
        cG = lambda o, t: np.conj(Gm(o, t, idf))  # noqa: E731
        xpy1 = cA(l1m1) * iul_ucl[0, c] + cB(l1m1) * iudl_ucl[0, c]
        xmy1 = cA(l1mm1) * iul_ucl[0, c] + cB(l1mm1) * iudl_ucl[0, c]
        z1 = cA(l1m) * iul_ucl[0, c] + cB(l1m) * iudl_ucl[0, c]
        for o in range(nlp1):
            xpy1 += cG(o, l1m1) * iulol_ucl[0, o, c]
            xmy1 += cG(o, l1mm1) * iulol_ucl[0, o, c]
            z1 += cG(o, l1m) * iulol_ucl[0, o, c]
        xpy2 = xmy2 = z2 = 0j
        if l > 0 and l - m - 2 >= 0:
            xpy2 = cA(lm1m1) * iul_ucl[1, c] + cB(lm1m1) * iudl_ucl[1, c]
            for o in range(nlm1):
                xpy2 += cG(o, lm1m1) * iulol_ucl[1, o, c]
        if l > 0 and l + m - 2 >= 0:
            xmy2 = cA(lm1mm1) * iul_ucl[1, c] + cB(lm1mm1) * iudl_ucl[1, c]
            for o in range(nlm1):
                xmy2 += cG(o, lm1mm1) * (iulol_ucl[1, o, c] if exact else iudl_ucl[1, c])
        if l > 0 and l + m - 1 >= 0 and l - m - 1 >= 0:
            z2 = cA(lm1m) * iul_ucl[1, c] + cB(lm1m) * iudl_ucl[1, c]
            for o in range(nlm1):
                z2 += cG(o, lm1m) * iulol_ucl[1, o, c]
        pxpy, pxmy, pz = flm1 * xpy1 + flm2 * xpy2, flm3 * xmy1 + flm4 * xmy2, flm5 * z1 + flm6 * z2
        mxvc, myvc, mzvc = -(pxpy + pxmy) * 1j / 2.0, -(pxpy - pxmy) / 2.0, -1j * pz
        mcv[icg] = [(mxcv + np.conj(mxvc)) / 2, (mycv + np.conj(myvc)) / 2, (mzcv + np.conj(mzvc)) / 2]
    return mcv


def momentum_elements(abc, Aeig, kl, indgk, nvk, ipwint, gindex, i_g0, k2cartes, nLO_at, mult, vv_tables,
                      corind, core_off, cv_tables, nst, nend, mst, mend):
    """One irreducible k of calc_head's first loop (pygw.py:2840-2932):
    mmatvv[ie2-mst, ie1-nst, 3] = (p_mt + p_is + conj(p'_mt + p'_is))/2 with
    p = <ie1|p|ie2>, p' = <ie2|p|ie1> (Re(p_mt + p_is) on the diagonal), and
    mmatcv[ie2-mst, icg, 3] from calcmmatcv of each core state's atom."""
    alfa, beta, gama = abc
    vv = np.zeros((mend - mst, nend - nst, 3), dtype=complex)
    for ie2 in range(mst, mend):
        for ie1 in range(nst, nend):
            p = calcmmatvv_mt(ie1, ie2, alfa, beta, gama, nLO_at, mult, *vv_tables) + \
                calcmmatvv_is(ie1, ie2, kl, Aeig, indgk, nvk, ipwint, gindex, i_g0, k2cartes)
            if ie1 == ie2:
                vv[ie2 - mst, ie1 - nst] = p.real
            else:
                ph = calcmmatvv_mt(ie2, ie1, alfa, beta, gama, nLO_at, mult, *vv_tables) + \
                    calcmmatvv_is(ie2, ie1, kl, Aeig, indgk, nvk, ipwint, gindex, i_g0, k2cartes)
                vv[ie2 - mst, ie1 - nst] = (p + np.conj(ph)) / 2.0
    ncg = corind.shape[0]
    cv = np.zeros((mend - mst, ncg, 3), dtype=complex)
    for ie2 in range(mst, mend):
        for iat in sorted(set(int(a) for a in corind[:, 0])):
            cv[ie2 - mst] += calcmmatcv(iat, ie2, corind, core_off, alfa, beta, gama, *cv_tables, nLO_at)
    return vv, cv


def head_from_elements(mmatvv, mmatcv, Ebnd, eig_core, corind, kcw, k_ind, weight, nomx, numin, nbmaxpol, ncg_p,
                       vol, Eg=1.0, kwfer=None, iop_drude=0, eta_head=0.0, omega=None, dUl=None):
    """The second half of calc_head (pygw.py:2940-3029): head[nom_nl] from
    mmatvv [nirk, nbmaxpol-numin, nomx+1, 3], mmatcv [nirk, ..., ncg, 3],
    Ebnd [nirk, nbnd] (no core), eig_core[icg] per core state, kcw
    [n_o, n_e, N_k, nom_nl] (Fortran semantics), k_ind[irk] = ik."""
    nirk = mmatvv.shape[0]
    fspin = 2.0
    nom_nl = kcw.shape[3]
    head = np.zeros(nom_nl, dtype=complex)
    nvbm, ncbm = nomx + 1, numin
    mst, nst = numin, 0
    c0_head = 0.0
    for irk in range(nirk):
        ik = int(k_ind[irk])
        coef = 4.0 * np.pi * weight[irk] * fspin / vol
        if Eg <= 0 and kwfer is not None:
            for ie in range(ncbm, nvbm):
                v = mmatvv[irk, ie - mst, ie - nst, :]
                c0_head += coef * np.vdot(v, v).real / 3.0 * kwfer[irk, ie]
        for icg in range(ncg_p):
            edif = Ebnd[irk, ncbm:nbmaxpol] - eig_core[icg]
            mm = mmatcv[irk, (ncbm - mst):(nbmaxpol - mst), icg, :]
            mm2 = np.sum(mm * np.conj(mm), axis=1).real / 3.0
            head -= coef * np.conj(np.dot(mm2 / edif ** 2, kcw[icg, :(nbmaxpol - ncbm), ik, :]))
        mm2 = np.sum(np.abs(mmatvv[irk, (ncbm - mst):(nbmaxpol - mst), 0:nvbm, :]) ** 2, axis=2) / 3.0
        termvv = np.zeros((nbmaxpol - ncbm, nvbm))
        for ie2 in range(nbmaxpol - ncbm):
            for ie1 in range(nvbm):
                edif = Ebnd[irk, ncbm + ie2] - Ebnd[irk, ie1]
                termvv[ie2, ie1] = 0.0 if abs(edif) < 1e-5 else mm2[ie2, ie1] / edif ** 2
        for iom in range(nom_nl):
            kk = kcw[ncg_p:(ncg_p + nvbm), :(nbmaxpol - ncbm), ik, iom]
            head[iom] -= coef * np.sum(termvv * kk.T)
    if Eg <= 0 and iop_drude == 1:
        w = c0_head / (np.asarray(omega) + eta_head) ** 2
        head += np.dot(w, dUl) if dUl is not None else w
    return head


def wing_pm(mmatvv_k, mmatcv_k, enk, ncg_p, nomx, numin, nbmaxpol, vol):
    """calc_eps's pm at q = 0 (pygw.py:3122-3141) for one k: pm[ie12], ie12 =
    (core or valence row) x (empty band), = sum_c p_c coefw/sqrt(3)/(e1 - e2)
    with coefw = 2 sqrt(pi/Vol), 0 where |e1 - e2| <= 1e-10; enk holds the
    core energies first, then the bands."""
    coefw = 2.0 * np.sqrt(np.pi / vol)
    nvbm, ncbm = nomx + 1, numin
    mmc = np.sum(mmatcv_k[:(nbmaxpol - numin)], axis=2) * (coefw / np.sqrt(3.0))
    mmv = np.sum(mmatvv_k[:(nbmaxpol - numin)], axis=2) * (coefw / np.sqrt(3.0))
    pm = np.zeros((ncg_p + nvbm) * (nbmaxpol - ncbm), dtype=complex)
    ie12 = 0
    for ie1 in range(ncg_p):
        for ie2 in range(ncbm, nbmaxpol):
            edif = enk[ie1] - enk[ie2 + ncg_p]
            if abs(edif) > 1e-10:
                pm[ie12] = mmc[ie2 - ncbm, ie1] / edif
            ie12 += 1
    for ie1 in range(nvbm):
        for ie2 in range(ncbm, nbmaxpol):
            edif = enk[ie1 + ncg_p] - enk[ie2 + ncg_p]
            if abs(edif) > 1e-10:
                pm[ie12] = mmv[ie2 - ncbm, ie1] / edif
            ie12 += 1
    return pm
