"""ORACLE (test infrastructure only): the LAPW augmentation coefficients and
their band-basis products, restated from PyGW:

* ``gap2_set_lapwcoef`` (lapwcoeff.f90:3-222) with ``lohns``
  (lapwcoeff.f90:421-449): A/B/C_lm(k+G) for every plane wave and atom;
* the products with the eigenvectors that turn them into alfa, beta, gama
  (pygw.py:2577-2583 at k, 2601-2607 at k - q; la.matmul of Aeig with the
  coefficients reshaped to [ngk, ...]).

Spherical Bessel functions and Y_lm come from scipy (library routines), not
from PyGW's sphbes.f90, so the restatement is pinned against PyGW's compiled
gap2_set_lapwcoef (oracle/ref.py) and by the plane-wave expansion identity.

Arrays keep the f2py shapes (Fortran semantics, 0-based indices here):
alm/blm[ngk, nt^2, ndf], clm[ngk, nLOmax, nt^2, ndf], umt[nat, nt, 5],
abcelo[nat, lomax+1, nloat, 3], rotloc[nat, 3, 3], trotij[ndf, 3, 3],
pos[3, ndf], nLO_at/nlo[lomax+1, nat], lapw[nt, nat], gindex[ng, 3].
"""
from __future__ import annotations

import numpy as np
import scipy.special as sp

PI_F = 3.14159265358979     # the value of pi lapwcoeff.f90 carries


def sphbes(n, x):
    """j_l(x) and j_l'(x) for l = 0..n (sphbes.f90:2-42 semantics, x = 0
    included: j_0 = 1, j_1' = 1/3, the rest 0)."""
    ls = np.arange(n + 1)
    return sp.spherical_jn(ls, x), sp.spherical_jn(ls, x, derivative=True)


def ylm(v, lmax):
    """Complex Y_lm(v) with the Condon-Shortley phase at index l*l + l + m
    (as sphbes.f90:132-271 orders them); the zero vector is read as the z
    axis."""
    v = np.asarray(v, dtype=np.float64)
    th = np.arctan2(np.hypot(v[0], v[1]), v[2])
    ph = np.arctan2(v[1], v[0]) if (v[0] != 0.0 or v[1] != 0.0) else 0.0
    out = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            out[l * l + l + m] = sp.sph_harm_y(l, m, th, ph)
    return out


def lohns(iat, l0, jlo, mult, nlo, lomax):
    """lapwcoeff.f90:421-449: the 1-based range [i1, i2] of the local-orbital
    rows (after the nvk plane waves) that belong to local orbital jlo (1-based)
    of angular momentum l0 on atom type iat (0-based): (2 l0 + 1) mult(iat)
    rows per local orbital, atoms and l in order."""
    i2 = 1
    for i in range(iat):
        for l in range(lomax + 1):
            i2 += nlo[l, i] * (2 * l + 1) * mult[i]
    for l in range(l0):
        i2 += nlo[l, iat] * (2 * l + 1) * mult[iat]
    i1 = i2
    for _ in range(jlo):
        i1 = i2
        i2 += (2 * l0 + 1) * mult[iat]
    return i1, i2 - 1


def gap2_set_lapwcoef(kl, indgk, iop, ifrot, nvk, gindex, abcelo, rmt, pos, mult, umt, rotloc, trotij, vol,
                      k2cartes, nLO_at, nlo, lapw, nLOmax):
    """lapwcoeff.f90:3-222.  Returns alm, blm [ngk, nt^2, ndf] and clm [ngk,
    nLOmax, nt^2, ndf] (complex128).  indgk holds 0-based rows of gindex; the
    first nvk are the LAPW plane waves, the rest the local-orbital rows."""
    kl = np.asarray(kl, dtype=np.float64)
    indgk = np.asarray(indgk)
    ngk = indgk.size
    nat = len(mult)
    ndf = int(np.sum(mult))
    nt = umt.shape[1]
    lmax = nt - 1
    lomax = nLO_at.shape[0] - 1
    kcar = k2cartes @ kl
    G = np.asarray(gindex)[indgk].astype(np.float64)     # [ngk, 3]
    kpg = kcar[None, :] + G @ k2cartes.T
    rk = np.sqrt(np.sum(kpg ** 2, axis=1))
    # radial matching coefficients al/bl[n, l, iat] for the plane waves
    al = np.zeros((nvk, nt, nat))
    bl = np.zeros((nvk, nt, nat))
    for n in range(nvk):
        for iat in range(nat):
            fj, dfj = sphbes(nt - 1, rmt[iat] * rk[n])
            for l in range(nt):
                p, pe, dp, dpe = umt[iat, l, 0:4]
                if lapw[l, iat] == 1:
                    al[n, l, iat] = rk[n] * dfj[l] * pe - fj[l] * dpe
                    bl[n, l, iat] = fj[l] * dp - rk[n] * dfj[l] * p
                else:
                    al[n, l, iat] = fj[l] / p / rmt[iat] ** 2
    alm = np.zeros((ngk, nt * nt, ndf), dtype=np.complex128)
    blm = np.zeros_like(alm)
    clm = np.zeros((ngk, nLOmax, nt * nt, ndf), dtype=np.complex128)
    idf = 0
    for iat in range(nat):
        rhoatm = 4.0 * PI_F * np.sqrt(1.0 / vol) * rmt[iat] ** 2
        for _ in range(mult[iat]):
            arg = 2.0 * PI_F * ((G + kl[None, :]) @ pos[:, idf])
            phs = np.cos(arg) + 1j * np.sin(arg)
            yl = np.zeros((ngk, nt * nt), dtype=np.complex128)
            for ig in range(ngk):
                vec = kpg[ig]
                if ifrot:
                    vec = rotloc[iat] @ (vec @ trotij[idf])
                yl[ig] = ylm(vec, lmax)
            cil = 1.0 + 0j
            for l in range(nt):
                for m in range(-l, l + 1):
                    ilm = l * l + l + m
                    comc = cil * rhoatm * phs * np.conj(yl[:, ilm])
                    if iop != 1:
                        comc = np.conj(comc)
                    alm[:nvk, ilm, idf] = al[:, l, iat] * comc[:nvk]
                    blm[:nvk, ilm, idf] = bl[:, l, iat] * comc[:nvk]
                    if l > lomax:
                        continue
                    ilo0 = lapw[l, iat]
                    for jlo in range(ilo0, nLO_at[l, iat] + 1):
                        i1, i2 = lohns(iat, l, jlo - ilo0 + 1, mult, nlo, lomax)
                        alo, blo, clo = abcelo[iat, l, jlo, 0:3]
                        rows = np.arange(nvk + i1 - 1, nvk + i2)
                        alm[rows, ilm, idf] = alo * comc[rows]
                        blm[rows, ilm, idf] = blo * comc[rows]
                        if jlo >= 1:
                            clm[rows, jlo - 1, ilm, idf] = clo * comc[rows]
                cil *= 1j
            idf += 1
    return alm, blm, clm


def band_coefficients(Aeig, alm, blm, clm):
    """alfa/beta [nbmax, nt^2, ndf], gama [nbmax, nLOmax, nt^2, ndf] =
    Aeig [nbmax, ngk] times the coefficients (pygw.py:2577-2583)."""
    nb, ngk = Aeig.shape
    alfa = (Aeig @ alm.reshape(ngk, -1)).reshape((nb,) + alm.shape[1:])
    beta = (Aeig @ blm.reshape(ngk, -1)).reshape((nb,) + blm.shape[1:])
    gama = (Aeig @ clm.reshape(ngk, -1)).reshape((nb,) + clm.shape[1:])
    return alfa, beta, gama
