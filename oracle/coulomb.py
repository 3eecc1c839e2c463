"""ORACLE (test infrastructure only): numpy/scipy FP64 restatement of PyGW's
v^(1/2) basis of the Coulomb matrix and of the matrix elements in it, for
checking the GPU path.

  eigh_lower    linalg.eigh(Vmat) at the end of Coulomb (pygw.py:2440):
                LAPACK on the lower triangle, ascending eigenvalues
  immax_of      the q = 0 eigenvector with the largest overlap with the
                constant function (pygw.py:2447-2454)
  coul_setev    Coul_setev, Real_sqrt_Vcoul = False (pygw.py:18,
                2505-2545): kept = ev > evtol, at q = 0 the immax vector
                dropped (iop_coul_x = 0) or always kept; decreasing order;
                barcvm[:, m] = V[:, j] sqrt|ev_j|
  minm          la.matmul(mmat, conj(barcvm)) of calc_minm (pygw.py:2666),
                with n_in = loctmatsize the minc product (pygw.py:2684)
"""
import numpy as np
import scipy.linalg


def eigh_lower(vmat):
    return scipy.linalg.eigh(np.asarray(vmat), lower=True)


def immax_of(wi0, V):
    wi0new = np.dot(wi0, np.conj(V))
    return int(np.argmax(np.abs(wi0new)))


def coul_setev(vmat, evtol, wi0=None, iop_coul_x=0):
    """Returns (ev, V, barcev, barcvm, immax); immax = -1 without wi0."""
    ev, V = eigh_lower(vmat)
    mbsize = len(ev)
    kept = ev > evtol
    immax = -1
    if wi0 is not None:
        immax = immax_of(wi0, V)
        kept[immax] = iop_coul_x != 0
    matsize = int(kept.sum())
    barcev = np.zeros(matsize)
    barcvm = np.zeros((mbsize, matsize), dtype=complex, order="F")
    im = matsize - 1
    for jm in range(mbsize):
        if kept[jm]:
            barcev[im] = ev[jm]
            barcvm[:, im] = V[:, jm] * np.sqrt(abs(ev[jm]))
            im -= 1
    return ev, V, barcev, barcvm, immax


def minm(mmat, barcvm, n_in=None):
    n_in = mmat.shape[-1] if n_in is None else n_in
    return np.matmul(mmat[..., :n_in], np.conj(barcvm[:n_in, :]))


def hermitian_with_spectrum(lam, rng):
    """Q diag(lam) Q^H with Q the unitary factor of a complex Gaussian
    matrix (test input with a known spectrum)."""
    n = len(lam)
    Q, R = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    Q = Q * (np.diag(R) / np.abs(np.diag(R)))
    A = (Q * lam) @ Q.conj().T
    return (A + A.conj().T) / 2, Q


def minm_is(nstart, nend, mstart, mend, Aeigk, Aeigq, iumklap, mpwipw, nvi, nvj, indgk_i, indgk_j,
            indgq_inverse, gindex, i_g0, Vol):
    """calc_minm_IS (for_vxcnn.f90:1030-1134), called at pygw.py:2661: the
    interstitial part of the matrix elements,
      mmat[ie1, ie2, iGp] = 1/sqrt(Vol) sum_{ivi, ivj} Aeigq[mstart+ie2, ivj]
                            mpwipw[iGp, K(ivi, ivj)] Aeigk[nstart+ie1, ivi],
    K(ivi, ivj) the product-basis plane wave G_ki - G_kj + umklap (through
    i_g0 and indgq_inverse; no term where there is none).  Index arrays are
    0-based (-1 = none), i_g0 is [2n1+1, 2n2+1, 2n3+1]."""
    n = (np.array(i_g0.shape) - 1) // 2
    gi = gindex[np.asarray(indgk_i)[:nvi]]
    gj = gindex[np.asarray(indgk_j)[:nvj]]
    iGv = gi[:, None, :] - gj[None, :, :] + np.asarray(iumklap)[None, None, :]       # [nvi, nvj, 3]
    inside = np.all(np.abs(iGv) <= n, axis=-1)
    c = np.where(inside[..., None], iGv + n, 0)
    idg = np.where(inside, i_g0[c[..., 0], c[..., 1], c[..., 2]], -1)
    ok = (idg >= 0) & (idg < len(indgq_inverse))
    K = np.where(ok, np.asarray(indgq_inverse)[np.where(ok, idg, 0)], -1)             # [nvi, nvj]
    Ak = Aeigk[nstart:nend, :nvi]
    Aq = Aeigq[mstart:mend, :nvj]
    ngq = mpwipw.shape[0]
    out = np.zeros((nend - nstart, mend - mstart, ngq), dtype=complex)
    valid = (K >= 0).T
    for iGp in range(ngq):
        tmat = np.where(valid, mpwipw[iGp][np.where(K >= 0, K, 0).T], 0.0)          # [nvj, nvi]
        tmat2 = tmat @ Ak.T                                                          # [nvj, ndim]
        mnn = Aq @ tmat2                                                             # [mdim, ndim]
        out[:, :, iGp] = mnn.T / np.sqrt(Vol)
    return out


def pw_olap(indgq, nrow, ncol, ipwint, gindex, i_g0):
    """cmp_pw_olap2 (for_Coulomb.f90:306-325): out[j, i] = ipwint[G_i - G_j]
    for j < nrow, i < ncol, G_i = gindex[indgq[i]]; 0 where the difference
    leaves the i_g0 box (the Fortran reuses the previous pair's index there)."""
    n = (np.array(i_g0.shape) - 1) // 2
    g = gindex[np.asarray(indgq)]
    d = g[None, :ncol, :] - g[:nrow, None, :]                      # [j, i, 3] = G_i - G_j
    inside = np.all(np.abs(d) <= n, axis=-1)
    c = np.where(inside[..., None], d + n, 0)
    idg = np.where(inside, i_g0[c[..., 0], c[..., 1], c[..., 2]], -1)
    return np.where(idg >= 0, np.asarray(ipwint)[np.where(idg >= 0, idg, 0)], 0.0)


def pw_basis(indgq, ngq, ngq_barc, ipwint, gindex, i_g0):
    """OrthogonalizedBasisInterstitials (pygw.py:2126-2214) with PyGW's
    defaults (Real_sqrt_olap = True, FORT = True): mpwipw = olap^(-1/2) tmat."""
    olap = pw_olap(indgq, ngq, ngq, ipwint, gindex, i_g0)
    if np.sum(np.abs(olap.imag)) < 1e-7:
        olap = np.array(olap.real, dtype=float)
    eps, sgi = scipy.linalg.eigh(olap)
    sgiH = np.matmul(sgi, np.dot(np.diag(1 / np.sqrt(np.abs(eps))), np.conj(sgi.T)))
    tmat = pw_olap(indgq, ngq, ngq_barc, ipwint, gindex, i_g0)
    return np.dot(sgiH, tmat)


def coul_pw_block(mpwipw, gqlen, q_is_zero):
    """The plane-wave block of Coulomb (pygw.py:2419-2430):
    mpwipw diag(4 pi / |q+G|^2) mpwipw^H, |q+G_0| = 1e100 at q = 0."""
    gq = np.array(gqlen, dtype=float)
    if q_is_zero:
        gq[0] = 1e100
    Vq = 4 * np.pi / gq ** 2
    Vmm = Vq[None, :] * np.conj(mpwipw)
    return np.dot(mpwipw, Vmm.T)


def ylm(v, lmax):
    """Complex spherical harmonics Y_lm(v) (Condon-Shortley phase), index
    l*l + l + m, from scipy; v = 0 (or on the z axis) takes phi = 0, as
    ylm (sphbes.f90:132-271) does."""
    import scipy.special as sp
    v = np.asarray(v, dtype=float)
    r, rxy = np.linalg.norm(v), np.hypot(v[0], v[1])
    th = np.arctan2(rxy, v[2]) if r > 0 else 0.0
    ph = np.arctan2(v[1], v[0]) if rxy > 0 else 0.0
    return np.array([sp.sph_harm_y(l, m, th, ph) for l in range(lmax + 1) for m in range(-l, l + 1)])


def mixed_coulomb(vq, iat, q_vq, jlam, gindex, indgq, gqlen0, G_unique, ngq_barc, big_l, k2cartes, rotloc,
                  trotij, mult, pos, vol):
    """Mixed_Coulomb (for_Coulomb.f90:327-408), iat 0-based: for the
    product functions (irm, L, M) of the mult[iat] equivalent atoms,
      vtemp[im, ig] = Y_LM(R (q + G_ig)) e^{-2 pi i pos . G_ig} (4 pi / sqrt(Vol))
                      (-i)^L jlam[irm, ipw] (4 pi / gqlen0[ipw]^2 if q_vq)
    with ipw the unique-length group of ig (G_unique its starts), R =
    rotloc[iat] trotij[idf]^T k2cartes; at q = 0 the G = 0 column stays 0."""
    big_l = np.asarray(big_l)
    lmaxb = int(big_l.max())
    per = int(np.sum(2 * big_l + 1))
    vtemp = np.zeros((per * mult[iat], ngq_barc), dtype=complex)
    ipw0 = 1 if np.sum(np.abs(vq)) < 1e-7 else 0
    ngs = len(gqlen0)
    grp = np.zeros(ngq_barc, dtype=int) - 1
    for ipw in range(ngs):
        i1 = G_unique[ipw + 1] if ipw < ngs - 1 else ngq_barc
        grp[max(G_unique[ipw], ipw0):i1] = ipw
    G = np.asarray(gindex)[np.asarray(indgq)[:ngq_barc]]
    idf0 = int(np.sum(mult[:iat]))
    prefac = 4 * np.pi / np.sqrt(vol)
    im = 0
    for ieq in range(mult[iat]):
        idf = idf0 + ieq
        rot = rotloc[iat] @ (trotij[idf].T @ k2cartes)
        sph = np.array([ylm(rot @ (np.asarray(vq) + g), lmaxb) for g in G])          # [ngq_barc, (L+1)^2]
        expqr = np.exp(-2j * np.pi * (G @ pos[:, idf]))
        for irm, L in enumerate(big_l):
            for M in range(-L, L + 1):
                cv = prefac * (-1j) ** L * np.where(grp >= 0, jlam[irm, np.maximum(grp, 0)], 0.0)
                if q_vq:
                    with np.errstate(divide="ignore"):      # |q+G| = 0 only in the dropped G = 0 column
                        cv = cv * np.where(grp >= 0, 4 * np.pi / gqlen0[np.maximum(grp, 0)] ** 2, 0.0)
                vtemp[im] = sph[:, L * L + L + M] * expqr * cv
                im += 1
    return vtemp


# --- muffin-tin block of the Coulomb matrix (for_Coulomb.f90:1-283) ---------

def lattice_vectors(rmax, shift, rbs):
    """genr_str / genr_str_size (for_Coulomb.f90:418-491): the vectors
    R + shift (R = rbs . (i1, i2, i3), rbs columns = basis vectors) with
    1e-6 < |R + shift| <= rmax, in the order of qsort4 (length, ties within
    1e-6 in generation order).  Returns [n, 4] (x, y, z, length)."""
    rbs = np.asarray(rbs, dtype=float)
    lrbs = np.sqrt((rbs * rbs).sum(axis=0))
    imax = int(rmax / lrbs.min()) + 1
    rng = np.arange(-imax, imax + 1)
    I = np.array(np.meshgrid(rng, rng, rng, indexing="ij")).reshape(3, -1)     # i1 slowest, i3 fastest
    rps = rbs @ I + np.asarray(shift, dtype=float)[:, None]
    rl = np.sqrt((rps * rps).sum(axis=0))
    keep = (rl <= rmax) & (rl > 1e-6)
    v = np.vstack([rps[:, keep], rl[keep]]).T
    # qsort4's order: by length, equal (within 1e-6) lengths by index
    import functools
    def cmp(a, b):
        if abs(v[a, 3] - v[b, 3]) < 1e-6:
            return -1 if a < b else (1 if a > b else 0)
        return -1 if v[a, 3] < v[b, 3] else 1
    order = sorted(range(len(v)), key=functools.cmp_to_key(cmp))
    return v[order]


def ylm_many(v, lmax):
    """ylm for many vectors v [n, 3] -> [n, (lmax+1)^2] (scipy)."""
    import scipy.special as sp
    v = np.asarray(v, dtype=float).reshape(-1, 3)
    r, rxy = np.linalg.norm(v, axis=1), np.hypot(v[:, 0], v[:, 1])
    th = np.where(r > 0, np.arctan2(rxy, v[:, 2]), 0.0)
    ph = np.where(rxy > 0, np.arctan2(v[:, 1], v[:, 0]), 0.0)
    return np.stack([sp.sph_harm_y(l, m, th, ph) for l in range(lmax + 1) for m in range(-l, l + 1)], axis=1)


def ewald_summation(pq, lam_max, pos, br2, rbas, alat, vol, ortho, r_cf, g_cf, eta):
    """Ewald_Summation (for_Coulomb.f90:1-164): the structure constants
    sgm[lm, i + j(j+1)/2] (atom pairs i <= j, 0-based) of
      sum_R Y_lm(R + r_ij) e^{-i q.R} / |R + r_ij|^(l+1)
    as a real-space sum (erfc terms) plus a reciprocal-space sum over q + G
    (Gaussian terms), divided by Gamma(l + 1/2); the i = j self term of l = 0
    is -1/(eta pi)."""
    from scipy.special import erfc
    pos = np.asarray(pos, dtype=float)
    ndf = pos.shape[1]
    nlm = (lam_max + 1) ** 2
    lidx = np.array([l for l in range(lam_max + 1) for _ in range(2 * l + 1)])
    sgm = np.zeros((nlm, ndf * (ndf + 1) // 2), dtype=complex)
    rbs = np.asarray(rbas, dtype=float).T
    pq = np.asarray(pq, dtype=float)
    pref = 4.0 * np.pi * np.sqrt(np.pi) / vol
    for i in range(ndf):
        for j in range(i, ndf):
            d = pos[:, i] - pos[:, j]
            r_aa = d * np.asarray(alat) if ortho else d @ np.asarray(rbas, dtype=float)
            v = lattice_vectors(r_cf, r_aa, rbs)
            rp, rl = v[:, :3], v[:, 3]
            eq = np.exp(-1j * (rp @ pq))
            gausr = np.exp(-rl ** 2 / eta ** 2)
            gam = np.zeros((len(rl), lam_max + 1))
            gam[:, 0] = np.sqrt(np.pi) * erfc(rl / eta) / rl
            for l in range(1, lam_max + 1):
                gam[:, l] = gam[:, l - 1] * (l - 0.5) / rl + rl ** (l - 2) * gausr / eta ** (2 * l - 1)
            s1 = np.sum((gam[:, lidx] * eq[:, None]) * ylm_many(rp, lam_max), axis=0)
            v = lattice_vectors(g_cf, pq, np.asarray(br2, dtype=float))
            gpq, gl = v[:, :3], v[:, 3]
            eg = np.exp(1j * ((gpq - pq) @ r_aa))
            pg = pref * np.exp(-(eta * gl) ** 2 / 4.0) * eg / gl ** 2
            gt = np.stack([(-0.5j * gl) ** l for l in range(lam_max + 1)], axis=1)
            s2 = np.sum((pg[:, None] * gt[:, lidx]) * ylm_many(gpq, lam_max), axis=0)
            gl_ = np.sqrt(np.pi) * np.cumprod(np.r_[1.0, np.arange(1, lam_max + 1) - 0.5])
            col = i + j * (j + 1) // 2
            sgm[:, col] = (s1 + s2) / gl_[lidx]
            if i == j:
                sgm[0, col] -= 1.0 / (eta * np.pi)
    return sgm


def tildeg_table(lmax):
    """calc_tildeg (pygw.py:1938-1962): the packed table tilg."""
    from scipy.special import binom
    out = []
    for l1 in range(lmax + 1):
        p0 = (-1) ** l1 * 8 * np.pi * np.sqrt(np.pi)
        for l2 in range(l1 + 1):
            p1 = p0 / np.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * (l1 + l2) + 1))
            for m1 in range(-l1, l1 + 1):
                for m2 in range(l2 + 1):
                    out.append(p1 * np.sqrt(binom(l1 + m1 + l2 + m2, l1 + m1) * binom(l1 - m1 + l2 - m2, l1 - m1)))
    return np.array(out)


def gettildeg(l1, l2, m1, m2, tilg):
    """gettildeg (for_Coulomb.f90:634-670), 0-based index into tilg."""
    par = abs(l1 - l2) % 2
    if l1 < l2:
        j1, j2, mj1, mj2, fact = l2, l1, m2, m1, 1 - 2 * par
    else:
        j1, j2, mj1, mj2, fact = l1, l2, m1, m2, 1
    if mj2 < 0:
        mj1, mj2 = -mj1, -mj2
    i1 = j2 + mj2 + 1
    i2 = (j2 + 1) * (mj1 + j1)
    i3 = (j2 + 1) * j1 * j2 + (j2 - 1) * j2 // 2
    i4 = (j1 + 1) * (j1 + 2) * (3 * j1 - 1) * j1 // 12
    return fact * tilg[i1 + i2 + i3 + i4 - 1]


def djmm_size(lmax):
    return sum((2 * l + 1) * (l + 1) for l in range(lmax + 1))


def getdjmm(idf, l, m1, m2, djmm):
    """getdjmm (for_Coulomb.f90:672-705): D^l_{m1 m2} of atom idf from the
    packed djmm [dimdj, ndf] (m2 >= 0 stored; m2 < 0 by (-1)^(m1-m2) conj)."""
    sg = 1 if m2 >= 0 else -1
    i = l * (l + 1) * (4 * l - 1) // 6 + (l + 1) * (sg * m1 + l) + abs(m2)
    if m2 >= 0:
        return djmm[i, idf]
    return (1 - 2 * (abs(m1 - m2) % 2)) * np.conj(djmm[i, idf])


def mt_coulomb(vq, iat, loctmatsize, big_l, nmix, im_start, tilg, rrint, djmm, rtlij, sgm, mult):
    """MT_Coulomb (for_Coulomb.f90:166-283), iat 0-based: the rows of the
    product functions of atom type iat's equivalent atoms,
      V[im, jm] = rtlij[jrm, irm, jat, iat] sum_{m3, m4} (+-) sgm(l1+l2, m3+m4)
                  tildeg(l1, l2, m3, m4) conj(D^l1_{-m3, m1}) D^l2_{m4, m2}
    plus, on the same atom and (l, m), 4 pi rrint[iat, (irm, jrm)]/(2 l + 1);
    at q = 0 the L = 0 x L = 0 entries stay 0.  big_l [max_nmix, nat],
    im_start [nat + 1] 0-based offsets."""
    big_l = np.asarray(big_l)
    q0 = np.sum(np.abs(vq)) < 1e-7
    V = np.zeros((loctmatsize, loctmatsize), dtype=complex)
    nat = len(mult)
    idf = int(np.sum(mult[:iat])) - 1
    im = im_start[iat] - 1
    for ieq in range(mult[iat]):
        idf += 1
        for irm in range(nmix[iat]):
            l1 = int(big_l[irm, iat])
            for m1 in range(-l1, l1 + 1):
                im += 1
                jdf = -1
                for jat in range(nat):
                    jm = im_start[jat] - 1
                    for jeq in range(mult[jat]):
                        jdf += 1
                        ii, jj = min(idf, jdf), max(idf, jdf)
                        col = ii + jj * (jj + 1) // 2
                        for jrm in range(nmix[jat]):
                            l2 = int(big_l[jrm, jat])
                            for m2 in range(-l2, l2 + 1):
                                jm += 1
                                if q0 and l1 == 0 and l2 == 0:
                                    continue
                                L = l1 + l2
                                s = 0j
                                for m3 in range(-l1, l1 + 1):
                                    for m4 in range(-l2, l2 + 1):
                                        if idf <= jdf:
                                            stc = sgm[L * L + L + m3 + m4, col]
                                            mn = 1 - 2 * (abs(m3) % 2)
                                        else:
                                            stc = np.conj(sgm[L * L + L - m3 - m4, col])
                                            mn = 1 - 2 * (abs(m4 + l1 + l2) % 2)
                                        tg = gettildeg(l1, l2, m3, m4, tilg)
                                        d3 = np.conj(getdjmm(idf, l1, -m3, m1, djmm))
                                        d4 = getdjmm(jdf, l2, m4, m2, djmm)
                                        s += mn * stc * tg * d3 * d4
                                V[im, jm] = rtlij[jrm, irm, jat, iat] * s
                                if m1 == m2 and l1 == l2 and idf == jdf:
                                    i1, j1 = min(irm, jrm), max(irm, jrm)
                                    V[im, jm] += 4 * np.pi * rrint[iat, j1 * (j1 + 1) // 2 + i1] / (2 * l1 + 1)
    return V


# --- radial integrals, the plane-wave overlaps and the q = 0 term ------------

def rint13g(rel, a, b, x, y, dh, nrpt, rnot):
    """RINT13g (radials.f90:140-187): the integral over [0, r_n] of
    A X + c2 B Y (c2 = 1/137.0359895^2 if rel else the single-precision
    1e-22) on the logarithmic mesh r_j = rnot e^{(j-1) dh}: Simpson's rule in
    the mesh index from point 2 (odd nrpt) or 3 (even nrpt, with a trapezoid
    over the first interval), each term weighted by r_j (dr = r dh), plus
    rnot f(rnot)/3 for [0, rnot]."""
    import math
    c2 = 1.0 / 137.0359895 ** 2 if rel else float(np.float32(1e-22))
    f = lambda j: a[j - 1] * x[j - 1] + c2 * b[j - 1] * y[j - 1]  # noqa: E731  1-based
    d = math.exp(dh)    # libm's, as the compiled routine takes it (numpy's may differ in the last bit)
    j = 3 - nrpt % 2
    j1 = j - 1
    r = rnot * d ** (j - 1)
    r1 = r / d
    z4 = z2 = 0.0
    while True:
        z4 += r * f(j)
        r *= d
        j += 1
        if j >= nrpt:
            break
        z2 += r * f(j)
        r *= d
        j += 1
    p1 = rnot * f(1)
    p2 = r1 * f(j1)
    s = 2 * z2 + 4 * z4 + r * f(j) + p2
    s = (dh * s + p1) / 3.0
    if j1 > 1:
        s += 0.5 * dh * (p1 + p2)
    return s


def radial_mesh(r0, dh, npt):
    """strc.radial_mesh (gwienfile.py:218-222): r0 e^{dh j}, j < npt."""
    return r0 * np.exp(dh) ** np.arange(npt)


def calc_olap_pw_product(gqlen, lmax, rx, dh, rel, big_l, ul, us):
    """Calc_Olap_PW_Product (pygw.py:2110-2124) on the mesh rx: jlam[irm, ig] =
    rint13g(rel, ul[irm], us[irm], j_L(k r) r, j_L(k r) r, dh, npt, rx[0])
    with k = gqlen[ig], L = big_l[irm]; the spherical Bessel functions from
    scipy."""
    from scipy.special import spherical_jn
    npt, r0 = len(rx), rx[0]
    jlam = np.zeros((len(big_l), len(gqlen)))
    for irm, L in enumerate(big_l):
        assert L <= lmax
        for ig, k in enumerate(gqlen):
            jl = spherical_jn(int(L), k * rx) * rx
            jlam[irm, ig] = rint13g(rel, ul[irm], us[irm], jl, jl, dh, npt, r0)
    return jlam


def pw_get_short_length(glen):
    """pw_get_short_length (for_Coulomb.f90:493-513)."""
    gind4 = np.zeros(len(glen), dtype=int)
    igl, prev = 0, 0.0
    for i, g in enumerate(glen):
        if abs(g - prev) > 1e-10:
            igl += 1
            prev = g
        gind4[i] = igl
    return igl, gind4


def pw_get_short_and_phase(ngs, gind4, gindex, glen, pos):
    """pw_get_short_and_phase (for_Coulomb.f90:515-582): glen0[g] the
    first length of group g, phase[i, j, g] = the sum over the plane waves
    2.. of group g of e^{-2 pi i G.(pos_i - pos_j)} (1 each for i = j)."""
    pos = np.asarray(pos, dtype=float)
    ndf = pos.shape[1]
    glen0 = np.zeros(ngs)
    jgl = 0
    for i in range(len(glen)):
        if gind4[i] != jgl:
            jgl = gind4[i]
            glen0[jgl - 1] = glen[i]
    G = np.asarray(gindex, dtype=float)[1:]
    grp = np.asarray(gind4)[1:] - 1
    phase = np.zeros((ndf, ndf, ngs), dtype=complex)
    for i in range(ndf):
        for j in range(ndf):
            e = np.ones(len(G), dtype=complex) if i == j else np.exp(-2j * np.pi * (G @ (pos[:, i] - pos[:, j])))
            np.add.at(phase[i, j], grp, e)
    return glen0, phase


def zero_momentum_coulomb(glen, gindex, pos, vol, mbsize, rel, meshes, big_l, ul, us, atm, iprod):
    """Zero_Momentum_Coulomb (pygw.py:2055-2104) after its plane-wave
    generation: the L = 0 x L = 0 entries
      V[im, jm] = sum_g phase[idf, jdf, g] 16 pi^2/(Vol glen0_g^4) s_im(g) s_jm(g),
    s(g) = rint13g(rel, u, us, sin(glen0_g r), sin(glen0_g r)) on the atom's mesh."""
    ngs, gind4 = pw_get_short_length(glen)
    glen0, phase = pw_get_short_and_phase(ngs, gind4, gindex, glen, pos)
    sing = {}
    for iat in range(len(big_l)):
        r0, dh, npt = meshes[iat]
        rx = radial_mesh(r0, dh, npt)
        for irm, L in enumerate(big_l[iat]):
            if L == 0:
                sing[(iat, irm)] = np.array([rint13g(rel, ul[iat][irm], us[iat][irm], np.sin(g * rx), np.sin(g * rx),
                                                     dh, npt, r0) for g in glen0])
    c = 16 * np.pi ** 2 / vol / glen0 ** 4
    V = np.zeros((mbsize, mbsize), dtype=complex)
    for idf, iat in enumerate(atm):
        for irm, L in enumerate(big_l[iat]):
            if L != 0:
                continue
            for jdf, jat in enumerate(atm):
                for jrm, L2 in enumerate(big_l[jat]):
                    if L2 != 0:
                        continue
                    V[iprod[(idf, irm)], iprod[(jdf, jrm)]] = np.sum(phase[idf, jdf] * c * sing[(iat, irm)] *
                                                                     sing[(jat, jrm)])
    return V
