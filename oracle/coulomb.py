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
