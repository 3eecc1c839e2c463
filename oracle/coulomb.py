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
