"""The matrix elements of one (k, q) pair, as PyGW's ``calc_minm`` builds
them (pygw.py:2547-2700), on the gfx950 backend.

``MatrixElements`` holds the per-run and per-q inputs on the device (the
LAPW plan, the muffin-tin plan and its radial/Gaunt constants, the
interstitial basis mpwipw with its G tables, and barcvm of the v^(1/2)
basis) and for each k runs

  gap2_set_lapwcoef at k (iop 1) and k - q (iop 2) -> alfa/beta/gama and
  alfp/betp/gamp (pygw.py:2567-2606), calc_minm_MT and calc_minm_IS into the
  columns of one concatenated mmat (pygw.py:2655-2663), the core part
  (calc_minc for 'selfe', calc_minc2 for 'polarization', pygw.py:2671-2680)
  and the products with conj(barcvm) (pygw.py:2666, 2684).

Every stage is a library entry point (lapw.hip, minm_mt.hip, minm_is.hip,
coulomb.hip); nothing here computes on the host.  The k-point data are
given as arrays (kl, indgk, nv, the eigenvectors) in place of PyGW's
``kqm``/``ks`` objects; ``Aeigq`` is what PyGW passes for k - q, the
conjugated eigenvectors (pygw.py:2585) with any symmetry phase applied.
"""
from __future__ import annotations

import numpy as np
import torch

from . import coulomb as C
from . import lapw as L
from .polarization import _dev, _require_device


class MatrixElements:
    """calc_minm for the k-points of one q.

    lapw_plan: ``lapw.LapwPlan``; mt_plan: ``coulomb.MinmMTPlan``; mt_consts:
    ``coulomb.mt_constants(mt_plan, s3r, vpos, cgcoef, corind)``; gindex [ng,
    3] and i_g0 [2n1+1, 2n2+1, 2n3+1] (PyGW's G list and its inverse map,
    0-based, -1 = none); mpwipw [ngq, ngq_barc] and indgq_inverse (the
    interstitial basis of this q); vol the cell volume; barcvm the device
    [mbsize, matsize] basis of ``coulomb.coul_setev`` (mbsize = loctmatsize
    + ngq); ql the q-point (semi-cartesian, as kl)."""

    def __init__(self, lapw_plan, mt_plan, mt_consts, gindex, i_g0, mpwipw, indgq_inverse, vol, barcvm, ql, *,
                 device="cuda"):
        d = _require_device(device)
        self.lapw_plan, self.mt_plan, self.consts = lapw_plan, mt_plan, mt_consts
        g = np.asarray(gindex)
        self.gindex_host = g
        self.gindex_lapw = torch.as_tensor(np.ascontiguousarray(g.ravel(order="F"), dtype=np.int32), device=d)
        self.gindex = _dev(g, torch.int32, d).contiguous()
        self.i_g0 = _dev(np.asarray(i_g0), torch.int32, d).contiguous()
        self.mpwipw = _dev(mpwipw, torch.complex128, d).contiguous()
        self.indgq_inverse = _dev(indgq_inverse, torch.int32, d).contiguous()
        self.vol, self.ql = float(vol), np.asarray(ql, dtype=np.float64)
        self.loct, self.ngq = mt_plan.loctmatsize, self.mpwipw.shape[0]
        self.mbsize = self.loct + self.ngq
        if barcvm.shape[0] != self.mbsize:
            raise ValueError("barcvm must have loctmatsize + ngq rows")
        self.barcvm, self.matsize = barcvm, barcvm.shape[1]
        self.device = d
        self.ws_is = None       # calc_minm_is's workspace, kept across k

    def _coefficients(self, kl, indgk, nv, A, iop, conj):
        c = L.gap2_set_lapwcoef(self.lapw_plan, kl, indgk, iop, True, nv, self.gindex_lapw)
        return L.band_coefficients(A, *c, conj=conj)

    def _mmat(self, out, kil, kjl, indgk_k, nvk, Aeigk, indgk_q, nvq, Aeigq, nst, nend, mst, mend, bk, bq):
        """mmat_mt and mmat_it of bands nst..nend-1 at k and mst..mend-1 at
        k - q into out [nend-nst, mend-mst, mbsize] (pygw.py:2655-2663)."""
        umk = np.rint(np.asarray(kil, dtype=np.float64) - np.asarray(kjl, dtype=np.float64) - self.ql).astype(np.int64)
        C.calc_minm_mt(self.mt_plan, self.ql, nst, nend, mst, mend, *bk, *bq, self.consts, None, None, out=out,
                       col0=0)
        n = C.minm_is_workspace(self.ngq, self.mpwipw.shape[1], nvk, nvq, nend - nst, mend - mst)
        if self.ws_is is None or self.ws_is.numel() < n:
            self.ws_is = torch.empty(n, dtype=torch.uint8, device=self.device)
        C.calc_minm_is(nst, nend, mst, mend, Aeigk, Aeigq, umk, self.mpwipw, nvk, nvq, indgk_k, indgk_q,
                       self.indgq_inverse, self.gindex, self.i_g0, self.vol, out=out, col0=self.loct,
                       workspace=self.ws_is, device=self.device)

    def _bands(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq):
        A_k = _dev(Aeigk, torch.complex128, self.device)
        A_q = _dev(Aeigq, torch.complex128, self.device)
        bk = self._coefficients(kil, indgk_k, nvk, A_k, 1, False)
        bq = self._coefficients(kjl, indgk_q, nvq, A_q, 2, False)
        return A_k, A_q, bk, bq

    def calc_minm(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, band_limits, mode):
        """(minm, minc) of pygw.py:2547-2700 for one k: minm [nend-nst,
        mend-mst, matsize] = mmat conj(barcvm); minc = mmat_c
        conj(barcvm[:loctmatsize]) ([nend-nst, cend-cst, matsize] for
        mode 'selfe', [cend-cst, mend-mst, matsize] for 'polarization'),
        None when cend = 0.  band_limits = (nst, nend, mst, mend, cst, cend)."""
        if mode not in ("selfe", "polarization"):
            raise ValueError("mode must be 'selfe' or 'polarization'")
        nst, nend, mst, mend, cst, cend = (int(x) for x in band_limits)
        A_k, A_q, bk, bq = self._bands(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq)
        mmat = torch.empty((nend - nst, mend - mst, self.mbsize), dtype=torch.complex128, device=self.device)
        self._mmat(mmat, kil, kjl, indgk_k, nvk, A_k, indgk_q, nvq, A_q, nst, nend, mst, mend, bk, bq)
        minm = C.calc_minm_vhalf(mmat, self.barcvm)
        minc = None
        if cend > 0:
            if mode == "selfe":
                mc = C.calc_minc(self.mt_plan, kil, nst, nend, cst, cend, *bk, self.consts, None, None, None)
            else:
                mc = C.calc_minc2(self.mt_plan, kjl, cst, cend, mst, mend, *bq, self.consts, None, None, None)
            minc = C.calc_minm_vhalf(mc, self.barcvm, self.loct)
        return minm, minc

    def polarization_minm(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, nvbm, ncbm, nbmaxpol, ncg_p,
                          *, out=None, complex64=False):
        """The minm that calc_eps contracts (pygw.py:3096-3101): band_limits
        (0, nvbm, ncbm, nbmaxpol, 0, ncg_p), the core rows (calc_minc2) first
        and the occupied bands after them, [ncg_p + nvbm, nbmaxpol - ncbm,
        matsize].  The core rows are written into the same mmat with zero
        interstitial columns, so one product with conj(barcvm) covers both
        (the zero columns add nothing to minc = mmat_c conj(barcvm[:loct]))."""
        n2 = nbmaxpol - ncbm
        A_k, A_q, bk, bq = self._bands(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq)
        mmat = torch.empty((ncg_p + nvbm, n2, self.mbsize), dtype=torch.complex128, device=self.device)
        self._mmat(mmat[ncg_p:], kil, kjl, indgk_k, nvk, A_k, indgk_q, nvq, A_q, 0, nvbm, ncbm, nbmaxpol, bk, bq)
        if ncg_p > 0:
            core = mmat[:ncg_p]
            core[..., self.loct:].zero_()
            C.calc_minc2(self.mt_plan, kjl, 0, ncg_p, ncbm, nbmaxpol, *bq, self.consts, None, None, None, out=core,
                         col0=0)
        if out is not None:
            return C.calc_minm_vhalf(mmat, self.barcvm, complex64=complex64, out=out)
        return C.calc_minm_vhalf(mmat, self.barcvm, complex64=complex64)

    def selfe_minm(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, ibgw, nbgw, nbands, ncg_c):
        """The minm that calc_selfc contracts (pygw.py:3339-3345): band_limits
        (ibgw, nbgw, 0, nbands, 0, ncg_c) and the core columns (calc_minc)
        before the valence ones, [nbgw - ibgw, ncg_c + nbands, matsize]."""
        minm, minc = self.calc_minm(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq,
                                    (ibgw, nbgw, 0, nbands, 0, ncg_c), "selfe")
        return minm if minc is None else torch.cat((minc, minm), dim=1)
