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
        self.ql_dev = torch.as_tensor(self.ql, device=d)
        self.ws = {}            # per stream slot: calc_minm_mt's and calc_minm_is's workspaces, kept across k
        self.streams = []       # side streams of polarization_minm_many
        self._kdev = {}         # k-points on the device (the MT and core kernels read them there)
        self.mmat_buf = None    # polarization_minm's mmat, reused across k
        self.mmat_many = None   # polarization_minm_many's
        self.split_basis = None  # barcvm's split planes, made on the first split-precision call
        self.ws_vh = None

    def _kd(self, kl):
        """kl on the device, uploaded once per distinct k-point (a host-to-
        device copy per call would wait for the stream)."""
        key = tuple(float(x) for x in np.asarray(kl, dtype=np.float64))
        t = self._kdev.get(key)
        if t is None:
            t = self._kdev[key] = torch.as_tensor(np.asarray(key), device=self.device)
        return t

    def _coefficients(self, kl, indgk, nv, A, iop, conj):
        c = L.gap2_set_lapwcoef(self.lapw_plan, kl, indgk, iop, True, nv, self.gindex_lapw)
        return L.band_coefficients(A, *c, conj=conj)

    def _workspace(self, slot, name, n):
        w = self.ws.get((slot, name))
        if w is None or w.numel() < n:
            w = self.ws[(slot, name)] = torch.empty(n, dtype=torch.uint8, device=self.device)
        return w

    def _mmat(self, out, kil, kjl, indgk_k, nvk, Aeigk, indgk_q, nvq, Aeigq, nst, nend, mst, mend, bk, bq, slot=0):
        """mmat_mt and mmat_it of bands nst..nend-1 at k and mst..mend-1 at
        k - q into out [nend-nst, mend-mst, mbsize] (pygw.py:2655-2663); bk /
        bq hold the band coefficients of exactly those bands; slot names the
        workspaces (one set per stream)."""
        umk = np.rint(np.asarray(kil, dtype=np.float64) - np.asarray(kjl, dtype=np.float64) - self.ql).astype(np.int64)
        n1, n2 = nend - nst, mend - mst
        ws = self._workspace(slot, "mt", C.minm_mt_workspace(self.mt_plan, n1, n2))
        C.calc_minm_mt(self.mt_plan, self.ql_dev, 0, n1, 0, n2, *bk, *bq, self.consts, None, None, out=out, col0=0,
                       workspace=ws)
        ws = self._workspace(slot, "is", C.minm_is_workspace(self.ngq, self.mpwipw.shape[1], nvk, nvq, n1, n2))
        C.calc_minm_is(nst, nend, mst, mend, Aeigk, Aeigq, umk, self.mpwipw, nvk, nvq, indgk_k, indgk_q,
                       self.indgq_inverse, self.gindex, self.i_g0, self.vol, out=out, col0=self.loct,
                       workspace=ws, device=self.device)

    def _bands(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, rows_k, rows_q):
        """The eigenvectors on the device and the band coefficients of the
        bands rows_k = (first, end) at k and rows_q at k - q only (the
        reference forms them for all bands, pygw.py:2577-2606; the matrix
        elements read no others)."""
        A_k = _dev(Aeigk, torch.complex128, self.device)
        A_q = _dev(Aeigq, torch.complex128, self.device)
        bk = self._coefficients(kil, indgk_k, nvk, A_k[rows_k[0]:rows_k[1]], 1, False)
        bq = self._coefficients(kjl, indgk_q, nvq, A_q[rows_q[0]:rows_q[1]], 2, False)
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
        A_k, A_q, bk, bq = self._bands(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, (nst, nend), (mst, mend))
        mmat = torch.empty((nend - nst, mend - mst, self.mbsize), dtype=torch.complex128, device=self.device)
        self._mmat(mmat, kil, kjl, indgk_k, nvk, A_k, indgk_q, nvq, A_q, nst, nend, mst, mend, bk, bq)
        minm = C.calc_minm_vhalf(mmat, self.barcvm)
        minc = None
        if cend > 0:
            if mode == "selfe":
                mc = C.calc_minc(self.mt_plan, self._kd(kil), 0, nend - nst, cst, cend, *bk, self.consts, None, None,
                                 None)
            else:
                mc = C.calc_minc2(self.mt_plan, self._kd(kjl), cst, cend, 0, mend - mst, *bq, self.consts, None, None,
                                  None)
            minc = C.calc_minm_vhalf(mc, self.barcvm, self.loct)
        return minm, minc

    def _vhalf(self, mmat, precision, complex64, out, n_in=None):
        if precision == "fp64":
            if out is not None:
                return C.calc_minm_vhalf(mmat, self.barcvm, n_in, complex64=complex64, out=out)
            return C.calc_minm_vhalf(mmat, self.barcvm, n_in, complex64=complex64)
        if precision != "mixed":
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if self.split_basis is None:
            self.split_basis = C.SplitBasis(self.barcvm)
        n = C.minm_vhalf_mixed_workspace(mmat.numel() // mmat.shape[-1], n_in or mmat.shape[-1])
        if self.ws_vh is None or self.ws_vh.numel() < n:
            self.ws_vh = torch.empty(n, dtype=torch.uint8, device=self.device)
        return C.calc_minm_vhalf_mixed(mmat, self.split_basis, n_in, complex64=complex64, out=out,
                                       workspace=self.ws_vh)

    def polarization_minm(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, nvbm, ncbm, nbmaxpol, ncg_p,
                          *, out=None, complex64=False, precision="fp64"):
        """The minm that calc_eps contracts (pygw.py:3096-3101): band_limits
        (0, nvbm, ncbm, nbmaxpol, 0, ncg_p), the core rows (calc_minc2) first
        and the occupied bands after them, [ncg_p + nvbm, nbmaxpol - ncbm,
        matsize].  The core rows are written into the same mmat with zero
        interstitial columns, so one product with conj(barcvm) covers both
        (the zero columns add nothing to minc = mmat_c conj(barcvm[:loct])).
        precision "mixed": the product with conj(barcvm) in split precision
        (the operand of the split-precision polarization; ~1e-7 of each
        row's size), everything before it FP64."""
        shape = (ncg_p + nvbm, nbmaxpol - ncbm, self.mbsize)
        if self.mmat_buf is None or tuple(self.mmat_buf.shape) != shape:
            self.mmat_buf = torch.empty(shape, dtype=torch.complex128, device=self.device)
        self._pol_mmat(self.mmat_buf, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, nvbm, ncbm, nbmaxpol, ncg_p)
        return self._vhalf(self.mmat_buf, precision, complex64, out)

    def polarization_minm_many(self, ks, nvbm, ncbm, nbmaxpol, ncg_p, *, out=None, complex64=False,
                               precision="fp64", streams=4):
        """polarization_minm for several k-points at once: ks a list of
        (kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq); returns [len(ks),
        ncg_p + nvbm, nbmaxpol - ncbm, matsize].  The mmat of every k is built
        first, the k-points dealt round robin to `streams` side streams (their
        launches of 20-120 bands each leave most of the chip idle alone), and
        the product with conj(barcvm) runs once over all their rows on the
        caller's stream, after every side stream's work."""
        shape = (len(ks), ncg_p + nvbm, nbmaxpol - ncbm, self.mbsize)
        if self.mmat_many is None or tuple(self.mmat_many.shape) != shape:
            self.mmat_many = torch.empty(shape, dtype=torch.complex128, device=self.device)
        ns = max(1, min(int(streams), len(ks)))
        while len(self.streams) < ns:
            self.streams.append(torch.cuda.Stream(device=self.device))      # default priority
        cur = torch.cuda.current_stream(self.device)
        start = torch.cuda.Event()
        start.record(cur)
        for i in range(ns):
            self.streams[i].wait_event(start)
        for j, k in enumerate(ks):
            with torch.cuda.stream(self.streams[j % ns]):
                self._pol_mmat(self.mmat_many[j], *k, nvbm, ncbm, nbmaxpol, ncg_p, slot=j % ns)
        for i in range(ns):
            done = torch.cuda.Event()
            done.record(self.streams[i])
            cur.wait_event(done)
        return self._vhalf(self.mmat_many, precision, complex64, out)

    def _pol_mmat(self, mmat, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, nvbm, ncbm, nbmaxpol, ncg_p,
                  slot=0):
        """calc_eps's mmat of one k into mmat [ncg_p + nvbm, nbmaxpol - ncbm,
        mbsize]: the core rows (MT columns only, the rest zero), then the
        occupied bands."""
        n2 = nbmaxpol - ncbm
        A_k, A_q, bk, bq = self._bands(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, (0, nvbm),
                                       (ncbm, nbmaxpol))
        self._mmat(mmat[ncg_p:], kil, kjl, indgk_k, nvk, A_k, indgk_q, nvq, A_q, 0, nvbm, ncbm, nbmaxpol, bk, bq,
                   slot)
        if ncg_p > 0:
            core = mmat[:ncg_p]
            core[..., self.loct:].zero_()
            C.calc_minc2(self.mt_plan, self._kd(kjl), 0, ncg_p, 0, n2, *bq, self.consts, None, None, None, out=core,
                         col0=0)

    def selfe_minm(self, kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, ibgw, nbgw, nbands, ncg_c, *,
                   complex64=False, precision="fp64"):
        """The minm that calc_selfc contracts (pygw.py:3339-3345): band_limits
        (ibgw, nbgw, 0, nbands, 0, ncg_c) and the core columns (calc_minc)
        before the valence ones, [nbgw - ibgw, ncg_c + nbands, matsize];
        precision "mixed" as in polarization_minm (the operand of the
        split-precision MWM)."""
        nst, nend = int(ibgw), int(nbgw)
        A_k, A_q, bk, bq = self._bands(kil, indgk_k, nvk, Aeigk, kjl, indgk_q, nvq, Aeigq, (nst, nend), (0, nbands))
        mmat = torch.empty((nend - nst, nbands, self.mbsize), dtype=torch.complex128, device=self.device)
        self._mmat(mmat, kil, kjl, indgk_k, nvk, A_k, indgk_q, nvq, A_q, nst, nend, 0, nbands, bk, bq)
        minm = self._vhalf(mmat, precision, complex64, None)
        if ncg_c <= 0:
            return minm
        mc = C.calc_minc(self.mt_plan, self._kd(kil), 0, nend - nst, 0, ncg_c, *bk, self.consts, None, None, None)
        minc = self._vhalf(mc, precision, complex64, None, self.loct)
        return torch.cat((minc, minm), dim=1)
