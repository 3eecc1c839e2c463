"""The v^(1/2) basis of the Coulomb matrix and its application to the
product-basis matrix elements, on the gfx950 backend.

* ``coul_setev`` replaces ``linalg.eigh(Vmat)`` at the end of
  ``Coulomb`` (pygw.py:2440), the q = 0 choice of ``immax``
  (pygw.py:2447-2454) and ``Coul_setev`` (pygw.py:2505-2545, the
  ``Real_sqrt_Vcoul = False`` branch, pygw.py:18): eigenvalues above
  ``evtol`` kept in decreasing order, ``barcvm[:, m] = v_j sqrt|ev_j|``.
* ``calc_minm_is`` replaces ``fvxcn.calc_minm_is`` (pygw.py:2661,
  for_vxcnn.f90:1030-1134), the interstitial part of mmat.
* ``calc_minm_vhalf`` replaces ``minm = la.matmul(mmat, conj(barcvm))``
  (pygw.py:2666) and, with ``n_in = loctmatsize``, ``minc = la.matmul(
  mmat_c, conj(barcvm[:loctmatsize, :]))`` (pygw.py:2684).

Eigenvectors are defined up to a phase per column (and a unitary mixing
within a degenerate eigenvalue): barcvm differs from LAPACK's by such
factors, which every physical quantity downstream (eps^-1 in the rotated
basis, MWM, Sigma_c) does not see.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .polarization import _dev, _require_device, _stream_ptr


@dataclass
class CoulombBasis:
    ev: torch.Tensor        # float64 [mbsize], ascending (the eigh order)
    barcev: torch.Tensor    # float64 [matsize], decreasing
    barcvm: torch.Tensor    # complex128 [mbsize, matsize] (Fortran-strided view)
    matsize: int
    immax: int              # q = 0: eigenvector with the largest overlap with 1 (eigh order); else -1
    vmat: torch.Tensor | None  # complex128 [mbsize, mbsize] eigenvectors in eigh order, if asked for
    sweeps: int


def coul_setev(vmat, evtol=1e-8, wi0=None, iop_coul_x=0, *, eigenvectors=False, device="cuda"):
    """vmat: complex128 [mbsize, mbsize] Hermitian Coulomb matrix (its lower
    triangle is used, like eigh).  wi0: complex128 [mbsize] overlaps <chi|1>
    of the basis with the constant function (q = 0 only, pygw.py:2448-2450),
    or None for q != 0."""
    device = _require_device(device)
    v = _dev(vmat, torch.complex128, device).contiguous()
    mb = v.shape[0]
    if v.dim() != 2 or v.shape[1] != mb:
        raise ValueError("vmat must be square")
    w = None if wi0 is None else _dev(wi0, torch.complex128, device).contiguous()
    if w is not None and w.numel() != mb:
        raise ValueError("wi0 must have mbsize entries")
    ev = torch.empty(mb, dtype=torch.float64, device=device)
    barcev = torch.empty(mb, dtype=torch.float64, device=device)
    cols = torch.empty((mb, mb), dtype=torch.complex128, device=device)   # column m at cols[m]
    vecs = torch.empty((mb, mb), dtype=torch.complex128, device=device) if eigenvectors else None
    ms, im, sw = C.c_int(0), C.c_int(-1), C.c_int(0)
    _lib.call("pygw_coul_setev", C.c_void_p(v.data_ptr()), mb, float(evtol),
              None if w is None else C.c_void_p(w.data_ptr()), int(iop_coul_x), C.c_void_p(ev.data_ptr()),
              C.c_void_p(barcev.data_ptr()), C.c_void_p(cols.data_ptr()),
              None if vecs is None else C.c_void_p(vecs.data_ptr()), C.byref(ms), C.byref(im), C.byref(sw),
              None, 0, _stream_ptr(device))
    m = ms.value
    return CoulombBasis(ev=ev, barcev=barcev[:m], barcvm=cols[:m].T, matsize=m,
                        immax=im.value if w is not None else -1,
                        vmat=None if vecs is None else vecs.T, sweeps=sw.value)


def calc_minm_vhalf(mmat, barcvm, n_in=None, *, complex64=False, out=None):
    """minm[..., m] = sum_{i < n_in} mmat[..., i] conj(barcvm[i, m]).

    mmat: complex128 [..., ld] device tensor, rows contiguous (the last axis
    may be longer than n_in, e.g. mmat_c against the first loctmatsize rows
    of barcvm); barcvm: complex128 [mbsize, matsize] with unit stride along
    mbsize (the view ``CoulombBasis.barcvm``).  Returns complex128 (or
    complex64 when ``complex64``) [..., matsize]."""
    if not (mmat.is_cuda and barcvm.is_cuda):
        raise RuntimeError("calc_minm_vhalf needs device tensors")
    if mmat.dtype != torch.complex128 or barcvm.dtype != torch.complex128:
        raise TypeError("mmat and barcvm must be complex128")
    mb, ms = barcvm.shape
    if barcvm.stride(0) != 1:
        raise ValueError("barcvm must have unit stride along mbsize (Fortran order)")
    ldb = barcvm.stride(1) if ms > 1 else mb
    ld = mmat.shape[-1]
    n_in = ld if n_in is None else int(n_in)
    if n_in > mb or n_in > ld:
        raise ValueError("n_in exceeds the inner dimension")
    if mmat.stride(-1) != 1:
        raise ValueError("mmat rows must be contiguous")
    rows = mmat.reshape(-1, ld)
    if rows.data_ptr() != mmat.data_ptr() or (rows.shape[0] > 1 and rows.stride(0) != ld):
        raise ValueError("mmat rows must be equally spaced")
    nrows = rows.shape[0]
    dt = torch.complex64 if complex64 else torch.complex128
    if out is None:
        out = torch.empty(mmat.shape[:-1] + (ms,), dtype=dt, device=mmat.device)
    elif out.dtype != dt or not out.is_contiguous() or out.numel() != nrows * ms:
        raise ValueError("out must be a contiguous tensor of the result's size and type")
    if ms == 0 or nrows == 0:      # nothing kept (every eigenvalue at or below evtol): empty, as numpy gives
        return out
    _lib.call("pygw_minm_vhalf", C.c_void_p(mmat.data_ptr()), nrows, n_in, ld, C.c_void_p(barcvm.data_ptr()),
              ldb, ms, int(complex64), C.c_void_p(out.data_ptr()), _stream_ptr(mmat.device))
    return out


class SplitBasis:
    """barcvm's FP16 hi/lo planes and column scales for the split-precision
    product (pygw_minm_vhalf_split_basis), built once per basis."""

    def __init__(self, barcvm):
        if not barcvm.is_cuda or barcvm.dtype != torch.complex128 or barcvm.stride(0) != 1:
            raise ValueError("barcvm must be a complex128 device tensor with unit stride along mbsize")
        self.mbsize, self.matsize = barcvm.shape
        ldb = barcvm.stride(1) if self.matsize > 1 else self.mbsize
        sz = C.c_size_t(0)
        _lib.call("pygw_minm_vhalf_split_basis_bytes", self.mbsize, self.matsize, C.byref(sz))
        self.buf = torch.empty(sz.value, dtype=torch.uint8, device=barcvm.device)
        _lib.call("pygw_minm_vhalf_split_basis", C.c_void_p(barcvm.data_ptr()), self.mbsize, ldb, self.matsize,
                  C.c_void_p(self.buf.data_ptr()), sz.value, _stream_ptr(barcvm.device))
        self.device = barcvm.device


def calc_minm_vhalf_mixed(mmat, basis, n_in=None, *, complex64=True, out=None, workspace=None):
    """minm = mmat conj(barcvm) (pygw.py:2666, 2684) in split precision
    (pygw_minm_vhalf_mixed): basis = SplitBasis(barcvm); mmat as for
    calc_minm_vhalf.  workspace: an optional uint8 device tensor of at least
    ``minm_vhalf_mixed_workspace(n_rows, n_in)`` bytes."""
    if not mmat.is_cuda or mmat.dtype != torch.complex128:
        raise TypeError("mmat must be a complex128 device tensor")
    ld = mmat.shape[-1]
    n_in = ld if n_in is None else int(n_in)
    if n_in > basis.mbsize or n_in > ld or mmat.stride(-1) != 1:
        raise ValueError("n_in exceeds the inner dimension, or mmat rows are not contiguous")
    rows = mmat.reshape(-1, ld)
    if rows.data_ptr() != mmat.data_ptr() or (rows.shape[0] > 1 and rows.stride(0) != ld):
        raise ValueError("mmat rows must be equally spaced")
    nrows, ms = rows.shape[0], basis.matsize
    dt = torch.complex64 if complex64 else torch.complex128
    if out is None:
        out = torch.empty(mmat.shape[:-1] + (ms,), dtype=dt, device=mmat.device)
    elif out.dtype != dt or not out.is_contiguous() or out.numel() != nrows * ms:
        raise ValueError("out must be a contiguous tensor of the result's size and type")
    if nrows == 0:
        return out
    need = minm_vhalf_mixed_workspace(nrows, n_in)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=mmat.device)
    _lib.call("pygw_minm_vhalf_mixed", C.c_void_p(mmat.data_ptr()), nrows, n_in, ld, C.c_void_p(basis.buf.data_ptr()),
              basis.mbsize, ms, int(complex64), C.c_void_p(out.data_ptr()), C.c_void_p(workspace.data_ptr()),
              workspace.numel(), _stream_ptr(mmat.device))
    return out


def minm_vhalf_mixed_workspace(n_rows, n_in):
    sz = C.c_size_t(0)
    _lib.call("pygw_minm_vhalf_mixed_workspace", int(n_rows), int(n_in), C.byref(sz))
    return sz.value


def calc_minm_is(nstart, nend, mstart, mend, Aeigk, Aeigq, iumklap, mpwipw, nvi, nvj, indgk_i, indgk_j,
                 indgq_inverse, gindex, i_g0, Vol, *, out=None, col0=0, form="auto", workspace=None,
                 device="cuda"):
    """fvxcn.calc_minm_is (pygw.py:2661, for_vxcnn.f90:1030-1134): the
    interstitial part of the matrix elements, mmat_it[ie1, ie2, G] =
    1/sqrt(Vol) sum_{i, j} Aeigq[mstart+ie2, j] mpwipw[G, K(G_i - G_j +
    umklap)] Aeigk[nstart+ie1, i], with PyGW's 0-based index arrays (i_g0
    [2n1+1, 2n2+1, 2n3+1], -1 = none).  Returns complex128 [nend-nstart,
    mend-mstart, ngq], or writes it into out[..., col0:col0 + ngq] of a
    contiguous complex128 [nend-nstart, mend-mstart, ld] (the columns of
    mmat = concatenate((mmat_mt, mmat_it), -1) after loctmatsize).
    form "auto" runs the association the library prefers for these sizes
    (the pair density first when it costs fewer products, else T first as a
    dense GEMM), with workspace (a uint8 device tensor of at least
    ``minm_is_workspace`` bytes, reused across calls) or else one the call
    allocates; "dense" and "gather" force T first, as a dense GEMM or
    gathered, through the size of the workspace given to the call."""
    device = _require_device(device)
    i32 = lambda x: _dev(x, torch.int32, device).contiguous()
    ak = _dev(Aeigk, torch.complex128, device).contiguous()
    aq = _dev(Aeigq, torch.complex128, device).contiguous()
    mp = _dev(mpwipw, torch.complex128, device).contiguous()
    n1, n2, n3 = [(s - 1) // 2 for s in i_g0.shape]
    gi, gj, inv, gx, g0 = i32(indgk_i), i32(indgk_j), i32(indgq_inverse), i32(gindex), i32(i_g0)
    ndim, mdim, ngq = nend - nstart, mend - mstart, mp.shape[0]
    if out is None:
        out = torch.empty((ndim, mdim, ngq), dtype=torch.complex128, device=device)
        ld = ngq
    else:
        if out.dtype != torch.complex128 or not out.is_contiguous() or out.shape[:2] != (ndim, mdim):
            raise ValueError("out must be a contiguous complex128 [nend-nstart, mend-mstart, ld] tensor")
        ld = out.shape[2]
    umk = (C.c_int * 3)(*[int(u) for u in np.asarray(iumklap)])
    ws, wsb = None, 0
    if form in ("dense", "gather"):
        kt = (int(nvi) * int(nvj) * 4 + 255) // 256 * 256 + ngq * int(nvj) * ndim * 16   # kmap and T
        wsb = kt + (mp.shape[1] * int(nvj) * ndim * 16 if form == "dense" else 0)
        if form == "dense" and _rho_cheaper(ngq, mp.shape[1], int(nvj), ndim, mdim) and \
                wsb >= minm_is_workspace(ngq, mp.shape[1], nvi, nvj, ndim, mdim):
            raise ValueError("the dense form cannot be selected for these sizes")
        workspace = torch.empty(wsb, dtype=torch.uint8, device=device)
    elif form != "auto":
        raise ValueError("form must be 'auto', 'dense' or 'gather'")
    if workspace is not None:
        ws, wsb = C.c_void_p(workspace.data_ptr()), workspace.numel()
    _lib.call("pygw_minm_is", int(nstart), int(nend), int(mstart), int(mend), C.c_void_p(ak.data_ptr()),
              ak.shape[0], ak.shape[1], C.c_void_p(aq.data_ptr()), aq.shape[0], aq.shape[1], umk,
              C.c_void_p(mp.data_ptr()), ngq, mp.shape[1], int(nvi), int(nvj), C.c_void_p(gi.data_ptr()),
              C.c_void_p(gj.data_ptr()), C.c_void_p(inv.data_ptr()), inv.numel(), C.c_void_p(gx.data_ptr()),
              gx.shape[0], C.c_void_p(g0.data_ptr()), n1, n2, n3, float(Vol), C.c_void_p(out.data_ptr()), ld,
              int(col0), ws, wsb, _stream_ptr(device))
    return out


def _rho_cheaper(ngq, ngq_barc, nvj, ndim, mdim):
    return ndim * ngq_barc * mdim * (nvj + ngq) < ngq * ndim * nvj * (ngq_barc + mdim)


def minm_is_workspace(ngq, ngq_barc, nvi, nvj, ndim, mdim):
    """Bytes of the workspace calc_minm_is prefers for these sizes."""
    sz = C.c_size_t(0)
    _lib.call("pygw_minm_is_workspace", int(ngq), int(ngq_barc), int(nvi), int(nvj), int(ndim), int(mdim),
              C.byref(sz))
    return sz.value


def pw_basis(indgq, ngq, ngq_barc, ipwint, gindex, i_g0, *, device="cuda"):
    """OrthogonalizedBasisInterstitials (pygw.py:2126-2214, Real_sqrt_olap =
    True): mpwipw = olap^(-1/2) tmat with olap[j, i] = ipwint[G_i - G_j]
    (i, j < ngq) and tmat the same for i < ngq_barc.  Returns complex128
    [ngq, ngq_barc]; raises PygwError (PYGW_E_NOTPD) when olap is singular
    (linearly dependent plane waves, e.g. a repeated G in indgq)."""
    device = _require_device(device)
    i32 = lambda x: _dev(x, torch.int32, device).contiguous()
    iq, gx, g0 = i32(indgq), i32(gindex), i32(i_g0)
    ip = _dev(ipwint, torch.complex128, device).contiguous()
    n1, n2, n3 = [(s - 1) // 2 for s in i_g0.shape]
    out = torch.empty((ngq, ngq_barc), dtype=torch.complex128, device=device)
    sz = C.c_size_t(0)
    _lib.call("pygw_pw_basis_workspace", int(ngq), int(ngq_barc), C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_pw_basis", C.c_void_p(iq.data_ptr()), int(ngq), int(ngq_barc), C.c_void_p(ip.data_ptr()),
              ip.numel(), C.c_void_p(gx.data_ptr()), C.c_void_p(g0.data_ptr()), n1, n2, n3,
              C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()), sz.value, _stream_ptr(device))
    return out


def coul_pw_block(mpwipw, gqlen, q_is_zero, *, vmat=None, row0=0):
    """Vmat[row0:, row0:] = mpwipw diag(4 pi / gqlen^2) mpwipw^H (pygw.py:
    2419-2430; gqlen[0] = 1e100 when q_is_zero).  Into vmat (contiguous
    complex128 [mbsize, mbsize], row0 = loctmatsize) or a new [ngq, ngq]."""
    if mpwipw.dtype != torch.complex128 or not mpwipw.is_contiguous():
        raise TypeError("mpwipw must be a contiguous complex128 tensor")
    ngq, nb = mpwipw.shape
    g = _dev(gqlen, torch.float64, mpwipw.device).contiguous()
    if g.numel() != nb:
        raise ValueError("gqlen must have ngq_barc entries")
    if vmat is None:
        vmat = torch.zeros((ngq, ngq), dtype=torch.complex128, device=mpwipw.device)
    elif vmat.dtype != torch.complex128 or not vmat.is_contiguous() or vmat.shape[0] < row0 + ngq:
        raise ValueError("vmat must be a contiguous complex128 [mbsize, mbsize] with row0 + ngq <= mbsize")
    _lib.call("pygw_coul_pw_block", C.c_void_p(mpwipw.data_ptr()), ngq, nb, C.c_void_p(g.data_ptr()),
              int(bool(q_is_zero)), C.c_void_p(vmat.data_ptr()), vmat.shape[1], int(row0), _stream_ptr(mpwipw.device))
    return vmat


def mixed_coulomb(vq, iat, q_vq, jlam, gindex, indgq, gqlen0, G_unique, ngq_barc, big_l, k2cartes, rotloc, trotij,
                  mult, pos, vol, *, out=None, row0=0, workspace=None, device="cuda"):
    """fCl.mixed_coulomb (pygw.py:2396, for_Coulomb.f90:327-408) for atom iat
    (0-based): the rows of Vmatit of its equivalent atoms' product
    functions, complex128 [sum_ieq sum_irm (2L+1), ngq_barc], or written
    into out[row0:, :ngq_barc] (Vmatit[istr:iend, :]).  pos is PyGW's
    [3, ndf].  workspace: optional uint8 device tensor of at least
    pygw_mixed_coulomb_workspace bytes (else one is taken from torch)."""
    device = _require_device(device)
    f64 = lambda x: _dev(x, torch.float64, device).contiguous()
    i32 = lambda x: _dev(x, torch.int32, device).contiguous()
    jl, g0, gu, gx, iq = f64(jlam), f64(gqlen0), i32(G_unique), i32(gindex), i32(indgq)
    hd = lambda x, dt: np.ascontiguousarray(np.asarray(x), dtype=dt)
    vqh, blh, k2h, rlh, trh, muh = (hd(vq, np.float64), hd(big_l, np.int32), hd(k2cartes, np.float64),
                                    hd(rotloc, np.float64), hd(trotij, np.float64), hd(mult, np.int32))
    psh = hd(np.asarray(pos).T, np.float64)
    rows = int(muh[iat]) * int(np.sum(2 * blh + 1))
    if out is None:
        out = torch.empty((rows, ngq_barc), dtype=torch.complex128, device=device)
    elif out.dtype != torch.complex128 or not out.is_contiguous() or out.shape[0] < row0 + rows:
        raise ValueError("out must be a contiguous complex128 Vmatit with room for the rows")
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    sz = C.c_size_t(0)
    _lib.call("pygw_mixed_coulomb_workspace", max(int(muh[iat]), 1), int(ngq_barc), C.byref(sz))
    if workspace is None or workspace.numel() < sz.value:
        workspace = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_mixed_coulomb", p(vqh), int(iat), int(bool(q_vq)), C.c_void_p(jl.data_ptr()), jl.shape[1],
              C.c_void_p(gx.data_ptr()), C.c_void_p(iq.data_ptr()), C.c_void_p(g0.data_ptr()),
              C.c_void_p(gu.data_ptr()), int(ngq_barc), p(blh), len(blh), p(k2h), p(rlh), p(trh), p(muh), p(psh),
              len(muh), float(vol), C.c_void_p(out.data_ptr()), out.shape[1], int(row0),
              C.c_void_p(workspace.data_ptr()), workspace.numel(), _stream_ptr(device))
    return out


def coul_mixed_block(mpwipw, vmatit, vmat):
    """Vmat[loct:, :loct] = mpwipw Vmatit^T and Vmat[:loct, loct:] its
    conjugate transpose (pygw.py:2409-2416); loct = Vmatit's rows."""
    ngq, nb = mpwipw.shape
    loct = vmatit.shape[0]
    for x in (mpwipw, vmatit, vmat):
        if x.dtype != torch.complex128 or not x.is_contiguous():
            raise TypeError("complex128 contiguous tensors expected")
    if vmatit.shape[1] != nb or vmat.shape[0] < loct + ngq:
        raise ValueError("shapes do not match")
    _lib.call("pygw_coul_mixed_block", C.c_void_p(mpwipw.data_ptr()), ngq, nb, C.c_void_p(vmatit.data_ptr()), loct,
              C.c_void_p(vmat.data_ptr()), vmat.shape[1], _stream_ptr(mpwipw.device))
    return vmat


# --- the muffin-tin block and the q = 0 term ---------------------------------

def calc_tildeg(lmax):
    """calc_tildeg (pygw.py:1938-1962): the packed table of
    tilde g(l1, l2, m1, m2) = (-1)^l1 8 pi^(3/2) / sqrt((2 l1+1)(2 l2+1)(2(l1+l2)+1))
                              sqrt(C(l1+m1+l2+m2, l1+m1) C(l1-m1+l2-m2, l1-m1))
    for l2 <= l1 <= lmax, -l1 <= m1 <= l1, 0 <= m2 <= l2 (that loop order);
    (lmax+1)(lmax+2)(lmax+3)(3 lmax+2)/12 entries, read by MT_Coulomb's
    gettildeg.  Host setup, as in the reference."""
    from math import comb
    out = []
    for l1 in range(lmax + 1):
        p0 = (-1) ** l1 * 8.0 * np.pi * np.sqrt(np.pi)
        for l2 in range(l1 + 1):
            p1 = p0 / np.sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * (l1 + l2) + 1))
            for m1 in range(-l1, l1 + 1):
                for m2 in range(l2 + 1):
                    out.append(p1 * np.sqrt(float(comb(l1 + m1 + l2 + m2, l1 + m1) * comb(l1 - m1 + l2 - m2, l1 - m1))))
    return np.array(out)


def _fortran_flat(x, dtype, device):
    return _dev(np.asarray(x).ravel(order="F"), dtype, device)


def ewald_summation(pq, lam_max, pos, br2, rbas, alat, vol, ortho, r_cf, g_cf, eta, *, out=None, device="cuda"):
    """fCl.ewald_summation (pygw.py:2350, for_Coulomb.f90:1-164): the
    structure constants sgm, complex128 [(lam_max+1)^2, ndf(ndf+1)/2] (a
    Fortran-strided view; column i + j(j+1)/2 holds the atom pair i <= j).
    pq: cartesian q; pos: [3, ndf] fractional positions (strc.vpos); br2:
    reciprocal vectors as columns; rbas: lattice vectors as rows (all host
    arrays; the call synchronises the current stream)."""
    device = _require_device(device)
    pos = np.asarray(pos, dtype=np.float64)
    ndf = pos.shape[1]
    nlm, npair = (lam_max + 1) ** 2, ndf * (ndf + 1) // 2
    if out is None:
        out = torch.empty((npair, nlm), dtype=torch.complex128, device=device)
    elif out.dtype != torch.complex128 or not out.is_contiguous() or tuple(out.shape) != (npair, nlm):
        raise ValueError("out must be a contiguous complex128 [npair, (lam_max+1)^2] tensor")
    h = [np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F")) for x in (pq, pos, br2, rbas, alat)]
    p = [x.ctypes.data_as(C.c_void_p) for x in h]
    sz = C.c_size_t(0)
    _lib.call("pygw_ewald_workspace", p[0], int(lam_max), p[1], p[2], p[3], p[4], int(bool(ortho)), float(r_cf),
              float(g_cf), ndf, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_ewald_summation", p[0], int(lam_max), p[1], p[2], p[3], p[4], float(vol), int(bool(ortho)),
              float(r_cf), float(g_cf), float(eta), ndf, C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()),
              ws.numel(), _stream_ptr(device))
    return out.T


def mt_coulomb(vq, iat, loctmatsize, big_l, nmix, im_start, tilg, rrint, djmm, rtlij, sgm, mult, lam_max, *,
               device="cuda"):
    """fCl.mt_coulomb (pygw.py:2394, for_Coulomb.f90:166-283) with iat
    0-based: the rows of atom type iat's product functions of the
    muffin-tin Coulomb matrix, complex128 [loctmatsize, loctmatsize] (a
    Fortran-strided view, every other row 0); iat=None gives every row at
    once, the sum over iat that the reference accumulates.  big_l
    [max_nmix, nat] and rtlij [max_nmix, max_nmix, nat, nat] as PyGW builds
    them (pygw.py:2361-2374), rrint [nat, nijrm], djmm [dimdj, ndf], sgm
    from ewald_summation (lam_max >= 2 max L)."""
    device = _require_device(device)
    big_l = np.asarray(big_l)
    max_nmix, nat = big_l.shape
    ims = np.ascontiguousarray(np.asarray(im_start), dtype=np.int32)
    if ims.shape != (nat + 1,):
        raise ValueError("im_start must hold nat + 1 offsets")
    sg = sgm if isinstance(sgm, torch.Tensor) else torch.as_tensor(np.asarray(sgm))
    if tuple(sg.shape)[0] != (lam_max + 1) ** 2:
        raise ValueError("sgm must have (lam_max+1)^2 rows")
    sg = sg.to(device=device, dtype=torch.complex128).T.contiguous()   # Fortran (nlm, npair)
    djmm = np.asarray(djmm)
    nm_h = np.asarray(nmix)
    if nm_h.shape != (nat,) or nm_h.min() < 1 or nm_h.max() > max_nmix:
        raise ValueError("nmix must give 1..max_nmix functions per atom type")
    if np.shape(rrint) != (nat, max_nmix * (max_nmix + 1) // 2) or np.shape(rtlij) != (max_nmix, max_nmix, nat, nat):
        raise ValueError("rrint must be [nat, nijrm] and rtlij [max_nmix, max_nmix, nat, nat]")
    if djmm.ndim != 2 or djmm.shape[1] != int(np.sum(mult)):
        raise ValueError("djmm must be [dimdj, ndf]")
    per = np.array([np.sum(2 * big_l[: nm_h[a], a] + 1) for a in range(nat)]) * np.asarray(mult)
    if not np.array_equal(np.diff(ims), per) or ims[0] != 0 or ims[-1] != loctmatsize:
        raise ValueError("im_start does not match big_l, nmix and mult")
    i32 = lambda x: _fortran_flat(x, torch.int32, device)  # noqa: E731
    f64 = lambda x: _fortran_flat(x, torch.float64, device)  # noqa: E731
    bl, nm, imd, mu = i32(big_l), i32(nmix), i32(ims), i32(mult)
    tg, rr, rt, vqd = f64(tilg), f64(rrint), f64(rtlij), f64(vq)
    dj = _fortran_flat(djmm, torch.complex128, device)
    out = torch.empty((loctmatsize, loctmatsize), dtype=torch.complex128, device=device)
    _lib.call("pygw_mt_coulomb", C.c_void_p(vqd.data_ptr()), -1 if iat is None else int(iat), int(loctmatsize),
              C.c_void_p(bl.data_ptr()), C.c_void_p(nm.data_ptr()), C.c_void_p(imd.data_ptr()),
              ims.ctypes.data_as(C.c_void_p), C.c_void_p(mu.data_ptr()), nat, max_nmix, C.c_void_p(tg.data_ptr()),
              tg.numel(), C.c_void_p(rr.data_ptr()), C.c_void_p(dj.data_ptr()), djmm.shape[0], C.c_void_p(rt.data_ptr()),
              C.c_void_p(sg.data_ptr()), int(lam_max), C.c_void_p(out.data_ptr()), _stream_ptr(device))
    return out.T


def radial_mesh(r0, dh, npt):
    """strc.radial_mesh (gwienfile.py:218-222): r0 e^{dh j} as numpy forms it."""
    return r0 * np.exp(dh) ** np.arange(npt)


def radial_olap(mode, rel, ul, us, big_l, rx, dh, k, *, lmax=None, device="cuda"):
    """out[f, ik] = rd.rint13g(rel, ul[f], us[f], F, F, dh, npt, rx[0]) with
    F = j_L(k rx) rx, L = big_l[f] (mode 0) or F = sin(k rx) (mode 1) on the
    logarithmic mesh rx (radials.f90:140-187); float64 [n_fun, n_k]."""
    device = _require_device(device)
    u = _dev(ul, torch.float64, device).contiguous()
    s = _dev(us, torch.float64, device).contiguous()
    r = _dev(rx, torch.float64, device).contiguous()
    npt = r.numel()
    lf = np.ascontiguousarray(np.asarray(big_l), dtype=np.int32)
    if u.shape != s.shape or u.dim() != 2 or u.shape[0] != lf.size or u.shape[1] != npt:
        raise ValueError("ul and us must be [n_fun, npt] with one L per function")
    if mode == 1:
        lmax = 0
    elif lmax is None:
        lmax = int(lf.max())
    kk = _dev(k, torch.float64, device).contiguous()
    lfd = _dev(lf, torch.int32, device)
    out = torch.empty((lf.size, kk.numel()), dtype=torch.float64, device=device)
    sz = C.c_size_t(0)
    _lib.call("pygw_radial_olap_workspace", int(mode), int(lmax), int(npt), kk.numel(), C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_radial_olap", int(mode), int(bool(rel)), C.c_void_p(u.data_ptr()), C.c_void_p(s.data_ptr()),
              C.c_void_p(lfd.data_ptr()), lf.size, int(lmax), int(npt), C.c_void_p(r.data_ptr()), float(dh),
              C.c_void_p(kk.data_ptr()), kk.numel(), C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()),
              ws.numel(), _stream_ptr(device))
    return out


def calc_olap_pw_product(gqlen, Lmax, rx, dh, rel, big_l, ul_product, us_product, *, device="cuda"):
    """Calc_Olap_PW_Product (pygw.py:2110-2124) for one atom type on its mesh
    rx: jlam[irm, ig] = <j_L(|q+G| r) r | u_irm> by RINT13g, float64
    [nmix, n_g]."""
    return radial_olap(0, rel, ul_product, us_product, big_l, rx, dh, gqlen, lmax=Lmax, device=device)


def pw_get_short_length(glen):
    """pw_get_short_length (for_Coulomb.f90:493-513): ngs and gind4[ipw]
    (1-based group of ipw; runs of lengths within 1e-10 of the run's first,
    the leading G = 0 in group 0).  Host setup."""
    glen = np.asarray(glen, dtype=np.float64)
    gind4 = np.zeros(glen.size, dtype=np.int64)
    igl, prev = 0, 0.0
    for i, g in enumerate(glen):
        if abs(g - prev) > 1e-10:
            igl += 1
            prev = g
        gind4[i] = igl
    return igl, gind4


def zero_momentum_coulomb(glen, gindex, pos, vol, mbsize, rel, meshes, big_l, ul_product, us_product, atm, iprod,
                          *, out=None, device="cuda"):
    """Zero_Momentum_Coulomb (pygw.py:2055-2104) after its plane-wave
    generation: glen and gindex [n_g, 3] the plane waves sorted by length
    (fkp.generate_pw), pos [3, ndf] (strc.vpos), meshes[iat] = (r0, dh, npt),
    big_l / ul_product / us_product per atom type, atm[idf] the type of atom
    idf, iprod[(idf, irm)] the Vmat index of the L = 0 function irm of atom
    idf.  Returns complex128 [mbsize, mbsize] (a Fortran-strided view, or
    written into the Fortran-strided `out`) with the L = 0 x L = 0 entries
    set and every other entry 0."""
    device = _require_device(device)
    glen = np.asarray(glen, dtype=np.float64)
    ngs, gind4 = pw_get_short_length(glen)
    if ngs == 0:
        raise ValueError("no plane wave of nonzero length")
    # the runs of pw_get_short_and_phase: plane waves 1.. (the first, G = 0, skipped)
    g_start = np.searchsorted(gind4[1:], np.arange(1, ngs + 2)) + 1
    first = np.searchsorted(gind4, np.arange(1, ngs + 1))
    glen0 = glen[first]
    sing_rows, f_sing, f_atom, f_row = [], [], [], []
    row_of = {}
    for iat in range(len(big_l)):
        bl = np.asarray(big_l[iat])
        l0 = np.flatnonzero(bl == 0)
        if l0.size == 0:
            continue
        r0, dh, npt = meshes[iat]
        s = radial_olap(1, rel, np.asarray(ul_product[iat])[l0], np.asarray(us_product[iat])[l0], np.zeros(l0.size),
                        radial_mesh(r0, dh, npt), dh, glen0, device=device)
        for a, irm in enumerate(l0):
            row_of[(iat, int(irm))] = len(sing_rows)
            sing_rows.append(s[a])
    for idf, iat in enumerate(atm):
        for irm in np.flatnonzero(np.asarray(big_l[iat]) == 0):
            f_sing.append(row_of[(iat, int(irm))])
            f_atom.append(idf)
            f_row.append(int(iprod[(idf, int(irm))]))
    sing = torch.stack(sing_rows).contiguous()
    gx = _fortran_flat(np.asarray(gindex), torch.int32, device)
    i32 = lambda x: _dev(np.asarray(x), torch.int32, device).contiguous()  # noqa: E731
    gs, fs, fa, fr = i32(g_start), i32(f_sing), i32(f_atom), i32(f_row)
    g0 = _dev(glen0, torch.float64, device)
    ps = _fortran_flat(pos, torch.float64, device)
    ndf = np.asarray(pos).shape[1]
    if out is None:
        out = torch.zeros((mbsize, mbsize), dtype=torch.complex128, device=device).T
    elif out.dtype != torch.complex128 or out.stride() != (1, mbsize) or tuple(out.shape) != (mbsize, mbsize):
        raise ValueError("out must be a Fortran-strided complex128 [mbsize, mbsize] view")
    sz = C.c_size_t(0)
    _lib.call("pygw_zero_momentum_workspace", ndf, ngs, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_zero_momentum_coulomb", C.c_void_p(gx.data_ptr()), len(glen), C.c_void_p(gs.data_ptr()),
              C.c_void_p(g0.data_ptr()), ngs, C.c_void_p(ps.data_ptr()), ndf, C.c_void_p(sing.data_ptr()),
              sing.shape[0], C.c_void_p(fs.data_ptr()), C.c_void_p(fa.data_ptr()), C.c_void_p(fr.data_ptr()), len(f_row), float(vol),
              int(mbsize), C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()), ws.numel(), _stream_ptr(device))
    return out


# --- the muffin-tin and core parts of the matrix elements --------------------

def _w3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol (Racah's sum; integer arithmetic until the last step)."""
    from math import factorial as f, isqrt  # noqa: F401
    if m1 + m2 + m3 or not abs(j1 - j2) <= j3 <= j1 + j2 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    num = f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) \
        * f(j2 - m2) * f(j3 + m3) * f(j3 - m3)
    den = f(j1 + j2 + j3 + 1)
    tmin, tmax = max(0, j2 - j3 - m1, j1 - j3 + m2), min(j1 + j2 - j3, j1 - m1, j2 + m2)
    # sum_t (-1)^t / D_t over a common denominator L = lcm-free product bound: use exact fractions of ints
    from fractions import Fraction
    s = sum(Fraction((-1) ** t, f(t) * f(j3 - j2 + t + m1) * f(j3 - j1 + t - m2) * f(j1 + j2 - j3 - t)
                     * f(j1 - t - m1) * f(j2 - t + m2)) for t in range(tmin, tmax + 1))
    return (-1) ** ((j1 - j2 - m3) % 2) * float(s) * (num / den) ** 0.5


def cmp_all_gaunt(maxj):
    """gn.cmp_all_gaunt (pygw.py:545, gaunt.f90:1-34): for l2 <= l1 <= maxj,
    l3 = l1-l2 .. l1+l2, m1 = -l1 .. l1, m2 = 0 .. l2 (that order) the Gaunt
    coefficient <Y_l3,m1+m2 | Y_l2m2 Y_l1m1> (0 for odd l1 + l2 + l3 or
    |m1 + m2| > l3), the table getcgcoef reads.  Host setup, once per run."""
    out = []
    for l1 in range(maxj + 1):
        for l2 in range(l1 + 1):
            for l3 in range(l1 - l2, l1 + l2 + 1):
                even = (l1 + l2 + l3) % 2 == 0
                c = (2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * np.pi)
                w0 = _w3j(l3, l2, l1, 0, 0, 0) if even else 0.0
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(l2 + 1):
                        m3 = m1 + m2
                        if not even or abs(m3) > l3:
                            out.append(0.0)
                        else:   # <Y_l3m3|Y_l2m2 Y_l1m1> = (-1)^m3 sqrt(c) 3j(l3 l2 l1; 000) 3j(l3 l2 l1; -m3 m2 m1)
                            out.append((-1) ** (m3 % 2) * c ** 0.5 * w0 * _w3j(l3, l2, l1, -m3, m2, m1))
    return np.array(out)


class MinmMTPlan:
    """The atom / product-function / radial-channel layout of the muffin-tin
    matrix elements (pygw_minm_mt_plan): built once per run from PyGW's
    strc.mult, nmix, big_l [max_nmix, nat], in1.nLO_at [lomax+1, nat],
    pb.ncore and in1.lmax; kept in host memory and on the device."""

    def __init__(self, mult, nmix, big_l, nLO_at, ncore, lmax, *, device="cuda"):
        device = _require_device(device)
        f = lambda a: np.ascontiguousarray(np.asarray(a).ravel(order="F"), dtype=np.int32)  # noqa: E731
        self.mult, self.nmix, self.ncore = f(mult), f(nmix), f(ncore)
        bl, nlo = np.asarray(big_l), np.asarray(nLO_at)
        self.big_l, self.nLO_at = f(bl), f(nlo)
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        n = C.c_int(0)
        args = (len(self.mult), p(self.mult), p(self.nmix), p(self.big_l), bl.shape[0], p(self.nLO_at),
                nlo.shape[0] - 1, p(self.ncore), int(lmax))
        _lib.call("pygw_minm_mt_plan", *args, None, C.byref(n))
        self.host = np.zeros(n.value, dtype=np.int32)
        _lib.call("pygw_minm_mt_plan", *args, p(self.host), C.byref(n))
        self.dev = torch.as_tensor(self.host, device=device)
        self.lmax, self.device = int(lmax), device
        self.loctmatsize = int(self.host[5])
        self.nlomax = int(nlo.max()) if nlo.size else 0

    def ptrs(self):
        return self.host.ctypes.data_as(C.c_void_p), C.c_void_p(self.dev.data_ptr())


def _bands(a, device):
    """A band-coefficient array in PyGW's index order ([nbmax, (lmax+1)^2,
    ndf] or with nLOmax) stored on the device in Fortran order, as the
    kernels read it."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.complex128).permute(*reversed(range(a.dim()))).contiguous()
    return torch.as_tensor(np.asarray(a, dtype=np.complex128).ravel(order="F"), device=device)


@dataclass
class MTConstants:
    """The per-run inputs of calc_minm_MT / calc_minc / calc_minc2 on the
    device (s3r, vpos, cgcoef and corind in the layouts the kernels read):
    built once by ``mt_constants`` and passed in place of s3r, so a k-loop
    does not upload them again on every call."""
    s3r: torch.Tensor
    n_mix: int
    n_fnc: int
    pos: torch.Tensor
    cg: torch.Tensor
    corind: torch.Tensor
    ncg_rows: int
    corind_host: np.ndarray


def mt_constants(plan, s3r, vpos, cgcoef, corind=None):
    """Upload s3r [n_mix, n_fnc, n_fnc, nat], vpos [3, ndf], cgcoef and corind
    [ncg, 5] (checked against the plan as calc_minc checks them) once."""
    device = plan.device
    s3 = np.asarray(s3r, dtype=np.float64)
    co = np.zeros((0, 5), dtype=np.int64) if corind is None else np.asarray(corind)
    if co.size:
        _check_corind(plan, co)
    return MTConstants(s3r=_dev(s3.ravel(order="F"), torch.float64, device), n_mix=s3.shape[0], n_fnc=s3.shape[1],
                       pos=_dev(np.asarray(vpos, dtype=np.float64).ravel(order="F"), torch.float64, device),
                       cg=_dev(np.asarray(cgcoef, dtype=np.float64), torch.float64, device),
                       corind=_dev(co.ravel(order="F"), torch.int32, device), ncg_rows=co.shape[0],
                       corind_host=co)


def _consts(plan, s3r, vpos, cgcoef, corind=None):
    return s3r if isinstance(s3r, MTConstants) else mt_constants(plan, s3r, vpos, cgcoef, corind)


def _check_corind(plan, co, rows=None):
    ty = plan.host[plan.host[7]:plan.host[7] + 4 * int(plan.host[1])].reshape(-1, 4)[:, 0]
    sel = co if rows is None else co[rows[0]:rows[1]]
    if co.ndim != 2 or co.shape[1] != 5 or sel.size == 0 or (sel[:, 1] < 0).any() or \
            (sel[:, 1] >= len(ty)).any() or (ty[sel[:, 1]] != sel[:, 0]).any() or (sel[:, 2] < 0).any() or \
            (sel[:, 2] >= np.asarray(plan.ncore)[sel[:, 0]]).any() or (np.abs(sel[:, 4]) > sel[:, 3]).any():
        raise ValueError("corind rows must be (atom type, atom of that type, core index < ncore, l, |m| <= l)")


def minm_mt_workspace(plan, n1, n2):
    """Bytes of calc_minm_mt's workspace for an n1 x n2 band window."""
    ph, _ = plan.ptrs()
    sz = C.c_size_t(0)
    _lib.call("pygw_minm_mt_workspace", ph, int(n1), int(n2), C.byref(sz))
    return sz.value


def calc_minm_mt(plan, ql, nstart, nend, mstart, mend, alfa, beta, gama, alfp, betp, gamp, s3r, vpos, cgcoef, *,
                 out=None, col0=0, workspace=None):
    """fvxcn.calc_minm_mt (pygw.py:2655, for_vxcnn.f90:839-1028):
    mmat_mt[ie1, ie2, im] for bands nstart..nend-1 at k and mstart..mend-1 at
    k - q, complex128 [nend-nstart, mend-mstart, loctmatsize], or written into
    out[..., col0:col0 + loctmatsize] of a contiguous [n1, n2, ld] (the
    concatenated mmat).  alfa/beta [nbmax, (lmax+1)^2, ndf], gama [nbmax,
    nLOmax, (lmax+1)^2, ndf] and the primed arrays as PyGW builds them
    (numpy, Fortran semantics), s3r [n_mix, n_fnc, n_fnc, nat], vpos [3, ndf],
    cgcoef = cmp_all_gaunt(maxnt); or s3r = mt_constants(...) (vpos and
    cgcoef then unused)."""
    device = plan.device
    k = _consts(plan, s3r, vpos, cgcoef)
    n1, n2, loct = nend - nstart, mend - mstart, plan.loctmatsize
    al, ap = alfa.shape, alfp.shape
    d = [_bands(x, device) for x in (alfa, beta, alfp, betp)]
    nlo = plan.nlomax
    ga = gp = None
    if nlo > 0:
        ga, gp = _bands(gama, device), _bands(gamp, device)
    qd = _dev(ql if isinstance(ql, torch.Tensor) else np.asarray(ql, dtype=np.float64), torch.float64, device)
    if out is None:
        out = torch.empty((n1, n2, loct), dtype=torch.complex128, device=device)
        ld = loct
    else:
        if out.dtype != torch.complex128 or not out.is_contiguous() or tuple(out.shape[:2]) != (n1, n2):
            raise ValueError("out must be a contiguous complex128 [nend-nstart, mend-mstart, ld] tensor")
        ld = out.shape[2]
    ph, pd = plan.ptrs()
    sz = C.c_size_t(0)
    _lib.call("pygw_minm_mt_workspace", ph, n1, n2, C.byref(sz))
    if workspace is None or workspace.numel() < sz.value:
        workspace = torch.empty(sz.value, dtype=torch.uint8, device=device)
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    nbg = gama.shape[0] if nlo > 0 else 0
    mbg = gamp.shape[0] if nlo > 0 else 0
    _lib.call("pygw_minm_mt", ph, pd, vp(qd), int(nstart), int(nend), int(mstart), int(mend), vp(d[0]), vp(d[1]),
              vp(ga), al[0], nbg, vp(d[2]), vp(d[3]), vp(gp), ap[0], mbg, nlo, vp(k.s3r), k.n_mix, k.n_fnc,
              vp(k.pos), vp(k.cg), k.cg.numel(), vp(out), ld, int(col0), vp(workspace), workspace.numel(),
              _stream_ptr(device))
    return out


def _minc_common(plan, name, kv, b0, b1, c0, c1, a, b, g, s3r, corind, vpos, cgcoef, out, col0, shape):
    device = plan.device
    ad, bd = _bands(a, device), _bands(b, device)
    nlo = plan.nlomax
    gd = _bands(g, device) if nlo > 0 else None
    k = _consts(plan, s3r, vpos, cgcoef, corind)
    rows = (c0, c1) if name == "pygw_minc" else (b0, b1)
    _check_corind(plan, k.corind_host, rows)
    kd = _dev(kv if isinstance(kv, torch.Tensor) else np.asarray(kv, dtype=np.float64), torch.float64, device)
    if out is None:
        out = torch.empty((*shape, plan.loctmatsize), dtype=torch.complex128, device=device)
        ld = plan.loctmatsize
    else:
        if out.dtype != torch.complex128 or not out.is_contiguous() or tuple(out.shape[:2]) != shape:
            raise ValueError("out must be a contiguous complex128 tensor of the result's first two extents")
        ld = out.shape[2]
    ph, pd = plan.ptrs()
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    nbg = g.shape[0] if nlo > 0 else 0
    _lib.call(name, ph, pd, vp(kd), int(b0), int(b1), int(c0), int(c1), vp(ad), vp(bd), vp(gd), a.shape[0],
              nbg, nlo, vp(k.s3r), k.n_mix, k.n_fnc, vp(k.corind), k.ncg_rows, vp(k.pos), vp(k.cg), k.cg.numel(),
              vp(out), ld, int(col0), _stream_ptr(device))
    return out


def calc_minc(plan, kl, nstart, nend, cstart, cend, alfa, beta, gama, s3r, corind, vpos, cgcoef, *, out=None,
              col0=0):
    """fvxcn.calc_minc (pygw.py:2677, for_vxcnn.f90:1136-1273): mmat_c[ie1, ic,
    im] = <u_im | psi_ie1,k psi*_core>, complex128 [nend-nstart, cend-cstart,
    loctmatsize]; corind [ncg, 5] (atom type, atom, core index, l, m); or
    s3r = mt_constants(..., corind) (corind, vpos, cgcoef then unused)."""
    return _minc_common(plan, "pygw_minc", kl, nstart, nend, cstart, cend, alfa, beta, gama, s3r, corind, vpos,
                        cgcoef, out, col0, (nend - nstart, cend - cstart))


def calc_minc2(plan, kmql, cstart, cend, mstart, mend, alfp, betp, gamp, s3r, corind, vpos, cgcoef, *, out=None,
               col0=0):
    """fvxcn.calc_minc2 (pygw.py:2680, for_vxcnn.f90:1275-1397): mmat_c[ic, ie2,
    im] = <u_im | psi_core psi*_ie2,k-q>, complex128 [cend-cstart,
    mend-mstart, loctmatsize]."""
    return _minc_common(plan, "pygw_minc2", kmql, cstart, cend, mstart, mend, alfp, betp, gamp, s3r, corind, vpos,
                        cgcoef, out, col0, (cend - cstart, mend - mstart))


def fwi0(vol, mult, big_l, rtl):
    """<chi_im | 1> of the muffin-tin product functions (pygw.py:2216-2228):
    sqrt(4 pi/Vol) rtl[iat, irm] at the L = 0 functions (M = 0), 0 elsewhere;
    the product functions ordered atom, function, M as the matrix elements
    order them.  big_l: per atom type the L of each function; rtl [nat,
    max_nmix] = <r^L | u_irm>.  Host arithmetic, as in PyGW."""
    fct = np.sqrt(4 * np.pi / vol)
    out = []
    for iat, m in enumerate(np.asarray(mult)):
        for _ in range(int(m)):
            for irm, L in enumerate(big_l[iat]):
                block = np.zeros(2 * int(L) + 1)
                if L == 0:
                    block[0] = fct * rtl[iat][irm]
                out.append(block)
    return np.concatenate(out) if out else np.zeros(0)


def singular_coefficients(gqlen, vol, nkdiv):
    """singc1, singc2 of the q = 0 Coulomb singularity (pygw.py:2456-2474):
    intf_n - sum over the q-mesh and G (without q = G = 0) of
    e^{-alpha Q^2}/Q^n / N_k, n = 1, 2, alpha = (Vol/(6 pi^2))^(1/3),
    intf1 = Vol/(4 pi^2 alpha), intf2 = Vol sqrt(pi/alpha)/(4 pi^2).
    gqlen[iq] = |q + G| over the q-point's Coulomb plane waves, iq = 0 the
    Gamma point with its G = 0 first.  Host arithmetic, as in PyGW."""
    alpha = (vol / (6 * np.pi ** 2)) ** (1.0 / 3.0)
    ankp = int(np.prod(nkdiv))
    sf1 = sf2 = 0.0
    for iq, q in enumerate(gqlen):
        Q1 = np.asarray(q, dtype=np.float64)[1 if iq == 0 else 0:]
        e = np.exp(-alpha * Q1 ** 2)
        sf1 += np.sum(e / Q1) / ankp
        sf2 += np.sum(e / Q1 ** 2) / ankp
    intf = vol / (4 * np.pi ** 2)
    return intf / alpha - sf1, intf * np.sqrt(np.pi / alpha) - sf2
