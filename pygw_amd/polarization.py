"""Host side of the gfx950 polarization path, mirroring PyGW's call sites.

Each function takes the arguments of the reference routine it replaces and
runs the HIP kernels of libpygw_hip.so on the current torch device/stream.
torch only provides device memory and streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

COMPAT_REFERENCE = 3   # reproduce PyGW exactly (both switches)
COMPAT_EXACT = 0       # exact corner weights (see include/pygw_hip.h)


def _dev(x, dtype, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def _stream_ptr(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _require_device(device):
    if not torch.cuda.is_available():
        raise RuntimeError("the gfx950 polarization path needs a HIP device")
    return torch.device(device)


def bz_calcqdepw_par2(enk, ef, kqid, ncg, nomx, numin, omega, tetc, *, layout="fortran",
                      compat=COMPAT_REFERENCE, device="cuda", check=True, workspace=None):
    """Polarization weights kcw (ft.bz_calcqdepw_par2, for_tetrahedra.f90:390).

    enk [nbnd, nkp] (band index first, as PyGW's Fortran-ordered array),
    kqid [nkp] (= kqm.kqid[:, iq]), tetc [4, ntet]; 0-based indices.
    layout="fortran": returns a [n_o, n_e, nkp, nom] tensor with Fortran
    strides (PyGW's kcw); layout="native": [nkp, nom, n_o*n_e] contiguous.
    With check=True the call synchronises and raises on geometry errors;
    the check_weight warning count is stored in `bz_calcqdepw_par2.nwarn`.
    workspace: optional uint8 device tensor (pygw_tetra_weights_workspace
    gives the full size; a smaller one runs the band pairs in more chunks).
    """
    device = _require_device(device)
    enk_t = _dev(np.asarray(enk).T if not isinstance(enk, torch.Tensor) else enk.T, torch.float64, device)
    enk_t = enk_t.contiguous()  # [nkp, nbnd] row-major == Fortran (nbnd, nkp)
    nkp, nbnd = enk_t.shape
    tetc_np = np.asarray(tetc) if not isinstance(tetc, torch.Tensor) else tetc.cpu().numpy()
    tetc_t = torch.as_tensor(np.ascontiguousarray(tetc_np.T), dtype=torch.int32, device=device)  # [ntet, 4]
    ntet = tetc_t.shape[0]
    kq_t = _dev(kqid, torch.int32, device).contiguous()
    om_t = _dev(omega, torch.float64, device).contiguous()
    nom = om_t.numel()
    no, ne = ncg + nomx + 1, nbnd - ncg - numin
    if no <= 0 or ne <= 0:
        raise ValueError("empty band window")
    buf = torch.empty(no * ne * nkp * nom, dtype=torch.float64, device=device)
    lay = 0 if layout == "fortran" else 1
    info = (C.c_int * 2)()
    _lib.call("pygw_tetra_weights", C.c_void_p(enk_t.data_ptr()), float(ef), C.c_void_p(kq_t.data_ptr()),
              C.c_void_p(tetc_t.data_ptr()), int(ncg), int(nomx), int(numin), C.c_void_p(om_t.data_ptr()), nom, nkp,
              nbnd, ntet, lay, int(compat), C.c_void_p(buf.data_ptr()),
              None if workspace is None else C.c_void_p(workspace.data_ptr()),
              0 if workspace is None else workspace.numel(),
              C.cast(info, C.c_void_p) if check else None, _stream_ptr(device))
    if check:
        bz_calcqdepw_par2.nwarn = info[1]
        bz_calcqdepw_par2.status = info[0]
        if info[0] & 4:
            raise _lib.PygwError("pygw_tetra_weights", -10)
    if lay == 0:
        return buf.as_strided((no, ne, nkp, nom), (1, no, no * ne, no * ne * nkp))
    return buf.view(nkp, nom, no * ne)


bz_calcqdepw_par2.nwarn = 0
bz_calcqdepw_par2.status = 0


def _fortran_view(buf, pb, nom):
    return buf.as_strided((pb, pb, nom), (1, pb, pb * pb))


def chi0_accumulate(M, kcw_native, coef=-2.0, *, eps=None, precision="fp64", k0=0, device="cuda"):
    """chi0 contraction of calc_eps (pygw.py:3161-3200).

    M: complex [n_kc, n_pairs, PB] (minm2 of the k-points k0..k0+n_kc-1,
    pair p = i_o*n_e + i_e); kcw_native: [N_k, n_om, n_pairs] (weights in
    layout "native", all k).  precision "fp64" (M complex128) or "mixed"
    (M complex64, FP16 hi+lo split on MFMA).  Returns eps as a complex128
    [PB, PB, n_om] tensor with Fortran strides; with eps given (same
    layout), the contribution is added to it.
    """
    device = _require_device(device)
    dt = {"fp64": torch.complex128, "mixed": torch.complex64}[precision]
    M = M.to(device=device, dtype=dt).contiguous() if isinstance(M, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(M), dtype=dt, device=device)
    nkc, npair, pb = M.shape
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    if kc.shape[2] != npair or kc.shape[0] < k0 + nkc:
        raise ValueError("kcw_native must be [N_k, n_om, n_pairs] covering k0..k0+n_kc-1")
    nom = kc.shape[1]
    beta = 0
    if eps is None:
        buf = torch.empty(pb * pb * nom, dtype=torch.complex128, device=device)
    else:
        if eps.stride() != (1, pb, pb * pb) or eps.shape != (pb, pb, nom):
            raise ValueError("eps must be a Fortran-strided [PB, PB, n_om] complex128 tensor")
        buf = eps.as_strided((pb * pb * nom,), (1,))
        beta = 1
    prec = {"fp64": 0, "mixed": 1}[precision]
    ws, wsb = None, 0
    if prec == 1:  # fragment planes + scales, from torch's caching allocator
        sz = C.c_size_t(0)
        _lib.call("pygw_chi0_workspace", pb, npair, nom, nkc, prec, C.byref(sz))
        wbuf = torch.empty(sz.value, dtype=torch.uint8, device=device)
        ws, wsb = C.c_void_p(wbuf.data_ptr()), sz.value
    _lib.call("pygw_chi0_accumulate", C.c_void_p(M.data_ptr()), npair * pb, npair, pb,
              C.c_void_p(kc.data_ptr()), nom, int(k0), nkc, float(coef), prec, beta, C.c_void_p(buf.data_ptr()),
              ws, wsb, _stream_ptr(device))
    return _fortran_view(buf, pb, nom)


def omega_basis(kcw_native, *, balance=True, device="cuda", workspace=None):
    """Frequency basis of one q's weights kcw_native [N_k, n_om, n_pairs],
    from the eigenvectors U of the Gram matrix of the weights as an
    (N_k N_p) x n_om matrix, in decreasing order of eigenvalue.  balance:
    every frequency's column first scaled to unit norm (D the norms), so that
    every frequency keeps the same relative accuracy.  Returns (proj, expand,
    sv): proj = D^-1 U and expand = D U ([n_om, n_om] float64, column r for
    basis function r; proj = expand = U without balance) and the singular
    values sv [n_om] of the (scaled) matrix, all on the device."""
    device = _require_device(device)
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    nk, nom, npair = kc.shape
    Pm = torch.empty(nom, nom, dtype=torch.float64, device=device)
    Em = torch.empty(nom, nom, dtype=torch.float64, device=device)
    sv = torch.empty(nom, dtype=torch.float64, device=device)
    if workspace is None:
        sz = C.c_size_t(0)
        _lib.call("pygw_omega_basis_workspace", nk, nom, npair, C.byref(sz))
        workspace = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_omega_basis", C.c_void_p(kc.data_ptr()), nk, nom, npair, int(bool(balance)),
              C.c_void_p(Pm.data_ptr()), C.c_void_p(Em.data_ptr()), C.c_void_p(sv.data_ptr()), None,
              C.c_void_p(workspace.data_ptr()), workspace.numel(), _stream_ptr(device))
    return Pm, Em, sv


def omega_rank(sv, tol=1e-7, margin=1):
    """Basis size for the compressed contraction: the singular values above
    tol * sv[0] plus `margin` more (at most n_om).  The callers with
    omega_rank="auto" contract frequency by frequency instead when this is
    more than 3/4 of n_om (COMPRESS_MAX_FRACTION: too little to gain).  On the C5 weights nine
    singular values lie above 1e-7 of the first, the rest on a plateau near
    1.5e-8; the relative chi0 change is 1e-8 at rank 8 and 5e-10 at ranks
    10-12 (FP64, every frequency): one function past the threshold (rank
    10) is as accurate as more, each more costs 1/rank of the contraction."""
    s = sv.detach().cpu().numpy() if isinstance(sv, torch.Tensor) else np.asarray(sv)
    if s.size == 0 or not s[0] > 0:
        return 1
    return int(min(s.size, np.count_nonzero(s > tol * s[0]) + margin))


omega_rank_of = omega_rank
COMPRESS_MAX_FRACTION = 0.75


def omega_project(kcw_native, proj, rank, *, out=None, device="cuda"):
    """kcw_r[k, r, p] = sum_w proj[w, r] kcw[k, w, p] for r < rank (proj:
    the basis U of omega_basis, or PyGW's dUl [n_om, n_l])."""
    device = _require_device(device)
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    nk, nom, npair = kc.shape
    P = _dev(proj, torch.float64, device)
    if P.shape[0] != nom or P.shape[1] < rank:
        raise ValueError("proj must be [n_om, >= rank]")
    if out is None:
        out = torch.empty(nk, rank, npair, dtype=torch.float64, device=device)
    _lib.call("pygw_omega_project", C.c_void_p(kc.data_ptr()), nk, nom, npair, C.c_void_p(P.data_ptr()),
              P.stride(0), P.stride(1), int(rank), C.c_void_p(out.data_ptr()), _stream_ptr(device))
    return out


def omega_expand(X, expand, rank, *, eps=None, device="cuda"):
    """eps[:, :, w] (+)= sum_r expand[w, r] X[:, :, r]: X the Fortran-strided
    [PB, PB, rank] result of chi0_accumulate on projected weights, expand
    [n_om, >= rank] (the basis U, or PyGW's Ul transposed: eps = dot(eps, Ul),
    pygw.py:3224-3226).  Returns eps Fortran-strided [PB, PB, n_om]; with eps
    given, the expansion is added to it."""
    device = _require_device(device)
    pb = X.shape[0]
    if X.stride() != (1, pb, pb * pb) or X.shape[2] < rank:
        raise ValueError("X must be a Fortran-strided [PB, PB, >= rank] complex128 tensor")
    E = _dev(expand, torch.float64, device)
    nom = E.shape[0]
    beta = 0
    if eps is None:
        buf = torch.empty(pb * pb * nom, dtype=torch.complex128, device=device)
    else:
        if eps.stride() != (1, pb, pb * pb) or eps.shape != (pb, pb, nom):
            raise ValueError("eps must be a Fortran-strided [PB, PB, n_om] complex128 tensor")
        buf = eps.as_strided((pb * pb * nom,), (1,))
        beta = 1
    _lib.call("pygw_omega_expand", C.c_void_p(X.data_ptr()), C.c_void_p(E.data_ptr()), E.stride(0), E.stride(1), pb,
              int(rank), nom, beta, C.c_void_p(buf.data_ptr()), _stream_ptr(device))
    return _fortran_view(buf, pb, nom)


def chi0_compressed(M, kcw_native, coef=-2.0, *, rank=None, basis=None, precision="mixed", k0=0,
                    device="cuda"):
    """chi0 of calc_eps through a compressed frequency basis: project the
    weights on the q's own (balanced) basis (omega_basis; rank from
    omega_rank unless given), contract with `rank` basis functions in place
    of the n_om frequencies, expand back.  basis: (proj, expand, sv) of
    omega_basis.  Returns (eps [PB, PB, n_om] Fortran-strided, rank, sv)."""
    device = _require_device(device)
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    Pm, Em, sv = omega_basis(kc, device=device) if basis is None else basis
    if rank is None:
        rank = omega_rank(sv)
    kr = omega_project(kc, Pm, rank, device=device)
    X = chi0_accumulate(M, kr, coef, precision=precision, k0=k0, device=device)
    return omega_expand(X, Em, rank, device=device), rank, sv


def eps_inverse(eps, *, device="cuda", check=True):
    """In place: eps[:,:,w] <- ((eps+eps^H)/2 + 1)^-1 - 1 (pygw.py:3229-3271)."""
    device = _require_device(device)
    pb, pb2, nom = eps.shape
    if pb != pb2 or eps.stride() != (1, pb, pb * pb) or eps.dtype != torch.complex128:
        raise ValueError("eps must be a Fortran-strided [PB, PB, n_om] complex128 tensor")
    info = (C.c_int * 1)()
    sz = C.c_size_t(0)
    _lib.call("pygw_eps_inverse_workspace", pb, nom, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)   # from torch's caching allocator
    _lib.call("pygw_eps_inverse", C.c_void_p(eps.data_ptr()), pb, nom, C.c_void_p(ws.data_ptr()), sz.value,
              C.cast(info, C.c_void_p) if check else None, _stream_ptr(device))
    return eps


def calc_mwm(M_selfe, W, wkq, *, precision="fp64", device="cuda"):
    """MWM of calc_selfc (pygw.py:3353-3358) for all k at once.

    M_selfe: complex [n_k, nb1*nb2, PB]; W: Fortran [PB, PB, n_om] (eps after
    eps_inverse).  precision "fp64" or "mixed" (M as complex64, FP16 hi+lo
    MFMA).  Returns complex128 [n_k, n_om, nb1*nb2]."""
    device = _require_device(device)
    dt = {"fp64": torch.complex128, "mixed": torch.complex64}[precision]
    M = M_selfe.to(device=device, dtype=dt).contiguous() if isinstance(M_selfe, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(M_selfe), dtype=dt, device=device)
    nk, rpk, pb = M.shape
    if W.shape[0] != pb or W.stride() != (1, pb, pb * pb) or W.dtype != torch.complex128:
        raise ValueError("W must be a Fortran-strided [PB, PB, n_om] complex128 tensor")
    nom = W.shape[2]
    out = torch.empty((nk, nom, rpk), dtype=torch.complex128, device=device)
    prec = {"fp64": 0, "mixed": 1}[precision]
    sz = C.c_size_t(0)
    _lib.call("pygw_mwm_workspace", nk, rpk, pb, nom, prec, C.byref(sz))
    wbuf = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=device)
    _lib.call("pygw_mwm", C.c_void_p(M.data_ptr()), nk, rpk, pb, C.c_void_p(W.data_ptr()), nom, float(wkq),
              prec, C.c_void_p(out.data_ptr()), C.c_void_p(wbuf.data_ptr()), sz.value, _stream_ptr(device))
    return out


def _fortran_vec(t, pb, nom):
    """A Fortran-strided [PB, n_om] complex128 view of t (any layout accepted
    as input; returns a new tensor when a copy is needed)."""
    if t.shape == (pb, nom) and t.dtype == torch.complex128 and t.stride() == (1, pb):
        return t
    buf = torch.empty(pb * nom, dtype=torch.complex128, device=t.device)
    v = buf.as_strided((pb, nom), (1, pb))
    v.copy_(t)
    return v


def chi0_wings(M, kcw_native, pm, coef=-2.0, *, epsw1=None, k0=0, device="cuda"):
    """q = 0 wing of calc_eps (pygw.py:3177-3193): epsw1[I, iom] (+)=
    sum_k sum_p coef*kcw[k,iom,p]*M[k,p,I]*pm[k,p], i.e. PyGW's
    dot(tmat, pm) summed over k.  M [n_k, n_pairs, PB] complex128 or
    complex64, kcw_native [N_k, n_om, n_pairs] read from k-point k0, pm
    [n_k, n_pairs] complex.  Returns (or accumulates into) epsw1, a
    Fortran-strided [PB, n_om] complex128 tensor."""
    device = _require_device(device)
    if not isinstance(M, torch.Tensor):
        M = torch.as_tensor(np.ascontiguousarray(M), device=device)
    M = M.to(device)
    if M.dtype not in (torch.complex64, torch.complex128):
        M = M.to(torch.complex128)
    M = M.contiguous()
    nk, npair, pb = M.shape
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    nom = kc.shape[1]
    if kc.shape[2] != npair or kc.shape[0] < k0 + nk:
        raise ValueError("kcw_native must be [N_k >= k0 + n_k, n_om, n_pairs]")
    pmt = _dev(pm, torch.complex128, device).contiguous()
    if pmt.shape != (nk, npair):
        raise ValueError("pm must be [n_k, n_pairs]")
    beta = 1
    if epsw1 is None:
        epsw1 = torch.empty(pb * nom, dtype=torch.complex128, device=device).as_strided((pb, nom), (1, pb))
        beta = 0
    elif epsw1.shape != (pb, nom) or epsw1.stride() != (1, pb) or epsw1.dtype != torch.complex128:
        raise ValueError("epsw1 must be a Fortran-strided [PB, n_om] complex128 tensor")
    sz = C.c_size_t(0)
    _lib.call("pygw_chi0_wings_workspace", pb, nom, nk, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_chi0_wings", C.c_void_p(M.data_ptr()), npair * pb, npair, pb,
              1 if M.dtype == torch.complex64 else 0, C.c_void_p(kc.data_ptr()), nom, int(k0), nk, float(coef),
              C.c_void_p(pmt.data_ptr()), C.c_void_p(epsw1.data_ptr()), beta, C.c_void_p(ws.data_ptr()), sz.value,
              _stream_ptr(device))
    return epsw1


def eps_inverse_head(eps, head, epsw1, *, device="cuda", check=True):
    """q = 0 inversion (pygw.py:3202-3263): eps <- ((eps+eps^H)/2 + 1)^-1 - 1
    with the head/wings update between the inverse and the -1.  eps is
    Fortran [PB, PB, n_om] complex128 (in place); head [n_om] and epsw1
    [PB, n_om] are updated.  Returns (eps, head, epsw1, epsw2, emac) with
    emac [n_om, 2] = (1 + head - ww, 1 + head) of the incoming head."""
    device = _require_device(device)
    pb, pb2, nom = eps.shape
    if pb != pb2 or eps.stride() != (1, pb, pb * pb) or eps.dtype != torch.complex128:
        raise ValueError("eps must be a Fortran-strided [PB, PB, n_om] complex128 tensor")
    h = _dev(head, torch.complex128, device).contiguous().clone()
    if h.shape != (nom,):
        raise ValueError("head must be [n_om]")
    w1 = _fortran_vec(_dev(epsw1, torch.complex128, device), pb, nom)
    w2 = torch.empty(pb * nom, dtype=torch.complex128, device=device).as_strided((pb, nom), (1, pb))
    emac = torch.empty((nom, 2), dtype=torch.complex128, device=device)
    info = (C.c_int * 1)()
    _lib.call("pygw_eps_inverse_head", C.c_void_p(eps.data_ptr()), pb, nom, C.c_void_p(h.data_ptr()),
              C.c_void_p(w1.data_ptr()), C.c_void_p(w2.data_ptr()), C.c_void_p(emac.data_ptr()), None, 0,
              C.cast(info, C.c_void_p) if check else None, _stream_ptr(device))
    return eps, h, w1, w2, emac


def calc_mwm_head(M_selfe, mwm, head, epsw1, epsw2, coefs1, coefs2, ncg_c, nb1, nb2, *, device="cuda"):
    """q = 0 terms of calc_selfc (pygw.py:3359-3364) added in place to mwm
    (the [n_k, n_om, nb1*nb2] output of calc_mwm) at the rows
    (b, b + ncg_c): coefs2*head + coefs1*M.epsw2 + coefs1*conj(M).epsw1."""
    device = _require_device(device)
    M = M_selfe if isinstance(M_selfe, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(M_selfe))
    M = M.to(device)
    if M.dtype not in (torch.complex64, torch.complex128):
        M = M.to(torch.complex128)
    M = M.contiguous()
    nk, rpk, pb = M.shape
    nom = mwm.shape[1]
    if rpk != nb1 * nb2 or mwm.shape != (nk, nom, rpk) or not mwm.is_contiguous() or mwm.dtype != torch.complex128:
        raise ValueError("mwm must be the contiguous [n_k, n_om, nb1*nb2] complex128 output of calc_mwm")
    hd = _dev(head, torch.complex128, device).contiguous()
    w1 = _fortran_vec(_dev(epsw1, torch.complex128, device), pb, nom)
    w2 = _fortran_vec(_dev(epsw2, torch.complex128, device), pb, nom)
    _lib.call("pygw_mwm_head", C.c_void_p(M.data_ptr()), nk, nb1, nb2, pb, int(ncg_c),
              1 if M.dtype == torch.complex64 else 0, C.c_void_p(hd.data_ptr()), C.c_void_p(w1.data_ptr()),
              C.c_void_p(w2.data_ptr()), nom, float(coefs1), float(coefs2), C.c_void_p(mwm.data_ptr()),
              _stream_ptr(device))
    return mwm


def compute_selfc_inside(enk, mwm, omega, womeg, omega_p=None, womeg_p=None, *, device="cuda"):
    """Sigma_c contribution of one q for all k (Compute_selfc_inside,
    mcommon.py:80-165).  enk [n_k, nb2] (energies at k-q), mwm complex
    [n_k, n_om, nb1, nb2].  With omega_p/womeg_p the spline refinement onto
    the precise mesh is used (iopfreq = 4 branch), else the coarse
    all_fr_convolutions.  Returns complex128 [n_k, nb1, n_om]."""
    device = _require_device(device)
    mw = mwm.to(device=device, dtype=torch.complex128).contiguous() if isinstance(mwm, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(mwm), dtype=torch.complex128, device=device)
    nk, nom, nb1, nb2 = mw.shape
    e = _dev(enk, torch.float64, device).contiguous()
    om = _dev(omega, torch.float64, device).contiguous()
    wo = _dev(womeg, torch.float64, device).contiguous()
    out = torch.empty((nk, nb1, nom), dtype=torch.complex128, device=device)
    s = _stream_ptr(device)
    if omega_p is None:
        _lib.call("pygw_selfc_convolve", C.c_void_p(e.data_ptr()), nk, nb1, nb2, C.c_void_p(mw.data_ptr()),
                  C.c_void_p(om.data_ptr()), nom, None, C.c_void_p(om.data_ptr()), C.c_void_p(wo.data_ptr()), nom,
                  C.c_void_p(out.data_ptr()), s)
        return out
    op = _dev(omega_p, torch.float64, device).contiguous()
    wp = _dev(womeg_p, torch.float64, device).contiguous()
    nomp = op.numel()
    mp = torch.empty((nk, nb1, nb2, nomp), dtype=torch.complex128, device=device)
    _lib.call("pygw_selfc_spline", C.c_void_p(mw.data_ptr()), nk, nb1, nb2, C.c_void_p(om.data_ptr()), nom,
              C.c_void_p(op.data_ptr()), nomp, C.c_void_p(mp.data_ptr()), s)
    _lib.call("pygw_selfc_convolve", C.c_void_p(e.data_ptr()), nk, nb1, nb2, C.c_void_p(mw.data_ptr()),
              C.c_void_p(om.data_ptr()), nom, C.c_void_p(mp.data_ptr()), C.c_void_p(op.data_ptr()),
              C.c_void_p(wp.data_ptr()), nomp, C.c_void_p(out.data_ptr()), s)
    return out


def pade_quasiparticles(sigc, omega, ene, sigx=None, vxc=None, npade=32, *, device="cuda"):
    """Pade continuation (iop_ac = 0) and G0W0 quasiparticle energies for
    every state (mcommon.py:167-210, 315-343).  sigc complex [n_state, n_om].
    Returns float64 [n_state, 6]: Re S(e), Im S(e), Re S', Im S', Z, E_qp."""
    device = _require_device(device)
    sg = sigc.to(device=device, dtype=torch.complex128).contiguous() if isinstance(sigc, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(sigc), dtype=torch.complex128, device=device)
    ns, nom = sg.shape
    om = _dev(omega, torch.float64, device).contiguous()
    en = _dev(ene, torch.float64, device).contiguous()
    sx = _dev(sigx, torch.float64, device).contiguous() if sigx is not None else None
    vx = _dev(vxc, torch.float64, device).contiguous() if vxc is not None else None
    out = torch.empty((ns, 6), dtype=torch.float64, device=device)
    _lib.call("pygw_pade_qp", C.c_void_p(sg.data_ptr()), ns, nom, C.c_void_p(om.data_ptr()), min(npade, nom),
              C.c_void_p(en.data_ptr()), C.c_void_p(sx.data_ptr()) if sx is not None else None,
              C.c_void_p(vx.data_ptr()) if vx is not None else None, C.c_void_p(out.data_ptr()), _stream_ptr(device))
    return out


def padecof(gn, zn, *, device="cuda"):
    """Thiele reciprocal differences (fpade.padecof, for_pade.f90:3-25).
    gn complex [n] or [n_state, n] on the nodes zn [n] (n <= 64); returns
    complex128 coefficients of the same shape."""
    device = _require_device(device)
    g = _dev(gn, torch.complex128, device)
    one = g.dim() == 1
    g = g.reshape(1, -1).contiguous() if one else g.contiguous()
    z = _dev(zn, torch.complex128, device).contiguous()
    ns, n = g.shape
    if z.numel() != n:
        raise ValueError("gn and zn must have the same number of points")
    coef = torch.empty_like(g)
    _lib.call("pygw_padecof", C.c_void_p(z.data_ptr()), C.c_void_p(g.data_ptr()), n, ns, C.c_void_p(coef.data_ptr()),
              _stream_ptr(device))
    return coef[0] if one else coef


def padeg(x, zn, coef, *, device="cuda"):
    """The continued fraction of padecof's coefficients at the complex points
    x [m] (fpade.padeg, for_pade.f90:30-51); coef [n] or [n_state, n].
    Returns complex128 [m] or [n_state, m]."""
    device = _require_device(device)
    c = _dev(coef, torch.complex128, device)
    one = c.dim() == 1
    c = c.reshape(1, -1).contiguous() if one else c.contiguous()
    z = _dev(zn, torch.complex128, device).contiguous()
    xv = _dev(x, torch.complex128, device).reshape(-1).contiguous()
    ns, n = c.shape
    if z.numel() != n:
        raise ValueError("coef and zn must have the same number of points")
    out = torch.empty((ns, xv.numel()), dtype=torch.complex128, device=device)
    _lib.call("pygw_padeg", C.c_void_p(xv.data_ptr()), xv.numel(), C.c_void_p(z.data_ptr()), C.c_void_p(c.data_ptr()),
              n, ns, C.c_void_p(out.data_ptr()), _stream_ptr(device))
    return out[0] if one else out


def bz_calcqdepw_par3(enk, ef, kqid, ncg, nomx, numin, dUl, omega, tetc, *, layout="fortran",
                      compat=COMPAT_REFERENCE, device="cuda", check=True):
    """SVD-basis weights (ft.bz_calcqdepw_par3, for_tetrahedra.f90:466-542):
    the weights of bz_calcqdepw_par2 projected on dUl [nom, n_l]."""
    device = _require_device(device)
    kn = bz_calcqdepw_par2(enk, ef, kqid, ncg, nomx, numin, omega, tetc, layout="native", compat=compat,
                           device=device, check=check)
    nkp, nom, npair = kn.shape
    du = torch.as_tensor(np.asfortranarray(dUl, dtype=np.float64).reshape(-1, order="F"), device=device)
    nl = np.asarray(dUl).shape[1]
    nbnd = np.asarray(enk).shape[0]
    no, ne = ncg + nomx + 1, nbnd - ncg - numin
    out = torch.empty(no * ne * nkp * nl, dtype=torch.float64, device=device)
    lay = 0 if layout == "fortran" else 1
    _lib.call("pygw_tetra_weights_svd", C.c_void_p(kn.data_ptr()), nkp, nom, no, ne, C.c_void_p(du.data_ptr()), nl,
              lay, C.c_void_p(out.data_ptr()), _stream_ptr(device))
    if lay == 0:
        return out.as_strided((no, ne, nkp, nl), (1, no, no * ne, no * ne * nkp))
    return out.view(nkp, nl, no * ne)


def autotune_chi0(M, kcw_native, coef=-2.0, *, precision="fp64", k0=0, device="cuda"):
    """Pick the chi0 schedule s in {1, 2, 4} for this (precision, PB, n_pairs,
    n_om) by timing the candidates on the given inputs (FP64: s Matsubara
    frequencies per tile block; mixed: FP32 tile sums folded into FP64 every
    64*s chunks); the choice is cached in the library for later calls.
    Returns (best, {candidate: ms})."""
    device = _require_device(device)
    dt = {"fp64": torch.complex128, "mixed": torch.complex64}[precision]
    M = M.to(device=device, dtype=dt).contiguous() if isinstance(M, torch.Tensor) else \
        torch.as_tensor(np.ascontiguousarray(M), dtype=dt, device=device)
    nkc, npair, pb = M.shape
    kc = _dev(kcw_native, torch.float64, device).contiguous()
    nom = kc.shape[1]
    scratch = torch.empty(pb * pb * nom, dtype=torch.complex128, device=device)
    best = C.c_int(0)
    times = (C.c_float * 3)()
    _lib.call("pygw_chi0_autotune", C.c_void_p(M.data_ptr()), npair * pb, npair, pb, C.c_void_p(kc.data_ptr()),
              nom, int(k0), nkc, float(coef), {"fp64": 0, "mixed": 1}[precision], C.c_void_p(scratch.data_ptr()),
              C.byref(best), C.cast(times, C.c_void_p), _stream_ptr(device))
    return best.value, dict(zip((1, 2, 4), list(times)))


def set_chi0_schedule(precision, pb, n_pairs, n_om, nw):
    """Force (nw = 1, 2, 4) or clear (0) the chi0 schedule of a shape."""
    _lib.call("pygw_chi0_set_schedule", {"fp64": 0, "mixed": 1}[precision], pb, n_pairs, n_om, nw)


class QPlan:
    """Device buffers for the per-q steps of one workload (weights -> chi0 ->
    eps^-1), allocated once.  run() launches with explicit workspaces and
    no allocation, host synchronisation or memset, so it can be captured in
    a HIP graph (torch.cuda.CUDAGraph) and replayed for every q.
    Per-q inputs: set_q(kqid column) copies into a fixed buffer; M is taken
    by pointer (one graph per M buffer)."""

    def __init__(self, enk, ef, ncg, nomx, numin, omega, tetc, pb, precision="fp64", compat=COMPAT_REFERENCE,
                 device="cuda"):
        self.device = d = _require_device(device)
        enk = np.asarray(enk)
        self.nbnd, self.nk = enk.shape
        self.ef, self.ncg, self.nomx, self.numin = float(ef), int(ncg), int(nomx), int(numin)
        self.no, self.ne = ncg + nomx + 1, self.nbnd - ncg - numin
        self.np = self.no * self.ne
        self.pb, self.precision, self.compat = int(pb), precision, int(compat)
        self.enk = torch.as_tensor(np.ascontiguousarray(enk.T), dtype=torch.float64, device=d)
        self.tetc = torch.as_tensor(np.ascontiguousarray(np.asarray(tetc).T), dtype=torch.int32, device=d)
        self.ntet = self.tetc.shape[0]
        self.om = torch.as_tensor(np.asarray(omega, dtype=np.float64), device=d)
        self.nom = self.om.numel()
        self.kq = torch.zeros(self.nk, dtype=torch.int32, device=d)
        self.kcw = torch.empty((self.nk, self.nom, self.np), dtype=torch.float64, device=d)
        self.eps_buf = torch.empty(self.pb * self.pb * self.nom, dtype=torch.complex128, device=d)
        self.eps = _fortran_view(self.eps_buf, self.pb, self.nom)
        sz = C.c_size_t(0)
        _lib.call("pygw_tetra_weights_workspace", self.nk, self.np, self.nom, C.byref(sz))
        self.ws_tw = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        _lib.call("pygw_chi0_workspace", self.pb, self.np, self.nom, self.nk, {"fp64": 0, "mixed": 1}[precision],
                  C.byref(sz))
        self.ws_c0 = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        _lib.call("pygw_eps_inverse_workspace", self.pb, self.nom, C.byref(sz))
        self.ws_ei = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)

    def set_q(self, kqid_col):
        self.kq.copy_(torch.as_tensor(np.asarray(kqid_col, dtype=np.int32)))

    def run(self, M):
        s = _stream_ptr(self.device)
        if M.shape != (self.nk, self.np, self.pb) or not M.is_contiguous():
            raise ValueError("M must be a contiguous [N_k, n_pairs, PB] device tensor")
        _lib.call("pygw_tetra_weights", C.c_void_p(self.enk.data_ptr()), self.ef, C.c_void_p(self.kq.data_ptr()),
                  C.c_void_p(self.tetc.data_ptr()), self.ncg, self.nomx, self.numin, C.c_void_p(self.om.data_ptr()),
                  self.nom, self.nk, self.nbnd, self.ntet, 1, self.compat, C.c_void_p(self.kcw.data_ptr()),
                  C.c_void_p(self.ws_tw.data_ptr()), self.ws_tw.numel(), None, s)
        prec = {"fp64": 0, "mixed": 1}[self.precision]
        _lib.call("pygw_chi0_accumulate", C.c_void_p(M.data_ptr()), self.np * self.pb, self.np, self.pb,
                  C.c_void_p(self.kcw.data_ptr()), self.nom, 0, self.nk, -2.0, prec, 0,
                  C.c_void_p(self.eps_buf.data_ptr()), C.c_void_p(self.ws_c0.data_ptr()), self.ws_c0.numel(), s)
        _lib.call("pygw_eps_inverse", C.c_void_p(self.eps_buf.data_ptr()), self.pb, self.nom,
                  C.c_void_p(self.ws_ei.data_ptr()), self.ws_ei.numel(), None, s)
        return self.eps

    def capture(self, M):
        """Capture run(M) in a HIP graph; returns the torch.cuda.CUDAGraph."""
        g = torch.cuda.CUDAGraph()
        self.run(M)                     # warm-up outside the graph (module load)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self.run(M)
        return g


class Chi0Batch:
    """Split-precision chi0 for a batch of q-points with their own PB and
    pair counts, in one launch (the q loop of G0W0.Compute,
    pygw.py:3654-3718, around the contraction of calc_eps,
    pygw.py:3161-3200).

    Ms[q]: complex64 [K_q, PB_q] device tensors (the window pairs of all k
    of q concatenated); kcws[q]: float64 [n_om, K_q] device tensors (their
    weights, layout of the kernel-native kcw with the pairs gathered).  The
    eps of every q (complex128, Fortran [PB_q, PB_q, n_om]) and the
    workspace are allocated once; run() only launches kernels and can be
    captured in a HIP graph (capture()).  With inverse=True run() continues
    with eps^-1 of every q (pygw_eps_inverse, explicit workspaces).

    omega_rank (None, "auto" or an int): contract every q in its own balanced
    frequency basis (omega_basis of its weights, recomputed by every run())
    with a common number of basis functions, chosen once from the weights
    given here ("auto": the largest omega_rank over the q's), and expand
    each q's result back to the n_om frequencies."""

    def __init__(self, Ms, kcws, coef=-2.0, inverse=False, omega_rank=None, device="cuda"):
        self.device = d = _require_device(device)
        if len(Ms) != len(kcws) or not Ms:
            raise ValueError("need one kcw per M")
        self.Ms = [m.to(device=d, dtype=torch.complex64).contiguous() for m in Ms]
        self.kcws = [w.to(device=d, dtype=torch.float64).contiguous() for w in kcws]
        self.nq = len(Ms)
        self.nom = self.kcws[0].shape[0]
        self.pb = np.array([m.shape[1] for m in self.Ms], dtype=np.int32)
        self.np = np.array([m.shape[0] for m in self.Ms], dtype=np.int32)
        for m, w in zip(self.Ms, self.kcws):
            if w.shape != (self.nom, m.shape[0]):
                raise ValueError("kcw[q] must be [n_om, K_q]")
        self.eps_buf = [torch.empty(int(p) * int(p) * self.nom, dtype=torch.complex128, device=d) for p in self.pb]
        self.eps = [_fortran_view(b, int(p), self.nom) for b, p in zip(self.eps_buf, self.pb)]
        ptr = lambda ts: (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])  # noqa: E731
        self.omega_rank, self.rank = omega_rank, self.nom
        kin, ein, nom_c = self.kcws, self.eps_buf, self.nom
        if omega_rank is not None:
            nq, nom = self.nq, self.nom
            self._wp = ptr(self.kcws)
            self._ones = np.ones(nq, dtype=np.int32)
            sz = C.c_size_t(0)
            _lib.call("pygw_omega_basis_many_workspace", nq, nom, C.byref(sz))
            self.ws_ob = torch.empty(sz.value, dtype=torch.uint8, device=d)
            self.Up = torch.empty((nq, nom, nom), dtype=torch.float64, device=d)
            self.Ue = torch.empty((nq, nom, nom), dtype=torch.float64, device=d)
            self.sv = torch.empty((nq, nom), dtype=torch.float64, device=d)
            self._bases()
            self.rank = (max(omega_rank_of(self.sv[q]) for q in range(nq)) if omega_rank == "auto"
                         else int(omega_rank))
            if omega_rank == "auto" and self.rank > COMPRESS_MAX_FRACTION * nom:
                self.omega_rank, self.rank = None, nom        # too little to gain: frequency by frequency
        if self.omega_rank is not None:
            self.kr = [torch.empty((self.rank, int(k)), dtype=torch.float64, device=d) for k in self.np]
            self.x_buf = [torch.empty(int(p) * int(p) * self.rank, dtype=torch.complex128, device=d) for p in self.pb]
            self._project()
            kin, ein, nom_c = self.kr, self.x_buf, self.rank
        sz = C.c_size_t(0)
        _lib.call("pygw_chi0_batch_workspace", self.nq, C.c_void_p(self.pb.ctypes.data),
                  C.c_void_p(self.np.ctypes.data), nom_c, C.byref(sz))
        self.ws = torch.empty(sz.value, dtype=torch.uint8, device=d)
        self._mp, self._kp, self._ep = ptr(self.Ms), ptr(kin), ptr(ein)
        self._nom_c = nom_c
        self.inverse = bool(inverse)
        self.ws_inv = []
        self._side = [torch.cuda.Stream(device=d) for _ in range(4)] if self.inverse else []  # default priority
        if self.inverse:  # eps^-1 of every q after the contraction (explicit workspaces)
            si = C.c_size_t(0)
            for p in self.pb:
                _lib.call("pygw_eps_inverse_workspace", int(p), self.nom, C.byref(si))
                self.ws_inv.append(torch.empty(si.value, dtype=torch.uint8, device=d))
        self.plan = C.c_void_p(0)
        _lib.call("pygw_chi0_batch_plan", self.nq, C.cast(self._mp, C.c_void_p), C.c_void_p(self.pb.ctypes.data),
                  C.c_void_p(self.np.ctypes.data), C.cast(self._kp, C.c_void_p), self._nom_c, float(coef),
                  C.cast(self._ep, C.c_void_p), C.c_void_p(self.ws.data_ptr()), sz.value, _stream_ptr(d),
                  C.byref(self.plan))

    def _bases(self):
        """The balanced frequency basis of every q's weights (one launch of
        n_q eigen-decompositions)."""
        _lib.call("pygw_omega_basis_many", self.nq, C.cast(self._wp, C.c_void_p), C.c_void_p(self._ones.ctypes.data),
                  C.c_void_p(self.np.ctypes.data), self.nom, 1, C.c_void_p(self.Up.data_ptr()),
                  C.c_void_p(self.Ue.data_ptr()), C.c_void_p(self.sv.data_ptr()), None,
                  C.c_void_p(self.ws_ob.data_ptr()), self.ws_ob.numel(), _stream_ptr(self.device))

    def _project(self):
        s = _stream_ptr(self.device)
        for q in range(self.nq):
            _lib.call("pygw_omega_project", C.c_void_p(self.kcws[q].data_ptr()), 1, self.nom, int(self.np[q]),
                      C.c_void_p(self.Up[q].data_ptr()), self.nom, 1, self.rank, C.c_void_p(self.kr[q].data_ptr()), s)

    def run(self, schedule=0, inverse=None):
        """Launch the batch (schedule 0 = autotuned / default fold interval);
        inverse (default: the constructor's) continues with eps^-1 per q.
        With omega_rank: bases and projections of every q first, the
        expansions after the contraction."""
        s = _stream_ptr(self.device)
        vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        if self.omega_rank is not None:
            self._bases()
            self._project()
        _lib.call("pygw_chi0_batch_run", self.plan, int(schedule), s)
        if self.omega_rank is not None:
            for q in range(self.nq):
                _lib.call("pygw_omega_expand", vp(self.x_buf[q]), vp(self.Ue[q]), self.nom, 1, int(self.pb[q]),
                          self.rank, self.nom, 0, vp(self.eps_buf[q]), s)
        if not (self.inverse if inverse is None else inverse):
            return self.eps
        return self.invert()

    def invert(self, streams=4):
        """eps^-1 - 1 of every q in place (the eps of the last run()).  The
        q's are independent: they are dealt (largest PB first) to `streams`
        side streams that fork from and join the current stream, so small
        inversions overlap (also when captured in a HIP graph)."""
        if not self.ws_inv:
            raise ValueError("Chi0Batch built without inverse=True has no inversion workspace")
        main = torch.cuda.current_stream(self.device)
        ns = max(1, min(int(streams), self.nq))
        if ns == 1:
            side = [main]
        else:
            if len(self._side) < ns:
                self._side += [torch.cuda.Stream(device=self.device) for _ in range(ns - len(self._side))]
            side = self._side[:ns]
            for st in side:
                st.wait_stream(main)
        order = np.argsort(-self.pb, kind="stable")
        for n, q in enumerate(order):
            st = side[n % ns]
            b, p, w = self.eps_buf[q], int(self.pb[q]), self.ws_inv[q]
            _lib.call("pygw_eps_inverse", C.c_void_p(b.data_ptr()), p, self.nom, C.c_void_p(w.data_ptr()),
                      w.numel(), None, C.c_void_p(st.cuda_stream))
        if ns > 1:
            for st in side:
                main.wait_stream(st)
        return self.eps

    def autotune(self):
        best = C.c_int(0)
        times = (C.c_float * 3)()
        _lib.call("pygw_chi0_batch_autotune", self.plan, C.byref(best), C.cast(times, C.c_void_p),
                  _stream_ptr(self.device))
        return best.value, dict(zip((1, 2, 4), list(times)))

    def capture(self, schedule=0, inverse=None):
        """Capture run() in a HIP graph; returns the torch.cuda.CUDAGraph."""
        g = torch.cuda.CUDAGraph()
        self.run(schedule, inverse)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self.run(schedule, inverse)
        return g

    def close(self):
        if self.plan:
            _lib.call("pygw_chi0_batch_free", self.plan)
            self.plan = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
