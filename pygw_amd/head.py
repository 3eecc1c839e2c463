"""The head of the dielectric matrix at q = 0 and the momentum matrix
elements behind it, on the gfx950 backend (calc_head, pygw.py:2808-3037).

* ``HeadTables`` uploads what calc_head takes from the run's setup once:
  the atoms (strc.mult), in1.nLO_at, the radial integrals of the momentum
  operator (ks.iul_ul & co and Give_fortran_ilocals' local-orbital tables,
  pygw.py:1338-1380) and of the core states (ks.iul_ucl & co, concatenated
  over atom types, the first local-orbital slot dropped as calc_head passes
  them, pygw.py:2918-2920), and core.corind.
* ``momentum_elements`` replaces the per-irk double loop over bands of
  calc_head (f_q_0.calcmmatvv_mt2 / calcmmatvv_mt / calcmmatvv_is, averaged
  over the two orders, and calcmmatcv per atom; pygw.py:2870-2932).
* ``head_from_elements`` replaces the head sums (pygw.py:2940-3025).
* ``calc_head`` runs the whole routine: LAPW coefficients without local
  rotations (ifrot False, pygw.py:2857), alfa/beta/gama, the elements and
  the head; it returns (head, mmatcv, mmatvv, mst) as PyGW does.
* ``wing_pm`` replaces calc_eps's pm of the q = 0 wing (pygw.py:3122-3141).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, lapw
from .polarization import _require_device, _stream_ptr


def _f64(a, device):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F")), device=device)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a).ravel(order="F"), dtype=np.int32)


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _bands(x, device):
    """Band coefficients in PyGW's index order, stored Fortran-ordered on the device."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.complex128).permute(*reversed(range(x.dim()))).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.complex128).ravel(order="F"), device=device)


class HeadTables:
    """vv_tables: iul_ul, iul_udl, iudl_ul, iudl_udl [2, lmax, nat];
    iulol_ul, iulol_udl, iul_ulol, iudl_ulol [2, nlo, lox, nat]; iulol_ulol
    [2, nlo, nlo, lox, nat].  cv_tables: iul_ucl, iudl_ucl, iucl_ul,
    iucl_udl [2, ncore_total]; iulol_ucl, iucl_ulol [2, nloat,
    ncore_total]; core_off[iat] = the first core of atom type iat in them.
    corind [ncg, 5] (atom type, atom, core index, l, m)."""

    def __init__(self, mult, nLO_at, vv_tables, core_off=None, cv_tables=None, corind=None, *, device="cuda"):
        device = _require_device(device)
        self.device = device
        self.mult, self.nLO_at = _i32(mult), _i32(nLO_at)
        self.nrows = np.asarray(nLO_at).shape[0]
        vv = [np.asarray(t, dtype=np.float64) for t in vv_tables]
        self.lmax = vv[0].shape[1]
        self.nlo, self.lox = vv[4].shape[1], vv[4].shape[2]
        self.vv = [_f64(t, device) for t in vv]
        self.vv_ptrs = (C.c_void_p * 9)(*[t.data_ptr() for t in self.vv])
        self.ncg = 0
        if cv_tables is not None:
            cv = [np.asarray(t, dtype=np.float64) for t in cv_tables]
            self.nct, self.nloat = cv[0].shape[1], cv[4].shape[1]
            self.cv = [_f64(t, device) for t in cv]
            self.cv_ptrs = (C.c_void_p * 6)(*[t.data_ptr() for t in self.cv])
            self.core_off = _i32(core_off)
            co = np.asarray(corind)
            self.ncg = co.shape[0]
            self.corind = torch.as_tensor(_i32(co), device=device)

    def layout_args(self):
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        return len(self.mult), p(self.mult), p(self.nLO_at), self.nrows, self.lmax


class PlaneWaves:
    """The interstitial inputs of calcmmatvv_is that do not change with k:
    pw.ipwint [ng] complex, pw.gindex [ng, 3], the index map i_g0 [2n1+1,
    2n2+1, 2n3+1] (0-based rows of gindex) and kqm.k2cartes."""

    def __init__(self, ipwint, gindex, i_g0, k2cartes, *, device="cuda"):
        device = _require_device(device)
        self.ipwint = torch.as_tensor(np.asarray(ipwint, dtype=np.complex128), device=device)
        g = np.asarray(gindex)
        self.ng = g.shape[0]
        self.gindex = torch.as_tensor(_i32(g), device=device)
        i0 = np.asarray(i_g0)
        self.i_g0 = torch.as_tensor(_i32(i0), device=device)
        self.n_g0 = np.array([(i0.shape[c] - 1) // 2 for c in range(3)], dtype=np.int32)
        self.k2cartes = np.ascontiguousarray(np.asarray(k2cartes, dtype=np.float64).ravel(order="F"))


def momentum_elements(tables, pw, abc, Aeig, kl, indgk, nvk, nst, nend, mst, mend, *, mmatvv=None, mmatcv=None,
                      exact=False, workspace=None):
    """One irreducible k of calc_head's loop: mmatvv [mend-mst, nend-nst, 3]
    = (p + conj(p'))/2 of the muffin-tin + interstitial elements and mmatcv
    [mend-mst, ncg, 3] from calcmmatcv of each core state's atom (exact:
    local orbitals at l-1 with iulol_ucl where the reference takes iudl_ucl,
    for_q_0.f90:648).  abc = (alfa, beta, gama) in PyGW's index order
    (numpy, or the views lapw.band_coefficients returns); Aeig [nbmax, ngk]."""
    device = tables.device
    alfa, beta, gama = abc
    a, b = _bands(alfa, device), _bands(beta, device)
    nbmax, ntnt = alfa.shape[0], alfa.shape[1]
    if ntnt != (tables.lmax + 1) ** 2:
        raise ValueError("alfa's second extent must be (lmax+1)^2 of the radial tables")
    # gama is read only where nLO_at has local orbitals (PyGW passes a
    # (1, 1, 1, 1) placeholder when there are none)
    nlomax = gama.shape[1] if int(tables.nLO_at.max(initial=0)) > 0 else 0
    g = _bands(gama, device) if nlomax else None
    nbg = gama.shape[0] if nlomax else 0
    A = Aeig if isinstance(Aeig, torch.Tensor) else torch.as_tensor(np.asarray(Aeig, dtype=np.complex128))
    A = A.to(device=device, dtype=torch.complex128)
    if A.stride(1) != 1:
        A = A.contiguous()
    ig = indgk if isinstance(indgk, torch.Tensor) else torch.as_tensor(np.asarray(indgk, dtype=np.int32))
    ig = ig.to(device=device, dtype=torch.int32).contiguous()
    n1, n2 = nend - nst, mend - mst
    if mmatvv is None:
        mmatvv = torch.empty((n2, n1, 3), dtype=torch.complex128, device=device)
    sz = C.c_size_t(0)
    _lib.call("pygw_mmat_vv_workspace", int(tables.mult.sum()), tables.lmax, nlomax, int(nvk), nst, nend, mst, mend,
              C.byref(sz))
    if workspace is None or workspace.numel() < sz.value:
        workspace = torch.empty(max(sz.value, 1), dtype=torch.uint8, device=device)
    k = np.ascontiguousarray(np.asarray(kl, dtype=np.float64))
    stream = _stream_ptr(device)
    _lib.call("pygw_mmat_vv", *tables.layout_args(), _vp(a), _vp(b), _vp(g), nbmax, nbg, nlomax,
              C.cast(tables.vv_ptrs, C.c_void_p), tables.nlo, tables.lox, _vp(A), A.stride(0),
              k.ctypes.data_as(C.c_void_p), _vp(ig), int(nvk), _vp(pw.ipwint), _vp(pw.gindex), pw.ng, _vp(pw.i_g0),
              pw.n_g0.ctypes.data_as(C.c_void_p), pw.k2cartes.ctypes.data_as(C.c_void_p), nst, nend, mst, mend,
              _vp(mmatvv), _vp(workspace), workspace.numel(), stream)
    if tables.ncg:
        if mmatcv is None:
            mmatcv = torch.empty((n2, tables.ncg, 3), dtype=torch.complex128, device=device)
        _lib.call("pygw_mmat_cv", *tables.layout_args(), tables.core_off.ctypes.data_as(C.c_void_p), _vp(a), _vp(b),
                  _vp(g), nbmax, nbg, nlomax, C.cast(tables.cv_ptrs, C.c_void_p), tables.nloat, tables.nct,
                  _vp(tables.corind), tables.ncg, mst, mend, 1 if exact else 0, _vp(mmatcv), stream)
    return mmatvv, mmatcv


def head_from_elements(mmatvv, mmatcv, Ebnd, eig_core, kcw, k_ind, weight, nomx, numin, nbmaxpol, ncg_p, vol, *,
                       Eg=1.0, kwfer=None, iop_drude=0, eta_head=0.0, omega=None, dUl=None, workspace=None):
    """head [nom_nl] complex128 from mmatvv [nirk, nbmaxpol-numin, nomx+1, 3],
    mmatcv [nirk, nbmaxpol-numin, ncg, 3], Ebnd [nirk, >= nbmaxpol] (no
    core), eig_core [ncg] (per core state), kcw Fortran [ncg_p+nomx+1,
    nbmaxpol-numin, N_k, nom_nl] (the layout bz_calcqdepw_par2 returns),
    k_ind [nirk] (irk -> k of kcw), weight [nirk]; the Drude term with
    Eg <= 0 and iop_drude 1 (omega [n_om], dUl [n_om, nom_nl] in SVD mode)."""
    device = mmatvv.device
    d = lambda x, dt: x.to(device=device, dtype=dt) if isinstance(x, torch.Tensor) else \
        torch.as_tensor(np.asarray(x), dtype=dt, device=device)  # noqa: E731
    vv = mmatvv.contiguous()
    nirk = vv.shape[0]
    cv = mmatcv.contiguous() if mmatcv is not None else None
    ncg = cv.shape[2] if cv is not None else 0
    E = d(Ebnd, torch.float64).contiguous()
    ec = d(eig_core, torch.float64).contiguous() if ncg else None
    if isinstance(kcw, torch.Tensor):
        kc = kcw.to(device=device, dtype=torch.float64)
        if kc.stride() != (1, kc.shape[0], kc.shape[0] * kc.shape[1], kc.shape[0] * kc.shape[1] * kc.shape[2]):
            kc = kc.permute(3, 2, 1, 0).contiguous().permute(3, 2, 1, 0)
    else:
        k = np.asfortranarray(np.asarray(kcw, dtype=np.float64))
        kc = torch.as_tensor(k.ravel(order="F"), device=device).view(k.shape[::-1]).permute(3, 2, 1, 0)
    n_o, n_e, nk, nom = kc.shape
    if n_o != ncg_p + nomx + 1 or n_e != nbmaxpol - numin:
        raise ValueError("kcw must be [ncg_p + nomx + 1, nbmaxpol - numin, N_k, nom_nl]")
    ki = d(k_ind, torch.int32).contiguous()
    w = d(weight, torch.float64).contiguous()
    drude = Eg <= 0 and iop_drude == 1
    kw = d(kwfer, torch.float64).contiguous() if drude else None
    om = d(omega, torch.float64).contiguous() if drude else None
    du = d(dUl, torch.float64).contiguous() if (drude and dUl is not None) else None
    head = torch.empty(nom, dtype=torch.complex128, device=device)
    sz = C.c_size_t(0)
    _lib.call("pygw_calc_head_workspace", nirk, n_o, n_e, C.byref(sz))
    if workspace is None or workspace.numel() < sz.value:
        workspace = torch.empty(sz.value, dtype=torch.uint8, device=device)
    _lib.call("pygw_calc_head", _vp(vv), _vp(cv), _vp(E), E.shape[1], _vp(ec), ncg, _vp(kc), nk, _vp(ki), _vp(w), nirk,
              nomx, numin, nbmaxpol, ncg_p, nom, float(vol), 1 if Eg <= 0 else 0, _vp(kw), int(iop_drude),
              float(eta_head), _vp(om), _vp(du), om.numel() if om is not None else 0, _vp(head), _vp(workspace),
              workspace.numel(), _stream_ptr(device))
    return head


def calc_head(tables, pw, lapw_plan, kls, indgk, nvk, Aeig, Ebnd, eig_core, kcw, k_ind, weight, nomx, numin,
              nbmaxpol, ncg_p, vol, *, Eg=1.0, kwfer=None, iop_drude=0, eta_head=0.0, omega=None, dUl=None,
              gindex=None, exact=False):
    """calc_head (pygw.py:2808-3037) over the irreducible k-points irk =
    0..len(kls)-1: kls[irk] semi-cartesian, indgk[irk] / nvk[irk] /
    Aeig[irk] [nbmax, ngk] as ks holds them; the LAPW coefficients without
    local rotations (gindex, default pw's, is the G list of indgk).
    Returns (head, mmatcv, mmatvv, mst) as PyGW's calc_head."""
    nst, nend, mst, mend = 0, nomx + 1, numin, nbmaxpol
    nirk = len(kls)
    dev = tables.device
    mmatvv = torch.empty((nirk, mend - mst, nend - nst, 3), dtype=torch.complex128, device=dev)
    mmatcv = torch.empty((nirk, mend - mst, max(tables.ncg, 0), 3), dtype=torch.complex128, device=dev) \
        if tables.ncg else None
    ws = None
    g = gindex if gindex is not None else pw.gindex
    for irk in range(nirk):
        coef = lapw.gap2_set_lapwcoef(lapw_plan, kls[irk], indgk[irk], 1, False, nvk[irk], g)
        abc = lapw.band_coefficients(Aeig[irk], *coef)
        momentum_elements(tables, pw, abc, Aeig[irk], kls[irk], indgk[irk], nvk[irk], nst, nend, mst, mend,
                          mmatvv=mmatvv[irk], mmatcv=mmatcv[irk] if mmatcv is not None else None, exact=exact,
                          workspace=ws)
    head = head_from_elements(mmatvv, mmatcv, Ebnd, eig_core, kcw, k_ind, weight, nomx, numin, nbmaxpol, ncg_p, vol,
                              Eg=Eg, kwfer=kwfer, iop_drude=iop_drude, eta_head=eta_head, omega=omega, dUl=dUl)
    return head, mmatcv, mmatvv, mst


def wing_pm(mmatvv, mmatcv, Ebnd, eig_core, kii_ind, nomx, numin, nbmaxpol, ncg_p, vol):
    """pm [n_k, (ncg_p + nomx + 1)(nbmaxpol - numin)] of calc_eps's q = 0 wing
    for the k-points whose irreducible index is kii_ind[k] (pygw.py:3122-3141):
    the input of polarization.chi0_wings."""
    device = mmatvv.device
    d = lambda x, dt: x.to(device=device, dtype=dt) if isinstance(x, torch.Tensor) else \
        torch.as_tensor(np.asarray(x), dtype=dt, device=device)  # noqa: E731
    vv = mmatvv.contiguous()
    cv = mmatcv.contiguous() if mmatcv is not None else None
    ncg = cv.shape[2] if cv is not None else 0
    E = d(Ebnd, torch.float64).contiguous()
    ec = d(eig_core, torch.float64).contiguous() if ncg else None
    ki = d(kii_ind, torch.int32).contiguous()
    nk = ki.numel()
    pm = torch.empty((nk, (ncg_p + nomx + 1) * (nbmaxpol - numin)), dtype=torch.complex128, device=device)
    _lib.call("pygw_wing_pm", _vp(vv), _vp(cv), _vp(E), E.shape[1], _vp(ec), ncg, vv.shape[0], _vp(ki), nk, nomx, numin,
              nbmaxpol, ncg_p, float(vol), _vp(pm), _stream_ptr(device))
    return pm
