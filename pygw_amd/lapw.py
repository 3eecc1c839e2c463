"""The LAPW augmentation coefficients and the band coefficients built from
them, on the gfx950 backend.

* ``LapwPlan`` holds what PyGW passes to ``lapwc.gap2_set_lapwcoef`` that
  does not change with k: radf.abcelo, strc.rmt, strc.vpos, strc.mult,
  radf.umt, strc.rotloc, latgen.trotij, latgen.Vol, kqm.k2cartes,
  in1.nLO_at, in1.nlo, in1.lapw (pygw.py:2571).
* ``gap2_set_lapwcoef`` replaces ``lapwc.gap2_set_lapwcoef`` (lapwcoeff.f90:
  3-222) for one k-point: alm/blm [ngk, nt^2, ndf], clm [ngk, nLOmax, nt^2,
  ndf] as device tensors (views of Fortran-ordered storage, the layout the
  reference returns).
* ``band_coefficients`` replaces the three ``la.matmul(Aeig, reshape(...))``
  that turn them into alfa, beta, gama (pygw.py:2577-2583, 2601-2607,
  2863-2869); the results feed ``coulomb.calc_minm_mt`` / ``calc_minc``
  without a copy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .polarization import _require_device, _stream_ptr


def _i32(a):
    return np.ascontiguousarray(np.asarray(a).ravel(order="F"), dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


class LapwPlan:
    """The k-independent inputs of gap2_set_lapwcoef in one device table
    (pygw_lapw_plan).  Shapes as f2py sees them: abcelo [nat, lomax+1,
    nloat, 3], rmt [nat], pos [3, ndf], mult [nat], umt [nat, nt, 5], rotloc
    [nat, 3, 3], trotij [ndf, 3, 3], k2cartes [3, 3], nLO_at/nlo [lomax+1,
    nat], lapw [nt, nat]."""

    def __init__(self, abcelo, rmt, pos, mult, umt, rotloc, trotij, vol, k2cartes, nLO_at, nlo, lapw, *,
                 device="cuda"):
        device = _require_device(device)
        abcelo, umt, nLO_at = np.asarray(abcelo), np.asarray(umt), np.asarray(nLO_at)
        nat, nt = len(np.asarray(mult)), umt.shape[1]
        lomax, nloat = nLO_at.shape[0] - 1, abcelo.shape[2]
        if abcelo.shape != (nat, lomax + 1, nloat, 3) or umt.shape[0] != nat or umt.shape[2] < 4 or \
                np.asarray(lapw).shape != (nt, nat) or np.asarray(nlo).shape != nLO_at.shape:
            raise ValueError("array shapes do not match (nat, nt, lomax, nloat)")
        keep = [_i32(mult), _i32(lapw), _i32(nlo), _i32(nLO_at), _f64(umt), _f64(abcelo), _f64(rmt), _f64(pos),
                _f64(rotloc), _f64(trotij), _f64(k2cartes)]
        p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        args = (nat, p(keep[0]), nt, lomax, nloat, p(keep[1]), p(keep[2]), p(keep[3]), p(keep[4]), p(keep[5]),
                p(keep[6]), p(keep[7]), p(keep[8]), p(keep[9]), float(vol), p(keep[10]))
        n = C.c_int(0)
        _lib.call("pygw_lapw_plan", *args, None, C.byref(n))
        self.host = np.zeros(n.value, dtype=np.float64)
        _lib.call("pygw_lapw_plan", *args, p(self.host), C.byref(n))
        self.dev = torch.as_tensor(self.host, device=device)
        self.device, self.nt, self.ndf = device, nt, int(self.host[4])
        self.nlomax = int(self.host[5])

    def ptrs(self):
        return self.host.ctypes.data_as(C.c_void_p), C.c_void_p(self.dev.data_ptr())


def gap2_set_lapwcoef(plan, kl, indgk, iop, ifrot, nvk, gindex, *, out=None):
    """lapwc.gap2_set_lapwcoef(kl, indgk, iop, ifrot, nvk, gindex, ...) for
    one k-point (kl semi-cartesian, indgk 0-based rows of gindex [ng, 3], the
    first nvk of them plane waves, the rest local-orbital rows).  Returns
    (alm, blm, clm) with clm None when there are no local orbitals; out =
    (alm, blm, clm) storage from an earlier call of the same size is reused."""
    device = plan.device
    ig = indgk if isinstance(indgk, torch.Tensor) else torch.as_tensor(np.asarray(indgk, dtype=np.int32))
    ig = ig.to(device=device, dtype=torch.int32).contiguous()
    if isinstance(gindex, torch.Tensor) and gindex.dtype == torch.int32 and gindex.dim() == 1:
        gd, ng = gindex.to(device), gindex.numel() // 3      # already Fortran-flattened [ng, 3]
    else:
        g = np.asarray(gindex)
        gd, ng = torch.as_tensor(_i32(g), device=device), g.shape[0]
    ngk, ntnt, ndf, nlo = ig.numel(), plan.nt * plan.nt, plan.ndf, plan.nlomax
    if out is None:
        a = torch.empty((ndf, ntnt, ngk), dtype=torch.complex128, device=device)
        b = torch.empty_like(a)
        c = torch.empty((ndf, ntnt, nlo, ngk), dtype=torch.complex128, device=device) if nlo > 0 else None
    else:
        a, b, c = (x.permute(*reversed(range(x.dim()))) if x is not None else None for x in out)
        if a.shape != (ndf, ntnt, ngk) or (nlo > 0 and c.shape != (ndf, ntnt, nlo, ngk)):
            raise ValueError("out does not match this k-point's sizes")
    k = np.ascontiguousarray(np.asarray(kl, dtype=np.float64))
    ph, pd = plan.ptrs()
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _lib.call("pygw_lapw_coef", ph, pd, k.ctypes.data_as(C.c_void_p), int(iop), 1 if ifrot else 0, int(nvk), vp(ig),
              ngk, vp(gd), int(ng), vp(a), vp(b), vp(c), _stream_ptr(device))
    return a.permute(2, 1, 0), b.permute(2, 1, 0), (c.permute(3, 2, 1, 0) if c is not None else None)


def band_coefficients(Aeig, alm, blm, clm, *, conj=False, out=None):
    """alfa = Aeig alm, beta = Aeig blm, gama = Aeig clm (pygw.py:2577-2583)
    with Aeig [nbmax, ngk] (conj: its conjugate, as PyGW forms Aeigq from
    ks.all_As at k - q, pygw.py:2587).  alm/blm/clm as gap2_set_lapwcoef
    returns them; alfa/beta [nbmax, nt^2, ndf], gama [nbmax, nLOmax, nt^2,
    ndf] (None without local orbitals), views of Fortran-ordered storage;
    out = (alfa, beta, gama) from an earlier call of the same sizes is
    reused."""
    device = alm.device
    A = Aeig if isinstance(Aeig, torch.Tensor) else torch.as_tensor(np.asarray(Aeig, dtype=np.complex128))
    A = A.to(device=device, dtype=torch.complex128)
    if A.stride(1) != 1:
        A = A.contiguous()
    nb, ngk = A.shape
    sa = alm.permute(2, 1, 0)
    if alm.shape[0] != ngk or not sa.is_contiguous():
        raise ValueError("alm must come from gap2_set_lapwcoef with Aeig's ngk rows")
    ntnt, ndf = alm.shape[1], alm.shape[2]
    nlo = clm.shape[1] if clm is not None else 0
    if out is None:
        alfa = torch.empty((ndf, ntnt, nb), dtype=torch.complex128, device=device)
        beta = torch.empty_like(alfa)
        gama = torch.empty((ndf, ntnt, nlo, nb), dtype=torch.complex128, device=device) if nlo > 0 else None
    else:
        alfa, beta, gama = (x.permute(*reversed(range(x.dim()))) if x is not None else None for x in out)
        if alfa.shape != (ndf, ntnt, nb) or not alfa.is_contiguous() or not beta.is_contiguous() or \
                (nlo > 0 and (gama.shape != (ndf, ntnt, nlo, nb) or not gama.is_contiguous())):
            raise ValueError("out does not match these band coefficients")
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    _lib.call("pygw_lapw_band", vp(A), A.stride(0), 1 if conj else 0, nb, ngk, vp(alm), vp(blm), vp(clm),
              ntnt * ndf, nlo, vp(alfa), vp(beta), vp(gama), _stream_ptr(device))
    return alfa.permute(2, 1, 0), beta.permute(2, 1, 0), (gama.permute(3, 2, 1, 0) if gama is not None else None)
