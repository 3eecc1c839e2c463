"""ORACLE (test infrastructure only): ctypes wrappers of tetra_oracle.c.

Follows for_tetrahedra.f90:390-542 and for_tetrahedra3.f90:3-743; see the
header of tetra_oracle.c for the list of deviations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of tetra_oracle.c (the sanitizer build of
# `make -C oracle sanitize`, run on the host only)
_LIB = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

ORC_F_F7A = 1
ORC_F_ZEROVOL = 2


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        _lib.orc_convw1t_simple.argtypes = [dp, dp, dp, C.c_double, dp, C.c_int,
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib.orc_convw1t_simple.restype = C.c_int
        _lib.orc_generictetra_simple.argtypes = [dp, dp, dp, dp, dp, C.c_int, C.POINTER(C.c_int)]
        _lib.orc_generictetra_simple.restype = C.c_int
        _lib.orc_bz_calcqdepw.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                          C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
        _lib.orc_bz_calcqdepw.restype = C.c_int
        _lib.orc_sorteq.argtypes = [dp, ip, C.POINTER(C.c_int)]
        _lib.orc_sorteq.restype = C.c_int
    return _lib


def convw1t(e, f, ef, omega):
    """Weights of one (occupied e, empty f) tetrahedron pair: [4, nom],
    plus (status, nwarn, flags, nnod).  Follows convw1t_par."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    w = np.zeros((om.size, 4))
    nw, fl, nn = C.c_int(0), C.c_int(0), C.c_int(-1)
    st = lib().orc_convw1t_simple(w, e, f, float(ef), om, om.size, C.byref(nw), C.byref(fl), C.byref(nn))
    return w.T.copy(), st, nw.value, fl.value, nn.value


def generictetra(corners, e, f, omega):
    corners = np.ascontiguousarray(corners, dtype=np.float64).reshape(4, 3)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    out = np.zeros((om.size, 4))
    nw = C.c_int(0)
    st = lib().orc_generictetra_simple(out, corners.ravel().copy(), np.ascontiguousarray(e, dtype=np.float64),
                                       np.ascontiguousarray(f, dtype=np.float64), om, om.size, C.byref(nw))
    return out.T.copy(), st


def sorteq(v):
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    idx = np.zeros(4, dtype=np.int32)
    cls = C.c_int(0)
    st = lib().orc_sorteq(v, idx, C.byref(cls))
    return v, idx, cls.value, st


def bz_calcqdepw(enk, ef, kqid_col, ncg, nomx, numin, omega, tetc, dUl=None, nthreads=1):
    """kcw Fortran-ordered [n_o, n_e, N_k, nom or nl] (bz_calcqdepw_par2/par3).
    Returns (kcw, status, flags); status >= 0 is the warning count."""
    enk = np.asfortranarray(enk, dtype=np.float64)
    nbnd, nkp = enk.shape
    tetc = np.asfortranarray(tetc, dtype=np.int32)
    ntet = tetc.shape[1]
    kq = np.ascontiguousarray(kqid_col, dtype=np.int32)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    no, ne = ncg + nomx + 1, nbnd - ncg - numin
    nl = 0
    dptr = None
    if dUl is not None:
        dUl = np.asfortranarray(dUl, dtype=np.float64)
        nl = dUl.shape[1]
        dptr = dUl.ctypes.data
    nout = nl if dUl is not None else om.size
    kcw = np.zeros((no, ne, nkp, nout), dtype=np.float64, order="F")
    fl = C.c_int(0)
    st = lib().orc_bz_calcqdepw(kcw.ctypes.data, enk.ctypes.data, float(ef), kq.ctypes.data, ncg, nomx, numin,
                                om.ctypes.data, tetc.ctypes.data, nkp, om.size, nbnd, ntet, dptr, nl,
                                int(nthreads), C.byref(fl))
    return kcw, st, fl.value


PERTURB_SEEDS = tuple(range(1, 9))


def with_spread(enk, ef, kqid_col, ncg, nomx, numin, omega, tetc, compat=(1, 1), seeds=PERTURB_SEEDS,
                nthreads=8):
    """Oracle weights plus their sensitivity to one-ulp changes of atan/log.

    The closed forms of the Matsubara weights (for_tetrahedra3.f90:1-137) are
    differences of atan/log values divided by high powers of corner-Delta
    differences; near degenerate classes a one-ulp change of one atan moves
    the weight by many orders of magnitude more than one ulp of the weight.
    Each seed re-runs the oracle with every atan/log result scaled by
    (1 +- 2^-52), the sign a hash of (argument, seed); the spread is the
    largest deviation over the seeds.  Several seeds are needed: one seed can
    move two nearly equal atan values the same way and leave their
    difference unchanged.  `compat` = (orc_compat_f7b, orc_compat_nnod78).
    Returns (kcw, spread, flags)."""
    L = lib()
    a = C.c_int.in_dll(L, "orc_compat_f7b")
    b = C.c_int.in_dll(L, "orc_compat_nnod78")
    p = C.c_ulonglong.in_dll(L, "orc_perturb_seed")
    old = (a.value, b.value, p.value)
    a.value, b.value = compat
    try:
        p.value = 0
        k0, st, fl = bz_calcqdepw(enk, ef, kqid_col, ncg, nomx, numin, omega, tetc, nthreads=nthreads)
        if st < 0:
            raise RuntimeError(f"oracle bz_calcqdepw failed: {st}")
        spread = np.zeros_like(k0)
        for s in seeds:
            p.value = int(s)
            k1, _, _ = bz_calcqdepw(enk, ef, kqid_col, ncg, nomx, numin, omega, tetc, nthreads=nthreads)
            np.maximum(spread, np.abs(k1 - k0), out=spread)
    finally:
        a.value, b.value, p.value = old
    return k0, spread, fl


def within_spread(got, kcw, spread, factor=8.0, floor=1e-13):
    """max over elements of |got - kcw| / (factor*spread + floor*max_w|kcw|);
    <= 1 means the device weights agree with the oracle to within what one-ulp
    differences of atan/log explain."""
    scale = np.abs(kcw).max(axis=tuple(range(kcw.ndim - 1)), keepdims=True)
    tol = factor * spread + floor * scale
    dev = np.abs(got - kcw)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(tol > 0, dev / np.where(tol > 0, tol, 1.0), np.where(dev == 0, 0.0, np.inf))
    return float(r.max()) if r.size else 0.0
