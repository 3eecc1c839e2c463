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
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
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
