"""ORACLE (test infrastructure only): PyGW's own Fortran, compiled in place
from the reference sources by `make -C oracle ref` into oracle/_ref/.

Fortran calling convention: every argument by reference, arrays
column-major, f2py's hidden dimensions passed explicitly at the end
(for_tetrahedra.f90:390, fnc.f90:139-176, for_pade.f90:3-51).

The Fortran `stop` statements terminate the calling process, so callers
only pass inputs for which the reference is defined (the C restatement
reports those cases through its status/flags first).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DIR = os.path.join(_HERE, "_ref")


def available(name="libref_tetra.so") -> bool:
    return os.path.exists(os.path.join(_DIR, name))


_cache = {}


def _load(name):
    if name not in _cache:
        _cache[name] = C.CDLL(os.path.join(_DIR, name), mode=os.RTLD_LAZY)
    return _cache[name]


def _i(x):
    return C.byref(C.c_int(int(x)))


def _d(x):
    return C.byref(C.c_double(float(x)))


def bz_calcqdepw_par2(enk, ef, kqid_col, ncg, nomx, numin, omega, tetc, omp=False):
    lib = _load("libref_tetra_omp.so" if omp else "libref_tetra.so")
    enk = np.asfortranarray(enk, dtype=np.float64)
    nbnd, nkp = enk.shape
    tetc = np.asfortranarray(tetc, dtype=np.int32)
    ntet = tetc.shape[1]
    kq = np.ascontiguousarray(kqid_col, dtype=np.int32)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    no, ne = ncg + nomx + 1, nbnd - ncg - numin
    kcw = np.zeros((no, ne, nkp, om.size), dtype=np.float64, order="F")
    lib.bz_calcqdepw_par2_(C.c_void_p(kcw.ctypes.data), C.c_void_p(enk.ctypes.data), _d(ef),
                           C.c_void_p(kq.ctypes.data), _i(ncg), _i(nomx), _i(numin), C.c_void_p(om.ctypes.data),
                           C.c_void_p(tetc.ctypes.data), _i(nkp), _i(om.size), _i(nbnd), _i(ntet))
    return kcw


def bz_calcqdepw_par3(enk, ef, kqid_col, ncg, nomx, numin, dUl, omega, tetc):
    lib = _load("libref_tetra.so")
    enk = np.asfortranarray(enk, dtype=np.float64)
    nbnd, nkp = enk.shape
    tetc = np.asfortranarray(tetc, dtype=np.int32)
    ntet = tetc.shape[1]
    kq = np.ascontiguousarray(kqid_col, dtype=np.int32)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    dUl = np.asfortranarray(dUl, dtype=np.float64)
    nl = dUl.shape[1]
    no, ne = ncg + nomx + 1, nbnd - ncg - numin
    kcw = np.zeros((no, ne, nkp, nl), dtype=np.float64, order="F")
    lib.bz_calcqdepw_par3_(C.c_void_p(kcw.ctypes.data), C.c_void_p(enk.ctypes.data), _d(ef),
                           C.c_void_p(kq.ctypes.data), _i(ncg), _i(nomx), _i(numin), C.c_void_p(dUl.ctypes.data),
                           C.c_void_p(om.ctypes.data), C.c_void_p(tetc.ctypes.data), _i(nkp), _i(om.size),
                           _i(nbnd), _i(ntet), _i(nl))
    return kcw


def convw1t_par(e, f, ef, omega):
    """weight(4, nom) of one tetrahedron pair (for_tetrahedra3.f90:3-136)."""
    lib = _load("libref_tetra.so")
    om = np.ascontiguousarray(omega, dtype=np.float64)
    w = np.zeros((4, om.size), dtype=np.float64, order="F")
    e = np.ascontiguousarray(e, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    dbg = C.c_int(0)
    lib.convw1t_par_(C.c_void_p(w.ctypes.data), C.c_void_p(e.ctypes.data), C.c_void_p(f.ctypes.data), _d(ef),
                     C.c_void_p(om.ctypes.data), _i(om.size), C.byref(dbg))
    return w


def all_fr_convolutions(enk, mwm, omega, womeg):
    """sc_p(nb1, nom) from mwm(nom, nb1, nb2) (fnc.f90:139-159)."""
    lib = _load("libref_fnc.so")
    mwm = np.asfortranarray(mwm, dtype=np.complex128)
    nom, nb1, nb2 = mwm.shape
    enk = np.ascontiguousarray(enk, dtype=np.float64)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    wo = np.ascontiguousarray(womeg, dtype=np.float64)
    out = np.zeros((nb1, nom), dtype=np.complex128, order="F")
    lib.all_fr_convolutions_(C.c_void_p(out.ctypes.data), C.c_void_p(enk.ctypes.data), C.c_void_p(mwm.ctypes.data),
                             C.c_void_p(om.ctypes.data), C.c_void_p(wo.ctypes.data), _i(nb1), _i(nb2), _i(nom))
    return out


def few_fr_convolutions(enk, mwm, mwm_iom, omega, omega_iom, womeg):
    """sum_ie2 of fr_convolution2 on the precise mesh (fnc.f90:161-176).
    mwm: (nb2, nom_precise)."""
    lib = _load("libref_fnc.so")
    mwm = np.asfortranarray(mwm, dtype=np.complex128)
    nb2, nom = mwm.shape
    enk = np.ascontiguousarray(enk, dtype=np.float64)
    mi = np.ascontiguousarray(mwm_iom, dtype=np.complex128)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    wo = np.ascontiguousarray(womeg, dtype=np.float64)
    res = np.zeros(1, dtype=np.complex128)
    lib.few_fr_convolutions_(C.c_void_p(res.ctypes.data), C.c_void_p(enk.ctypes.data), C.c_void_p(mwm.ctypes.data),
                             C.c_void_p(mi.ctypes.data), C.c_void_p(om.ctypes.data), _d(omega_iom),
                             C.c_void_p(wo.ctypes.data), _i(nom), _i(nb2))
    return res[0]


def padecof(y, z):
    lib = _load("libref_pade.so")
    y = np.ascontiguousarray(y, dtype=np.complex128)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    pt = np.zeros(y.size, dtype=np.complex128)
    lib.padecof_(C.c_void_p(pt.ctypes.data), C.c_void_p(y.ctypes.data), C.c_void_p(z.ctypes.data), _i(y.size))
    return pt


def padeg(x, z, apar):
    lib = _load("libref_pade.so")
    x = np.ascontiguousarray(x, dtype=np.complex128)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    a = np.ascontiguousarray(apar, dtype=np.complex128)
    g = np.zeros(x.size, dtype=np.complex128)
    lib.padeg_(C.c_void_p(g.ctypes.data), C.c_void_p(x.ctypes.data), C.c_void_p(z.ctypes.data),
               C.c_void_p(a.ctypes.data), _i(z.size), _i(x.size))
    return g


def _tet_args(ebnd, atet, wtet):
    e = np.asfortranarray(ebnd, dtype=np.float64)           # Ebnd(nk, nbnd)
    a = np.ascontiguousarray(atet, dtype=np.int32)          # [ntet][4] = atet(4, ntet)
    w = np.ascontiguousarray(wtet, dtype=np.float64)
    return e, a, w


def idos(energy, ebnd, atet, wtet):
    """for_tetrahedra.f90:84-116 (ebnd [nk, nbnd], atet [ntet, 4] 0-based)."""
    f = _load("libref_tetra.so").idos_
    f.restype = C.c_double
    e, a, w = _tet_args(ebnd, atet, wtet)
    return f(_d(energy), e.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
             _i(e.shape[0]), _i(e.shape[1]), _i(a.shape[0]))


def dostet(energy, ebnd, atet, wtet):
    """for_tetrahedra.f90:118-141."""
    f = _load("libref_tetra.so").dostet_
    f.restype = C.c_double
    e, a, w = _tet_args(ebnd, atet, wtet)
    return f(_d(energy), e.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
             _i(e.shape[0]), _i(e.shape[1]), _i(a.shape[0]))
