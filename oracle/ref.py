"""ORACLE (test infrastructure only): PyGW's own Fortran, compiled in place
from the reference sources by `make -C oracle ref` into oracle/_ref/.

Fortran calling convention: every argument by reference, arrays
column-major, f2py's hidden dimensions passed explicitly at the end
(for_tetrahedra.f90:390, fnc.f90:139-176, for_pade.f90:3-51,
for_Coulomb.f90:306,327, sphbes.f90:132).

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


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ylm(v, lmax):
    """sphbes.f90:132-271: complex Y_lm(v), index l*l + l + m."""
    f = _load("libref_coul.so").ylm_
    y = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    vv = np.asarray(v, dtype=np.float64).copy()
    f(_p(y), _p(vv), _i(lmax))
    return y


def cmp_pw_olap2(indgq, ipwint, gindex, i_g0, ngq, ngq_barc):
    """for_Coulomb.f90:306-325 (0-based indgq, i_g0 [2n1+1, 2n2+1, 2n3+1])."""
    f = _load("libref_coul.so").cmp_pw_olap2_
    out = np.zeros((ngq, ngq_barc), dtype=np.complex128, order="F")
    iq = np.asarray(indgq, dtype=np.int32)
    ip = np.asarray(ipwint, dtype=np.complex128)
    gx = np.asfortranarray(np.asarray(gindex, dtype=np.int32))
    g0 = np.asfortranarray(np.asarray(i_g0, dtype=np.int32))
    n1, n2, n3 = [(s - 1) // 2 for s in g0.shape]
    f(_p(out), _p(iq), _p(ip), _p(gx), _p(g0), _i(ngq), _i(ngq_barc), _i(gx.shape[0]), _i(len(iq)), _i(n1), _i(n2),
      _i(n3))
    return np.ascontiguousarray(out)


def mixed_coulomb(vq, iat, q_vq, jlam, gindex, indgq, gqlen, gqlen0, G_unique, ngq_barc, sloctmatsize, big_l,
                  k2cartes, rotloc, trotij, mult, pos, vol):
    """for_Coulomb.f90:327-408 with PyGW's argument meaning (pygw.py:2396):
    iat 1-based, indgq / G_unique 0-based, jlam [nmix, ngs], rotloc [nat, 3,
    3], trotij [ndf, 3, 3], pos [3, ndf]; returns vtemp [sloctmatsize,
    ngq_barc]."""
    f = _load("libref_coul.so").mixed_coulomb_
    out = np.zeros((sloctmatsize, ngq_barc), dtype=np.complex128, order="F")
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))
    jl, gx, iq, gl, g0l, gu = (F(jlam, np.float64), F(gindex, np.int32), F(indgq, np.int32), F(gqlen, np.float64),
                               F(gqlen0, np.float64), F(G_unique, np.int32))
    bl, k2, rl, tr, mu, ps = (F(big_l, np.int32), F(k2cartes, np.float64), F(rotloc, np.float64),
                              F(trotij, np.float64), F(mult, np.int32), F(pos, np.float64))
    v = F(vq, np.float64)
    f(_p(out), _p(v), _i(iat), _i(1 if q_vq else 0), _p(jl), _p(gx), _p(iq), _p(gl), _p(g0l), _p(gu), _i(ngq_barc),
      _i(sloctmatsize), _p(bl), _p(k2), _p(rl), _p(tr), _p(mu), _p(ps), _d(vol), _i(jl.shape[0]), _i(rl.shape[0]),
      _i(gx.shape[0]), _i(jl.shape[1]), _i(len(gu)), _i(ps.shape[1]), _i(len(iq)))
    return np.ascontiguousarray(out)


def k_m_q_id(div):
    """kqid[ik, iq] of k_m_q_id (for_kpts.f90:21-45), 0-based, int32 [N_k, N_k]."""
    lib = _load("libref_kpts.so")
    d = np.ascontiguousarray(div, dtype=np.int32)
    nkp = int(np.prod(d))
    out = np.zeros((nkp, nkp), dtype=np.int32, order="F")
    lib.k_m_q_id_(C.c_void_p(out.ctypes.data), C.c_void_p(d.ctypes.data), _i(nkp))
    return out


def k2indx(k, div):
    """k2indx (for_kpts.f90:2-7) of one integer k-point."""
    lib = _load("libref_kpts.so")
    kk = np.ascontiguousarray(k, dtype=np.int32)
    d = np.ascontiguousarray(div, dtype=np.int32)
    out = C.c_int(0)
    lib.k2indx_(C.byref(out), C.c_void_p(kk.ctypes.data), C.c_void_p(d.ctypes.data))
    return out.value


def ewald_summation(pq, lam_max, pos, br2, rbas, alat, vol, ortho, r_cf, g_cf, eta):
    """for_Coulomb.f90:1-164: sgm [(lam_max+1)^2, ndf(ndf+1)/2] (column
    i + j(j+1)/2 for the 0-based atom pair i <= j)."""
    f = _load("libref_coul.so").ewald_summation_
    ps = np.asfortranarray(np.asarray(pos, dtype=np.float64))
    ndf = ps.shape[1]
    out = np.zeros(((lam_max + 1) ** 2, ndf * (ndf + 1) // 2), dtype=np.complex128, order="F")
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))  # noqa: E731
    q, b2, rb, al = F(pq), F(br2), F(rbas), F(alat)
    orth = C.c_int(1 if ortho else 0)
    f(_p(out), _p(q), _i(lam_max), _p(ps), _p(b2), _p(rb), _p(al), _d(vol), C.byref(orth), _d(r_cf), _d(g_cf), _d(eta),
      _i(ndf))
    return np.ascontiguousarray(out)


def mt_coulomb(vq, iat, loctmatsize, big_l, nmix, im_start, tilg, rrint, djmm, rtlij, sgm, mult, lam_max):
    """for_Coulomb.f90:166-283 with PyGW's arguments (pygw.py:2394): iat
    1-based, big_l [max_nmix, nat] Fortran, im_start [nat+1] 0-based offsets,
    rrint [nat, nijrm], djmm [dimdj, ndf], rtlij [max_nmix, max_nmix, nat,
    nat] Fortran, sgm from ewald_summation.  Returns Vmat [loct, loct]."""
    f = _load("libref_coul.so").mt_coulomb_
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))  # noqa: E731
    out = np.zeros((loctmatsize, loctmatsize), dtype=np.complex128, order="F")
    bl, nm, ims, tg = F(big_l, np.int32), F(nmix, np.int32), F(im_start, np.int32), F(tilg, np.float64)
    rr, dj, rt, sg = F(rrint, np.float64), F(djmm, np.complex128), F(rtlij, np.float64), F(sgm, np.complex128)
    mu, v = F(mult, np.int32), F(vq, np.float64)
    nat, ndf = bl.shape[1], dj.shape[1]
    f(_p(out), _p(v), _i(iat), _i(loctmatsize), _p(bl), _p(nm), _p(ims), _p(tg), _p(rr), _p(dj), _p(rt), _p(sg),
      _p(mu), _i(nat), _i(ndf), _i(bl.shape[0]), _i(dj.shape[0]), _i(rr.shape[1]), _i(tg.size), _i(lam_max))
    return np.ascontiguousarray(out)


def rint13g(rel, a, b, x, y, dh, nrpt, rnot):
    """radials.f90:140-187."""
    f = _load("libref_rad.so").rint13g_
    f.restype = C.c_double
    F = lambda z: np.ascontiguousarray(np.asarray(z, dtype=np.float64))  # noqa: E731
    A, B, X, Y = F(a), F(b), F(x), F(y)
    r = C.c_int(1 if rel else 0)
    return f(C.byref(r), _p(A), _p(B), _p(X), _p(Y), _d(dh), _i(nrpt), _d(rnot))


def spher_bessel(lmax, x):
    """for_Coulomb.f90 spher_bessel (sphbes.f90): j_l(x), l <= lmax -> [lmax+1, nx]."""
    f = _load("libref_coul.so").spher_bessel_
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
    out = np.zeros((lmax + 1, xx.size), dtype=np.float64, order="F")
    f(_p(out), _i(lmax), _p(xx), _i(xx.size))
    return np.ascontiguousarray(out)


def pw_get_short_and_phase(glen, gindex, pos, mult):
    """for_Coulomb.f90:493-582: the unique |G| (G = 0 dropped) and
    phase[idf, jdf, igl] = sum over G of that length of exp(-2 pi i G.(pos_i - pos_j))."""
    lib = _load("libref_coul.so")
    gl = np.ascontiguousarray(np.asarray(glen, dtype=np.float64))
    nk = gl.size
    ngs, gind4 = C.c_int(0), np.zeros(nk, dtype=np.int32)
    lib.pw_get_short_length_(C.byref(ngs), _p(gind4), _p(gl), _i(nk))
    ngs = ngs.value
    ps = np.asfortranarray(np.asarray(pos, dtype=np.float64))
    mu = np.ascontiguousarray(np.asarray(mult, dtype=np.int32))
    gx = np.asfortranarray(np.asarray(gindex, dtype=np.int32))
    ndf = ps.shape[1]
    glen0 = np.zeros(ngs, dtype=np.float64)
    phase = np.zeros((ndf, ndf, ngs), dtype=np.complex128, order="F")
    lib.pw_get_short_and_phase_(_p(glen0), _p(phase), _i(ngs), _p(gind4), _p(gx), _p(gl), _p(ps), _p(mu), _i(nk),
                                _i(mu.size), _i(ndf))
    return glen0, np.ascontiguousarray(phase), gind4


def cmp_all_gaunt(maxj):
    """gaunt.f90:1-34."""
    f = _load("libref_gaunt.so").cmp_all_gaunt_
    n = (maxj + 1) * (maxj + 2) * (maxj + 3) * (16 * maxj * maxj + 29 * maxj + 10) // 60
    out = np.zeros(n, dtype=np.float64)
    f(_p(out), _i(maxj))
    return out


def getcgcoef(l1, l2, l3, m1, m2, cgcoef):
    """gaunt.f90:36-101."""
    f = _load("libref_gaunt.so").getcgcoef_
    f.restype = C.c_double
    cg = np.ascontiguousarray(cgcoef, dtype=np.float64)
    return f(_i(l1), _i(l2), _i(l3), _i(m1), _i(m2), _p(cg))


def _mt_common(s3r, corind, pos, mult, nmix, big_l, nLO_at, ncore, cgcoef):
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))  # noqa: E731
    return (F(s3r, np.float64), F(corind, np.int32), F(pos, np.float64), F(mult, np.int32), F(nmix, np.int32),
            F(big_l, np.int32), F(nLO_at, np.int32), F(ncore, np.int32), F(cgcoef, np.float64))


def calc_minc(kl, nstart, nend, cstart, cend, alfa, beta, gama, s3r, corind, pos, mult, nmix, big_l, nLO_at, ncore,
              cgcoef, lmax, loctmatsize):
    """for_vxcnn.f90:1136-1273 (no zgemm); arrays as PyGW passes them."""
    f = _load("libref_vxcnn.so").calc_minc_
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.complex128))  # noqa: E731
    al, be, ga = F(alfa), F(beta), F(gama)
    s3, co, ps, mu, nm, bl, nlo, nco, cg = _mt_common(s3r, corind, pos, mult, nmix, big_l, nLO_at, ncore, cgcoef)
    out = np.zeros((nend - nstart, cend - cstart, loctmatsize), dtype=np.complex128, order="F")
    k = np.asfortranarray(np.asarray(kl, dtype=np.float64))
    f(_p(out), _p(k), _i(nstart), _i(nend), _i(cstart), _i(cend), _p(al), _p(be), _p(ga), _p(s3), _p(co), _p(ps),
      _p(mu), _p(nm), _p(bl), _p(nlo), _p(nco), _p(cg), _i(lmax), _i(loctmatsize), _i(al.shape[0]), _i(mu.size),
      _i(al.shape[1]), _i(ga.shape[1]), _i(al.shape[2]), _i(s3.shape[0]), _i(s3.shape[1]), _i(bl.shape[0]),
      _i(nlo.shape[0] - 1), _i(ga.shape[0]), _i(ga.shape[2]), _i(ga.shape[3]), _i(co.shape[0]))
    return np.ascontiguousarray(out)


def calc_minc2(kmql, cstart, cend, mstart, mend, alfp, betp, gamp, s3r, corind, pos, mult, nmix, big_l, nLO_at,
               ncore, cgcoef, lmax, loctmatsize):
    """for_vxcnn.f90:1275-1397 (no zgemm)."""
    f = _load("libref_vxcnn.so").calc_minc2_
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.complex128))  # noqa: E731
    al, be, ga = F(alfp), F(betp), F(gamp)
    s3, co, ps, mu, nm, bl, nlo, nco, cg = _mt_common(s3r, corind, pos, mult, nmix, big_l, nLO_at, ncore, cgcoef)
    out = np.zeros((cend - cstart, mend - mstart, loctmatsize), dtype=np.complex128, order="F")
    k = np.asfortranarray(np.asarray(kmql, dtype=np.float64))
    f(_p(out), _p(k), _i(cstart), _i(cend), _i(mstart), _i(mend), _p(al), _p(be), _p(ga), _p(s3), _p(co), _p(ps),
      _p(mu), _p(nm), _p(bl), _p(nlo), _p(nco), _p(cg), _i(lmax), _i(loctmatsize), _i(al.shape[0]), _i(mu.size),
      _i(al.shape[1]), _i(ga.shape[1]), _i(al.shape[2]), _i(s3.shape[0]), _i(s3.shape[1]), _i(bl.shape[0]),
      _i(nlo.shape[0] - 1), _i(ga.shape[0]), _i(ga.shape[2]), _i(ga.shape[3]), _i(co.shape[0]))
    return np.ascontiguousarray(out)


def gap2_set_lapwcoef(kl, indgk, iop, ifrot, nvk, gindex, abcelo, rmt, pos, mult, umt, rotloc, trotij, vol,
                      k2cartes, nLO_at, nlo, lapw, nLOmax):
    """lapwcoeff.f90:3-222 (with sphbes/ylm of sphbes.f90); arrays as PyGW
    passes them, logical ifrot as a 4-byte 0/1."""
    f = _load("libref_lapw.so").gap2_set_lapwcoef_
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))  # noqa: E731
    ig, gi = F(indgk, np.int32), F(gindex, np.int32)
    ab, rm, ps, mu, um = F(abcelo, np.float64), F(rmt, np.float64), F(pos, np.float64), F(mult, np.int32), \
        F(umt, np.float64)
    rl, tr, kc = F(rotloc, np.float64), F(trotij, np.float64), F(k2cartes, np.float64)
    nl, no, lw = F(nLO_at, np.int32), F(nlo, np.int32), F(lapw, np.int32)
    k = F(kl, np.float64)
    ngk, nat, nt, ndf = ig.size, mu.size, um.shape[1], ps.shape[1]
    alm = np.zeros((ngk, nt * nt, ndf), dtype=np.complex128, order="F")
    blm = np.zeros_like(alm, order="F")
    clm = np.zeros((ngk, max(nLOmax, 1), nt * nt, ndf), dtype=np.complex128, order="F")
    f(_p(alm), _p(blm), _p(clm), _p(k), _p(ig), _i(iop), _i(1 if ifrot else 0), _i(nvk), _p(gi), _p(ab), _p(rm),
      _p(ps), _p(mu), _p(um), _p(rl), _p(tr), _d(vol), _p(kc), _p(nl), _p(no), _p(lw), _i(nLOmax), _i(nat), _i(nt),
      _i(ngk), _i(gi.shape[0]), _i(ndf), _i(nl.shape[0] - 1), _i(ab.shape[2]))
    return np.ascontiguousarray(alm), np.ascontiguousarray(blm), np.ascontiguousarray(clm[:, :nLOmax])


def _q0_tables(nLO_at, mult, tables):
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))  # noqa: E731
    return F(nLO_at, np.int32), F(mult, np.int32), [F(t, np.float64) for t in tables]


def calcmmatvv_mt(n1, n2, alfa, beta, gama, nLO_at, mult, *tables):
    """for_q_0.f90:10-234 (0-based n1, n2; the nine radial tables in PyGW's
    order iul_ul, iul_udl, iudl_ul, iudl_udl, iulol_ul, iulol_udl,
    iul_ulol, iudl_ulol, iulol_ulol)."""
    f = _load("libref_q0.so").calcmmatvv_mt_
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.complex128))  # noqa: E731
    al, be, ga = F(alfa), F(beta), F(gama)
    nl, mu, t = _q0_tables(nLO_at, mult, tables)
    out = np.zeros(3, dtype=np.complex128)
    f(_p(out), _i(n1 + 1), _i(n2 + 1), _p(al), _p(be), _p(ga), _p(nl), _p(mu), *[_p(x) for x in t],
      _i(mu.size), _i(t[0].shape[1]), _i(al.shape[0]), _i(al.shape[1]), _i(al.shape[2]), _i(ga.shape[0]),
      _i(ga.shape[1]), _i(ga.shape[2]), _i(ga.shape[3]), _i(t[4].shape[1]), _i(t[4].shape[2] - 1),
      _i(nl.shape[0] - 1))
    return out


def calcmmatvv_mt2(ca1, cb1, cg1, a2, b2, g2, nLO_at, mult, *tables):
    """for_q_0.f90:236-458."""
    f = _load("libref_q0.so").calcmmatvv_mt2_
    F = lambda a: np.asfortranarray(np.asarray(a, dtype=np.complex128))  # noqa: E731
    x = [F(a) for a in (ca1, cb1, cg1, a2, b2, g2)]
    nl, mu, t = _q0_tables(nLO_at, mult, tables)
    out = np.zeros(3, dtype=np.complex128)
    f(_p(out), *[_p(a) for a in x], _p(nl), _p(mu), *[_p(a) for a in t], _i(mu.size), _i(t[0].shape[1]),
      _i(x[0].shape[0]), _i(x[0].shape[1]), _i(x[2].shape[0]), _i(x[2].shape[1]), _i(x[2].shape[2]),
      _i(t[4].shape[1]), _i(t[4].shape[2] - 1), _i(nl.shape[0] - 1))
    return out


def calcmmatvv_is(i, j, kl, Ak, indgk, nvk, ipwint, gindex, i_g0, k2cartes):
    """for_q_0.f90:460-507 (0-based i, j)."""
    f = _load("libref_q0.so").calcmmatvv_is_
    A = np.asfortranarray(np.asarray(Ak, dtype=np.complex128))
    ig = np.asfortranarray(np.asarray(indgk, dtype=np.int32))
    pw = np.asfortranarray(np.asarray(ipwint, dtype=np.complex128))
    gi = np.asfortranarray(np.asarray(gindex, dtype=np.int32))
    i0 = np.asfortranarray(np.asarray(i_g0, dtype=np.int32))
    kc = np.asfortranarray(np.asarray(k2cartes, dtype=np.float64))
    k = np.asfortranarray(np.asarray(kl, dtype=np.float64))
    out = np.zeros(3, dtype=np.complex128)
    n = [(i0.shape[c] - 1) // 2 for c in range(3)]
    f(_p(out), _i(i + 1), _i(j + 1), _p(k), _p(A), _p(ig), _i(nvk), _p(pw), _p(gi), _p(i0), _p(kc), _i(n[0]),
      _i(n[1]), _i(n[2]), _i(ig.size), _i(gi.shape[0]), _i(A.shape[0]), _i(A.shape[1]))
    return out


def calcmmatcv(iat, iv, corind, alfa, beta, gama, iul_ucl, iudl_ucl, iucl_ul, iucl_udl, iulol_ucl, iucl_ulol,
               nLO_at):
    """for_q_0.f90:510-683 (0-based iat, iv); the core tables of atom type
    iat alone: iul_ucl & co [2, ncg_a], iulol_ucl / iucl_ulol [2, nloat,
    ncg_a]."""
    f = _load("libref_q0.so").calcmmatcv_
    F = lambda a, dt: np.asfortranarray(np.asarray(a, dtype=dt))  # noqa: E731
    co = F(corind, np.int32)
    al, be, ga = F(alfa, np.complex128), F(beta, np.complex128), F(gama, np.complex128)
    t = [F(x, np.float64) for x in (iul_ucl, iudl_ucl, iucl_ul, iucl_udl, iulol_ucl, iucl_ulol)]
    nl = F(nLO_at, np.int32)
    out = np.zeros((co.shape[0], 3), dtype=np.complex128, order="F")
    f(_p(out), _i(iat + 1), _i(iv + 1), _p(co), _p(al), _p(be), _p(ga), *[_p(x) for x in t], _p(nl),
      _i(co.shape[0]), _i(nl.shape[1]), _i(t[4].shape[1]), _i(al.shape[0]), _i(al.shape[2]), _i(al.shape[1]),
      _i(ga.shape[0]), _i(ga.shape[1]), _i(ga.shape[2]), _i(ga.shape[3]), _i(t[0].shape[1]), _i(nl.shape[0] - 1))
    return np.ascontiguousarray(out)
