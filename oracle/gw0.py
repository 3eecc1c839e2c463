"""ORACLE (test infrastructure only): numpy FP64 restatement of PyGW's GW0
post-processing (self-consistent Sigma_c on fixed mwm), for checking the
GPU path.

  intdos1t / dos1t     one-tetrahedron integrated DOS and DOS
                                        (for_tetrahedra2.f90:108-212)
  idos / dostet        tetrahedron sums over bands  (for_tetrahedra.f90:84-141)
  calc_fermi           Fermi level, gap, band edges (mcommon.py:34-66)
  band_analys          last valence / first conduction band index
                                        (mcommon.py:212-232)
  quasiparticles       Compute_quasiparticles, iop_ac = 0 (mcommon.py:167-210)
  compute_selfc        Compute_selfc: sum over q of Compute_selfc_inside
                                        (postprocess.py:167-186, mcommon.py:80-165)
  gw0                  the GW0 iteration of calceqp (postprocess.py:206-321)

Array conventions: bands / eqp are [n_irk, n_band] (C order), atet is
[n_tet, 4] (0-based irreducible k indices), mwm is [n_q, n_irk, n_om, nb1,
nb2] (the stack of PyGW's mwm.{iq}.{irk}.npy files), sigc is [n_irk, nb1,
n_om].
"""
from __future__ import annotations

import numpy as np
from scipy import optimize

from . import chain


def intdos1t(et, ei):
    """Integrated DOS of one tetrahedron at ei; et sorted ascending
    (for_tetrahedra2.f90:108-160)."""
    e1, e2, e3, e4 = et
    if e4 <= ei:
        return 1.0
    idx = sum(1 for e in et if e <= ei)
    if idx == 0:
        return 0.0
    if idx == 1:
        e21, e31, e41, eij = e2 - e1, e3 - e1, e4 - e1, ei - e1
        return eij * eij * eij / (e21 * e31 * e41)
    if idx == 2:
        e21, e31, e41, e32, e42, eij = e2 - e1, e3 - e1, e4 - e1, e3 - e2, e4 - e2, ei - e2
        br = e21 * e21 + 3.0 * eij * e21 + 3.0 * eij * eij - (e31 + e42) * eij * eij * eij / (e32 * e42)
        return br / (e31 * e41)
    e43, e42, e41, eij = e4 - e3, e4 - e2, e4 - e1, e4 - ei
    return 1.0 - eij * eij * eij / (e41 * e42 * e43)


def dos1t(et, ei):
    """DOS of one tetrahedron at ei; et sorted ascending (Bloechl et al.,
    PRB 49, 16223 (1994); for_tetrahedra2.f90:162-212)."""
    e1, e2, e3, e4 = et
    idx = 4 if e4 <= ei else sum(1 for e in et if e <= ei)
    if idx in (0, 4):
        return 0.0
    if idx == 1:
        e21, e31, e41, eij = e2 - e1, e3 - e1, e4 - e1, ei - e1
        return 3.0 * eij * eij / (e21 * e31 * e41)
    if idx == 2:
        e21, e31, e41, e32, e42, eij = e2 - e1, e3 - e1, e4 - e1, e3 - e2, e4 - e2, ei - e2
        br = e21 + 2.0 * eij - (e31 + e42) * eij * eij / (e32 * e42)
        return 3.0 * br / (e31 * e41)
    e43, e42, e41, eij = e4 - e3, e4 - e2, e4 - e1, e4 - ei
    return 3.0 * eij * eij / (e41 * e42 * e43)


def idos(energy, ebnd, atet, wtet):
    """sum over bands and tetrahedra of wtet * intdos1t (for_tetrahedra.f90:84-116);
    ebnd [n_irk, n_band]."""
    s = 0.0
    for ib in range(ebnd.shape[1]):
        for it in range(atet.shape[0]):
            s += intdos1t(np.sort(ebnd[atet[it], ib]), energy) * wtet[it]
    return s


def dostet(energy, ebnd, atet, wtet):
    """sum over tetrahedra and bands of wtet * dos1t (for_tetrahedra.f90:118-141)."""
    s = 0.0
    for it in range(atet.shape[0]):
        for ib in range(ebnd.shape[1]):
            s += dos1t(np.sort(ebnd[atet[it], ib]), energy) * wtet[it]
    return s


def calc_fermi(ebnd, atet, wtet, nval, nspin=1, idos_fn=None, dos_fn=None):
    """(EF, Eg, evbm, ecbm, eDos) as calc_Fermi (mcommon.py:34-66): band
    counting with an idos check for an even electron count, else brentq on
    idos, then mid-gap if the DOS at EF is below 1e-4."""
    idos_fn = idos_fn or (lambda e: idos(e, ebnd, atet, wtet))
    dos_fn = dos_fn or (lambda e: dostet(e, ebnd, atet, wtet))
    e_dos = 1.0
    if nspin == 1 and nval % 2 == 0:
        nvm = int(nval / 2.0 + 0.51)
        evbm = ebnd[:, nvm - 1].max()
        ecbm = ebnd[:, nvm].min()
        ef = 0.5 * (evbm + ecbm)
        ocint = idos_fn(ef) * 2.0 / nspin
        if ecbm >= evbm and abs(ocint - nval) < 1e-6:
            eg = ecbm - evbm
            e_dos = 0
    if e_dos != 0:
        evbm, ecbm = ebnd.min(), ebnd.max()
        ocmax = nspin * idos_fn(ecbm) * 2.0 / nspin
        if ocmax <= nval:
            raise ValueError("not enough bands")
        ef = optimize.brentq(lambda x: nspin * idos_fn(x) * 2.0 / nspin - nval, evbm, ecbm)
        e_dos = nspin * dos_fn(ef) * 2.0 / nspin
        if e_dos < 1e-4:
            flat = ebnd.ravel()
            evbm = flat[flat < ef].max()
            ecbm = flat[flat > ef].min()
            ef = (evbm + ecbm) / 2.0
            eg = ecbm - evbm
        else:
            eg = -e_dos
            evbm, ecbm = ef, ef
    return ef, eg, evbm, ecbm, e_dos


def band_analys(bande, ef, nbmax):
    """(nomax, numin) of Band_Analys (mcommon.py:212-232)."""
    bands = bande[:, :nbmax]
    nocc = (bands < ef).sum(axis=1)
    return int(nocc.max()) - 1, int(nocc.min())


def quasiparticles(bands, ebnd0, sigc, sigx, vxc, omega, iop_es, iop_gw0=1, npts=32):
    """Compute_quasiparticles with iop_ac = 0 (mcommon.py:167-210): per state
    the Pade value S and slope S' at enk = bands, Z = 1/(1 - Re S') (1 and
    S' = 0 when outside [0, 1]), s = Re S + sigx and the update of iop_es /
    iop_gw0.  bands, ebnd0 [n_irk, nb]; sigc [n_irk, nb, n_om]; vxc the
    diagonal [n_irk, nb].  Returns (eqp, eqp_im)."""
    nirk, nb = bands.shape
    eqp = np.zeros((nirk, nb))
    eqp_im = np.zeros((nirk, nb))
    for irk in range(nirk):
        for ie in range(nb):
            enk0, enk, vx = ebnd0[irk, ie], bands[irk, ie], vxc[irk, ie]
            sig, dsig = chain.analytic_continuation(omega, sigc[irk, ie], enk, npts)
            z = 1.0 / (1.0 - dsig.real)
            s = sig.real + sigx[irk, ie]
            if z > 1.0 or z < 0:
                z = 1.0
            if iop_es == -1:
                if iop_gw0 == 1:
                    delta = s - vx + enk0 - enk
                elif iop_gw0 == 2:
                    delta = z * (s - vx) + enk0 - enk
                else:
                    delta = z * (s - vx + enk0 - enk)
            elif iop_es == 0:
                delta = z * (s - vx)
            else:
                raise ValueError("iop_es")
            eqp[irk, ie] = enk + delta
            eqp_im[irk, ie] = sig.imag * z
    return eqp, eqp_im


def selfc_inside(enk, mwm, omega, womeg, omega_p, womeg_p):
    """Compute_selfc_inside for one (q, irk), spline branch (mcommon.py:126-144):
    enk [nb2] (core energies, then bands at k-q), mwm [n_om, nb1, nb2]."""
    return chain.selfc_spline(enk, mwm, omega, omega_p, womeg_p)


def compute_selfc(bands, core, jrk, mwm, omega, womeg, omega_p, womeg_p):
    """Compute_selfc (postprocess.py:167-186): sigc[irk] = sum over q (in q
    order) of Compute_selfc_inside with enk = core energies then
    bands[jrk[q, irk], :nb2 - ncg] (mcommon.py:94-101)."""
    nq, nirk, nom, nb1, nb2 = mwm.shape
    ncg = len(core)
    sigc = np.zeros((nirk, nb1, nom), dtype=complex)
    for iq in range(nq):
        dsig = np.zeros_like(sigc)
        for irk in range(nirk):
            enk = np.concatenate([np.asarray(core, dtype=float), bands[jrk[iq, irk], :nb2 - ncg]])
            dsig[irk] += selfc_inside(enk, mwm[iq, irk], omega, womeg, omega_p, womeg_p)
        sigc += dsig
    return sigc


def gw0(ebnd, core, jrk, mwm, sigc, sigx, vxc, omega, womeg, omega_p, womeg_p, atet, wtet, nval, ibgw,
        iop_gw0=1, iop_esgw0=1, mix=1.0, nmax_sc=10, eps_sc=1e-3, shift_semicore=False, npts=32):
    """The GW0 branch of calceqp (postprocess.py:206-321) for one spin:
    G0W0 quasiparticles and Fermi level, then, for an insulator, iterations
    of Sigma_c on the quasiparticle bands (mixed by `mix`) until the k
    dependent gap changes by less than eps_sc.  ebnd [n_irk, n_band] KS
    energies (E_F = 0), bands ibgw .. ibgw + nb1 - 1 are the GW bands; core
    energies are shifted with the valence as the reference does.  Returns
    (eqp0, eqp or None, history) where history holds (isc, ediff, EF) per
    iteration."""
    nirk, nbgw = ebnd.shape[0], sigc.shape[1]
    nst, nend = ibgw, ibgw + nbgw
    core = np.array(core, dtype=float)
    bands = ebnd.copy()
    ef_qp = 0.0
    eqp0, _ = quasiparticles(bands[:, nst:nend], ebnd[:, nst:nend], sigc, sigx, vxc, omega, 0, iop_gw0, npts)
    ef, eg, evbm, ecbm, edos = calc_fermi(eqp0, atet, wtet, nval - ibgw * 2)
    if iop_esgw0 == 1:
        eqp0 = eqp0 - ef
        ef_qp += ef
        ef = 0.0
    nomax, numin = band_analys(eqp0, ef, nbgw)
    eqp = eqp0.copy()
    history = []
    if nomax >= numin:
        return eqp0, None, history
    egk = bands[:, numin].copy() - bands[:, nomax]
    sigc = sigc.copy()
    ediff = 0.0
    for isc in range(nmax_sc):
        de_vb = (eqp[:, nst] - bands[:, nst]).sum() / bands.shape[0] if shift_semicore else 0.0
        bands[:, :nst] += de_vb
        bands[:, nst:nend] = eqp
        if iop_esgw0 == 1 and ef != 0:
            bands -= ef
            core = core + de_vb - ef
            ef_qp += ef
            ef = 0.0
        egk_new = bands[:, numin] - bands[:, nomax]
        ediff = np.abs(egk - egk_new).max()
        egk = egk_new
        history.append((isc, ediff, ef))
        if ediff < eps_sc and isc > 0:
            break
        snew = compute_selfc(bands, core, jrk, mwm, omega, womeg, omega_p, womeg_p)
        sigc = sigc * (1 - mix) + mix * snew
        eqp, _ = quasiparticles(bands[:, nst:nend], ebnd[:, nst:nend], sigc, sigx, vxc, omega, -1, iop_gw0, npts)
        ef, eg, evbm, ecbm, edos = calc_fermi(eqp, atet, wtet, nval - ibgw * 2)
    if ediff >= 5e-2:
        return eqp0, None, history
    return eqp0, eqp, history
