"""GW0 on stored mwm: the self-consistent Sigma_c iteration of PyGW's
post-processing (SCGW0.calceqp, postprocess.py:206-321) on the GPU.

PyGW writes mwm[iom, ie1, ie2] of every (q, irreducible k) to
data/mwm.{iq}.{irk}.npy (pygw.py:3368) and then, per GW0 iteration,
recomputes Sigma_c of every (q, k) with the current quasiparticle bands
(Compute_selfc, postprocess.py:143-204), mixes it into the previous one,
continues it to the real axis (Compute_quasiparticles, mcommon.py:167-210)
and finds the new Fermi level (calc_Fermi, mcommon.py:34-66).  Here the mwm
stack stays resident in HBM and each iteration is:

  pygw_gw0_enk        energies at k - q of every (q, k) item
  pygw_selfc_spline   (1 + w^2) mwm splined onto the precise mesh
  pygw_selfc_convolve G0 (x) W of every item
  pygw_selfc_qsum     sum over q, in q order, and the mixing
  pygw_pade_qp_gw0    Pade + quasiparticle update (iop_es = 0 or -1)
  pygw_tetra_idos     tetrahedron idos / DOS for the Fermi level

walking q in chunks so that the precise-mesh spline of a chunk fits the
given budget.  The loop logic around the kernels (band analysis, gap
change, shifts) is the reference's, on arrays of n_k x n_band numbers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .polarization import _dev, _require_device, _stream_ptr


def mwm_path(ddir, iq, irk):
    """PyGW's file name of one (q, irreducible k) mwm (pygw.py:3368)."""
    return os.path.join(ddir, f"mwm.{iq}.{irk}.npy")


def save_mwm(ddir, iq, irk, mwm):
    """Write mwm [n_om, nb1, nb2] complex128 as PyGW does (numpy .npy)."""
    np.save(mwm_path(ddir, iq, irk), np.ascontiguousarray(mwm, dtype=np.complex128))


def load_mwm_stack(ddir, q_list, n_irk, *, device="cuda"):
    """The mwm files of the given q and every irreducible k as one device
    stack [n_q, n_irk, n_om, nb1, nb2] complex128 (read without pickle)."""
    import torch
    device = _require_device(device)
    first = np.load(mwm_path(ddir, q_list[0], 0), allow_pickle=False)
    out = torch.empty((len(q_list), n_irk) + first.shape, dtype=torch.complex128, device=device)
    for i, iq in enumerate(q_list):
        for irk in range(n_irk):
            a = first if (i == 0 and irk == 0) else np.load(mwm_path(ddir, iq, irk), allow_pickle=False)
            if a.shape != first.shape or a.dtype != np.complex128:
                raise ValueError(f"{mwm_path(ddir, iq, irk)}: {a.shape} {a.dtype}")
            out[i, irk].copy_(torch.from_numpy(a))
    return out


class SelfEnergy:
    """Sigma_c of all (q, k) for given bands: sigc[irk] = sum over q of
    Compute_selfc_inside (postprocess.py:176-186, mcommon.py:80-165, spline
    branch), with mwm [n_q, n_irk, n_om, nb1, nb2] resident on the device.
    jrk [n_q, n_irk]: irreducible index of k - q; core [n_core] energies.
    cache: keep the precise-mesh splines of all q (None: when they fit in
    half of the free device memory); else walk q in chunks of chunk_bytes
    of splines, recomputing them on every call."""

    def __init__(self, mwm, jrk, core, omega, omega_p, womeg_p, *, chunk_bytes=8 << 30, cache=None, device="cuda"):
        import torch
        self.device = device = _require_device(device)
        self.mwm = mwm if isinstance(mwm, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(mwm), device=device)
        if self.mwm.dtype != torch.complex128 or not self.mwm.is_contiguous():
            raise ValueError("mwm must be a contiguous complex128 stack")
        self.nq, self.nirk, self.nom, self.nb1, self.nb2 = self.mwm.shape
        self.jrk = torch.as_tensor(np.ascontiguousarray(jrk, dtype=np.int32), device=device)
        if tuple(self.jrk.shape) != (self.nq, self.nirk):
            raise ValueError("jrk shape")
        self.core = _dev(np.asarray(core, dtype=np.float64), torch.float64, device).contiguous()
        self.ncore = int(np.asarray(core).size)
        self.om = _dev(omega, torch.float64, device).contiguous()
        self.omp = _dev(omega_p, torch.float64, device).contiguous()
        self.wop = _dev(womeg_p, torch.float64, device).contiguous()
        self.nomp = self.omp.numel()
        per_q = self.nirk * self.nb1 * self.nb2 * self.nomp * 16
        # The splines of mwm on the precise mesh do not depend on the bands:
        # when those of every q fit (by default: in half of the free device
        # memory) they are computed once, on the first call, and kept.
        if cache is None:
            free, _ = torch.cuda.mem_get_info(device)
            cache = self.nq * per_q <= free // 2
        self.cached = bool(cache)
        self.splined = False
        self.qc = self.nq if self.cached else max(1, min(self.nq, chunk_bytes // per_q))
        items = self.qc * self.nirk
        self.mwmp = torch.empty((items, self.nb1, self.nb2, self.nomp), dtype=torch.complex128, device=device)
        self.sc = torch.empty((items, self.nb1, self.nom), dtype=torch.complex128, device=device)
        self.enk = torch.empty((items, self.nb2), dtype=torch.float64, device=device)

    def __call__(self, bands, out=None, alpha=1.0, beta=0.0):
        """out = beta*out + alpha * sum_q Sigma_c(q) at `bands` [n_irk, n_band]
        (device or host); returns out [n_irk, nb1, n_om]."""
        import torch
        b = _dev(np.ascontiguousarray(bands) if not isinstance(bands, torch.Tensor) else bands, torch.float64,
                 self.device).contiguous()
        nband = b.shape[1]
        if out is None:
            out = torch.empty((self.nirk, self.nb1, self.nom), dtype=torch.complex128, device=self.device)
            beta = 0.0
        s = _stream_ptr(self.device)
        n = self.nirk * self.nb1 * self.nom
        for q0 in range(0, self.nq, self.qc):
            nq = min(self.qc, self.nq - q0)
            items = nq * self.nirk
            jr = self.jrk[q0:q0 + nq]
            _lib.call("pygw_gw0_enk", C.c_void_p(b.data_ptr()), self.nirk, nband,
                      C.c_void_p(self.core.data_ptr()) if self.ncore else None, self.ncore,
                      C.c_void_p(jr.data_ptr()), items, self.nb2, C.c_void_p(self.enk.data_ptr()), s)
            mw = self.mwm[q0:q0 + nq]
            if not self.splined:
                _lib.call("pygw_selfc_spline", C.c_void_p(mw.data_ptr()), items, self.nb1, self.nb2,
                          C.c_void_p(self.om.data_ptr()), self.nom, C.c_void_p(self.omp.data_ptr()), self.nomp,
                          C.c_void_p(self.mwmp.data_ptr()), s)
                self.splined = self.cached
            _lib.call("pygw_selfc_convolve", C.c_void_p(self.enk.data_ptr()), items, self.nb1, self.nb2,
                      C.c_void_p(mw.data_ptr()), C.c_void_p(self.om.data_ptr()), self.nom,
                      C.c_void_p(self.mwmp.data_ptr()), C.c_void_p(self.omp.data_ptr()),
                      C.c_void_p(self.wop.data_ptr()), self.nomp, C.c_void_p(self.sc.data_ptr()), s)
            _lib.call("pygw_selfc_qsum", C.c_void_p(self.sc.data_ptr()), nq, n, alpha,
                      beta if q0 == 0 else 1.0, C.c_void_p(out.data_ptr()), s)
        return out


def mix(sigc, snew, mix_sc, *, device="cuda"):
    """sigc <- sigc*(1 - mix) + mix*snew in place (postprocess.py:297)."""
    _lib.call("pygw_selfc_qsum", C.c_void_p(snew.data_ptr()), 1, snew.numel(), float(mix_sc), 1.0 - float(mix_sc),
              C.c_void_p(sigc.data_ptr()), _stream_ptr(_require_device(device)))
    return sigc


def quasiparticles(bands, ebnd0, sigc, sigx, vxc, omega, iop_es, iop_gw0=1, npade=32, *, device="cuda"):
    """Compute_quasiparticles with iop_ac = 0 (mcommon.py:167-210).  bands,
    ebnd0, sigx, vxc [n_irk, nb]; sigc [n_irk, nb, n_om].  Returns (eqp,
    eqp_im) as host arrays [n_irk, nb]."""
    import torch
    device = _require_device(device)
    sg = sigc if isinstance(sigc, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(sigc), device=device)
    sg = sg.to(device=device, dtype=torch.complex128).contiguous()
    nirk, nb, nom = sg.shape
    e = _dev(np.ascontiguousarray(bands, dtype=np.float64).reshape(-1), torch.float64, device)
    e0 = _dev(np.ascontiguousarray(ebnd0, dtype=np.float64).reshape(-1), torch.float64, device)
    sx = _dev(np.ascontiguousarray(sigx, dtype=np.float64).reshape(-1), torch.float64, device)
    vx = _dev(np.ascontiguousarray(vxc, dtype=np.float64).reshape(-1), torch.float64, device)
    om = _dev(omega, torch.float64, device).contiguous()
    out = torch.empty((nirk * nb, 6), dtype=torch.float64, device=device)
    _lib.call("pygw_pade_qp_gw0", C.c_void_p(sg.data_ptr()), nirk * nb, nom, C.c_void_p(om.data_ptr()),
              min(npade, nom), C.c_void_p(e.data_ptr()), C.c_void_p(e0.data_ptr()), C.c_void_p(sx.data_ptr()),
              C.c_void_p(vx.data_ptr()), int(iop_es), int(iop_gw0), C.c_void_p(out.data_ptr()), _stream_ptr(device))
    o = out.cpu().numpy()
    return o[:, 5].reshape(nirk, nb), (o[:, 1] * o[:, 4]).reshape(nirk, nb)


class TetraDos:
    """Tetrahedron integrated DOS / DOS of a band set at many energies per
    launch (ft.idos / ft.dostet, for_tetrahedra.f90:84-141)."""

    NE = 255

    def __init__(self, atet, wtet, *, device="cuda"):
        import torch
        self.device = device = _require_device(device)
        self.atet = torch.as_tensor(np.ascontiguousarray(atet, dtype=np.int32), device=device)
        self.wtet = _dev(np.asarray(wtet, dtype=np.float64), torch.float64, device).contiguous()
        self.ntet = self.atet.shape[0]
        self.ebnd = None

    def set_bands(self, ebnd):
        import torch
        self.ebnd = torch.as_tensor(np.ascontiguousarray(ebnd, dtype=np.float64), device=self.device)
        self.nk, self.nb = self.ebnd.shape

    def __call__(self, energies):
        import torch
        en = torch.as_tensor(np.atleast_1d(np.asarray(energies, dtype=np.float64)), device=self.device)
        idos = torch.empty_like(en)
        dos = torch.empty_like(en)
        _lib.call("pygw_tetra_idos", C.c_void_p(self.ebnd.data_ptr()), self.nk, self.nb,
                  C.c_void_p(self.atet.data_ptr()), C.c_void_p(self.wtet.data_ptr()), self.ntet,
                  C.c_void_p(en.data_ptr()), en.numel(), C.c_void_p(idos.data_ptr()), C.c_void_p(dos.data_ptr()),
                  _stream_ptr(self.device))
        return idos.cpu().numpy(), dos.cpu().numpy()

    def root(self, target, a, b):
        """The energy in [a, b] where 2*idos = target (the bracketing root of
        calc_Fermi's brentq), by multisection on NE energies per launch down
        to a few ulps of the bracket."""
        fa = 2.0 * self(a)[0][0] - target
        for _ in range(12):
            xs = np.linspace(a, b, self.NE + 2)[1:-1]
            f = 2.0 * self(xs)[0] - target
            xs = np.concatenate([[a], xs, [b]])
            fs = np.concatenate([[fa], f, [2.0 * self(b)[0][0] - target]])
            i = int(np.argmax(np.sign(fs[1:]) != np.sign(fs[0])))   # first sign change of the monotone idos
            if fs[i + 1] == 0.0:
                return xs[i + 1]
            a, b, fa = xs[i], xs[i + 1], fs[i]
            if b - a <= 4.0 * np.finfo(float).eps * max(abs(a), abs(b), 1e-300):
                break
        return 0.5 * (a + b)


def calc_fermi(ebnd, dos: TetraDos, nval, nspin=1):
    """(EF, Eg, evbm, ecbm, eDos) of calc_Fermi (mcommon.py:34-66) with the
    tetrahedron sums on the GPU."""
    ebnd = np.asarray(ebnd, dtype=np.float64)
    dos.set_bands(ebnd)
    e_dos = 1.0
    if nspin == 1 and nval % 2 == 0:
        nvm = int(nval / 2.0 + 0.51)
        evbm, ecbm = ebnd[:, nvm - 1].max(), ebnd[:, nvm].min()
        ef = 0.5 * (evbm + ecbm)
        ocint = dos(ef)[0][0] * 2.0 / nspin
        if ecbm >= evbm and abs(ocint - nval) < 1e-6:
            eg = ecbm - evbm
            e_dos = 0
    if e_dos != 0:
        evbm, ecbm = ebnd.min(), ebnd.max()
        if nspin * dos(ecbm)[0][0] * 2.0 / nspin <= nval:
            raise ValueError("calc_fermi: not enough bands")
        ef = dos.root(nval, evbm, ecbm)
        e_dos = dos(ef)[1][0] * 2.0
        if e_dos < 1e-4:
            flat = ebnd.ravel()
            evbm, ecbm = flat[flat < ef].max(), flat[flat > ef].min()
            ef = 0.5 * (evbm + ecbm)
            eg = ecbm - evbm
        else:
            eg = -e_dos
            evbm = ecbm = ef
    return ef, eg, evbm, ecbm, e_dos


def band_analys(bande, ef, nbmax):
    """(nomax, numin) of Band_Analys (mcommon.py:212-232)."""
    nocc = (np.asarray(bande)[:, :nbmax] < ef).sum(axis=1)
    return int(nocc.max()) - 1, int(nocc.min())


def gw0(ebnd, core, selfe: SelfEnergy, sigc, sigx, vxc, omega, atet, wtet, nval, ibgw, iop_gw0=1, iop_esgw0=1,
        mix_sc=1.0, nmax_sc=10, eps_sc=1e-3, shift_semicore=False, npade=32, *, device="cuda"):
    """The GW0 branch of calceqp (postprocess.py:206-321) for one spin.
    ebnd [n_irk, n_band] Kohn-Sham energies; sigc [n_irk, nb1, n_om] the
    G0W0 Sigma_c (device or host); selfe recomputes Sigma_c of all q at new
    bands.  Returns (eqp0, eqp or None, history [(isc, ediff, EF)])."""
    import torch
    device = _require_device(device)
    nbgw = sigc.shape[1]
    nst, nend = ibgw, ibgw + nbgw
    core = np.array(core, dtype=np.float64)
    ebnd = np.asarray(ebnd, dtype=np.float64)
    bands = ebnd.copy()
    dos = TetraDos(atet, wtet, device=device)
    sg = sigc if isinstance(sigc, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(sigc), device=device)
    sg = sg.to(device=device, dtype=torch.complex128).clone()
    eqp0, _ = quasiparticles(bands[:, nst:nend], ebnd[:, nst:nend], sg, sigx, vxc, omega, 0, iop_gw0, npade,
                             device=device)
    ef, eg, evbm, ecbm, edos = calc_fermi(eqp0, dos, nval - ibgw * 2)
    if iop_esgw0 == 1:
        eqp0 = eqp0 - ef
        ef = 0.0
    nomax, numin = band_analys(eqp0, ef, nbgw)
    eqp = eqp0.copy()
    history = []
    if nomax >= numin:
        return eqp0, None, history
    egk = bands[:, numin] - bands[:, nomax]
    ediff = 0.0
    snew = None
    for isc in range(nmax_sc):
        de_vb = (eqp[:, nst] - bands[:, nst]).sum() / bands.shape[0] if shift_semicore else 0.0
        bands[:, :nst] += de_vb
        bands[:, nst:nend] = eqp
        if iop_esgw0 == 1 and ef != 0:
            bands -= ef
            core = core + de_vb - ef
            ef = 0.0
        egk_new = bands[:, numin] - bands[:, nomax]
        ediff = float(np.abs(egk - egk_new).max())
        egk = egk_new
        history.append((isc, ediff, ef))
        if ediff < eps_sc and isc > 0:
            break
        selfe.core.copy_(torch.as_tensor(core))
        snew = selfe(bands, out=snew)
        mix(sg, snew, mix_sc, device=device)
        eqp, _ = quasiparticles(bands[:, nst:nend], ebnd[:, nst:nend], sg, sigx, vxc, omega, -1, iop_gw0, npade,
                                device=device)
        ef, eg, evbm, ecbm, edos = calc_fermi(eqp, dos, nval - ibgw * 2)
    if ediff >= 5e-2:
        return eqp0, None, history
    return eqp0, eqp, history
