"""The q loop of G0W0.Compute for the polarization -> eps^-1 -> Sigma_c chain.

PyGW runs, for every q of its q-list (pygw.py:3659-3711):

    kcw  = Polarization_weights(iq, ...)          pygw.py:3690  (bz_calcqdepw_par2)
    eps  = me.calc_eps(iq, ..., kcw, ...)          pygw.py:3699  (chi0, eps^-1 - 1)
    sc   = me.calc_selfc(iq, eps, ...)             pygw.py:3705  (MWM with wkq = 1/N_q,
                                                                  spline + convolution)
    sigc += sc                                     pygw.py:3706

and after the loop the Pade continuation and the quasiparticle equation
(mcommon.py:167-210, 315-343).  SelfEnergyQLoop holds every device buffer
and workspace of that chain for one band structure, allocated once, so that
add_q() launches the HIP kernels of libpygw_hip.so on the current stream
with no allocation, memset or host synchronisation (but for the basis sizes
of omega_rank / w_rank "auto", read on the host): a q step can be captured in a HIP
graph (capture()) and replayed with the next q's inputs copied into the
loop's input buffers (set_q()).

The per-q inputs are the k-q index column (kqm.kqid[:, iq]), M of the
polarization pairs of every k ([N_k][n_o*n_e][PB], pair p = i_o*n_e + i_e,
v^(1/2) folded in as calc_minm does at pygw.py:2666), M of the self-energy
rows of the Sigma_c k-points ([n_ks][nb1*nb2][PB], calc_selfc's minm2,
pygw.py:3339-3347) and the energies of the k-q states of those k-points
([n_ks][nb2]; the enk of Compute_selfc_inside, mcommon.py:94-101).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .polarization import COMPAT_REFERENCE, COMPRESS_MAX_FRACTION, _fortran_view, _require_device, _stream_ptr, \
    omega_rank

_PREC = {"fp64": 0, "mixed": 1}


class SelfEnergyQLoop:
    """Sigma_c(k, i omega) summed over the q-points of a q-list.

    enk [n_bnd, N_k] (PyGW's Fortran-ordered band energies, E_F = ef),
    ncg/nomx/numin the band windows of Polarization_weights, omega/womeg the
    Matsubara mesh and weights, tetc [4, 6 N_k], pb the product-basis size,
    n_q the length of the q-list (wkq = 1/n_q, pygw.py:3312), n_ks the
    Sigma_c k-points, nb1 = nbgw - ibgw self-energy bands, nb2 the bands of
    the k-q states (calc_selfc's minm2 rows are (ie1, ie2) = ie1*nb2 + ie2).
    With omega_p/womeg_p the spline branch on the precise mesh is used
    (iopfreq = 4, mcommon.py:126-144), else the coarse all_fr_convolutions.
    precision "fp64" or "mixed" (split-precision chi0 and MWM on MFMA; the
    weights, inversion, convolution and Pade are FP64 either way).
    omega_rank: None contracts chi0 frequency by frequency; an int r or
    "auto" contracts it in the q's own frequency basis (the leading r
    eigenvectors of the weights' Gram matrix with every frequency scaled to
    unit norm, "auto": polarization.omega_rank
    of its singular values, read on the host once per q) and expands it back
    to the n_om frequencies before the inversion, as PyGW's SVD mode does with
    its analytic basis (pygw.py:3206-3226).  w_rank: the same for the MWM
    contraction, in the frequency basis of W = eps^-1 - 1 itself (PyGW's SVD
    mode projects eps on its basis before calc_selfc, pygw.py:3273-3276);
    mwm is expanded back to the n_om frequencies for the convolution.
    """

    def __init__(self, enk, ef, ncg, nomx, numin, omega, womeg, tetc, pb, n_q, n_ks, nb1, nb2, *,
                 omega_p=None, womeg_p=None, precision="mixed", compat=COMPAT_REFERENCE, omega_rank=None,
                 w_rank=None, device="cuda"):
        d = _require_device(device)
        self.device = d = d if d.index is not None else torch.device(d.type, torch.cuda.current_device())
        if precision not in _PREC:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        enk = np.asarray(enk, dtype=np.float64)
        self.nbnd, self.nk = enk.shape
        self.ef, self.ncg, self.nomx, self.numin = float(ef), int(ncg), int(nomx), int(numin)
        self.no, self.ne = self.ncg + self.nomx + 1, self.nbnd - self.ncg - self.numin
        if self.no <= 0 or self.ne <= 0:
            raise ValueError("empty band window")
        self.np = self.no * self.ne
        self.pb, self.n_q, self.precision, self.compat = int(pb), int(n_q), precision, int(compat)
        self.wkq = 1.0 / self.n_q
        self.nks, self.nb1, self.nb2 = int(n_ks), int(nb1), int(nb2)
        self.rows = self.nb1 * self.nb2
        f64 = lambda x: torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=d)  # noqa: E731
        self.enk = f64(enk.T)                                                  # [N_k, n_bnd] = Fortran (n_bnd, N_k)
        self.tetc = torch.as_tensor(np.ascontiguousarray(np.asarray(tetc).T), dtype=torch.int32, device=d)
        self.ntet = self.tetc.shape[0]
        if self.tetc.shape[1] != 4 or self.ntet != 6 * self.nk:
            raise ValueError("tetc must be [4, 6 N_k]")
        self.om, self.wo = f64(omega), f64(womeg)
        self.nom = self.om.numel()
        self.spline = omega_p is not None
        if self.spline:
            self.omp, self.wop = f64(omega_p), f64(womeg_p)
            self.nomp = self.omp.numel()
        # per-q inputs (set_q copies into these; the M buffers are made on
        # the first set_q that passes M, since run_q may take M by pointer)
        self.kq = torch.zeros(self.nk, dtype=torch.int32, device=d)
        self.mdt = torch.complex128 if precision == "fp64" else torch.complex64
        self.M = self.Ms = None
        self.enk_kq = torch.zeros((self.nks, self.nb2), dtype=torch.float64, device=d)
        # per-q intermediates
        self.kcw = torch.empty((self.nk, self.nom, self.np), dtype=torch.float64, device=d)
        self.eps_buf = torch.empty(self.pb * self.pb * self.nom, dtype=torch.complex128, device=d)
        self.eps = _fortran_view(self.eps_buf, self.pb, self.nom)
        self.mwm = torch.empty((self.nks, self.nom, self.rows), dtype=torch.complex128, device=d)
        self.mwm_p = (torch.empty((self.nks, self.nb1, self.nb2, self.nomp), dtype=torch.complex128, device=d)
                      if self.spline else None)
        self.sc = torch.empty((self.nks, self.nb1, self.nom), dtype=torch.complex128, device=d)
        self.sigc = torch.zeros((self.nks, self.nb1, self.nom), dtype=torch.complex128, device=d)
        self.chi = None                     # optional copy of chi0 (keep_chi0)
        # workspaces of every entry point, sized once
        sz = C.c_size_t(0)
        _lib.call("pygw_tetra_weights_workspace", self.nk, self.np, self.nom, C.byref(sz))
        self.ws_tw = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        p = _PREC[precision]
        _lib.call("pygw_chi0_workspace", self.pb, self.np, self.nom, self.nk, p, C.byref(sz))
        self.ws_c0 = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        _lib.call("pygw_eps_inverse_workspace", self.pb, self.nom, C.byref(sz))
        self.ws_ei = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        _lib.call("pygw_mwm_workspace", self.nks, self.rows, self.pb, self.nom, p, C.byref(sz))
        self.ws_mw = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        # compressed frequency basis of the contraction
        for r in (omega_rank, w_rank):
            if r is not None and r != "auto" and not 1 <= int(r) <= self.nom:
                raise ValueError("omega_rank and w_rank must be None, 'auto' or 1..n_om")
        self.omega_rank, self.rank_used = omega_rank, None
        self.w_rank, self.w_rank_used = w_rank, None
        if omega_rank is not None or w_rank is not None:
            _lib.call("pygw_omega_basis_workspace", self.nk, self.nom, self.np, C.byref(sz))
            self.ws_ob = torch.empty(max(sz.value, 8), dtype=torch.uint8, device=d)
        if omega_rank is not None:
            rmax = self.nom if omega_rank == "auto" else int(omega_rank)
            self.Up = torch.empty((self.nom, self.nom), dtype=torch.float64, device=d)   # projection
            self.Ue = torch.empty((self.nom, self.nom), dtype=torch.float64, device=d)   # expansion
            self.sv = torch.empty(self.nom, dtype=torch.float64, device=d)
            self.kcw_r = torch.empty((self.nk, rmax, self.np), dtype=torch.float64, device=d)
            self.x_buf = torch.empty(self.pb * self.pb * rmax, dtype=torch.complex128, device=d)
        if w_rank is not None:
            rmax = self.nom if w_rank == "auto" else int(w_rank)
            self.Uwp = torch.empty((self.nom, self.nom), dtype=torch.float64, device=d)
            self.Uwe = torch.empty((self.nom, self.nom), dtype=torch.float64, device=d)
            self.svw = torch.empty(self.nom, dtype=torch.float64, device=d)
            self.xw_buf = torch.empty(self.pb * self.pb * rmax, dtype=torch.complex128, device=d)
            self.mwm_r = torch.empty((self.nks, rmax, self.rows), dtype=torch.complex128, device=d)
        self.n_added = 0

    # -- inputs ---------------------------------------------------------------
    def set_q(self, kqid_col, M=None, M_selfe=None, enk_kq=None):
        """Copy one q's inputs into the loop's buffers (device-to-device for
        device tensors).  Any left as None keeps its current contents (e.g.
        when M was written in place through `self.M`)."""
        if kqid_col is not None:
            k = kqid_col if isinstance(kqid_col, torch.Tensor) else torch.as_tensor(np.asarray(kqid_col, np.int32))
            if k.numel() != self.nk:
                raise ValueError(f"kqid column must have {self.nk} entries")
            self.kq.copy_(k.reshape(-1).to(torch.int32))
        if M is not None and self.M is None:
            self.M = torch.empty((self.nk, self.np, self.pb), dtype=self.mdt, device=self.device)
        if M_selfe is not None and self.Ms is None:
            self.Ms = torch.empty((self.nks, self.rows, self.pb), dtype=self.mdt, device=self.device)
        for dst, src, name in ((self.M, M, "M"), (self.Ms, M_selfe, "M_selfe"), (self.enk_kq, enk_kq, "enk_kq")):
            if src is None:
                continue
            src = src if isinstance(src, torch.Tensor) else torch.as_tensor(np.asarray(src))
            if tuple(src.shape) != tuple(dst.shape):
                raise ValueError(f"{name} must have shape {tuple(dst.shape)}, got {tuple(src.shape)}")
            dst.copy_(src)

    # -- one q -------------------------------------------------------------------
    def _operand(self, t, buf, shape, name):
        t = buf if t is None else t
        if t is None:
            raise ValueError(f"no {name}: pass it to set_q or run_q")
        if tuple(t.shape) != shape or t.dtype != self.mdt or not t.is_contiguous() or t.device != self.device:
            raise ValueError(f"{name} must be a contiguous {self.mdt} device tensor of shape {shape}")
        return t

    def run_q(self, events=None, keep_chi0=False, M=None, M_selfe=None):
        """Polarization_weights -> calc_eps -> calc_selfc for the q in the
        input buffers (M / M_selfe: device tensors of the loop's shapes and
        dtype used in place of the M buffers, by pointer), on the current
        stream; leaves this q's Sigma_c contribution in self.sc (and chi0 in
        self.chi with keep_chi0) and eps^-1 - 1 in self.eps.  events:
        optional list of 6 torch events recorded before the weights, chi0,
        inversion, MWM, convolution and after it.  Nothing is allocated
        (keep_chi0 copies into a buffer made on its first use)."""
        Mq = self._operand(M, self.M, (self.nk, self.np, self.pb), "M")
        Msq = self._operand(M_selfe, self.Ms, (self.nks, self.rows, self.pb), "M_selfe")
        s = _stream_ptr(self.device)
        cs = torch.cuda.current_stream(self.device)
        rec = (lambda i: events[i].record(cs)) if events is not None else (lambda i: None)
        vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        p = _PREC[self.precision]
        rec(0)
        _lib.call("pygw_tetra_weights", vp(self.enk), self.ef, vp(self.kq), vp(self.tetc), self.ncg, self.nomx,
                  self.numin, vp(self.om), self.nom, self.nk, self.nbnd, self.ntet, 1, self.compat, vp(self.kcw),
                  vp(self.ws_tw), self.ws_tw.numel(), None, s)
        rec(1)
        r = self.nom
        if self.omega_rank is not None:
            _lib.call("pygw_omega_basis", vp(self.kcw), self.nk, self.nom, self.np, 1, vp(self.Up), vp(self.Ue),
                      vp(self.sv), None, vp(self.ws_ob), self.ws_ob.numel(), s)
            r = omega_rank(self.sv) if self.omega_rank == "auto" else int(self.omega_rank)
            if self.omega_rank == "auto" and r > COMPRESS_MAX_FRACTION * self.nom:
                r = self.nom                        # too little to gain: frequency by frequency
        self.rank_used = r
        if r == self.nom and self.omega_rank != self.nom:
            _lib.call("pygw_chi0_accumulate", vp(Mq), self.np * self.pb, self.np, self.pb, vp(self.kcw), self.nom, 0,
                      self.nk, -2.0, p, 0, vp(self.eps_buf), vp(self.ws_c0), self.ws_c0.numel(), s)
        else:
            _lib.call("pygw_omega_project", vp(self.kcw), self.nk, self.nom, self.np, vp(self.Up), self.nom, 1, r,
                      vp(self.kcw_r), s)
            _lib.call("pygw_chi0_accumulate", vp(Mq), self.np * self.pb, self.np, self.pb, vp(self.kcw_r), r, 0,
                      self.nk, -2.0, p, 0, vp(self.x_buf), vp(self.ws_c0), self.ws_c0.numel(), s)
            _lib.call("pygw_omega_expand", vp(self.x_buf), vp(self.Ue), self.nom, 1, self.pb, r, self.nom, 0,
                      vp(self.eps_buf), s)
        if keep_chi0:
            if self.chi is None:
                self.chi = _fortran_view(torch.empty_like(self.eps_buf), self.pb, self.nom)
            self.chi.as_strided((self.eps_buf.numel(),), (1,)).copy_(self.eps_buf)
        rec(2)
        _lib.call("pygw_eps_inverse", vp(self.eps_buf), self.pb, self.nom, vp(self.ws_ei), self.ws_ei.numel(), None,
                  s)
        rec(3)
        r = self.nom
        if self.w_rank is not None:
            n2 = 2 * self.pb * self.pb        # W read as real [1][n_om][2 PB^2]
            _lib.call("pygw_omega_basis", vp(self.eps_buf), 1, self.nom, n2, 1, vp(self.Uwp), vp(self.Uwe),
                      vp(self.svw), None, vp(self.ws_ob), self.ws_ob.numel(), s)
            r = omega_rank(self.svw) if self.w_rank == "auto" else int(self.w_rank)
            if self.w_rank == "auto" and r > COMPRESS_MAX_FRACTION * self.nom:
                r = self.nom
        self.w_rank_used = r
        if r == self.nom and self.w_rank != self.nom:
            _lib.call("pygw_mwm", vp(Msq), self.nks, self.rows, self.pb, vp(self.eps_buf), self.nom, self.wkq, p,
                      vp(self.mwm), vp(self.ws_mw), self.ws_mw.numel(), s)
        else:
            _lib.call("pygw_omega_project", vp(self.eps_buf), 1, self.nom, n2, vp(self.Uwp), self.nom, 1, r,
                      vp(self.xw_buf), s)
            _lib.call("pygw_mwm", vp(Msq), self.nks, self.rows, self.pb, vp(self.xw_buf), r, self.wkq, p,
                      vp(self.mwm_r), vp(self.ws_mw), self.ws_mw.numel(), s)
            _lib.call("pygw_omega_expand_batched", vp(self.mwm_r), vp(self.Uwe), self.nom, 1, self.nks, self.rows, r,
                      self.nom, 0, vp(self.mwm), s)
        rec(4)
        if self.spline:
            _lib.call("pygw_selfc_spline", vp(self.mwm), self.nks, self.nb1, self.nb2, vp(self.om), self.nom,
                      vp(self.omp), self.nomp, vp(self.mwm_p), s)
            _lib.call("pygw_selfc_convolve", vp(self.enk_kq), self.nks, self.nb1, self.nb2, vp(self.mwm), vp(self.om),
                      self.nom, vp(self.mwm_p), vp(self.omp), vp(self.wop), self.nomp, vp(self.sc), s)
        else:
            _lib.call("pygw_selfc_convolve", vp(self.enk_kq), self.nks, self.nb1, self.nb2, vp(self.mwm), vp(self.om),
                      self.nom, None, vp(self.om), vp(self.wo), self.nom, vp(self.sc), s)
        rec(5)
        return self.sc

    def accumulate(self):
        """sigc += sc (pygw.py:3706), on the device."""
        vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        _lib.call("pygw_selfc_qsum", vp(self.sc), 1, self.sc.numel(), 1.0, 1.0, vp(self.sigc),
                  _stream_ptr(self.device))
        self.n_added += 1
        return self.sigc

    def add_q(self, kqid_col, M=None, M_selfe=None, enk_kq=None, events=None, keep_chi0=False):
        """One iteration of the q loop: set_q (kqid, enk_kq), run_q on M and
        M_selfe (by pointer when they are device tensors of the loop's dtype,
        else copied into the loop's buffers), sigc += sc."""
        direct = lambda t: isinstance(t, torch.Tensor) and t.dtype == self.mdt and t.is_contiguous() and \
            t.device == self.device  # noqa: E731
        dm, ds = direct(M), direct(M_selfe)
        self.set_q(kqid_col, None if dm else M, None if ds else M_selfe, enk_kq)
        self.run_q(events, keep_chi0, M if dm else None, M_selfe if ds else None)
        return self.accumulate()

    def reset(self):
        self.sigc.zero_()
        self.n_added = 0

    def capture(self, keep_chi0=False, M=None, M_selfe=None, accumulate=True):
        """Record run_q() (and sigc += sc) on the current input buffers (or
        on the given M / M_selfe, by pointer) in a HIP graph; replaying it
        after set_q() of the next q's kqid/energies (and M contents) runs
        that q.  Returns the torch.cuda.CUDAGraph."""
        if self.omega_rank == "auto" or self.w_rank == "auto":
            raise ValueError("graph capture needs fixed omega_rank / w_rank (the basis size is read on the host)")
        g = torch.cuda.CUDAGraph()
        self.run_q(keep_chi0=keep_chi0, M=M, M_selfe=M_selfe)   # first call outside the graph (module load)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self.run_q(keep_chi0=keep_chi0, M=M, M_selfe=M_selfe)
            if accumulate:
                vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
                _lib.call("pygw_selfc_qsum", vp(self.sc), 1, self.sc.numel(), 1.0, 1.0, vp(self.sigc),
                          _stream_ptr(self.device))
        return g

    # -- after the loop ----------------------------------------------------------
    def quasiparticles(self, ene, sigx=None, vxc=None, npade=32, sigc=None):
        """Pade continuation of sigc (iop_ac = 0, first min(32, n_om)
        points) and the G0W0 quasiparticle equation for every (k, band):
        returns float64 [n_ks*nb1, 6] (Re S, Im S, Re S', Im S', Z, E_qp),
        ene/sigx/vxc [n_ks*nb1] in the same (k, band) order."""
        from .polarization import pade_quasiparticles
        s = self.sigc if sigc is None else sigc
        return pade_quasiparticles(s.reshape(-1, self.nom), self.om, ene, sigx, vxc, npade=npade, device=self.device)
