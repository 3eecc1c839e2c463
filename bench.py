"""Benchmark of the gfx950 polarization path: seconds per q-point.

One step = one q-point of PyGW's GW loop: tetrahedron weights ->
chi0 contraction -> dielectric inversion (calc_eps, pygw.py:3039-3297 with
Polarization_weights, pygw.py:3416-3476) -> MWM -> spline + G0 (x) W
convolution (calc_selfc, pygw.py:3299-3413; mcommon.py:80-165), on inputs
already resident in HBM; Pade continuation + QP energies after the q loop
(mcommon.py:167-210).

Default workload (no flags): the full chain at 8x8x8 k (3072 tetrahedra),
20+120 bands, PB = 900, 64 Matsubara points, Sigma_c summed over a q-list of
8 distinct q-points (wkq = 1/8) through the package's q loop
(pygw_amd.qloop.SelfEnergyQLoop).  Every q of the list runs once on the FP64
path and once with the split-precision (FP16 hi+lo, 3-product MFMA)
contractions on the same complex64-rounded M (errors of chi0, eps^-1 and
Sigma_c per q and frequency; quasiparticle energies of both q sums beside
the Pade's own sensitivity); then the timed split-precision steps cycle over
the q-list.  value = split-precision seconds per q when every q passes
(chi0 and eps^-1 within 1e-6, Sigma_c within 1e-5, relative Frobenius per
frequency), else the FP64 seconds per q.

    python bench.py [--steps K] [--warmup W] [--config c5|c2|c3|c4|gw0|gamma|coulomb|minm_is|mt_coulomb] [--no-cpu]

--config c2 / c3: chi0 + eps^-1 only (FP64 and split precision) at
8x8x8 k, 16+48 bands, PB = 600 / 10x10x10 k, 32+200 bands, PB = 1200.
--config c4: a metal batch of 32 q (PB_q 300..1100, ragged per-k pair
windows) through the one-launch batched chi0, then eps^-1 per q.
--config gw0 / gamma: the GW0 iteration on stored mwm / the q = 0 extras.
--config coulomb: the v^(1/2) basis (Coul_setev) and minm = mmat conj(barcvm).
--config minm_is: the interstitial part of mmat (calc_minm_IS) over all k of a q.
--config mt_coulomb: the muffin-tin Coulomb block of a q (Ewald, jlam, MT_Coulomb).
Prints one JSON line.  --gpus N with N > 1 exits: this build runs one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12       # MI355X dense FP64 (vector and MFMA), FLOP/s
BASE_SEED_GW0 = 20261209  # synthetic mwm of the GW0 workload
PEAK_FP16 = 2.5e15        # MI355X dense FP16 MFMA, FLOP/s
PEAK_HBM = 8.0e12         # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5")
    ap.add_argument("--nq", type=int, default=None,
                    help="c2: q-points; c5: the first nq of its q-list (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="c4: time the split-precision batch only (no FP64 path, no precision check; profiling runs)")
    ap.add_argument("--cpu-kslab", type=int, default=2)
    ap.add_argument("--omega-rank", default="auto",
                    help="c5 split-precision chi0: 'auto' (the q's own frequency basis, rank from its singular "
                         "values), an integer rank, or 'off' (frequency by frequency)")
    ap.add_argument("--w-rank", default="auto",
                    help="c5 split-precision MWM: W in its own frequency basis ('auto', an integer rank) or 'off'")
    ap.add_argument("--uncompressed-steps", type=int, default=3,
                    help="c5: split-precision q steps timed frequency by frequency for comparison (0: none)")
    ap.add_argument("--save-sigma", default=None,
                    help="c5: write the FP64 and split-precision Sigma_c sums and the Sigma_c energies to this .npz")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 1 GiB device copy (scale for HBM byte counters)")
    ap.add_argument("--fp64-kfrac", type=float, default=None,
                    help="time the FP64 contraction on this fraction of k and extrapolate (large configs)")
    return ap.parse_args()


class Workload:
    def __init__(self, name, nq, mixed, fp64_kfrac):
        import torch

        from pygw_amd import mesh, synth

        self.base = base = synth.CONFIGS[name]
        div = base.div
        nq = nq or (8 if name == "c2" else 1)
        self.qs = [((i * div[0]) // max(nq, 1) if nq > 1 else 1, 0, 0) for i in range(nq)]
        self.bands = synth.band_energies(base)
        self.om, self.wom = synth.omega_mesh(base)
        self.tetc = mesh.tetra_corners(div)
        kqid = mesh.k_minus_q(div)
        self.items = []
        nk = base.nk
        self.nk64 = nk if fp64_kfrac is None else max(1, int(round(nk * fp64_kfrac)))
        for q in self.qs:
            cfg = synth.PolConfig(**{**base.__dict__, "q": q})
            kq = kqid[:, synth.q_index(cfg)]
            M64 = synth.device_matrix_elements(cfg, self.bands, kq, nkc=self.nk64, complex128=True)
            M32 = synth.device_matrix_elements(cfg, self.bands, kq, complex128=False) if mixed else None
            self.items.append((cfg, kq, M64, M32))
        self.orank, self.rank_used = None, None
        torch.cuda.synchronize()


def step(P, wl, it, precision, ev=None, keep=False):
    import torch
    s = torch.cuda.current_stream()
    cfg, kq, M64, M32 = it
    b = wl.bands
    rec = lambda i: ev[i].record(s) if ev is not None else None  # noqa: E731
    rec(0)
    kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, wl.om, wl.tetc, layout="native", check=False)
    rec(1)
    if precision == "fp64":
        eps = P.chi0_accumulate(M64, kcw)
    elif precision == "mixed" and wl.orank is not None:      # in the q's own frequency basis
        eps, wl.rank_used, _ = P.chi0_compressed(M32, kcw, precision="mixed",
                                                 rank=None if wl.orank == "auto" else int(wl.orank))
    else:
        eps = P.chi0_accumulate(M32, kcw, precision="mixed")
    rec(2)
    chi = eps.clone() if keep else None
    P.eps_inverse(eps, check=False)
    rec(3)
    return chi, eps


def timed(P, wl, precision, steps, warmup):
    import torch
    n = len(wl.items)
    for i in range(warmup):
        step(P, wl, wl.items[i % n], precision)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        step(P, wl, wl.items[i % n], precision, evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = np.array([[e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]), e[2].elapsed_time(e[3])] for e in evs]) / 1e3
    return wall / steps, st.mean(axis=0)


def rel_errors(P, wl):
    """Per-frequency relative Frobenius error of chi0 and eps^-1, mixed vs FP64
    (first q).  Both contractions see the same complex64-rounded M."""
    import torch
    cfg, kq, M64, M32 = wl.items[0]
    nk = M64.shape[0]
    it64 = (cfg, kq, M32[:nk].to(torch.complex128), M32)
    c64, e64 = step(P, wl, it64, "fp64", keep=True)
    itm = (cfg, kq, M64, M32[:nk].contiguous())
    cmx, emx = step(P, wl, itm, "mixed", keep=True)
    torch.cuda.synchronize()

    def rel(a, b):
        d = torch.linalg.matrix_norm((a - b).permute(2, 0, 1))
        return (d / torch.linalg.matrix_norm(b.permute(2, 0, 1))).cpu().numpy()
    one = torch.eye(c64.shape[0], dtype=c64.dtype, device=c64.device)[:, :, None]
    r_chi = rel(cmx, c64)
    r_inv = rel(emx + one, e64 + one)
    return {"chi0_max": float(r_chi.max()), "chi0_median": float(np.median(r_chi)),
            "eps_inv_max": float(r_inv.max()), "eps_inv_median": float(np.median(r_inv)),
            "k_points_compared": int(nk)}


def cpu_baseline(wl, kslab):
    """The CPU path of one q on this host: the C restatement of
    bz_calcqdepw_par2 (OpenMP) in full, plus the calc_eps numpy loop (per
    (k, w) ZGEMM, then inv per w) on a k-slab and a few frequencies,
    extrapolated linearly in N_k and n_om."""
    from oracle import chain
    from pygw_amd import synth

    cfg, kq, _, _ = wl.items[0]
    b, om, tetc = wl.bands, wl.om, wl.tetc
    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    kcw = port_weights(b, kq, om, tetc, nthr)
    t_w = time.perf_counter() - t
    M = synth.matrix_elements(cfg, b, kq, ks=range(kslab))
    no, ne = b.n_o, b.n_e
    t = time.perf_counter()
    eps = chain.calc_eps_accumulate(((i, M[i]) for i in range(kslab)), lambda ik: kcw[:, :, ik, :].reshape(no * ne, -1))
    t_chi = (time.perf_counter() - t) * cfg.nk / kslab
    nw = 4
    t = time.perf_counter()
    chain.eps_inverse(eps[:, :, :nw])
    t_inv = (time.perf_counter() - t) * cfg.n_om / nw
    return {"value": t_w + t_chi + t_inv, "unit": "s/q", "cores": nthr, "kind": "port", **host_info(),
            "sample": (f"one q of {cfg.name}: weights in full (C restatement of bz_calcqdepw_par2, OpenMP, {nthr} "
                       f"threads); chi0 = calc_eps numpy ZGEMM loop on {kslab} of {cfg.nk} k, extrapolated "
                       f"x{cfg.nk / kslab:.0f}; inverse on {nw} of {cfg.n_om} frequencies, extrapolated"),
            "stages_s": {"weights": t_w, "chi0_extrapolated": t_chi, "inverse_extrapolated": t_inv}}


# the 8 distinct q-points (integer coordinates on the 8x8x8 k-mesh, none at
# Gamma) of the c5 q-list: Sigma_c is summed over them, each once, with
# wkq = 1/8 (pygw.py:3312,3706)
C5_QLIST = ((1, 0, 0), (2, 0, 0), (3, 0, 0), (4, 0, 0), (1, 1, 0), (2, 1, 0), (2, 2, 0), (1, 1, 1))


def host_info():
    """CPU model and the host BLAS numpy calls (for the CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        b = np.__config__.CONFIG["Build Dependencies"]["blas"]
        blas = f"{b.get('name')} {b.get('version')}"
    except (AttributeError, KeyError, TypeError):
        blas = "unknown"
    return {"cpu_model": model, "logical_cpus_visible": os.cpu_count(), "blas": blas}


def port_weights(b, kq, om, tetc, nthr):
    """The C restatement of bz_calcqdepw_par2 (oracle/tetra_oracle.c, OpenMP
    over band pairs per tetrahedron as for_tetrahedra.f90:440)."""
    from oracle import tetra
    kcw, _, _ = tetra.bz_calcqdepw(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, nthreads=nthr)
    return kcw


class ChainWorkload:
    """Full chain at C5: 8x8x8 k-mesh, 20+120 bands, PB=900, 64 Matsubara
    points on PyGW's tan mesh (precise mesh x20 for the spline); Sigma_c of
    the 8 bands around the gap at the 64 k of the 4x4x4 sub-lattice, summed
    over the 8 q-points of C5_QLIST."""

    def __init__(self, nq=None):
        from pygw_amd import mesh, synth

        self.base = base = synth.CONFIGS["c5"]
        self.b = b = synth.band_energies(base)
        fm = mesh.FrequencyMesh(base.n_om, base.wmin_ev / mesh.H2EV, base.wmax_ev / mesh.H2EV, iop_multiple=20)
        self.om, self.wo, self.omp, self.wop = fm.omega, fm.womeg, fm.omega_precise, fm.womeg_precise
        self.tetc = mesh.tetra_corners(base.div)
        self.kqid = mesh.k_minus_q(base.div)
        div = np.array(base.div)
        pts = np.array([[i, j, k] for i in range(div[0]) for j in range(div[1]) for k in range(div[2])])
        self.klist = np.where((pts % 2 == 0).all(1))[0]          # 4x4x4 sub-lattice, 64 k
        self.nbgw = 8
        self.ibgw = b.n_o - self.nbgw // 2                        # 4 highest occupied + 4 lowest empty
        self.nb2 = b.enk.shape[0]
        self.qlist = C5_QLIST[:nq] if nq else C5_QLIST
        self.wkq = 1.0 / len(self.qlist)

    def loop(self, prec, omega_rank=None, w_rank=None):
        from pygw_amd.qloop import SelfEnergyQLoop
        b = self.b
        return SelfEnergyQLoop(b.enk, 0.0, b.ncg, b.nomx, b.numin, self.om, self.wo, self.tetc, self.base.pb,
                               len(self.qlist), len(self.klist), self.nbgw, self.nb2, omega_p=self.omp,
                               womeg_p=self.wop, precision=prec, omega_rank=omega_rank, w_rank=w_rank)

    def q_inputs(self, q, M=None, Ms=None):
        """Device inputs of one q, complex64 (the split-precision operand;
        the FP64 path takes the same values widened): kqid column, M of the
        polarization pairs, the self-energy rows and the k-q energies.  M /
        Ms: complex64 buffers to write into (the loop's input buffers), as
        PyGW's calc_minm produces M for the q just before calc_eps."""
        import torch

        from pygw_amd import synth
        cfg = synth.PolConfig(**{**self.base.__dict__, "q": q})
        kq = self.kqid[:, synth.q_index(cfg)]
        b, kl = self.b, self.klist
        M = synth.device_matrix_elements(cfg, b, kq, complex128=False, out=M)
        e2 = np.asfortranarray(np.hstack([b.enk[:, kl], b.enk[:, kq[kl]]]))
        b2 = synth.Bands(e2, b.ncg, b.nomx, b.numin)
        kq2 = np.arange(len(kl), 2 * len(kl), dtype=np.int32)
        kq2 = np.concatenate([kq2, kq2])
        Ms = synth.device_matrix_elements(cfg, b2, kq2, nkc=len(kl), complex128=False, out=Ms,
                                          rows=(self.ibgw, self.nbgw, 0, self.nb2), seed=cfg.seed + 99)
        enk_q = torch.as_tensor(np.ascontiguousarray(b.enk[:, kq[kl]].T), device="cuda")
        return kq, M, Ms, enk_q

    def ene(self):
        """Kohn-Sham energies of the Sigma_c states, (k, band) order."""
        b = self.b
        return np.ascontiguousarray(b.enk[self.ibgw:self.ibgw + self.nbgw][:, self.klist].T).reshape(-1)


def chain_cpu_baseline(wl, kslab=2, nw=4, nconv=1):
    """The CPU path of one q, timed on this host: the C restatement of
    bz_calcqdepw_par2 (OpenMP) in full, plus the numpy restatements of
    calc_eps (per-(k, w) ZGEMM through the host BLAS) on `kslab` k-points,
    of the inversion on `nw` frequencies, of MWM on `kslab` self-energy
    k-points and of the spline + convolution on `nconv` of them, each
    extrapolated linearly to the full q."""
    from oracle import chain
    from pygw_amd import synth

    base, b, om = wl.base, wl.b, wl.om
    cfg = synth.PolConfig(**{**base.__dict__, "q": wl.qlist[0]})
    kq = wl.kqid[:, synth.q_index(cfg)]
    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    kcw = port_weights(b, kq, om, wl.tetc, nthr)
    t_w = time.perf_counter() - t
    M = synth.matrix_elements(cfg, b, kq, ks=range(kslab))
    no, ne = b.n_o, b.n_e
    t = time.perf_counter()
    eps = chain.calc_eps_accumulate(((i, M[i]) for i in range(kslab)),
                                    lambda ik: kcw[:, :, ik, :].reshape(no * ne, -1))
    t_chi = (time.perf_counter() - t) * base.nk / kslab
    t = time.perf_counter()
    w = chain.eps_inverse(eps[:, :, :nw])
    t_inv = (time.perf_counter() - t) * base.n_om / nw
    W = np.concatenate([w] * (base.n_om // nw + 1), axis=2)[:, :, :base.n_om]
    rng = np.random.default_rng(3)
    rows = wl.nbgw * wl.nb2
    Ms = (rng.normal(size=(kslab, rows, base.pb)) + 1j * rng.normal(size=(kslab, rows, base.pb))) / np.sqrt(base.pb)
    t = time.perf_counter()
    mw = [chain.calc_mwm(Ms[i], W, wl.wkq) for i in range(kslab)]
    t_mwm = (time.perf_counter() - t) * len(wl.klist) / kslab
    t = time.perf_counter()
    for i in range(min(nconv, kslab)):
        chain.selfc_spline(b.enk[:, i], mw[i].reshape(base.n_om, wl.nbgw, wl.nb2), om, wl.omp, wl.wop)
    t_conv = (time.perf_counter() - t) * len(wl.klist) / min(nconv, kslab)
    return {"value": t_w + t_chi + t_inv + t_mwm + t_conv, "unit": "s/q", "cores": nthr, "kind": "port",
            **host_info(),
            "sample": (f"one q of c5: weights in full (C restatement of bz_calcqdepw_par2, OpenMP, {nthr} threads); "
                       f"calc_eps numpy ZGEMM loop on {kslab} of {base.nk} k (x{base.nk // kslab}); inverse on "
                       f"{nw} of {base.n_om} frequencies; MWM on {kslab} and spline+convolution on "
                       f"{min(nconv, kslab)} of {len(wl.klist)} self-energy k-points; all extrapolated linearly"),
            "stages_s": {"weights": t_w, "chi0_extrapolated": t_chi, "inverse_extrapolated": t_inv,
                         "mwm_extrapolated": t_mwm, "selfc_convolution_extrapolated": t_conv}}


def load_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload`, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        e = d[workload][kernel]
        return float(e["read_bytes"]) + float(e["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def chi0_launches(pb, npair, nom, nk):
    """Split-precision contraction launches per call: the library walks k in
    slabs of fragment planes of at most 1 GiB (its default workspace)."""
    import ctypes as C
    from pygw_amd import _lib
    s1, s2 = C.c_size_t(0), C.c_size_t(0)
    _lib.call("pygw_chi0_workspace", pb, npair, nom, 1, 1, C.byref(s1))
    _lib.call("pygw_chi0_workspace", pb, npair, nom, 2, 1, C.byref(s2))
    per_k = s2.value - s1.value
    ks = max(1, min(nk, (1 << 30) // per_k))
    return -(-nk // ks)


def _rel_per_w(a, b_, one=None):
    """Relative Frobenius error per frequency of Fortran [PB, PB, n_om]
    tensors (with one: of a + 1 against b + 1)."""
    import torch
    if one is not None:
        a, b_ = a + one, b_ + one
    d = torch.linalg.matrix_norm((a - b_).permute(2, 0, 1))
    return (d / torch.linalg.matrix_norm(b_.permute(2, 0, 1))).cpu().numpy()


def pade_conditioning(P, sig, om, ene, nsamp=4, noise=1e-7, seed=5, npade=32):
    """Per-state sensitivity of the Pade quasiparticle energy: the largest
    shift of E_qp over `nsamp` draws of complex Gaussian noise of relative
    size `noise` on Sigma_c(i w) (Hartree)."""
    import torch
    rng = np.random.default_rng(seed)
    base = P.pade_quasiparticles(torch.as_tensor(sig), om, ene, npade=npade).cpu().numpy()[:, 5]
    out = np.zeros_like(base)
    for _ in range(nsamp):
        z = sig * (noise * (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) / np.sqrt(2))
        qp = P.pade_quasiparticles(torch.as_tensor(sig + z), om, ene, npade=npade).cpu().numpy()[:, 5]
        out = np.maximum(out, np.abs(qp - base))
    return out


def qp_report(P, sig64, sigmx, om, ene, npade):
    """Quasiparticle energies of the FP64 and split-precision Sigma_c sums
    through an npade-point Thiele continued fraction, their difference
    beside the Pade's own sensitivity (1e-7 relative noise), and the
    difference relative to the quasiparticle correction E_qp - e."""
    import torch

    from pygw_amd import mesh
    nom = sig64.shape[-1]
    s64, smx = sig64.reshape(-1, nom), sigmx.reshape(-1, nom)
    q64 = P.pade_quasiparticles(torch.as_tensor(s64), om, ene, npade=npade).cpu().numpy()
    qmx = P.pade_quasiparticles(torch.as_tensor(smx), om, ene, npade=npade).cpu().numpy()
    d = np.abs(qmx[:, 5] - q64[:, 5])
    sens = pade_conditioning(P, s64, om, ene, npade=npade)
    shift = np.abs(q64[:, 5] - ene)
    ms = lambda x: {"median": float(np.median(x)), "max": float(np.max(x))}  # noqa: E731
    well = sens <= np.median(sens)
    h = mesh.H2EV * 1e3
    return {"pade_points": min(npade, nom), "states": int(d.size),
            "mixed_vs_fp64_abs_diff_meV": ms(d * h),
            "mixed_vs_fp64_diff_relative_to_qp_correction": ms(d / np.maximum(shift, 1e-300)),
            "pade_sensitivity_meV_for_1e-7_noise": ms(sens * h),
            "better_conditioned_half": {"median_diff_meV": float(np.median(d[well]) * h),
                                        "max_diff_meV": float(d[well].max() * h),
                                        "median_sensitivity_meV": float(np.median(sens[well]) * h)},
            "states_with_diff_below_sensitivity": int(np.sum(d <= sens)),
            "qp_correction_eV": ms(shift * mesh.H2EV)}


def run_chain(args):
    """The default workload: weights -> chi0 -> eps^-1 -> MWM -> Sigma_c at
    C5 through the package's q loop (pygw_amd.qloop.SelfEnergyQLoop), FP64
    and split precision.

    1. Precision pass: every q of the 8-point q-list once on both paths (the
       same complex64 M, widened for FP64): per q and frequency the relative
       Frobenius errors of chi0, eps^-1 and the q's Sigma_c term; Sigma_c
       summed over the q-list (each q once, wkq = 1/8) on both paths; Pade
       quasiparticle energies of both sums with each state's Pade
       sensitivity.  FP64 seconds per q are the stage times of this pass.
    2. Timed pass: `warmup` + `steps` split-precision q steps cycling over the
       q-list, inputs resident in HBM.  By default the split-precision chi0
       is contracted in the q's own frequency basis (--omega-rank auto: the
       leading eigenvectors of the weights' Gram matrix, as PyGW's SVD mode
       contracts in its analytic basis) and expanded to the 64 frequencies
       before the inversion; the FP64 path stays frequency by frequency, so
       the precision pass measures the basis truncation and the split
       together.  A few steps frequency by frequency are timed for
       comparison.
    value = split-precision seconds per q when every q's chi0 and eps^-1 are
    within 1e-6 and its Sigma_c within 1e-5 of FP64 (relative Frobenius,
    every frequency), else the FP64 seconds per q."""
    import torch

    from pygw_amd import mesh
    import ctypes as C

    from pygw_amd import _lib
    from pygw_amd import polarization as P

    wl = ChainWorkload(args.nq)
    base, b = wl.base, wl.b
    nql = len(wl.qlist)
    stage_names = ["polarization_weights", "chi0", "eps_inverse", "mwm", "selfc_convolution"]
    orank = None if args.omega_rank == "off" else (args.omega_rank if args.omega_rank == "auto"
                                                   else int(args.omega_rank))
    wrank = None if args.w_rank == "off" else (args.w_rank if args.w_rank == "auto" else int(args.w_rank))
    L64, Lmx = wl.loop("fp64"), wl.loop("mixed", orank, wrank)
    ranks = {"precision_pass": [], "timed": [], "w_precision_pass": [], "w_timed": []}
    # the mixed loop's input buffers take each q's M as it is produced
    Mbuf = torch.empty((base.nk, b.n_o * b.n_e, base.pb), dtype=torch.complex64, device="cuda")
    Msbuf = torch.empty((len(wl.klist), wl.nbgw * wl.nb2, base.pb), dtype=torch.complex64, device="cuda")
    one = torch.eye(base.pb, dtype=torch.complex128, device="cuda")[:, :, None]
    ev = lambda: [torch.cuda.Event(enable_timing=True) for _ in range(6)]  # noqa: E731
    t64, errs = [], {"chi0": [], "eps_inv": [], "sigma_c_q": []}
    for iq, q in enumerate(wl.qlist):
        kq, M, Ms, eq = wl.q_inputs(q, Mbuf, Msbuf)
        L64.set_q(kq, M, Ms, eq)                        # the complex64 values, widened
        Lmx.set_q(kq, enk_kq=eq)
        torch.cuda.synchronize()
        e64 = ev()
        L64.run_q(e64, keep_chi0=True)
        L64.accumulate()
        Lmx.run_q(keep_chi0=True, M=M, M_selfe=Ms)
        Lmx.accumulate()
        torch.cuda.synchronize()
        t64.append([e64[i].elapsed_time(e64[i + 1]) / 1e3 for i in range(5)])
        errs["chi0"].append(_rel_per_w(Lmx.chi, L64.chi))
        errs["eps_inv"].append(_rel_per_w(Lmx.eps, L64.eps, one))
        d = (Lmx.sc - L64.sc).abs().pow(2).sum(dim=(0, 1)).sqrt()
        errs["sigma_c_q"].append((d / L64.sc.abs().pow(2).sum(dim=(0, 1)).sqrt()).cpu().numpy())
        ranks["precision_pass"].append(Lmx.rank_used)
        ranks["w_precision_pass"].append(Lmx.w_rank_used)
        print(f"[bench] precision pass q={q}: fp64 {sum(t64[-1]):.3f} s; basis rank {Lmx.rank_used}, "
              f"W {Lmx.w_rank_used}; "
              f"chi0 {errs['chi0'][-1].max():.2e} "
              f"eps^-1 {errs['eps_inv'][-1].max():.2e} Sigma_c {errs['sigma_c_q'][-1].max():.2e}",
              file=sys.stderr, flush=True)
    sig64, sigmx = L64.sigc.cpu().numpy(), Lmx.sigc.cpu().numpy()
    if args.save_sigma:
        np.savez(args.save_sigma, sig64=sig64, sigmx=sigmx, omega=wl.om, ene=wl.ene())
    L64.chi = Lmx.chi = None
    del L64
    torch.cuda.empty_cache()
    emax = {k: float(np.max(v)) for k, v in errs.items()}
    gates = {"chi0": 1e-6, "eps_inv": 1e-6, "sigma_c_q": 1e-5}
    ok = all(emax[k] <= gates[k] for k in gates)
    # quasiparticle energies of the two q sums and the Pade's own sensitivity:
    # PyGW's rule (the first min(32, n_om) points, mcommon.py:330-338) and,
    # better conditioned on this 64-point tan mesh, all n_om points
    ene = wl.ene()
    nom = base.n_om
    qp_ref = qp_report(P, sig64, sigmx, wl.om, ene, 32)
    qp_all = qp_report(P, sig64, sigmx, wl.om, ene, nom)
    rel_sig = np.linalg.norm(sigmx - sig64, axis=(0, 1)) / np.linalg.norm(sig64, axis=(0, 1))
    dsig = float(np.abs(sigmx - sig64).max()) * mesh.H2EV * 1e3
    sig_scale = float(np.median(np.abs(sig64))) * mesh.H2EV
    # timed pass: split precision, q steps cycling over the q-list
    Lmx.reset()
    s = torch.cuda.current_stream()
    times, walls = [], []
    kern_ms, kern_n = 0.0, 0         # the contraction kernel alone, HIP events around each launch
    for it in range(args.warmup + args.steps):
        kq, M, Ms, eq = wl.q_inputs(wl.qlist[it % nql], Mbuf, Msbuf)   # not timed: the q's M is produced
        Lmx.set_q(kq, enk_kq=eq)
        torch.cuda.synchronize()
        if it >= args.warmup and ok:
            _lib.call("pygw_chi0_launch_timer", 64)
        e = ev()
        tw = time.perf_counter()
        Lmx.run_q(e, M=M, M_selfe=Ms)
        Lmx.accumulate()
        torch.cuda.synchronize()
        wall = time.perf_counter() - tw
        if it >= args.warmup and ok:
            tms, tn = C.c_double(0.0), C.c_int(0)
            _lib.call("pygw_chi0_launch_time", C.byref(tms), C.byref(tn))
            _lib.call("pygw_chi0_launch_timer", 0)
            kern_ms, kern_n = kern_ms + tms.value, kern_n + tn.value
        if it >= args.warmup:
            times.append([e[i].elapsed_time(e[i + 1]) / 1e3 for i in range(5)])
            walls.append(wall)
            ranks["timed"].append(Lmx.rank_used)
            ranks["w_timed"].append(Lmx.w_rank_used)
        print(f"[bench] mixed q={wl.qlist[it % nql]} {wall:.3f} s{' (warm-up)' if it < args.warmup else ''}",
              file=sys.stderr, flush=True)
    # for comparison: the same split-precision steps frequency by frequency
    unc = None
    if (orank is not None or wrank is not None) and args.uncompressed_steps > 0:
        Lmx.omega_rank = Lmx.w_rank = None
        tu = []
        for it in range(1 + args.uncompressed_steps):
            kq, M, Ms, eq = wl.q_inputs(wl.qlist[it % nql], Mbuf, Msbuf)
            Lmx.set_q(kq, enk_kq=eq)
            torch.cuda.synchronize()
            e = ev()
            Lmx.run_q(e, M=M, M_selfe=Ms)
            torch.cuda.synchronize()
            if it >= 1:
                tu.append([e[i].elapsed_time(e[i + 1]) / 1e3 for i in range(5)])
            print(f"[bench] mixed, frequency by frequency, q={wl.qlist[it % nql]}{' (warm-up)' if it < 1 else ''}",
                  file=sys.stderr, flush=True)
        Lmx.omega_rank, Lmx.w_rank = orank, wrank
        tu = np.array(tu).mean(0)
        unc = {"s_per_q": float(tu.sum()), "stages_s": dict(zip(stage_names, map(float, tu))),
               "steps": args.uncompressed_steps}
    del s
    tmx, t64m = np.array(times).mean(0), np.array(t64).mean(0)
    st_mx, st_64 = dict(zip(stage_names, map(float, tmx))), dict(zip(stage_names, map(float, t64m)))
    s_mx, s_64 = float(tmx.sum()), float(t64m.sum())
    nk, no, ne, pb = base.nk, b.n_o, b.n_e, base.pb
    f_chi = 4.0 * pb * pb * nk * no * ne * nom
    rmean = float(np.mean(ranks["timed"])) if orank is not None else float(nom)
    f_exec = f_chi * rmean / nom if ok else f_chi          # the contraction run: rank basis functions
    t_chi = st_mx["chi0"] if ok else st_64["chi0"]
    kern = "pygw_chi0_mixed" if ok else "pygw_chi0_f64"
    nl = chi0_launches(pb, no * ne, int(round(rmean)) if ok else nom, nk) if ok else 1
    t_kern = kern_ms / 1e3 / kern_n if kern_n else t_chi / nl
    peak = PEAK_FP16 if ok else PEAK_FP64
    # weights: HBM roofline on the algorithmic bytes (kcw written once, enk
    # and the index arrays read once) and the FP64-VALU count per (small
    # tetrahedron, frequency) of the reference's closed forms
    w_bytes = 8.0 * no * ne * nk * nom + 8.0 * b.enk.shape[0] * nk + 16.0 * base.ntet + 4.0 * nk
    w_flops = base.ntet * no * ne * nom * (200.0 + 8 * 20.0)
    t_w = st_mx["polarization_weights"]
    rec = {
        "metric": "chi0+eps^-1+Sigma_c seconds per q-point, mixed vs FP64 GPU at <=1e-6 rel err; % roofline",
        "value": s_mx if ok else s_64, "unit": "s/q", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (s_mx if ok else s_64) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp16x3-split/fp32-acc chi0+MWM, f64 weights/inverse/Sigma_c"
                  + (", chi0 in the q's frequency basis" if orank is not None else "")
                  + (", MWM in W's frequency basis" if wrank is not None else "")) if ok else "f64",
        "data": "synthetic",
        "config": {"workload": f"c5: 8x8x8 k-mesh (3072 tetrahedra), {no}+{ne} bands, PB={pb}, {nom} Matsubara "
                               f"points (tan mesh, precise mesh x20 = {wl.omp.size}), Sigma_c of {wl.nbgw} bands at "
                               f"{len(wl.klist)} k summed over a q-list of {nql} (wkq = 1/{nql}); one q per step",
                   "precision": "mixed" if ok else "fp64", "q_list": [list(q) for q in wl.qlist],
                   "timed_steps_cycle_q_list": True,
                   "chi0_frequency_basis": {"mode": "off" if orank is None else str(orank),
                                            "ranks_precision_pass": ranks["precision_pass"],
                                            "ranks_timed": ranks["timed"], "n_om": nom},
                   "mwm_frequency_basis": {"mode": "off" if wrank is None else str(wrank),
                                           "ranks_precision_pass": ranks["w_precision_pass"],
                                           "ranks_timed": ranks["w_timed"], "n_om": nom}},
        "stages_s": st_mx if ok else st_64,
        "roofline": {"bound": "mfma", "kernel": kern, "achieved": f_exec / nl / t_kern / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": f_exec / nl / t_kern / peak, "traffic": load_traffic("c5", kern),
                     "launches_per_q": nl, "flops_per_launch": f_exec / nl, "time_per_launch_s": t_kern,
                     "time_per_launch_source": ("HIP events around each launch of the kernel in the timed steps "
                                                f"({kern_n} launches)" if kern_n else "chi0 stage time / launches"),
                     "chi0_stage_time_per_launch_s": t_chi / nl,
                     "executed_flops_per_launch": (3 * f_exec if ok else f_exec) / nl,
                     "basis_functions": rmean if ok else nom,
                     "frequency_by_frequency_equivalent_tflops": f_chi / t_chi / 1e12},
        "weights_roofline": {"bound": "hbm", "kernels": "pygw_tw_* (per q)", "achieved": w_bytes / t_w / 1e9,
                             "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": w_bytes / t_w / PEAK_HBM,
                             "algorithmic_bytes": w_bytes, "traffic": load_traffic("c5", "pygw_tw_*"),
                             "fp64_valu_tflops": w_flops / t_w / 1e12,
                             "fp64_valu_frac": w_flops / t_w / PEAK_FP64,
                             "fp64_valu_count": "200 FP64 ops + 8 transcendentals (20 ops each) per (tetrahedron, "
                                                "pair, frequency)"},
        "fp64": {"s_per_q": s_64, "stages_s": st_64, "chi0_tflops": f_chi / st_64["chi0"] / 1e12,
                 "chi0_frac_fp64_peak": f_chi / st_64["chi0"] / PEAK_FP64},
        "mixed_speedup_vs_fp64": s_64 / s_mx,
        "mixed_frequency_by_frequency": unc,
        "rel_frobenius_error_mixed_vs_fp64": {
            "per_q_max": {k: [float(np.max(x)) for x in v] for k, v in errs.items()},
            "chi0_max": emax["chi0"], "chi0_median": float(np.median(np.concatenate(errs["chi0"]))),
            "eps_inv_max": emax["eps_inv"], "eps_inv_median": float(np.median(np.concatenate(errs["eps_inv"]))),
            "sigma_c_q_max": emax["sigma_c_q"], "sigma_c_summed_over_q_max": float(rel_sig.max()),
            "sigma_c_summed_abs_max_meV": dsig, "gates": gates, "passed": ok},
        "quasiparticles": {"pyGW_rule_first_32_points": qp_ref, "all_points": qp_all,
                           "sigma_c_median_abs_eV": sig_scale,
                           "note": "the synthetic Sigma_c is ~1e4-1e5 eV (v = 4 pi/|q+G|^2 taken literally in M): "
                                   "meV differences scale with it; the relative figures carry over"},
    }
    print("[bench] precision report done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        rec["cpu_baseline"] = chain_cpu_baseline(wl, args.cpu_kslab)
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)
    print(json.dumps(rec))


def batch_cpu_baseline(batch, Ks, om, tetc, kqid, kslab=16, nw=2):
    """One q of the batch (PB closest to the batch mean): the C restatement
    of the weights (OpenMP) in full, the numpy calc_eps restatement on
    the windows of `kslab` k-points and the inversion on `nw` frequencies;
    extrapolated linearly in k / frequency, then to the batch mean by the
    flop ratio sum_q PB_q^2 K_q / (n_q PB_i^2 K_i) (chi0) and PB^3 (inverse)."""
    from oracle import chain
    from pygw_amd import synth

    b, cfg = batch.bands, batch.cfg
    pbs = np.asarray(batch.pb, dtype=float)
    i = int(np.argmin(np.abs(pbs - pbs.mean())))
    qc = batch.q_config(i)
    kq = kqid[:, batch.q_index[i]]
    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    kcw = port_weights(b, kq, om, tetc, nthr)
    kind = "port"
    t_w = time.perf_counter() - t
    win = batch.windows[i]
    ks = list(range(kslab))
    M = synth.matrix_elements(qc, b, kq, ks=ks)
    ne = b.n_e
    t = time.perf_counter()
    rows = [win[win[:, 0] == k, 1] for k in ks]     # pair p = i_o*n_e + i_e
    eps = chain.calc_eps_accumulate(((k, M[k][rows[k]]) for k in ks),
                                    lambda k: kcw[rows[k] // ne, rows[k] % ne, k, :])
    k_in = sum(r.size for r in rows)
    t_chi = (time.perf_counter() - t) * Ks[i] / k_in
    t = time.perf_counter()
    chain.eps_inverse(eps[:, :, :nw])
    t_inv = (time.perf_counter() - t) * cfg.n_om / nw
    f_ratio = float(np.mean(pbs ** 2 * np.asarray(Ks))) / (pbs[i] ** 2 * Ks[i])
    i_ratio = float(np.mean(pbs ** 3)) / pbs[i] ** 3
    val = t_w + t_chi * f_ratio + t_inv * i_ratio
    return {"value": val, "unit": "s/q", "cores": nthr, "kind": kind, **host_info(),
            "sample": (f"q {i} of the batch (PB={int(pbs[i])}, K={Ks[i]} window pairs): weights in full "
                       f"(C restatement of bz_calcqdepw_par2, OpenMP, {nthr} threads); "
                       f"calc_eps numpy ZGEMM loop on the windows of {kslab} of {cfg.nk} k ({k_in} pairs, "
                       f"x{Ks[i] / k_in:.1f}); inverse on {nw} of {cfg.n_om} frequencies; scaled to the batch "
                       f"mean by PB^2 K (x{f_ratio:.3f}) and PB^3 (x{i_ratio:.3f})"),
            "stages_s": {"weights": t_w, "chi0_extrapolated": t_chi * f_ratio, "inverse_extrapolated": t_inv * i_ratio}}


def run_batch(args):
    """C4: n_q (default 32) q-points of a metal (3 bands crossing E_F) on
    the 8x8x8 k-mesh, PB_q ~ U{300..1100}, per-(q, k) windows of 40-260
    pairs, 48 frequencies.  One step = the whole batch: the weights of every
    q gathered to its window pairs, then one split-precision chi0
    launch over all q and eps^-1 of every q (both captured in HIP graphs,
    fold interval autotuned).  The FP64 path runs the same q one after
    another (chi0_accumulate on the complex64-rounded M, then eps^-1).
    value = mixed seconds per q (batch time / n_q) when every chi0 is within
    1e-6 of FP64 (relative Frobenius, every q and frequency)."""
    import torch

    from pygw_amd import mesh, synth
    from pygw_amd import polarization as P

    cfg = synth.CONFIGS["c4"]
    nq = args.nq or 32
    batch = synth.metal_batch(cfg, n_q=nq)
    b = batch.bands
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kqid = mesh.k_minus_q(cfg.div)
    nom, npair = cfg.n_om, b.n_o * b.n_e
    Ks = [int(w.shape[0]) for w in batch.windows]
    pbs = [int(p) for p in batch.pb]
    kqs = [kqid[:, batch.q_index[i]] for i in range(nq)]
    dev = "cuda"
    ar = torch.arange(nom, dtype=torch.int64, device=dev)[:, None]
    gidx = []  # flat index of (k, w, p) in the native weights [nk, nom, n_o*n_e], shaped [nom, K_q]
    for w in batch.windows:
        k = torch.as_tensor(w[:, 0], dtype=torch.int64, device=dev)[None, :]
        p = torch.as_tensor(w[:, 1], dtype=torch.int64, device=dev)[None, :]
        gidx.append(((k * nom + ar) * npair + p).reshape(-1))
    kcws = [torch.empty((nom, K), dtype=torch.float64, device=dev) for K in Ks]
    Ms = [synth.device_window_elements(batch, i) for i in range(nq)]
    print(f"[bench] c4 batch: {nq} q, PB {min(pbs)}..{max(pbs)}, K mean {np.mean(Ks):.0f}", file=sys.stderr,
          flush=True)

    def weights(i):
        kn = P.bz_calcqdepw_par2(b.enk, 0.0, kqs[i], b.ncg, b.nomx, b.numin, om, tetc, layout="native",
                                 check=False)
        torch.index_select(kn.view(-1), 0, gidx[i], out=kcws[i].view(-1))

    for i in range(nq):
        weights(i)
    orank = None if args.omega_rank == "off" else args.omega_rank
    B = P.Chi0Batch(Ms, kcws, inverse=True, omega_rank=orank)
    best, tune = B.autotune()
    g_chi = B.capture(best, inverse=False)
    g_inv = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g_inv):
        B.invert()
    torch.cuda.synchronize()
    print(f"[bench] autotune best={best} {tune}", file=sys.stderr, flush=True)
    s = torch.cuda.current_stream()
    times = []
    for it in range(args.warmup + args.steps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        torch.cuda.synchronize()
        ev[0].record(s)
        for i in range(nq):
            weights(i)
        ev[1].record(s)
        g_chi.replay()
        ev[2].record(s)
        g_inv.replay()
        ev[3].record(s)
        torch.cuda.synchronize()
        st = [ev[j].elapsed_time(ev[j + 1]) / 1e3 for j in range(3)]
        print(f"[bench] mixed batch step {it}: {sum(st):.3f} s{' (warm-up)' if it < args.warmup else ''}",
              file=sys.stderr, flush=True)
        if it >= args.warmup:
            times.append(st)
    tm = np.array(times).mean(0)
    if args.no_fp64:
        B.close()
        print(json.dumps({"metric": "c4 split-precision batch only (profiling run, precision unchecked)",
                          "s_per_q": float(tm.sum()) / nq, "stages_s_per_q": (tm / nq).tolist()}))
        return
    # FP64: the same q one after another on the complex64-rounded M
    t64 = np.zeros(3)
    for i in range(nq):
        M128 = Ms[i][None].to(torch.complex128)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        torch.cuda.synchronize()
        ev[0].record(s)
        weights(i)
        ev[1].record(s)
        e64 = P.chi0_accumulate(M128, kcws[i][None])
        ev[2].record(s)
        P.eps_inverse(e64, check=False)
        ev[3].record(s)
        torch.cuda.synchronize()
        t64 += [ev[j].elapsed_time(ev[j + 1]) / 1e3 for j in range(3)]
        del M128, e64
    print(f"[bench] fp64 per q {t64.sum() / nq:.4f} s", file=sys.stderr, flush=True)
    # precision: chi0 and eps^-1 of every q against FP64 on the same inputs
    g_chi.replay()
    torch.cuda.synchronize()
    r_chi, r_inv = [], []
    one = None
    for i in range(nq):
        ref64 = P.chi0_accumulate(Ms[i][None].to(torch.complex128), kcws[i][None])
        mx = B.eps[i].clone()
        d = torch.linalg.matrix_norm((mx - ref64).permute(2, 0, 1))
        r_chi.append((d / torch.linalg.matrix_norm(ref64.permute(2, 0, 1))).cpu().numpy())
        P.eps_inverse(ref64, check=False)
        P.eps_inverse(mx, check=False)
        one = torch.eye(pbs[i], dtype=torch.complex128, device=dev)[:, :, None]
        d = torch.linalg.matrix_norm((mx - ref64).permute(2, 0, 1))
        r_inv.append((d / torch.linalg.matrix_norm((ref64 + one).permute(2, 0, 1))).cpu().numpy())
        del ref64, mx, one
    r_chi, r_inv = np.concatenate(r_chi), np.concatenate(r_inv)
    ok = bool(r_chi.max() <= 1e-6)
    f_q = [4.0 * p * p * K * nom for p, K in zip(pbs, Ks)]
    f_full = float(sum(f_q))                        # frequency by frequency (the FP64 path)
    f_tot = f_full * (B.rank / nom)                 # the contraction run: B.rank basis functions per q
    s_mx, s_64 = float(tm.sum()) / nq, float(t64.sum()) / nq
    kern = "pygw_chi0_mixed_batch"
    rec = {
        "metric": "chi0+eps^-1 seconds per q-point, mixed vs FP64 GPU at <=1e-6 rel err; % roofline",
        "value": s_mx if ok else s_64, "unit": "s/q", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (tm.sum() if ok else t64.sum()) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16x3-split/fp32-acc chi0, f64 weights/inverse" if ok else "f64", "data": "synthetic",
        "config": {"workload": f"c4: metal ({b.n_o}+{b.n_e} bands, 3 crossing E_F), 8x8x8 k-mesh, {nq} q in one "
                               f"batch, PB_q {min(pbs)}..{max(pbs)} (U{{300..1100}}), per-k windows of 40-260 "
                               f"pairs (K_q mean {np.mean(Ks):.0f}), {nom} frequencies; one step = the batch",
                   "precision": "mixed" if ok else "fp64", "fold_schedule": best, "autotune_s": tune,
                   "chi0_frequency_basis": {"mode": "off" if orank is None else str(orank), "rank": B.rank,
                                            "n_om": nom}},
        "stages_s_per_q": {"polarization_weights": tm[0] / nq, "chi0": tm[1] / nq, "eps_inverse": tm[2] / nq},
        "roofline": {"bound": "mfma", "kernel": kern, "achieved": f_tot / tm[1] / 1e12, "peak": PEAK_FP16 / 1e12,
                     "unit": "TFLOP/s", "frac": f_tot / tm[1] / PEAK_FP16, "traffic": load_traffic("c4", kern),
                     "flops_per_launch": f_tot, "executed_flops_per_launch": 3 * f_tot,
                     "frequency_by_frequency_equivalent_tflops": f_full / tm[1] / 1e12,
                     "timed": "HIP events around the captured batch (bases, projections, contraction, mirrors, "
                              "expansions)"},
        "fp64": {"s_per_q": s_64, "stages_s_per_q": dict(zip(("polarization_weights", "chi0", "eps_inverse"),
                                                              map(float, t64 / nq))),
                 "chi0_tflops": f_full / t64[1] / 1e12, "chi0_frac_fp64_peak": f_full / t64[1] / PEAK_FP64},
        "mixed_speedup_vs_fp64": s_64 / s_mx,
        "rel_frobenius_error_mixed_vs_fp64": {"chi0_max": float(r_chi.max()), "chi0_median": float(np.median(r_chi)),
                                              "eps_inv_max": float(r_inv.max()),
                                              "eps_inv_median": float(np.median(r_inv))},
    }
    B.close()
    if not args.no_cpu:
        rec["cpu_baseline"] = batch_cpu_baseline(batch, Ks, om, tetc, kqid)
        print("[bench] cpu baseline done", file=sys.stderr, flush=True)
    print(json.dumps(rec))


def gw0_cpu_baseline(g, mwm_item, nb2):
    """PyGW's per-(q, k) Sigma_c of the GW0 loop (Compute_selfc_inside,
    spline branch, mcommon.py:126-144) restated on the host: scipy
    CubicSpline per (ie1, ie2) and the numpy convolution (fr_convolution2,
    fnc.f90:86-119, vectorised over ie2) per (ie1, iom), on one (q, k) item,
    extrapolated to all items of an iteration."""
    from scipy import interpolate

    from oracle import chain
    enk = g.ebnd[g.jrk[0, 0], :nb2]
    nom, nb1, _ = mwm_item.shape
    fa, fp = g.omega ** 2 + 1.0, g.omega_p ** 2 + 1.0
    t = time.perf_counter()
    for ie1 in range(nb1):
        mp = np.zeros((nb2, g.omega_p.size), dtype=complex)
        for ie2 in range(nb2):
            v = mwm_item[:, ie1, ie2]
            sr = interpolate.CubicSpline(g.omega, fa * v.real, bc_type=((2, 0.0), (1, 0.0)), extrapolate=True)
            si = interpolate.CubicSpline(g.omega, fa * v.imag, bc_type=((2, 0.0), (1, 0.0)), extrapolate=True)
            mp[ie2] = (sr(g.omega_p) + 1j * si(g.omega_p)) / fp
        for iom in range(nom):
            z = enk - 1j * g.omega[iom]                                   # [nb2]
            sc0 = ((mp - mwm_item[iom, ie1][:, None]) * g.womeg_p[None, :]
                   / (g.omega_p[None, :] ** 2 + (z ** 2)[:, None])).sum(1)
            m0 = mwm_item[iom, ie1]
            tail = np.where(np.abs(m0) > 1e-16, m0 / chain.PI_F * np.arctan(g.omega_p[-1] / z), 0.0)
            (z * sc0 / chain.PI_F + tail).sum()
    t1 = time.perf_counter() - t
    nitems = g.jrk.size
    return {"value": t1 * nitems, "unit": "s/iteration", "cores": 1, "kind": "port", **host_info(),
            "sample": (f"one (q, k) item of the {nitems} of an iteration: scipy CubicSpline per (ie1, ie2) and the "
                       f"numpy convolution per (ie1, iom), x{nitems}"),
            "item_s": t1}


def run_gw0(args):
    """One GW0 iteration on stored mwm (postprocess.py:281-301): Sigma_c of
    every (q, k) at the current bands, mixing, quasiparticles, Fermi level.
    4x4x4 mesh without symmetry (64 irreducible k), 64 q, 8 GW bands, nb2 =
    140, 64 Matsubara points, precise mesh x20; mwm resident in HBM."""
    import torch

    from pygw_amd import gw0, mesh, synth
    g = synth.gw0_system(div=(4, 4, 4), n_occ=20, n_emp=120, nbgw=8, n_core=0, n_om=64, iop_multiple=20,
                         mwm_array=False, seed=BASE_SEED_GW0)
    nq, nk = g.jrk.shape
    nom, nb1, nb2 = g.omega.size, 8, 140
    # synthetic mwm on the device (the pole form of synth.gw0_system)
    gen = torch.Generator(device="cuda").manual_seed(BASE_SEED_GW0)
    shp = (nq, nk, 1, nb1, nb2)
    a = torch.rand(shp, generator=gen, device="cuda", dtype=torch.float64) * (0.15 / (nb2 * nq))
    W = 0.3 + 1.2 * torch.rand(shp, generator=gen, device="cuda", dtype=torch.float64)
    b = a / 10.0 * (2 * torch.rand(shp, generator=gen, device="cuda", dtype=torch.float64) - 1)
    om = torch.as_tensor(g.omega, device="cuda").view(1, 1, nom, 1, 1)
    f = 2 * W / (om * om + W * W)
    mwm = torch.complex(-a * f, b * f).contiguous()
    del a, W, b, f
    torch.cuda.synchronize()
    print(f"[bench] gw0 mwm stack {tuple(mwm.shape)} {mwm.numel() * 16 / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    selfe = gw0.SelfEnergy(mwm, g.jrk, g.core, g.omega, g.omega_p, g.womeg_p)
    dos = gw0.TetraDos(g.atet, g.wtet)
    ibgw = g.ibgw
    bands = g.ebnd.copy()
    sig = selfe(bands)            # the G0W0 Sigma_c
    snew = torch.empty_like(sig)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def iteration(step_bands):
        ev[0].record()
        selfe(step_bands, out=snew)
        ev[1].record()
        gw0.mix(sig, snew, 0.5)
        eqp, _ = gw0.quasiparticles(step_bands[:, ibgw:ibgw + nb1], g.ebnd[:, ibgw:ibgw + nb1], sig, g.sigx, g.vxc,
                                    g.omega, -1, 1)
        ev[2].record()
        ef = gw0.calc_fermi(eqp, dos, g.nval - 2 * ibgw)[0]
        ev[3].record()
        torch.cuda.synchronize()
        return eqp, ef, [ev[0].elapsed_time(ev[1]) / 1e3, ev[1].elapsed_time(ev[2]) / 1e3,
                         ev[2].elapsed_time(ev[3]) / 1e3]

    for _ in range(max(1, args.warmup)):   # the first call also splines mwm onto the precise mesh
        eqp, ef, _ = iteration(bands)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stages = np.zeros(3)
    for _ in range(args.steps):
        eqp, ef, st = iteration(bands)
        stages += st
        bands[:, ibgw:ibgw + nb1] = eqp - ef
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    stages /= args.steps
    print(f"[bench] gw0 iteration {t:.3f} s", file=sys.stderr, flush=True)
    # convolution kernel: per (item, ie1, iom, ie2, precise point) one complex
    # term of fr_convolution2 (fnc.f90:86-119), counted as 20 FP64 flops
    nomp = g.omega_p.size
    terms = nq * nk * nb1 * nom * nb2 * nomp
    rec = {"metric": "seconds per GW0 iteration (Sigma_c of every (q, k) on stored mwm + Pade QP + Fermi level)",
           "value": t, "unit": "s/iteration", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"gw0: 4x4x4 k-mesh, every k irreducible (64), {nq} q, mwm [{nom}, {nb1}, {nb2}] "
                                  f"per (q, k) resident in HBM ({mwm.numel() * 16 / 1e9:.1f} GB), precise mesh "
                                  f"x20 = {nomp}, "
                                  + (f"precise-mesh splines of all q kept in HBM ({selfe.mwmp.numel() * 16 / 1e9:.0f} GB,"
                                     f" computed in the warm-up)" if selfe.cached else
                                     f"q chunks of {selfe.qc} (splines recomputed per chunk)")},
           "stages_s": {"selfc_all_q": stages[0], "mix_and_quasiparticles": stages[1], "fermi_level": stages[2]},
           "roofline": {"bound": "valu", "kernel": "pygw_selfc_conv", "achieved": terms * 20 / stages[0] / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": terms * 20 / stages[0] / PEAK_FP64,
                        "traffic": None, "flops_per_term": 20, "terms_per_iteration": terms,
                        "note": "FP64 vector work; time of the whole all-q stage (enk, spline, convolution, q sum)"}}
    if not args.no_cpu:
        item = mwm[0, 0].cpu().numpy()
        rec["cpu_baseline"] = gw0_cpu_baseline(g, item, nb2)
    print(json.dumps(rec))


def run_gamma(args):
    """The q = 0 step of the C5 chain with PyGW's head/wings terms: the wing
    accumulation beside chi0 (pygw.py:3192-3193), the inversion with the
    head/wings update (pygw.py:3202-3263) and the head/wing terms of MWM
    (pygw.py:3359-3364), timed beside the q != 0 stages they accompany
    (split-precision M; the q = 0 extras in FP64)."""
    import torch

    from pygw_amd import polarization as P
    wl = ChainWorkload()
    b, base = wl.b, wl.base
    kq, M, Ms, enk_q = wl.q_inputs((0, 0, 0))
    nk, npair = M.shape[0], M.shape[1]
    rng = np.random.default_rng(BASE_SEED_GW0 + 1)
    # synthetic momentum matrix elements / transition energies, and head
    pm = torch.as_tensor((rng.normal(size=(nk, npair)) + 1j * rng.normal(size=(nk, npair))) * 0.05, device="cuda")
    head0 = torch.as_tensor(2.0 + 3.0 / (1.0 + wl.om ** 2) + 0j, device="cuda")
    nom = base.n_om
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
    s = torch.cuda.current_stream()

    def step():
        ev[0].record(s)
        kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, wl.om, wl.tetc, layout="native",
                                  check=False)
        ev[1].record(s)
        eps = P.chi0_accumulate(M, kcw, precision="mixed")
        ev[2].record(s)
        w1 = P.chi0_wings(M, kcw, pm)
        ev[3].record(s)
        e2 = eps.clone()
        ev[4].record(s)
        P.eps_inverse(e2, check=False)
        ev[5].record(s)
        eps, head, w1, w2, emac = P.eps_inverse_head(eps, head0, w1, check=False)
        ev[6].record(s)
        mwm = P.calc_mwm(Ms, eps, wl.wkq, precision="mixed")
        ev[7].record(s)
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(s)
        P.calc_mwm_head(Ms, mwm, head, w1, w2, 1.0, 1.0, 0, wl.nbgw, wl.nb2)
        t1.record(s)
        torch.cuda.synchronize()
        return np.array([ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in (1, 2, 4, 5, 6)] +
                        [t0.elapsed_time(t1) / 1e3]), emac

    for _ in range(max(1, args.warmup)):
        step()
    st = np.zeros(6)
    for _ in range(args.steps):
        t, emac = step()
        st += t
    st /= args.steps
    chi0, wings, inv, inv_head, mwm, mwm_head = st
    extra = wings + (inv_head - inv) + mwm_head
    pb, K = base.pb, nk * npair
    f_w = 8.0 * pb * K * nom
    print(f"[bench] gamma extras {extra * 1e3:.2f} ms (chi0 {chi0:.3f} s)", file=sys.stderr, flush=True)
    rec = {"metric": "q = 0 head/wings extras, seconds per q = 0 step (wings + head/wings update + MWM head terms)",
           "value": extra, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": extra * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 (wings, head update, MWM head); fp16x3-split chi0/MWM beside them", "data": "synthetic",
           "config": {"workload": f"gamma: the c5 chain at q = 0 (8x8x8 k, {b.n_o}+{b.n_e} bands, PB={pb}, {nom} "
                                  f"Matsubara points, MWM rows of {wl.nbgw}x{wl.nb2} at {len(wl.klist)} k); "
                                  "synthetic momentum matrix elements and head"},
           "stages_s": {"chi0_mixed": chi0, "chi0_wings": wings, "eps_inverse": inv,
                        "eps_inverse_head": inv_head, "mwm_mixed": mwm, "mwm_head": mwm_head},
           "roofline": {"bound": "mfma", "kernel": "pygw_gm_wings", "achieved": f_w / wings / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_w / wings / PEAK_FP64,
                        "traffic": None, "flops_per_launch": f_w,
                        "note": "8 FLOP per (I, pair, frequency); time of the whole wing call"},
           "emac_w0": [complex(emac[0, 0]).real, complex(emac[0, 1]).real]}
    print(json.dumps(rec))


def run_coulomb(args):
    """The v^(1/2) basis of one q and the matrix elements in it at C5 size:
    Coul_setev on a 1000 x 1000 Coulomb matrix (900 eigenvalues kept at
    barcevtol = 1e-7, so PB = 900; pygw.py:2440, 2505-2545), then minm =
    mmat conj(barcvm) for all 512 k x 2400 pairs (pygw.py:2666), written as
    complex64 (the split-precision chi0's operand) and as complex128."""
    import torch

    from pygw_amd import coulomb, synth
    mb, nk, npair, evtol = 1000, 512, 2400, 1e-7
    vmat, lam = synth.coulomb_matrix(mb, 100)
    vd = torch.as_tensor(vmat, device="cuda")
    K = nk * npair
    g = torch.Generator(device="cuda")
    g.manual_seed(20261211)
    mm = torch.randn((K, mb), dtype=torch.complex128, device="cuda", generator=g)
    out64 = torch.empty((K, 900), dtype=torch.complex64, device="cuda")
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def step():
        ev[0].record(s)
        cb = coulomb.coul_setev(vd, evtol)
        ev[1].record(s)
        coulomb.calc_minm_vhalf(mm, cb.barcvm, complex64=True, out=out64[:, :cb.matsize])
        ev[2].record(s)
        o128 = coulomb.calc_minm_vhalf(mm, cb.barcvm)
        ev[3].record(s)
        torch.cuda.synchronize()
        del o128
        return np.array([ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(3)]), cb

    for _ in range(max(1, args.warmup)):
        step()
    st = np.zeros(3)
    for _ in range(args.steps):
        t, cb = step()
        st += t
    st /= args.steps
    t_ev, t_m32, t_m128 = st
    # the orthogonalised interstitial basis and the plane-wave block of Vmat (once per q)
    pws = synth.plane_wave_system(ngq=400, ngq_barc=700, gmax=10)
    pd = {k: torch.as_tensor(np.asarray(getattr(pws, k)), device="cuda") for k in ("indgq", "gindex", "i_g0", "ipwint")}
    gq = torch.as_tensor(pws.gqlen, device="cuda")
    vblk = torch.zeros((1000, 1000), dtype=torch.complex128, device="cuda")
    tp = []
    for _ in range(1 + args.steps):
        ev[0].record(s)
        mpw = coulomb.pw_basis(pd["indgq"], 400, 700, pd["ipwint"], pd["gindex"], pd["i_g0"])
        ev[1].record(s)
        coulomb.coul_pw_block(mpw, gq, False, vmat=vblk, row0=600)
        ev[2].record(s)
        torch.cuda.synchronize()
        tp.append([ev[0].elapsed_time(ev[1]) / 1e3, ev[1].elapsed_time(ev[2]) / 1e3])
    t_pw, t_blk = np.mean(tp[1:], axis=0)
    assert cb.matsize == 900
    ev_err = float(np.abs(cb.ev.cpu().numpy() - lam).max() / lam.max())
    f_m = 8.0 * K * mb * cb.matsize
    value = t_ev + t_m32
    print(f"[bench] coulomb setev {t_ev:.3f} s ({cb.sweeps} sweeps), minm c64 {t_m32:.3f} s "
          f"({f_m / t_m32 / 1e12:.1f} TF/s), c128 {t_m128:.3f} s", file=sys.stderr, flush=True)
    rec = {"metric": "Coul_setev + minm = mmat conj(barcvm) seconds per q (all k)", "value": value, "unit": "s/q",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": value * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (c64 output)",
           "data": "synthetic",
           "config": {"workload": f"coulomb: mbsize {mb} Hermitian Coulomb matrix (900 eigenvalues 4pi/g^2, 100 "
                                  f"below barcevtol {evtol}), PB = 900; minm for {nk} k x {npair} pairs"},
           "stages_s": {"coul_setev": t_ev, "minm_c64": t_m32, "minm_c128": t_m128,
                        "pw_basis_400x700": t_pw, "coul_pw_block_400x700": t_blk},
           "jacobi_sweeps": cb.sweeps, "max_eigenvalue_error_rel": ev_err,
           "roofline": {"bound": "mfma", "kernel": "pygw_minm_vhalf", "achieved": f_m / t_m32 / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_m / t_m32 / PEAK_FP64,
                        "traffic": load_traffic("coulomb", "pygw_minm_vhalf<float2>"), "flops_per_launch": f_m,
                        "note": "8 FLOP per (row, basis function, kept vector); complex64 output"}}
    if not args.no_cpu:
        rec["cpu_baseline"] = coulomb_cpu_baseline(vmat, mm[:2 * npair].cpu().numpy(), evtol, nk // 2, t_ev)
    print(json.dumps(rec))


def coulomb_cpu_baseline(vmat, mm_sample, evtol, scale, t_gpu_ev):
    """PyGW's own steps on the host: scipy eigh (LAPACK, lower) + the
    Coul_setev selection, then numpy matmul(mmat, conj(barcvm)) on 2 of 512
    k, extrapolated."""
    from oracle import coulomb as O
    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    _, _, _, bv, _ = O.coul_setev(vmat, evtol)
    t1 = time.perf_counter()
    O.minm(mm_sample, bv)
    t2 = time.perf_counter()
    return {"value": (t1 - t0) + (t2 - t1) * scale, "unit": "s/q", "cores": nthr, "kind": "port", **host_info(),
            "sample": f"scipy eigh of the {vmat.shape[0]}-square matrix in full; numpy matmul on 2 of 512 k (x{scale})",
            "stages_s": {"coul_setev": t1 - t0, "minm_extrapolated": (t2 - t1) * scale}}


def _unique_lengths(v):
    """Lengths of the vectors v [n, 3] in runs equal within 1e-10 (first of each run)."""
    ln = np.sort(np.linalg.norm(v, axis=1))
    return ln[np.r_[True, np.diff(ln) > 1e-10]]


def run_mt_coulomb(args):
    """The muffin-tin block of the Coulomb matrix of one q at PyGW's sizes
    (pygw.py:2348-2394): Ewald structure constants to lam_max = 16 for 8
    atoms (36 pairs), the radial overlaps jlam of every atom type against
    j_L(|q+G| r) r for the unique |q+G| (Calc_Olap_PW_Product), and
    MT_Coulomb over all rows (loctmatsize 896); the q = 0 term
    (Zero_Momentum_Coulomb) timed beside it."""
    import torch

    from pygw_amd import coulomb, synth
    S = synth.mt_bench_system()
    G = S.gindex[S.glen <= 5.0] @ S.br2.T
    gq = _unique_lengths(G + S.pq[None, :])
    lmaxb = int(max(int(b.max()) for b in S.big_l))
    meshes = [coulomb.radial_mesh(*m) for m in S.meshes]
    rx_d = [torch.as_tensor(r, device="cuda") for r in meshes]
    ul_d = [torch.as_tensor(u, device="cuda") for u in S.ul]
    us_d = [torch.as_tensor(u, device="cuda") for u in S.us]
    gq_d = torch.as_tensor(gq, device="cuda")
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    mt_args = (S.loctmatsize, S.big_l_f, S.nmix, S.im_start, S.tilg, S.rrint, S.djmm, S.rtlij)

    def step():
        ev[0].record(s)
        sgm = coulomb.ewald_summation(S.pq, S.lam_max, S.pos, S.br2, S.rbas, S.alat, S.vol, S.ortho, S.r_cf, S.g_cf,
                                      S.eta)
        ev[1].record(s)
        jl = [coulomb.calc_olap_pw_product(gq_d, lmaxb, rx_d[a], S.meshes[a][1], S.rel, S.big_l[a], ul_d[a],
                                           us_d[a]) for a in range(len(S.big_l))]
        ev[2].record(s)
        V = coulomb.mt_coulomb(S.pq, None, *mt_args, sgm, S.mult, S.lam_max)
        ev[3].record(s)
        torch.cuda.synchronize()
        return np.array([ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(3)]), (sgm, jl, V)

    for _ in range(max(1, args.warmup)):
        step()
    st = np.zeros(3)
    for _ in range(args.steps):
        t, out = step()
        st += t
    st /= args.steps
    t_ew, t_jl, t_mt = st
    zm_args = (S.glen, S.gindex, S.pos, S.vol, S.loctmatsize + 500, S.rel, S.meshes, S.big_l, S.ul, S.us, S.atm,
               S.iprod)
    coulomb.zero_momentum_coulomb(*zm_args)
    torch.cuda.synchronize()
    tz = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        coulomb.zero_momentum_coulomb(*zm_args)
        torch.cuda.synchronize()
        tz.append(time.perf_counter() - t0)
    value = float(st.sum())
    # the radial overlaps of one atom type (table, weights, overlap kernels), timed alone on the stream
    a0 = int(np.argmax([len(b) for b in S.big_l]))
    tj = []
    for _ in range(max(1, args.steps)):
        ev[0].record(s)
        coulomb.calc_olap_pw_product(gq_d, lmaxb, rx_d[a0], S.meshes[a0][1], S.rel, S.big_l[a0], ul_d[a0], us_d[a0])
        ev[1].record(s)
        torch.cuda.synchronize()
        tj.append(ev[0].elapsed_time(ev[1]) / 1e3)
    t_olap = float(np.mean(tj))
    npt, nf = S.meshes[a0][2], len(S.big_l[a0])
    # algorithmic bytes: the Bessel table written and read once, the radial functions, k, the mesh, the result
    b_olap = 2.0 * 8 * (lmaxb + 1) * npt * len(gq) + 2.0 * 8 * nf * npt + 8.0 * (len(gq) + npt) + 8.0 * nf * len(gq)
    print(f"[bench] mt_coulomb ewald {t_ew * 1e3:.2f} ms, jlam {t_jl * 1e3:.2f} ms ({len(gq)} |q+G|), "
          f"mt {t_mt * 1e3:.2f} ms, zero momentum {np.mean(tz) * 1e3:.2f} ms", file=sys.stderr, flush=True)
    rec = {"metric": "muffin-tin Coulomb block (Ewald + jlam + MT_Coulomb) seconds per q", "value": value,
           "unit": "s/q", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": value * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"mt_coulomb: 8 atoms (4 types x mult 2), lam_max {S.lam_max}, loctmatsize "
                                  f"{S.loctmatsize}, {len(gq)} unique |q+G|, 781-point meshes, Lmax {lmaxb}"},
           "stages_s": {"ewald_summation": t_ew, "calc_olap_pw_product_all_atoms": t_jl, "mt_coulomb": t_mt,
                        "zero_momentum_coulomb_q0_host_timed": float(np.mean(tz))},
           "roofline": {"bound": "hbm", "kernel": "pygw_radial_olap (pygw_rad_table with the weights + "
                                                  "pygw_rad_olap, one atom type)",
                        "achieved": b_olap / t_olap / 1e9, "peak": 8000.0, "unit": "GB/s",
                        "frac": b_olap / t_olap / 8e12, "traffic": None, "bytes_per_call": b_olap,
                        "call_s": t_olap,
                        "note": "Bessel table written and read once + radial functions + result; latency-bound "
                                "at this size (one atom type: 781 points x 1473 |q+G| x 4 L)"}}
    if not args.no_cpu:
        rec["cpu_baseline"] = mt_coulomb_cpu_baseline(S, gq, meshes, lmaxb, value)
    print(json.dumps(rec))


def mt_coulomb_cpu_baseline(S, gq, meshes, lmaxb, t_gpu):
    """PyGW's own sequence on the host with its compiled Fortran
    (oracle/_ref): fCl.ewald_summation, fCl.spher_bessel + rd.rint13g per
    (function, |q+G|) as Calc_Olap_PW_Product calls them, fCl.mt_coulomb per
    atom type."""
    from oracle import ref
    if not (ref.available("libref_coul.so") and ref.available("libref_rad.so")):
        return None
    t0 = time.perf_counter()
    sgm = ref.ewald_summation(S.pq, S.lam_max, S.pos, S.br2, S.rbas, S.alat, S.vol, S.ortho, S.r_cf, S.g_cf, S.eta)
    t1 = time.perf_counter()
    for a in range(len(S.big_l)):
        rx = meshes[a]
        r0, dh, npt = S.meshes[a]
        jln = ref.spher_bessel(lmaxb, np.outer(gq, rx).ravel()).reshape(lmaxb + 1, len(gq), npt) * rx
        for i, L in enumerate(S.big_l[a]):
            for g in range(len(gq)):
                ref.rint13g(S.rel, S.ul[a][i], S.us[a][i], jln[L, g], jln[L, g], dh, npt, r0)
    t2 = time.perf_counter()
    for a in range(len(S.big_l)):
        ref.mt_coulomb(S.pq, a + 1, S.loctmatsize, S.big_l_f, S.nmix, S.im_start, S.tilg, S.rrint, S.djmm, S.rtlij,
                       sgm, S.mult, S.lam_max)
    t3 = time.perf_counter()
    return {"value": t3 - t0, "unit": "s/q", "cores": 1, "kind": "reference", **host_info(),
            "sample": "the whole q: the reference's compiled Ewald_Summation, spher_bessel, RINT13g and MT_Coulomb "
                      "(oracle/_ref), called in PyGW's order from numpy",
            "stages_s": {"ewald_summation": t1 - t0, "calc_olap_pw_product_all_atoms": t2 - t1,
                         "mt_coulomb": t3 - t2}}


def run_minm_is(args):
    """The interstitial matrix elements of one q (calc_minm_IS per k,
    pygw.py:2661): 512 k, LAPW bases of 600 plane waves at k and k - q,
    400 product-basis plane waves (mpwipw against 700), the 20 occupied x
    120 empty band window of the polarization.  The same synthetic inputs
    serve every k (the work per k is the same)."""
    import torch

    from pygw_amd import coulomb, synth
    nk, rng_ = 512, (0, 20, 20, 140)
    s = synth.interstitial_system(seed=20261212, gmax=8, nvi=600, nvj=600, ngq=400, ngq_barc=700, nb=140)
    d = lambda x, dt: torch.as_tensor(np.asarray(x), dtype=dt, device="cuda")
    dev = synth.InterstitialSystem(gindex=d(s.gindex, torch.int32), i_g0=d(s.i_g0, torch.int32),
                                   indgk_i=d(s.indgk_i, torch.int32), indgk_j=d(s.indgk_j, torch.int32),
                                   indgq=s.indgq, indgq_inverse=d(s.indgq_inverse, torch.int32),
                                   mpwipw=d(s.mpwipw, torch.complex128), Aeigk=d(s.Aeigk, torch.complex128),
                                   Aeigq=d(s.Aeigq, torch.complex128), iumklap=s.iumklap, nvi=s.nvi, nvj=s.nvj,
                                   vol=s.vol)
    ndim, mdim, ngq = rng_[1] - rng_[0], rng_[3] - rng_[2], s.mpwipw.shape[0]
    out = torch.empty((ndim, mdim, ngq), dtype=torch.complex128, device="cuda")
    ws = torch.empty(coulomb.minm_is_workspace(ngq, s.mpwipw.shape[1], s.nvi, s.nvj, ndim, mdim),
                     dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def step():
        e0.record(st)
        for _ in range(nk):
            coulomb.calc_minm_is(*dev.args(*rng_), out=out, workspace=ws)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    for _ in range(max(1, args.warmup)):
        step()
    t = sum(step() for _ in range(args.steps)) / args.steps
    nb = s.mpwipw.shape[1]
    f_ref = 8.0 * ngq * ndim * s.nvj * (s.nvi + mdim)          # the Fortran's two products per plane wave
    rho = coulomb._rho_cheaper(ngq, nb, s.nvj, ndim, mdim)
    f_k = 8.0 * ndim * nb * mdim * (s.nvj + ngq) if rho else 8.0 * ngq * ndim * s.nvj * (nb + mdim)
    kernels = (("pygw_is_gemm_ab<2>", "pygw_is_gemm_ab<1>", "pygw_is_reduce", "pygw_is_imap", "pygw_is_imap_fill")
               if rho else ("pygw_is_gemm1p", "pygw_is_gemm2<4>", "pygw_is_scatter", "pygw_is_zero")) + ("pygw_is_map",)
    print(f"[bench] minm_is {t:.3f} s/q ({f_k * nk / t / 1e12:.1f} TF/s on the products run, "
          f"{f_ref * nk / t / 1e12:.1f} TF/s on the Fortran's count)", file=sys.stderr, flush=True)
    rec = {"metric": "interstitial matrix elements (calc_minm_IS) seconds per q (all k)", "value": t, "unit": "s/q",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"minm_is: {nk} k, nvi = nvj = {s.nvi} plane waves, ngq = {ngq} (mpwipw against "
                                  f"{nb}), bands {ndim} x {mdim}",
                      "association": "pair density first" if rho else "T first (dense)"},
           "roofline": {"bound": "mfma", "kernel": " + ".join(kernels[:2]), "achieved": f_k * nk / t / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_k * nk / t / PEAK_FP64,
                        "traffic": sum(load_traffic("minm_is", k) or 0.0 for k in kernels) or None,
                        "flops_per_k": f_k, "fortran_flops_per_k": f_ref,
                        "fortran_equivalent_tflops": f_ref * nk / t / 1e12,
                        "note": "8 FLOP per complex multiply-add of the two products; time of whole calls"}}
    if not args.no_cpu:
        from oracle import coulomb as O
        t0 = time.perf_counter()
        want = O.minm_is(*s.args(*rng_))
        t1 = time.perf_counter()
        err = float(np.abs(out.cpu().numpy() - want).max() / np.abs(want).max())
        rec["max_rel_error_vs_oracle"] = err
        rec["cpu_baseline"] = {"value": (t1 - t0) * nk, "unit": "s/q", "cores": int(os.environ.get(
            "OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port", **host_info(),
            "sample": f"the numpy restatement (two matmuls per product-basis plane wave, as the Fortran's zgemm pair) "
                      f"on 1 of {nk} k, x{nk}"}
    print(json.dumps(rec))


def _minm_mt_inputs(nvk=600, nb=140, seed=20261213):
    """A PyGW-sized muffin-tin system: 4 atom types of 2 atoms, lmax 10,
    local orbitals, product bases of 26-30 functions per type (L <= 3), 600
    plane waves per k-point, 140 bands; the LAPW inputs at k and k - q and
    the matrix-element inputs on the same atoms."""
    import dataclasses

    from pygw_amd import synth
    sk = synth.lapw_system(lmax=10, lomax=10, nvk=nvk, nbmax=nb, mult=(2, 2, 2, 2), apw=(True, False, True, False),
                           gmax=7, alat=11.0, kl=(0.125, 0.25, -0.375), seed=seed)
    sq = synth.lapw_system(lmax=10, lomax=10, nvk=nvk, nbmax=nb, mult=(2, 2, 2, 2), apw=(True, False, True, False),
                           gmax=7, alat=11.0, kl=(0.0, 0.125, -0.25), seed=seed)
    base = synth.minm_mt_system(big=True, lmax=10, nbmax=nb, mbmax=nb)
    nfnc = max(int(base.ncore[a]) + 2 * 11 + int(sk.nLO_at[:, a].sum()) for a in range(4))
    rng = np.random.default_rng(seed)
    s3r = rng.normal(size=(base.s3r.shape[0], nfnc, nfnc, 4))
    s3r = 0.5 * (s3r + s3r.transpose(0, 2, 1, 3))
    s = dataclasses.replace(base, mult=sk.mult, nLO_at=sk.nLO_at, s3r=s3r, vpos=sk.pos)
    return sk, sq, s


def run_minm(args):
    """The product-basis matrix elements of one q as calc_eps consumes them
    (calc_minm in polarization mode per k, pygw.py:2547-2700, 3096-3101):
    LAPW coefficients at k and k - q, band coefficients, calc_minm_MT,
    calc_minm_IS, calc_minc2 and minm = mmat conj(barcvm), through
    pygw_amd.minm.MatrixElements, for the 512 k of one q at the shape of the
    C5 chain's operand: 20 occupied x 120 empty bands (plus 36 core rows),
    mbsize 1000 = loctmatsize 896 (8 atoms, lmax 10) + 104 interstitial
    plane waves (against 200), PB 900 kept by coul_setev; 600 plane waves
    per k.  The same synthetic inputs at every k; the k-points run in
    batches of 8 (MatrixElements.polarization_minm_many: one product with
    conj(barcvm) per batch)."""
    import torch

    from pygw_amd import coulomb, lapw, synth
    from pygw_amd.minm import MatrixElements
    nk, nvbm, ncbm, nb = 512, 20, 20, 140
    ngq, ngq_barc, evtol = 104, 200, 1e-7
    sk, sq, s = _minm_mt_inputs(nb=nb)
    g, kc, nvk = sk.gindex, sk.k2cartes, sk.nvk
    nlo_rows = sk.indgk.size - nvk
    kil = np.asarray(sk.kl, dtype=np.float64)
    ql = np.array([0.125, 0.125, -0.125])
    kjl = kil - ql

    def rows(kl):      # the nvk shortest |k+G|, then the local-orbital rows
        order = np.argsort(np.linalg.norm((g + kl) @ kc.T, axis=1), kind="stable")
        return order[:nvk + nlo_rows].astype(np.int32)

    gm = int(np.abs(g).max())
    i_g0 = -np.ones((2 * gm + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gm, g[:, 1] + gm, g[:, 2] + gm] = np.arange(len(g))
    indgq = np.argsort(np.linalg.norm((g + ql) @ kc.T, axis=1), kind="stable")[:ngq_barc]
    inv = -np.ones(indgq.max() + 1, dtype=np.int64)
    inv[indgq] = np.arange(ngq_barc)
    rng = np.random.default_rng(20261214)
    mpwipw = (rng.normal(size=(ngq, ngq_barc)) + 1j * rng.normal(size=(ngq, ngq_barc))) / np.sqrt(ngq_barc)
    mb = s.loctmatsize + ngq
    vmat, _ = synth.coulomb_matrix(mb, 100)
    cb = coulomb.coul_setev(torch.as_tensor(vmat, device="cuda"), evtol)
    lp = lapw.LapwPlan(*sk.plan_args())
    mp = coulomb.MinmMTPlan(s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.lmax)
    consts = coulomb.mt_constants(mp, s.s3r, s.vpos, s.cgcoef, s.corind)
    me = MatrixElements(lp, mp, consts, g, i_g0, mpwipw, inv, sk.vol, cb.barcvm, ql)
    dev = lambda x, dt: torch.as_tensor(np.asarray(x), dtype=dt, device="cuda")  # noqa: E731
    ig_k, ig_q = dev(rows(kil), torch.int32), dev(rows(kjl), torch.int32)
    Ak = dev(sk.Aeig, torch.complex128)
    Aq = dev(np.conj(sq.Aeig[:, :nvk + nlo_rows]), torch.complex128)
    ncg = len(s.corind)
    out = torch.empty((ncg + nvbm, nb - ncbm, cb.matsize), dtype=torch.complex64, device="cuda")
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    kb = 8                                      # k-points per product with conj(barcvm)
    outb = torch.empty((kb,) + tuple(out.shape), dtype=torch.complex64, device="cuda")
    batch = [(kil, ig_k, nvk, Ak, kjl, ig_q, nvk, Aq)] * kb

    def step(prec):
        e0.record(st)
        for _ in range(nk // kb):
            me.polarization_minm_many(batch, nvbm, ncbm, nb, ncg, out=outb, complex64=True, precision=prec)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    # the split-precision product against the FP64 one on this k (relative
    # Frobenius error of the complex64 minm handed to the chain)
    ref = me.polarization_minm(kil, ig_k, nvk, Ak, kjl, ig_q, nvk, Aq, nvbm, ncbm, nb, ncg, precision="fp64")
    mix = me.polarization_minm(kil, ig_k, nvk, Ak, kjl, ig_q, nvk, Aq, nvbm, ncbm, nb, ncg, complex64=True,
                               precision="mixed")
    err = float(torch.linalg.norm(mix.to(torch.complex128) - ref) / torch.linalg.norm(ref))
    times = {}
    for prec in ("mixed", "fp64"):
        for _ in range(max(1, args.warmup)):
            step(prec)
        times[prec] = sum(step(prec) for _ in range(args.steps)) / args.steps
    t = times["mixed"]
    pairs = (ncg + nvbm) * (nb - ncbm)
    f_vh = 8.0 * pairs * mb * cb.matsize
    print(f"[bench] minm {t:.3f} s/q with the split-precision v^1/2 product ({times['fp64']:.3f} FP64; rel err "
          f"{err:.1e}; PB {cb.matsize}, mmat conj(barcvm) {f_vh * nk / 1e12:.1f} TFLOP per q)",
          file=sys.stderr, flush=True)
    rec = {"metric": "product-basis matrix elements of one q (calc_minm per k: LAPW coefficients, calc_minm_MT, "
                     "calc_minm_IS, calc_minc2, mmat conj(barcvm)) seconds per q (all k)", "value": t, "unit": "s/q",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64, the v^1/2 product fp16x3-split/fp32-acc (M stored complex64)",
           "data": "synthetic",
           "config": {"workload": f"minm: {nk} k, 4 atom types x 2 atoms, lmax 10, {nvk} plane waves + {nlo_rows} "
                                  f"local-orbital rows, bands {nvbm} x {nb - ncbm} + {ncg} core rows, mbsize {mb} "
                                  f"(loctmatsize {s.loctmatsize} + {ngq} plane waves against {ngq_barc}), PB "
                                  f"{cb.matsize}"},
           "fp64": {"s_per_q": times["fp64"]},
           "rel_frobenius_error_split_vs_fp64": err,
           "roofline": {"bound": "mfma", "kernel": "pygw_vh_mixed<true> (minm = mmat conj(barcvm), split precision)",
                        "flops_per_k": f_vh, "achieved_if_all_time": f_vh * nk / t / 1e12, "peak": PEAK_FP16 / 1e12,
                        "unit": "TFLOP/s", "frac": f_vh * nk / t / PEAK_FP16, "traffic": None,
                        "note": "the v^1/2 product's algorithmic FLOP over the whole call time: a lower bound of "
                                "its rate"}}
    if not args.no_cpu:
        from oracle import coulomb as OCo
        from oracle import lapw as OL
        from oracle import minm_mt as OM
        from oracle import ref
        import dataclasses
        Bh = cb.barcvm.cpu().numpy()
        st_ = {}
        t0 = time.perf_counter()
        lk = dataclasses.replace(sk, indgk=rows(kil).astype(np.int64), nvk=nvk)
        lq = dataclasses.replace(sk, kl=kjl, indgk=rows(kjl).astype(np.int64), nvk=nvk)
        Akh, Aqh = sk.Aeig, np.conj(sq.Aeig[:, :nvk + nlo_rows])
        wk = OL.band_coefficients(Akh[:nvbm], *ref.gap2_set_lapwcoef(*lk.coef_args(1, True)))
        wq = OL.band_coefficients(Aqh[ncbm:nb], *ref.gap2_set_lapwcoef(*lq.coef_args(2, True)))
        t1 = time.perf_counter()
        rest = (s.s3r, s.vpos, s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
        mt = OM.calc_minm_mt(ql, 0, nvbm, 0, nb - ncbm, *wk, *wq, *rest)
        mc = OM.calc_minc2(kjl, 0, ncg, 0, nb - ncbm, *wq, s.s3r, s.corind, *rest[1:])
        t2 = time.perf_counter()
        it = OCo.minm_is(0, nvbm, ncbm, nb, Akh, Aqh, np.zeros(3, dtype=int), mpwipw, nvk, nvk, rows(kil),
                         rows(kjl), inv, g, i_g0, sk.vol)
        t3 = time.perf_counter()
        mm_h = np.concatenate((np.concatenate((mc, np.zeros(mc.shape[:2] + (ngq,), complex)), -1),
                               np.concatenate((mt, it), -1)), 0)
        OCo.minm(mm_h, Bh)
        t4 = time.perf_counter()
        st_ = {"lapw_coefficients_and_band_products": (t1 - t0) * nk, "calc_minm_mt_and_minc2": (t2 - t1) * nk,
               "calc_minm_is": (t3 - t2) * nk, "v_half_product": (t4 - t3) * nk}
        rec["cpu_baseline"] = {"value": (t4 - t0) * nk, "unit": "s/q", "cores": int(os.environ.get(
            "OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port", **host_info(), "stages_s": st_,
            "sample": f"1 of {nk} k, x{nk}: PyGW's compiled gap2_set_lapwcoef (1 core) and numpy products, the "
                      f"numpy restatements of calc_minm_MT / calc_minc2 / calc_minm_IS, numpy matmul with "
                      f"conj(barcvm) (FP64)"}
    print(json.dumps(rec))


def run_minm_mt(args):
    """The muffin-tin and core matrix elements of one q as calc_minm builds
    them (pygw.py:2559-2680, polarization mode): per k, gap2_set_lapwcoef at
    k and at k - q, the band coefficients alfa/beta/gama and alfp/betp/gamp
    of all 140 bands, calc_minm_MT over the 20 occupied x 120 empty window
    and calc_minc2 over the core states; 512 k, the same synthetic inputs at
    every k (the work per k is the same)."""
    import torch

    from pygw_amd import coulomb, lapw
    nk, win = 512, (0, 20, 20, 140)
    sk, sq, s = _minm_mt_inputs()
    dev = lambda x, dt: torch.as_tensor(np.asarray(x), dtype=dt, device="cuda")  # noqa: E731
    plan = lapw.LapwPlan(*sk.plan_args())
    mplan = coulomb.MinmMTPlan(s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.lmax)
    ig_k, ig_q = dev(sk.indgk, torch.int32), dev(sq.indgk, torch.int32)
    gk = dev(np.asarray(sk.gindex).ravel(order="F"), torch.int32)
    gq = dev(np.asarray(sq.gindex).ravel(order="F"), torch.int32)
    Ak, Aq = dev(sk.Aeig, torch.complex128), dev(sq.Aeig, torch.complex128)
    ck = lapw.gap2_set_lapwcoef(plan, sk.kl, ig_k, 1, True, sk.nvk, gk)
    cq = lapw.gap2_set_lapwcoef(plan, sq.kl, ig_q, 2, True, sq.nvk, gq)
    bk = lapw.band_coefficients(Ak, *ck)
    bq = lapw.band_coefficients(Aq, *cq, conj=True)
    n1, n2, ncg, loct = win[1] - win[0], win[3] - win[2], len(s.corind), s.loctmatsize
    mm = torch.empty((n1, n2, loct), dtype=torch.complex128, device="cuda")
    mc = torch.empty((ncg, n2, loct), dtype=torch.complex128, device="cuda")
    import ctypes as C
    from pygw_amd import _lib
    sz = C.c_size_t(0)
    _lib.call("pygw_minm_mt_workspace", mplan.ptrs()[0], n1, n2, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    consts = coulomb.mt_constants(mplan, s.s3r, s.vpos, s.cgcoef, s.corind)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    tt = np.zeros(4)

    def step():
        times = np.zeros(4)
        for _ in range(nk):
            ev[0].record(st)
            lapw.gap2_set_lapwcoef(plan, sk.kl, ig_k, 1, True, sk.nvk, gk, out=ck)
            lapw.gap2_set_lapwcoef(plan, sq.kl, ig_q, 2, True, sq.nvk, gq, out=cq)
            ev[1].record(st)
            lapw.band_coefficients(Ak, *ck, out=bk)
            lapw.band_coefficients(Aq, *cq, conj=True, out=bq)
            ev[2].record(st)
            coulomb.calc_minm_mt(mplan, s.ql, *win, *bk, *bq, consts, None, None, out=mm, workspace=ws)
            ev[3].record(st)
            coulomb.calc_minc2(mplan, sq.kl, 0, ncg, win[2], win[3], *bq, consts, None, None, None, out=mc)
            ev[4].record(st)
            torch.cuda.synchronize()
            times += [ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(4)]
        return times

    for _ in range(max(1, args.warmup)):
        step()
    for _ in range(args.steps):
        tt += step()
    tt /= args.steps
    t = float(tt.sum())
    ngk, ntnt, ndf, nlo = ck[0].shape[0], ck[0].shape[1], ck[0].shape[2], plan.nlomax
    f_band = 2 * 8.0 * sk.Aeig.shape[0] * ngk * ntnt * ndf * (2 + nlo)      # both sides, per k
    print(f"[bench] minm_mt {t:.3f} s/q (coef {tt[0]:.3f}, band {tt[1]:.3f} = {f_band * nk / tt[1] / 1e12:.1f} TF/s, "
          f"minm_mt {tt[2]:.3f}, minc2 {tt[3]:.3f})", file=sys.stderr, flush=True)
    rec = {"metric": "muffin-tin + core matrix elements (LAPW coefficients -> alfa/beta/gama -> calc_minm_MT, "
                     "calc_minc2) seconds per q (all k)", "value": t, "unit": "s/q",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"minm_mt: {nk} k, 4 atom types x 2 atoms, lmax 10, {sk.nvk} plane waves + "
                                  f"{ngk - sk.nvk} local-orbital rows, {sk.Aeig.shape[0]} bands, window {n1} x {n2}, "
                                  f"loctmatsize {loct}, {ncg} core states"},
           "stages_s": {"lapw_coefficients": tt[0], "band_coefficients": tt[1], "calc_minm_mt": tt[2],
                        "calc_minc2": tt[3]},
           "roofline": {"bound": "mfma", "kernel": "pygw_lapw_band_k", "achieved": f_band * nk / tt[1] / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_band * nk / tt[1] / PEAK_FP64,
                        "traffic": None, "flops_per_k": f_band,
                        "note": "the band-coefficient products (8 FLOP per complex multiply-add, k and k - q); "
                                "time of whole calls"}}
    if not args.no_cpu:
        from oracle import lapw as OL
        from oracle import minm_mt as OM
        from oracle import ref
        t0 = time.perf_counter()
        wk = OL.band_coefficients(sk.Aeig, *ref.gap2_set_lapwcoef(*sk.coef_args(1, True)))
        wq = OL.band_coefficients(np.conj(sq.Aeig), *ref.gap2_set_lapwcoef(*sq.coef_args(2, True)))
        t1 = time.perf_counter()
        want = OM.calc_minm_mt(s.ql, *win, *wk, *wq, s.s3r, s.vpos, s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore,
                               s.cgcoef, s.lmax, s.loctmatsize)
        t2 = time.perf_counter()
        rec["max_rel_error_vs_oracle"] = float(np.abs(mm.cpu().numpy() - want).max() / np.abs(want).max())
        rec["cpu_baseline"] = {"value": (t2 - t0) * nk, "unit": "s/q", "cores": int(os.environ.get(
            "OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port", **host_info(),
            "stages_s": {"lapw_coefficients_and_band_products": (t1 - t0) * nk, "calc_minm_mt": (t2 - t1) * nk},
            "sample": f"1 of {nk} k, x{nk}: PyGW's compiled gap2_set_lapwcoef (1 core) and numpy products for the "
                      f"coefficients, the numpy restatement of calc_minm_MT (the compiled one needs BLAS zgemm, "
                      f"absent here); calc_minc2 not timed"}
    print(json.dumps(rec))


def run_head(args):
    """calc_head at q = 0 (pygw.py:2808-3037) for the 512 k of an 8x8x8 mesh
    (all irreducible): per k the LAPW coefficients (ifrot False), alfa/beta/
    gama of 140 bands, the momentum elements of the 20 occupied x 120 empty
    pairs (muffin-tin, lmax 10, 8 atoms, + interstitial over 600 plane waves,
    both orders) and of 36 core states; then the head over 64 frequencies
    from kcw [56, 120, 512, 64].  The same synthetic inputs at every k."""
    import torch

    from pygw_amd import head as H
    from pygw_amd import lapw, synth
    nk, nomx, numin, nbmaxpol, nom = 512, 19, 20, 140, 64
    s = synth.lapw_system(lmax=10, lomax=10, nvk=600, nbmax=nbmaxpol, mult=(2, 2, 2, 2),
                          apw=(True, False, True, False), gmax=10, alat=11.0, kl=(0.125, 0.25, -0.375), seed=20261214)
    assert np.abs(s.gindex[s.indgk[:s.nvk]]).max() <= 5
    base = synth.minm_mt_system(big=True, lmax=10)
    rng = np.random.default_rng(20261214)
    g = s.gindex
    gm = 10
    i_g0 = -np.ones((2 * gm + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gm, g[:, 1] + gm, g[:, 2] + gm] = np.arange(len(g))
    ipw = (rng.normal(size=len(g)) + 1j * rng.normal(size=len(g))) * np.exp(-0.1 * (g ** 2).sum(1))
    neg = i_g0[-g[:, 0] + gm, -g[:, 1] + gm, -g[:, 2] + gm]
    ipwint = 0.5 * (ipw + np.conj(ipw[neg]))
    nlo = max(int(s.nLO_at.max()), 1)
    R = lambda *sh: rng.normal(size=sh)  # noqa: E731
    vv_t = tuple(R(2, 10, 4) for _ in range(4)) + tuple(R(2, nlo, 11, 4) for _ in range(4)) + \
        (R(2, nlo, nlo, 11, 4),)
    ncore = np.asarray(base.ncore)
    core_off = np.concatenate([[0], np.cumsum(ncore)[:-1]])
    nct = int(ncore.sum())
    cv_t = tuple(R(2, nct) for _ in range(4)) + tuple(R(2, nlo, nct) for _ in range(2))
    corind = base.corind
    ncg = len(corind)
    tabs = H.HeadTables(s.mult, s.nLO_at, vv_t, core_off, cv_t, corind)
    pw = H.PlaneWaves(ipwint, g, i_g0, s.k2cartes)
    plan = lapw.LapwPlan(*s.plan_args())
    dev = lambda x, dt: torch.as_tensor(np.asarray(x), dtype=dt, device="cuda")  # noqa: E731
    ig, A = dev(s.indgk, torch.int32), dev(s.Aeig, torch.complex128)
    Ebnd = np.sort(np.concatenate([rng.uniform(-0.8, -0.05, (nk, nomx + 1)),
                                   rng.uniform(0.05, 2.0, (nk, nbmaxpol - nomx - 1))], axis=1), axis=1)
    eig_core = -20.0 - rng.uniform(0, 5, ncg)
    n_o, n_e = ncg + nomx + 1, nbmaxpol - numin
    kcw = (-0.01 * torch.rand((nom, nk, n_e, n_o), dtype=torch.float64, device="cuda")).permute(3, 2, 1, 0)
    Ed, ecd = dev(Ebnd, torch.float64), dev(eig_core, torch.float64)
    kind, wt = dev(np.arange(nk), torch.int32), dev(np.ones(nk), torch.float64)
    mvv = torch.empty((nk, n_e, nomx + 1, 3), dtype=torch.complex128, device="cuda")
    mcv = torch.empty((nk, n_e, ncg, 3), dtype=torch.complex128, device="cuda")
    coef = lapw.gap2_set_lapwcoef(plan, s.kl, ig, 1, False, s.nvk, pw.gindex)
    abc = lapw.band_coefficients(A, *coef)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def step():
        t = np.zeros(3)
        ev[0].record(st)
        for k in range(nk):
            lapw.gap2_set_lapwcoef(plan, s.kl, ig, 1, False, s.nvk, pw.gindex, out=coef)
            lapw.band_coefficients(A, *coef, out=abc)
        ev[1].record(st)
        for k in range(nk):
            H.momentum_elements(tabs, pw, abc, A, s.kl, ig, s.nvk, 0, nomx + 1, numin, nbmaxpol, mmatvv=mvv[k],
                                mmatcv=mcv[k])
        ev[2].record(st)
        head = H.head_from_elements(mvv, mcv, Ed, ecd, kcw, kind, wt, nomx, numin, nbmaxpol, ncg, s.vol)
        ev[3].record(st)
        torch.cuda.synchronize()
        t += [ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(3)]
        return t, head

    for _ in range(max(1, args.warmup)):
        step()
    tt = np.zeros(3)
    for _ in range(args.steps):
        dt, head = step()
        tt += dt
    tt /= args.steps
    t = float(tt.sum())
    print(f"[bench] head {t:.3f} s (coefficients {tt[0]:.3f}, momentum elements {tt[1]:.3f}, head sums {tt[2]:.4f})",
          file=sys.stderr, flush=True)
    f_is = 8.0 * s.nvk * s.nvk * nbmaxpol * nk          # the overlap product T = O conj(A)^T per k
    rec = {"metric": "q = 0 head (calc_head: LAPW coefficients, momentum matrix elements, head sums) seconds",
           "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"head: {nk} k (8x8x8, all irreducible), 4 atom types x 2 atoms, lmax 10, "
                                  f"{s.nvk} plane waves, bands {nomx + 1} x {nbmaxpol - numin}, {ncg} core states, "
                                  f"{nom} frequencies"},
           "stages_s": {"lapw_and_band_coefficients": tt[0], "momentum_elements": tt[1], "head_sums": tt[2]},
           "roofline": {"bound": "mfma", "kernel": "pygw_hd_is_t", "achieved": None, "peak": PEAK_FP64 / 1e12,
                        "unit": "TFLOP/s", "frac": None, "traffic": None, "flops_per_k": f_is / nk,
                        "note": "kernel rates from the rocprof summary (profiles/head_kernels.md)"}}
    if not args.no_cpu:
        from oracle import lapw as OL
        from oracle import ref
        # PyGW's per-pair calls (compiled for_q_0.f90) on a sample of pairs of k 0
        coefs = OL.band_coefficients(s.Aeig, *ref.gap2_set_lapwcoef(*s.coef_args(iop=1, ifrot=False)))
        al, be, ga = coefs
        vv0 = mvv[0].cpu().numpy()
        sample = [(ie1, ie2) for ie1 in (0, 7, 19) for ie2 in (20, 77, 139)]
        t0 = time.perf_counter()
        errs = []
        for ie1, ie2 in sample:
            p = ref.calcmmatvv_mt2(np.conj(al[ie1]), np.conj(be[ie1]), np.conj(ga[ie1]), al[ie2], be[ie2], ga[ie2],
                                   s.nLO_at, s.mult, *vv_t) + \
                ref.calcmmatvv_is(ie1, ie2, s.kl, s.Aeig, s.indgk, s.nvk, ipwint, g, i_g0, s.k2cartes)
            ph = ref.calcmmatvv_mt(ie2, ie1, al, be, ga, s.nLO_at, s.mult, *vv_t) + \
                ref.calcmmatvv_is(ie2, ie1, s.kl, s.Aeig, s.indgk, s.nvk, ipwint, g, i_g0, s.k2cartes)
            want = (p + np.conj(ph)) / 2
            errs.append(np.abs(vv0[ie2 - numin, ie1] - want).max() / np.abs(want).max())
        t1 = time.perf_counter()
        per_pair = (t1 - t0) / len(sample)
        t2 = time.perf_counter()
        ref.gap2_set_lapwcoef(*s.coef_args(iop=1, ifrot=False))
        t3 = time.perf_counter()
        rec["max_rel_error_vs_reference_sample"] = float(max(errs))
        rec["cpu_baseline"] = {"value": per_pair * (nomx + 1) * (nbmaxpol - numin) * nk + (t3 - t2) * nk,
                               "unit": "s", "cores": 1, "kind": "reference", **host_info(),
                               "sample": f"PyGW's compiled calcmmatvv_mt2 + calcmmatvv_is and their reverse (the "
                                         f"calls calc_head makes per pair) on {len(sample)} pairs of one k, x "
                                         f"{(nomx + 1) * (nbmaxpol - numin)} pairs x {nk} k, plus gap2_set_lapwcoef x "
                                         f"{nk}; the core elements and the head sums not timed"}
    print(json.dumps(rec))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over q-points is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from pygw_amd import polarization as P

    assert torch.cuda.is_available(), "bench needs a HIP device"
    if args.steps is None:
        args.steps = {"c5": 8, "c2": 8, "c4": 3, "gw0": 2, "gamma": 2, "coulomb": 2, "minm_is": 2, "mt_coulomb": 3, "minm_mt": 2, "head": 2}.get(args.config, 1)
    if args.copy_calibration:
        # one copy kernel that reads and writes exactly 1 GiB (one launch: under
        # 2 GiB), its source evicted from the 256 MB last-level cache by a 2 GiB
        # fill: the ratio of its FETCH_SIZE / WRITE_SIZE counters to 1 GiB
        # scales the kernels' counters
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        evict = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst, evict
    if args.config == "c5":
        run_chain(args)
        return
    if args.config == "c4":
        run_batch(args)
        return
    if args.config == "gw0":
        run_gw0(args)
        return
    if args.config == "gamma":
        run_gamma(args)
        return
    if args.config == "coulomb":
        run_coulomb(args)
        return
    if args.config == "minm_is":
        run_minm_is(args)
        return
    if args.config == "mt_coulomb":
        run_mt_coulomb(args)
        return
    if args.config == "minm_mt":
        run_minm_mt(args)
        return
    if args.config == "minm":
        run_minm(args)
        return
    if args.config == "head":
        run_head(args)
        return
    mixed = not args.no_mixed
    wl = Workload(args.config, args.nq, mixed, args.fp64_kfrac)
    wl.orank = None if args.omega_rank == "off" else args.omega_rank
    cfg, b = wl.base, wl.bands
    no, ne, pb, nom, nk = b.n_o, b.n_e, cfg.pb, cfg.n_om, cfg.nk
    steps = args.steps if args.config == "c2" else min(args.steps, 2)
    s64, st64 = timed(P, wl, "fp64", steps, args.warmup)
    print(f"[bench] fp64 {s64:.3f} s/step", file=sys.stderr, flush=True)
    kscale = nk / wl.nk64
    t_chi64 = st64[1] * kscale
    K = nk * no * ne
    f_chi = 4.0 * pb * pb * K * nom
    f_chi_launch = 4.0 * pb * pb * wl.nk64 * no * ne * nom
    kcw_bytes = 8.0 * no * ne * nk * nom + 8.0 * b.enk.size + 16.0 * 6 * nk + 4.0 * nk
    s_per_q64 = st64[0] + t_chi64 + st64[2]
    rec = {
        "metric": "chi0+eps^-1 seconds per q-point (FP64 GPU path; split-precision path alongside)",
        "value": s_per_q64, "unit": "s/q", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": s64 * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {'x'.join(map(str, cfg.div))} k-mesh, {no}+{ne} bands, {nom} Matsubara "
                               f"points, PB={pb}, {len(wl.items)} q-point(s)", "precision": "fp64",
                   "fp64_k_timed": wl.nk64},
        "stages_s": {"polarization_weights": st64[0], "chi0": t_chi64, "eps_inverse": st64[2]},
        "roofline": {"bound": "mfma", "kernel": "pygw_chi0_f64", "achieved": f_chi_launch / st64[1] / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_chi_launch / st64[1] / PEAK_FP64,
                     "traffic": None, "flops_per_launch": f_chi_launch},
        "weights_roofline": {"bound": "hbm", "kernel": "pygw_tw_gather", "achieved": kcw_bytes / st64[0] / 1e9,
                             "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": kcw_bytes / st64[0] / PEAK_HBM,
                             "bytes_per_launch": kcw_bytes},
    }
    if wl.nk64 != nk:
        rec["note"] = f"FP64 contraction timed on {wl.nk64} of {nk} k-points and scaled x{kscale:.2f} (cost is linear in N_k)"
    if mixed:
        smx, stmx = timed(P, wl, "mixed", steps, args.warmup)
        s_per_qmx = stmx.sum()
        f_mx = f_chi * (wl.rank_used / nom if wl.orank is not None else 1.0)   # the contraction run
        rec["mixed"] = {"s_per_q": s_per_qmx, "ms_per_step": smx * 1e3,
                        "stages_s": {"polarization_weights": stmx[0], "chi0": stmx[1], "eps_inverse": stmx[2]},
                        "speedup_vs_fp64": s_per_q64 / s_per_qmx, "chi0_speedup_vs_fp64": t_chi64 / stmx[1],
                        "roofline": {"bound": "mfma", "kernel": "pygw_chi0_mixed",
                                     "achieved": f_mx / stmx[1] / 1e12, "peak": PEAK_FP16 / 1e12,
                                     "unit": "TFLOP/s", "frac_algorithmic": f_mx / stmx[1] / PEAK_FP16,
                                     "frac_executed_3_products": 3 * f_mx / stmx[1] / PEAK_FP16,
                                     "frequency_by_frequency_equivalent_tflops": f_chi / stmx[1] / 1e12}}
        rec["mixed"]["rel_frobenius_error_vs_fp64"] = rel_errors(P, wl)
        if wl.orank is not None:
            rec["mixed"]["chi0_frequency_basis"] = {"mode": str(wl.orank), "rank": wl.rank_used, "n_om": nom}
            if args.uncompressed_steps > 0:
                sf, stf = timed(P, wl, "mixed_full", min(steps, args.uncompressed_steps), 1)
                rec["mixed"]["frequency_by_frequency"] = {
                    "s_per_q": float(stf.sum()),
                    "stages_s": {"polarization_weights": stf[0], "chi0": stf[1], "eps_inverse": stf[2]}}
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(wl, args.cpu_kslab)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
