"""Benchmark of the gfx950 polarization path: seconds per q-point.

One step = one q-point: tetrahedron weights -> chi0 contraction ->
dielectric inversion (calc_eps of pygw.py:3039-3297 with
Polarization_weights, pygw.py:3416-3476), on inputs (bands, meshes, the
matrix elements of every q) already resident in HBM.  With --config c5 the
step continues through MWM, the spline + G0 (x) W convolution and Pade
(calc_selfc, pygw.py:3299-3413; mcommon.py:80-210).

Default workload (no flags): 8x8x8 k-mesh, 16+48 bands, 64 Matsubara points
log-spaced 0.005-200 eV, PB = 600, the 8 q = (i/8,0,0), FP64 -- the
production-size FP64 case; the same steps are then repeated with the
split-precision contraction and the relative errors against FP64 reported.

    python bench.py [--steps K] [--warmup W] [--config c2|c3|c5] [--no-cpu]

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
PEAK_FP16 = 2.5e15        # MI355X dense FP16 MFMA, FLOP/s
PEAK_HBM = 8.0e12         # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--nq", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-kslab", type=int, default=2)
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
    """PyGW's own Fortran weights (oracle/_ref, OpenMP) for one q, plus the
    calc_eps numpy loop (per (k, w) ZGEMM, then inv per w) on a k-slab and a
    few frequencies, extrapolated linearly in N_k and n_om."""
    from oracle import chain, ref
    from pygw_amd import synth

    cfg, kq, _, _ = wl.items[0]
    b, om, tetc = wl.bands, wl.om, wl.tetc
    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    if ref.available("libref_tetra_omp.so"):
        kcw = ref.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, omp=True)
        kind = "reference"
    else:
        from oracle import tetra
        kcw, _, _ = tetra.bz_calcqdepw(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, nthreads=nthr)
        kind = "port"
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
    return {"value": t_w + t_chi + t_inv, "unit": "s/q", "cores": nthr, "kind": kind,
            "sample": (f"one q of {cfg.name}: weights in full (PyGW's Fortran bz_calcqdepw_par2, OpenMP, from "
                       f"oracle/_ref); chi0 = calc_eps numpy ZGEMM loop on {kslab} of {cfg.nk} k, extrapolated "
                       f"x{cfg.nk / kslab:.0f}; inverse on {nw} of {cfg.n_om} frequencies, extrapolated"
                       if kind == "reference" else "C restatement of the weights + numpy calc_eps loop"),
            "stages_s": {"weights": t_w, "chi0_extrapolated": t_chi, "inverse_extrapolated": t_inv}}


def run_chain(args):
    """Full chain at C5: weights -> chi0 -> eps^-1 -> MWM (8 GW bands x all
    bands, 64 output k = the 4x4x4 sub-lattice) -> spline onto the 20x
    precise tan mesh -> G0 (x) W convolution; Pade + QP after the q-loop.
    FP64 and split precision (chi0 and MWM) side by side."""
    import torch

    from pygw_amd import mesh, synth
    from pygw_amd import polarization as P

    base = synth.CONFIGS["c5"]
    b = synth.band_energies(base)
    fm = mesh.FrequencyMesh(base.n_om, base.wmin_ev / mesh.H2EV, base.wmax_ev / mesh.H2EV, iop_multiple=20)
    om, wo, omp, wop = fm.omega, fm.womeg, fm.omega_precise, fm.womeg_precise
    tetc = mesh.tetra_corners(base.div)
    kqid = mesh.k_minus_q(base.div)
    div = np.array(base.div)
    pts = np.array([[i, j, k] for i in range(div[0]) for j in range(div[1]) for k in range(div[2])])
    klist = np.where((pts % 2 == 0).all(1))[0]          # 4x4x4 sub-lattice, 64 k
    nbgw = 8
    ibgw = b.n_o - nbgw // 2                               # 4 highest occupied + 4 lowest empty
    nb2 = b.enk.shape[0]
    nq = args.nq or 1
    qs = [((i + 1) % div[0], 0, 0) for i in range(nq)]
    wkq = 1.0 / base.nk
    res = {}
    for prec in ("fp64", "mixed"):
        sig = torch.zeros((len(klist), nbgw, base.n_om), dtype=torch.complex128, device="cuda")
        times = []
        for q in qs:
            cfg = synth.PolConfig(**{**base.__dict__, "q": q})
            kq = kqid[:, synth.q_index(cfg)]
            M = synth.device_matrix_elements(cfg, b, kq, complex128=(prec == "fp64"))
            e2 = np.asfortranarray(np.hstack([b.enk[:, klist], b.enk[:, kq[klist]]]))
            b2 = synth.Bands(e2, b.ncg, b.nomx, b.numin)
            kq2 = np.arange(len(klist), 2 * len(klist), dtype=np.int32)
            kq2 = np.concatenate([kq2, kq2])
            Ms = synth.device_matrix_elements(cfg, b2, kq2, nkc=len(klist), complex128=(prec == "fp64"),
                                              rows=(ibgw, nbgw, 0, nb2), seed=cfg.seed + 99)
            enk_q = torch.as_tensor(np.ascontiguousarray(b.enk[:, kq[klist]].T), device="cuda")
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
            s = torch.cuda.current_stream()
            ev[0].record(s)
            kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native", check=False)
            ev[1].record(s)
            eps = P.chi0_accumulate(M, kcw, precision=prec)
            ev[2].record(s)
            P.eps_inverse(eps, check=False)
            ev[3].record(s)
            mwm = P.calc_mwm(Ms, eps, wkq, precision=prec)
            ev[4].record(s)
            sc = P.compute_selfc_inside(enk_q, mwm.view(len(klist), base.n_om, nbgw, nb2), om, wo, omp, wop)
            ev[5].record(s)
            sig += sc
            torch.cuda.synchronize()
            times.append([ev[i].elapsed_time(ev[i + 1]) / 1e3 for i in range(5)])
            del M, Ms, eps, kcw, mwm
        ene = np.ascontiguousarray(b.enk[ibgw:ibgw + nbgw][:, klist].T).reshape(-1)
        t0 = time.perf_counter()
        qp = P.pade_quasiparticles(sig.view(-1, base.n_om), om, ene).cpu().numpy()
        t_pade = time.perf_counter() - t0
        tm = np.array(times).mean(0)
        res[prec] = {"s_per_q": float(tm.sum()), "sig": sig.cpu().numpy(), "qp": qp,
                     "stages_s": dict(zip(["polarization_weights", "chi0", "eps_inverse", "mwm", "selfc_convolution"],
                                          map(float, tm))), "pade_qp_s": t_pade}
    s64, smx = res["fp64"], res["mixed"]
    rel_sig = np.linalg.norm(smx["sig"] - s64["sig"], axis=(0, 1)) / np.linalg.norm(s64["sig"], axis=(0, 1))
    dqp = np.abs(smx["qp"][:, 5] - s64["qp"][:, 5]) * mesh.H2EV * 1e3
    nk, no, ne, pb, nom = base.nk, b.n_o, b.n_e, base.pb, base.n_om
    f_chi = 4.0 * pb * pb * nk * no * ne * nom
    rec = {"metric": "chi0+eps^-1+Sigma_c seconds per q-point (full chain)", "value": s64["s_per_q"], "unit": "s/q",
           "n_gpus": 1, "steps": len(qs), "warmup": 0, "ms_per_step": s64["s_per_q"] * 1e3, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"c5: 8x8x8 k-mesh, {no}+{ne} bands, PB={pb}, {nom} tan-mesh Matsubara points, "
                                  f"Sigma_c of {nbgw} bands at {len(klist)} k, precise mesh x20 ({omp.size} points)",
                      "precision": "fp64"},
           "stages_s": s64["stages_s"],
           "roofline": {"bound": "mfma", "kernel": "pygw_chi0_f64", "achieved": f_chi / s64["stages_s"]["chi0"] / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                        "frac": f_chi / s64["stages_s"]["chi0"] / PEAK_FP64, "traffic": None},
           "mixed": {"s_per_q": smx["s_per_q"], "stages_s": smx["stages_s"],
                     "speedup_vs_fp64": s64["s_per_q"] / smx["s_per_q"],
                     "sigma_c_rel_frobenius_vs_fp64_max": float(rel_sig.max()),
                     "qp_shift_abs_diff_meV_max": float(dqp.max()), "qp_shift_abs_diff_meV_mean": float(dqp.mean())}}
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
    if args.config == "c5":
        run_chain(args)
        return
    mixed = not args.no_mixed
    if args.copy_calibration:
        # one copy kernel that reads and writes exactly 1 GiB: the ratio of its
        # FETCH_SIZE / WRITE_SIZE counters to 1 GiB scales the kernels' counters
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    wl = Workload(args.config, args.nq, mixed, args.fp64_kfrac)
    cfg, b = wl.base, wl.bands
    no, ne, pb, nom, nk = b.n_o, b.n_e, cfg.pb, cfg.n_om, cfg.nk
    steps = args.steps if args.config == "c2" else min(args.steps, 2)
    s64, st64 = timed(P, wl, "fp64", steps, args.warmup)
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
        rec["mixed"] = {"s_per_q": s_per_qmx, "ms_per_step": smx * 1e3,
                        "stages_s": {"polarization_weights": stmx[0], "chi0": stmx[1], "eps_inverse": stmx[2]},
                        "speedup_vs_fp64": s_per_q64 / s_per_qmx, "chi0_speedup_vs_fp64": t_chi64 / stmx[1],
                        "roofline": {"bound": "mfma", "kernel": "pygw_chi0_mixed",
                                     "achieved": f_chi / stmx[1] / 1e12, "peak": PEAK_FP16 / 1e12,
                                     "unit": "TFLOP/s", "frac_algorithmic": f_chi / stmx[1] / PEAK_FP16,
                                     "frac_executed_3_products": 3 * f_chi / stmx[1] / PEAK_FP16}}
        rec["mixed"]["rel_frobenius_error_vs_fp64"] = rel_errors(P, wl)
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(wl, args.cpu_kslab)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
