"""Benchmark of the gfx950 polarization path: seconds per q-point.

One step = one q-point of the workload: tetrahedron weights -> chi0
contraction -> dielectric inversion (calc_eps of pygw.py:3039-3297 with
Polarization_weights, pygw.py:3416-3476), inputs (bands, meshes, matrix
elements of every q) resident in HBM before the timed region.  Default
workload: 8x8x8 k-mesh, 16+48 bands, 64 Matsubara points, PB = 600, the 8 q
= (i/8,0,0), FP64 (the "plain FP64 on GPU at production size" case).

    python bench.py [--steps K] [--warmup W] [--config c2] [--no-cpu]

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
PEAK_HBM = 8.0e12         # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--nq", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kslab", type=int, default=2)
    return ap.parse_args()


def setup(cfg_name, nq):
    import torch

    from pygw_amd import mesh, synth

    base = synth.CONFIGS[cfg_name]
    div = base.div
    qs = [(i * div[0] // nq, 0, 0) for i in range(nq)]
    bands = synth.band_energies(base)
    om, _ = synth.omega_mesh(base)
    tetc = mesh.tetra_corners(div)
    kqid = mesh.k_minus_q(div)
    items = []
    for q in qs:
        cfg = synth.PolConfig(**{**base.__dict__, "q": q})
        kq = kqid[:, synth.q_index(cfg)]
        M = synth.device_matrix_elements(cfg, bands, kq, complex128=True)
        items.append((cfg, kq, M))
    torch.cuda.synchronize()
    return base, bands, om, tetc, items


def run_step(P, bands, om, tetc, kq, M, ev=None):
    import torch
    s = torch.cuda.current_stream()
    if ev is not None:
        ev[0].record(s)
    kcw = P.bz_calcqdepw_par2(bands.enk, 0.0, kq, bands.ncg, bands.nomx, bands.numin, om, tetc,
                              layout="native", check=False)
    if ev is not None:
        ev[1].record(s)
    eps = P.chi0_accumulate(M, kcw)
    if ev is not None:
        ev[2].record(s)
    P.eps_inverse(eps, check=False)
    if ev is not None:
        ev[3].record(s)
    return eps


def cpu_baseline(cfg, bands, om, tetc, kq, kslab):
    """PyGW's own Fortran weights (oracle/_ref, OpenMP) for one q, plus the
    calc_eps numpy loop (per (k, w) ZGEMM, then inv per w) on a k-slab and a
    few frequencies, extrapolated linearly in N_k and n_om."""
    from oracle import chain, ref

    nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"cores": nthr}
    t = time.perf_counter()
    if ref.available("libref_tetra_omp.so"):
        kcw = ref.bz_calcqdepw_par2(bands.enk, 0.0, kq, bands.ncg, bands.nomx, bands.numin, om, tetc, omp=True)
        kind = "reference"
    else:
        from oracle import tetra
        kcw, _, _ = tetra.bz_calcqdepw(bands.enk, 0.0, kq, bands.ncg, bands.nomx, bands.numin, om, tetc,
                                       nthreads=nthr)
        kind = "port"
    t_w = time.perf_counter() - t
    from pygw_amd import synth
    M = synth.matrix_elements(cfg, bands, kq, ks=range(kslab))
    no, ne = bands.n_o, bands.n_e
    t = time.perf_counter()
    eps = chain.calc_eps_accumulate(((i, M[i]) for i in range(kslab)),
                                    lambda ik: kcw[:, :, ik, :].reshape(no * ne, -1))
    t_chi = (time.perf_counter() - t) * cfg.nk / kslab
    nw = 4
    t = time.perf_counter()
    chain.eps_inverse(eps[:, :, :nw])
    t_inv = (time.perf_counter() - t) * cfg.n_om / nw
    total = t_w + t_chi + t_inv
    out.update({"value": total, "unit": "s/q", "kind": kind,
                "sample": (f"one q of {cfg.name}: weights timed in full ({'PyGW Fortran, OpenMP' if kind == 'reference' else 'C restatement'}); "
                           f"chi0 numpy ZGEMM loop of calc_eps on {kslab} of {cfg.nk} k extrapolated x{cfg.nk / kslab:.0f}; "
                           f"inverse on {nw} of {cfg.n_om} frequencies extrapolated"),
                "stages_s": {"weights": t_w, "chi0(extrap)": t_chi, "inverse(extrap)": t_inv}})
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding over q is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from pygw_amd import polarization as P

    assert torch.cuda.is_available(), "bench needs a HIP device"
    cfg, bands, om, tetc, items = setup(args.config, args.nq)
    no, ne, pb, nom, nk = bands.n_o, bands.n_e, cfg.pb, cfg.n_om, cfg.nk
    for i in range(args.warmup):
        c, kq, M = items[i % len(items)]
        run_step(P, bands, om, tetc, kq, M)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        c, kq, M = items[i % len(items)]
        run_step(P, bands, om, tetc, kq, M, evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = np.array([[e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]),
                    e[2].elapsed_time(e[3])] for e in evs]) / 1e3
    mean = st.mean(axis=0)
    t_w, t_chi, t_inv = mean
    ms_step = wall / args.steps * 1e3
    K = nk * no * ne
    f_chi = 4.0 * pb * pb * K * nom
    kcw_bytes = 8.0 * no * ne * nk * nom + 8.0 * bands.enk.size + 16.0 * 6 * nk + 4.0 * nk
    rec = {
        "metric": "chi0+eps^-1 seconds per q-point (FP64 GPU path)",
        "value": wall / args.steps, "unit": "s/q", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{cfg.name}: {'x'.join(map(str, cfg.div))} k-mesh, {no}+{ne} bands, "
                               f"{nom} Matsubara points, PB={pb}, {len(items)} q-points", "precision": "fp64"},
        "stages_s": {"polarization_weights": t_w, "chi0": t_chi, "eps_inverse": t_inv},
        "roofline": {"bound": "mfma", "kernel": "pygw_chi0_f64", "achieved": f_chi / t_chi / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": f_chi / t_chi / PEAK_FP64,
                     "traffic": None, "flops_per_launch": f_chi},
        "weights_roofline": {"bound": "hbm", "achieved": kcw_bytes / t_w / 1e9, "peak": PEAK_HBM / 1e9,
                             "unit": "GB/s", "frac": kcw_bytes / t_w / PEAK_HBM},
    }
    if not args.no_cpu:
        c, kq, M = items[1 % len(items)]
        rec["cpu_baseline"] = cpu_baseline(c, bands, om, tetc, kq, args.cpu_kslab)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
