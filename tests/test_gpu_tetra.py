"""GPU parity of the tetrahedron-weight kernel against the oracle.

The closed forms of PyGW's weights are ill conditioned for nearly equal
corner Delta's and for w close to the Taylor switch (10 max|Delta|): there
the reference itself loses many digits, and a device atan/log differing by
one ulp from the host libm moves the result by the same amount.  The
tolerance is therefore measured, not guessed: the oracle is re-run with
atan/log perturbed by +-1 ulp (eight seeds, oracle.tetra.with_spread), and
the device result must stay within 8x that spread plus 1e-13 of the largest
|kcw| of its frequency.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import tetra
from pygw_amd import mesh, synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _check(gpu, enk, kq, win, om, tetc, compat=(1, 1), name=""):
    o0, spread, _ = tetra.with_spread(enk, 0.0, kq, *win, om, tetc, compat=compat)
    ratio = tetra.within_spread(gpu, o0, spread)
    dev = np.abs(gpu - o0)
    print(f"[{name}] max|gpu-oracle|={dev.max():.3e} max|oracle|={np.abs(o0).max():.3e} "
          f"max dev/tol={ratio:.3f} exact-equal={np.mean(gpu == o0):.3f}", flush=True)
    assert ratio <= 1.0
    return o0


def _gpu(enk, kq, win, om, tetc, compat=3, layout="fortran"):
    from pygw_amd import polarization as P
    k = P.bz_calcqdepw_par2(enk, 0.0, kq, *win, om, tetc, compat=compat, layout=layout)
    return k.cpu().numpy(), P.bz_calcqdepw_par2.status


@pytest.mark.parametrize("name", ["tetra_insulator", "tetra_metal", "tetra_metal_b"])
def test_gpu_weights_golden(name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    win = tuple(int(x) for x in g["win"])
    print(f"case {name}", flush=True)
    kcw, st = _gpu(g["enk"], g["kqid"], win, g["omega"], g["tetc"])
    assert st & 1 == 0
    o0 = _check(kcw, g["enk"], g["kqid"], win, g["omega"], g["tetc"], name=name)
    np.testing.assert_array_equal(o0, g["kcw"])  # oracle == PyGW's Fortran


def test_gpu_weights_c1():
    cfg = synth.CONFIGS["c1"]
    b = synth.band_energies(cfg)
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, synth.q_index(cfg)]
    win = (b.ncg, b.nomx, b.numin)
    print("case c1", flush=True)
    kcw, st = _gpu(b.enk, kq, win, om, tetc)
    assert st == 0
    _check(kcw, b.enk, kq, win, om, tetc, name="c1")
    # kernel-native layout holds the same numbers, pair p = i_o*n_e + i_e
    nat, _ = _gpu(b.enk, kq, win, om, tetc, layout="native")
    no, ne = b.n_o, b.n_e
    ref = kcw.transpose(2, 3, 0, 1).reshape(cfg.nk, cfg.n_om, no * ne)
    np.testing.assert_array_equal(nat, ref)


@pytest.mark.parametrize("seed", [5, 6])
def test_gpu_weights_metal(seed):
    cfg = synth.small_config(div=(4, 4, 4), n_occ=4, n_emp=6, n_metal=3, n_om=10, seed=seed)
    b = synth.band_energies(cfg)
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, 3]
    win = (b.ncg, b.nomx, b.numin)
    print(f"case metal seed={seed}", flush=True)
    for compat, ccompat in ((3, (1, 1)), (0, (0, 0))):
        kcw, st = _gpu(b.enk, kq, win, om, tetc, compat=compat)
        assert st & 1 == 0
        _check(kcw, b.enk, kq, win, om, tetc, compat=ccompat, name=f"metal{seed}/compat{compat}")


def test_gpu_flat_band():
    div = (4, 3, 2)
    nk = int(np.prod(div))
    d = 0.41
    enk = np.asfortranarray(np.vstack([np.full(nk, -0.2), np.full(nk, -0.2 + d)]))
    om = np.array([0.001, 0.05, 0.3, 1.0, 3.0, 4.09, 4.11, 40.0])
    tetc = mesh.tetra_corners(div)
    kq = mesh.k_minus_q(div)[:, 7]
    print("case flat band", flush=True)
    kcw, st = _gpu(enk, kq, (0, 0, 1), om, tetc)
    from oracle import cubature
    exact = cubature.flat_band_kcw(d, om, nk)
    taylor = (-2 * d / om ** 2 + 2 * d ** 3 / om ** 4) / nk
    want = np.where(om > 10 * d, taylor, exact)
    np.testing.assert_allclose(kcw[0, 0], np.broadcast_to(want, (nk, om.size)), rtol=1e-12)


def test_gpu_weights_svd_basis():
    """bz_calcqdepw_par3: weights projected on an SVD basis (pygw.py:3478-3501)."""
    cfg = synth.small_config(div=(3, 3, 3), n_occ=3, n_emp=5, n_om=12, seed=41)
    b = synth.band_energies(cfg)
    fm = mesh.FrequencyMesh(12, 0.01, 5.0, iop_multiple=2)
    om, wo = fm.omega, fm.womeg
    _, dUl = mesh.svd_functions(om, wo, 0.005, 10.0, 40, 1e-6)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, 4]
    win = (b.ncg, b.nomx, b.numin)
    from pygw_amd import polarization as P
    print(f"case svd n_l={dUl.shape[1]}", flush=True)
    got = P.bz_calcqdepw_par3(b.enk, 0.0, kq, *win, dUl, om, tetc).cpu().numpy()
    o0, spread, _ = tetra.with_spread(b.enk, 0.0, kq, *win, om, tetc)
    want = o0 @ dUl
    tol = 8 * (spread @ np.abs(dUl)) + 1e-12 * np.abs(want).max()
    assert (np.abs(got - want) <= tol).all()
    L = tetra.lib()
    kl, st, _ = tetra.bz_calcqdepw(b.enk, 0.0, kq, *win, om, tetc, dUl=dUl)
    assert st >= 0
    np.testing.assert_allclose(kl, want, rtol=0, atol=1e-12 * np.abs(want).max())
    del L


def test_gpu_weights_pair_chunks():
    """A workspace of one 64-pair chunk (three chunks here) gives the same
    bits as the default single chunk; both match the oracle."""
    import torch
    from pygw_amd import _lib
    from pygw_amd import polarization as P
    cfg = synth.small_config(div=(4, 4, 4), n_occ=8, n_emp=20, n_metal=3, n_om=7, seed=17)
    b = synth.band_energies(cfg)
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, 5]
    win = (b.ncg, b.nomx, b.numin)
    assert b.n_o * b.n_e > 128
    sz = C.c_size_t(0)
    _lib.call("pygw_tetra_weights_workspace", cfg.nk, 64, om.size, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    print(f"case pair chunks np={b.n_o * b.n_e} ws={sz.value}", flush=True)
    full = P.bz_calcqdepw_par2(b.enk, 0.0, kq, *win, om, tetc).cpu().numpy()
    for layout in ("fortran", "native"):
        a = P.bz_calcqdepw_par2(b.enk, 0.0, kq, *win, om, tetc, layout=layout).cpu().numpy()
        c = P.bz_calcqdepw_par2(b.enk, 0.0, kq, *win, om, tetc, layout=layout, workspace=ws).cpu().numpy()
        np.testing.assert_array_equal(a, c)
    _check(full, b.enk, kq, win, om, tetc, name="pair chunks")
    with pytest.raises(_lib.PygwError):
        P.bz_calcqdepw_par2(b.enk, 0.0, kq, *win, om, tetc, workspace=ws[: sz.value // 2])
