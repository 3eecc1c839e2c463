"""GPU parity of the chi0 contraction and the dielectric inversion against
the numpy restatement of calc_eps (oracle/chain.py)."""
import numpy as np
import pytest

from oracle import chain
from pygw_amd import mesh, synth

pytestmark = pytest.mark.gpu


def _inputs(cfg, ks=None):
    b = synth.band_energies(cfg)
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, synth.q_index(cfg)]
    M = synth.matrix_elements(cfg, b, kq, ks=ks)
    return b, om, tetc, kq, M


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(2, 3, 5, 70, 2), (3, 4, 7, 64, 5), (2, 5, 3, 129, 3)])
def test_chi0_f64_vs_oracle(shape):
    """Ragged sizes: PB not a multiple of the 64 tile, pairs not a multiple
    of the 16-pair stage, odd frequency counts."""
    import torch
    from pygw_amd import polarization as P
    d, no, ne, pb, nom = shape
    cfg = synth.small_config(div=(d, d, 2), n_occ=no, n_emp=ne, n_om=nom, pb=pb, seed=17 + pb)
    b, om, tetc, kq, M = _inputs(cfg)
    print(f"case chi0 shape={shape}", flush=True)
    kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
    eps = P.chi0_accumulate(M, kcw).cpu().numpy()
    kh = kcw.cpu().numpy()
    want = chain.calc_eps_accumulate(((ik, M[ik]) for ik in range(cfg.nk)), lambda ik: kh[ik].T)
    assert _rel(eps, want) < 1e-13
    # accumulate over two k halves == one call
    h = cfg.nk // 2
    e2 = P.chi0_accumulate(M[:h], kcw)
    P.chi0_accumulate(M[h:], kcw, eps=e2, k0=h)
    assert _rel(e2.cpu().numpy(), want) < 1e-13
    torch.cuda.synchronize()


def test_eps_inverse_vs_oracle():
    import torch
    from pygw_amd import polarization as P
    cfg = synth.small_config(div=(3, 3, 2), n_occ=3, n_emp=6, n_om=4, pb=97, seed=3, vhalf="coulomb")
    b, om, tetc, kq, M = _inputs(cfg)
    print("case eps_inverse pb=97", flush=True)
    kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
    eps = P.chi0_accumulate(M, kcw)
    e_host = eps.cpu().numpy()
    P.eps_inverse(eps)
    got = eps.cpu().numpy()
    want = chain.eps_inverse(e_host)
    for w in range(cfg.n_om):
        a = (e_host[:, :, w] + e_host[:, :, w].conj().T) / 2 + np.eye(cfg.pb)
        cond = np.linalg.cond(a)
        err = _rel(got[:, :, w] + np.eye(cfg.pb), want[:, :, w] + np.eye(cfg.pb))
        print(f"  w={w} cond={cond:.2e} rel={err:.2e}", flush=True)
        assert err < 1e-15 * cond * 50
        np.testing.assert_allclose(a @ (got[:, :, w] + np.eye(cfg.pb)), np.eye(cfg.pb), atol=1e-12 * cond)
    torch.cuda.synchronize()


def test_eps_inverse_rejects_indefinite():
    import torch
    from pygw_amd import _lib
    from pygw_amd import polarization as P
    pb, nom = 40, 2
    a = np.zeros((pb, pb, nom), dtype=complex, order="F")
    a[:, :, 0] = np.diag(np.linspace(0.5, 2.0, pb))
    a[:, :, 1] = np.diag(np.linspace(0.5, 2.0, pb))
    a[7, 7, 1] = -3.0  # 1 + a has a negative pivot
    t = torch.as_tensor(a.reshape(-1, order="F"), device="cuda").as_strided((pb, pb, nom), (1, pb, pb * pb))
    with pytest.raises(_lib.PygwError):
        P.eps_inverse(t)


# PB 96 / 30 / 60: one partial 128 tile; PB 129 / 170: the last tile
# column is 1 / 42 columns wide
@pytest.mark.parametrize("shape", [(3, 4, 7, 96, 5), (2, 5, 3, 129, 3), (2, 4, 5, 30, 3), (2, 3, 5, 60, 2),
                                   (2, 3, 5, 170, 2)])
def test_chi0_mixed_vs_fp64(shape):
    """Split-precision contraction: relative Frobenius error per frequency
    <= 1e-6 against the FP64 oracle (the bar the precision scheme is built
    for), and the Hermitian mirror is exact."""
    import torch
    from pygw_amd import polarization as P
    d, no, ne, pb, nom = shape
    cfg = synth.small_config(div=(d, d, 2), n_occ=no, n_emp=ne, n_om=nom, pb=pb, seed=29 + pb, vhalf="coulomb")
    b, om, tetc, kq, M = _inputs(cfg)
    print(f"case chi0 mixed shape={shape}", flush=True)
    kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
    kh = kcw.cpu().numpy()
    m64 = M.astype(np.complex64)
    want = chain.calc_eps_accumulate(((ik, m64[ik].astype(complex)) for ik in range(cfg.nk)), lambda ik: kh[ik].T)
    got = P.chi0_accumulate(m64, kcw, precision="mixed").cpu().numpy()
    for w in range(nom):
        err = _rel(got[:, :, w], want[:, :, w])
        print(f"  w={w} rel_frob={err:.2e}", flush=True)
        assert err < 1e-6
    lower = np.tril(np.ones((pb, pb), bool), -1)
    tile = (np.arange(pb)[:, None] // 128) > (np.arange(pb)[None, :] // 128)
    mir = np.conj(got.transpose(1, 0, 2))
    assert np.array_equal(got[tile & lower], mir[tile & lower])
    # chunked accumulation over k equals one call to the same precision bar
    h = cfg.nk // 2
    e2 = P.chi0_accumulate(m64[:h], kcw, precision="mixed")
    P.chi0_accumulate(m64[h:], kcw, precision="mixed", eps=e2, k0=h)
    assert _rel(e2.cpu().numpy(), want) < 1e-6
    # the smallest workspace (one k-point of fragment planes) walks k in
    # slabs of one k-point: same result to the same bar
    import ctypes as C
    from pygw_amd import _lib
    sz = C.c_size_t(0)
    _lib.call("pygw_chi0_workspace", pb, no * ne, nom, 1, 1, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    Md = torch.as_tensor(m64, device="cuda")
    e3d = torch.empty(pb * pb * nom, dtype=torch.complex128, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (C.c_void_p(Md.data_ptr()), no * ne * pb, no * ne, pb, C.c_void_p(kcw.data_ptr()), nom, 0, cfg.nk,
            -2.0, 1, 0, C.c_void_p(e3d.data_ptr()), C.c_void_p(ws.data_ptr()))
    _lib.call("pygw_chi0_accumulate", *args, sz.value, stream)
    e3 = e3d.view(nom, pb, pb).permute(2, 1, 0).cpu().numpy()
    assert _rel(e3, want) < 1e-6
    # half of it has no room for the part slots of the split last round: that
    # round runs unsplit, to the same bar
    e3d.zero_()
    _lib.call("pygw_chi0_accumulate", *args, sz.value // 2, stream)
    e4 = e3d.view(nom, pb, pb).permute(2, 1, 0).cpu().numpy()
    assert _rel(e4, want) < 1e-6
    # below one k-point of planes the call is rejected before any launch
    with pytest.raises(_lib.PygwError):
        _lib.call("pygw_chi0_accumulate", *args, 4096, stream)
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chi0_schedules_and_autotune(precision):
    """Every tile schedule (1, 2, 4 frequencies per block) gives the same
    contraction; the autotuner picks one of them and caches it."""
    import torch
    from pygw_amd import polarization as P
    cfg = synth.small_config(div=(2, 3, 2), n_occ=3, n_emp=5, n_om=7, pb=80, seed=5)
    b, om, tetc, kq, M = _inputs(cfg)
    print(f"case schedules {precision}", flush=True)
    kcw = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
    kh = kcw.cpu().numpy()
    Mp = M.astype(np.complex64) if precision == "mixed" else M
    want = chain.calc_eps_accumulate(((ik, Mp[ik].astype(complex)) for ik in range(cfg.nk)), lambda ik: kh[ik].T)
    tol = 1e-13 if precision == "fp64" else 1e-6
    for nw in (1, 2, 4):
        P.set_chi0_schedule(precision, cfg.pb, b.n_o * b.n_e, cfg.n_om, nw)
        got = P.chi0_accumulate(Mp, kcw, precision=precision).cpu().numpy()
        print(f"  nw={nw} rel={_rel(got, want):.2e}", flush=True)
        assert _rel(got, want) < tol
    best, times = P.autotune_chi0(Mp, kcw, precision=precision)
    print(f"  autotune best={best} times={times}", flush=True)
    assert best in ((1, 2) if precision == "mixed" else (1, 2, 4))
    got = P.chi0_accumulate(Mp, kcw, precision=precision).cpu().numpy()
    assert _rel(got, want) < tol
    P.set_chi0_schedule(precision, cfg.pb, b.n_o * b.n_e, cfg.n_om, 0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_qplan_graph_replays(precision):
    """The per-q steps captured in a HIP graph: every replay (with eager work
    and synchronisation between replays) equals the eager result."""
    import torch
    from pygw_amd import polarization as P
    cfg = synth.small_config(div=(3, 2, 2), n_occ=3, n_emp=4, n_om=6, pb=72, seed=13, vhalf="coulomb")
    b, om, tetc, kq, M = _inputs(cfg)
    print(f"case graph {precision}", flush=True)
    plan = P.QPlan(b.enk, 0.0, b.ncg, b.nomx, b.numin, om, tetc, cfg.pb, precision=precision)
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    Md = torch.as_tensor(M, device="cuda").to(dt).contiguous()
    plan.set_q(kq)
    eager = plan.run(Md).clone()
    torch.cuda.synchronize()
    g = plan.capture(Md)
    for rep in range(3):
        plan.eps.zero_()
        _ = torch.randn(1000, device="cuda").sum()   # unrelated eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(plan.eps, eager), f"replay {rep} differs"
