"""GPU parity of the GW0 iteration on stored mwm (pygw_amd.gw0) against the
numpy restatement of PyGW's post-processing (oracle/gw0.py)."""
import numpy as np
import pytest

from oracle import gw0 as G
from pygw_amd import mesh, synth

pytestmark = pytest.mark.gpu


def _system(div=(2, 2, 1), n_om=12, mul=4, q_list=None):
    return synth.gw0_system(n_occ=2, n_emp=6, nbgw=4, div=div, n_om=n_om, iop_multiple=mul, q_list=q_list)


def test_tetra_idos_vs_oracle():
    from pygw_amd import gw0
    div = (3, 3, 2)
    rng = np.random.default_rng(3)
    nk = int(np.prod(div))
    e = np.sort(rng.uniform(-0.3, 0.3, (nk, 5)), axis=1)
    e[:, 2] = 0.1          # a flat band (degenerate tetrahedra)
    atet = mesh.tetra_corners(div).T.astype(np.int32)
    wtet = np.full(atet.shape[0], 1.0 / atet.shape[0])
    print("case idos 3x3x2 5 bands", flush=True)
    d = gw0.TetraDos(atet, wtet)
    d.set_bands(e)
    xs = np.concatenate([np.linspace(-0.35, 0.35, 41), e[[0, 5, 9], [0, 2, 4]]])
    i_g, d_g = d(xs)
    for x, ig, dg in zip(xs, i_g, d_g):
        assert abs(ig - G.idos(x, e, atet, wtet)) < 1e-13
        assert abs(dg - G.dostet(x, e, atet, wtet)) < 1e-10 * max(1.0, abs(dg))


def test_calc_fermi_vs_oracle():
    from pygw_amd import gw0
    div = (3, 3, 2)
    atet = mesh.tetra_corners(div).T.astype(np.int32)
    wtet = np.full(atet.shape[0], 1.0 / atet.shape[0])
    rng = np.random.default_rng(8)
    nk = int(np.prod(div))
    s = np.cos(2 * np.pi * mesh.kpoint_coords(div)).sum(axis=1) / 3.0
    ins = np.sort(np.array([-0.3, -0.2, 0.2, 0.3])[None, :] + 0.05 * s[:, None], axis=1)
    met = np.sort(rng.uniform(-0.3, 0.3, 5)[None, :] + 0.1 * s[:, None], axis=1)
    d = gw0.TetraDos(atet, wtet)
    print("case fermi", flush=True)
    for e, nval in ((ins, 4), (met, 5), (met, 4)):
        got = gw0.calc_fermi(e, d, nval)
        want = G.calc_fermi(e, atet, wtet, nval)
        print("  ", nval, got, want, flush=True)
        assert abs(got[0] - want[0]) < 1e-10
        for a, b in zip(got[1:], want[1:]):
            assert abs(a - b) < 1e-8 * max(1.0, abs(b))


@pytest.mark.parametrize("iop_es,iop_gw0", [(0, 1), (-1, 1), (-1, 2), (-1, 3)])
def test_quasiparticles_vs_oracle(iop_es, iop_gw0):
    from pygw_amd import gw0
    g = _system()
    rng = np.random.default_rng(4)
    nk, nb1 = g.ebnd.shape[0], 4
    a = 0.02 * rng.uniform(0.5, 1.5, (nk, nb1, 2, 1))
    w = 1j * g.omega[None, None, :]
    sig = a[:, :, 0] / (w + 0.8) + a[:, :, 1] / (w - 0.8)
    b0 = g.ebnd[:, :nb1]
    b = b0 + 0.01 * rng.normal(size=b0.shape)
    print(f"case qp iop_es={iop_es} iop_gw0={iop_gw0}", flush=True)
    got, got_im = gw0.quasiparticles(b, b0, sig, g.sigx, g.vxc, g.omega, iop_es, iop_gw0)
    want, want_im = G.quasiparticles(b, b0, sig, g.sigx, g.vxc, g.omega, iop_es, iop_gw0)
    assert np.abs(got - want).max() < 1e-11
    assert np.abs(got_im - want_im).max() < 1e-11


def test_selfenergy_all_q_vs_oracle():
    """Sigma_c summed over q at shifted bands, walking q in chunks of 3 (the
    last chunk partial) and in one chunk: both equal the oracle."""
    import torch
    from pygw_amd import gw0
    g = _system()
    rng = np.random.default_rng(9)
    bands = g.ebnd + 0.005 * rng.normal(size=g.ebnd.shape)
    nq, nk, nom, nb1, nb2 = g.mwm.shape
    want = G.compute_selfc(bands, g.core, g.jrk, g.mwm, g.omega, g.womeg, g.omega_p, g.womeg_p)
    per_q = nk * nb1 * nb2 * g.omega_p.size * 16
    print(f"case selfe nq={nq} nk={nk}", flush=True)
    for chunk, cache in ((3 * per_q, False), (1 << 30, False), (3 * per_q, True)):
        se = gw0.SelfEnergy(g.mwm, g.jrk, g.core, g.omega, g.omega_p, g.womeg_p, chunk_bytes=chunk, cache=cache)
        got = se(bands).cpu().numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"  chunk q={se.qc} cache={cache} rel={err:.2e}", flush=True)
        assert err < 1e-12
        # a second call (splines reused when cached) at other bands
        b2 = bands + 0.003
        want2 = G.compute_selfc(b2, g.core, g.jrk, g.mwm, g.omega, g.womeg, g.omega_p, g.womeg_p)
        assert np.abs(se(b2).cpu().numpy() - want2).max() < 1e-12 * np.abs(want2).max()
    # accumulate form: out = 0.25 out + 0.5 sum_q
    acc = torch.as_tensor(want, device="cuda").clone()
    se(bands, out=acc, alpha=0.5, beta=0.25)
    assert np.abs(acc.cpu().numpy() - 0.75 * want).max() < 1e-12 * np.abs(want).max()


def test_gw0_loop_vs_oracle(tmp_path):
    """The whole GW0 loop, with the mwm stack written to and read back from
    PyGW's mwm.{iq}.{irk}.npy files."""
    from pygw_amd import gw0
    g = _system()
    nq, nk = g.mwm.shape[:2]
    for iq in range(nq):
        for irk in range(nk):
            gw0.save_mwm(str(tmp_path), iq, irk, g.mwm[iq, irk])
    stack = gw0.load_mwm_stack(str(tmp_path), list(range(nq)), nk)
    assert np.array_equal(stack.cpu().numpy(), g.mwm)
    print("case gw0 loop", flush=True)
    sig0 = G.compute_selfc(g.ebnd, g.core, g.jrk, g.mwm, g.omega, g.womeg, g.omega_p, g.womeg_p)
    want = G.gw0(g.ebnd, g.core, g.jrk, g.mwm, sig0, g.sigx, g.vxc, g.omega, g.womeg, g.omega_p, g.womeg_p,
                 g.atet, g.wtet, g.nval, g.ibgw, nmax_sc=12, eps_sc=1e-3, mix=0.7)
    se = gw0.SelfEnergy(stack, g.jrk, g.core, g.omega, g.omega_p, g.womeg_p)
    s0 = se(g.ebnd)
    assert np.abs(s0.cpu().numpy() - sig0).max() < 1e-12 * np.abs(sig0).max()
    got = gw0.gw0(g.ebnd, g.core, se, s0, g.sigx, g.vxc, g.omega, g.atet, g.wtet, g.nval, g.ibgw, nmax_sc=12,
                  eps_sc=1e-3, mix_sc=0.7)
    print("  history", [(h[0], round(h[1], 6)) for h in got[2]], flush=True)
    assert np.abs(got[0] - want[0]).max() < 1e-10
    assert (got[1] is None) == (want[1] is None) and want[1] is not None
    assert np.abs(got[1] - want[1]).max() < 1e-9
    assert len(got[2]) == len(want[2])
    for a, b in zip(got[2], want[2]):
        assert a[0] == b[0] and abs(a[1] - b[1]) < 1e-9
