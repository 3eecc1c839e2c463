"""The matrix elements built from the LAPW coefficients, fed through the q
loop: for every k of a 2x2x2 mesh, gap2_set_lapwcoef at k and k - q, the
band coefficients, calc_minm_MT + calc_minm_IS into one mmat, the core part
(calc_minc2 for calc_eps, calc_minc for calc_selfc) and the products with
conj(barcvm) of coul_setev (pygw.py:2547-2700, 3096-3101, 3339-3345); then
weights -> chi0 -> eps^-1 -> MWM -> spline + convolution (pygw.py:3039-3413).
Against the same composition in the oracle (oracle/lapw.py, oracle/minm_mt.py,
oracle/coulomb.py, oracle/chain.py).

barcvm's columns are eigenvectors, defined up to a phase each (a unitary
mixing within a degenerate eigenvalue): the GPU's and the oracle's bases
differ by such a D, minm by conj(D) on the right and chi0 / eps^-1 by
D^H X D.  The test therefore compares what D leaves alone: minm barcvm^T
(= mmat conj(barcvm barcvm^H)), barcvm chi0 barcvm^H, barcvm (eps^-1 - 1)
barcvm^H, and Sigma_c itself.  The oracle's chain is fed the GPU's weights
(their own parity is tested against the oracle in test_gpu_qloop.py), so
every comparison is at rounding level."""
import dataclasses

import numpy as np
import pytest

from oracle import chain
from oracle import coulomb as OC
from oracle import lapw as OL
from oracle import minm_mt as OM
from pygw_amd import mesh, synth

pytestmark = pytest.mark.gpu

KS = (0, 5)         # the Sigma_c k-points
NB1 = 2             # Sigma_c bands: highest occupied, lowest empty


def _gpu_builder(s):
    from pygw_amd import coulomb as C
    from pygw_amd import lapw as L
    from pygw_amd.minm import MatrixElements
    ls = s.lapw
    lp = L.LapwPlan(*ls.plan_args())
    mp = C.MinmMTPlan(ls.mult, s.nmix, s.big_l, ls.nLO_at, s.ncore, s.lmax)
    consts = C.mt_constants(mp, s.s3r, ls.pos, s.cgcoef, s.corind if len(s.corind) else None)
    cb = C.coul_setev(s.vmat, 1e-10)
    assert cb.matsize == s.vmat.shape[0]
    return MatrixElements(lp, mp, consts, ls.gindex, s.i_g0, s.mpwipw, s.indgq_inverse, ls.vol, cb.barcvm, s.ql), cb


def _k(s, ik):
    jk = s.kqid[ik, s.iq]
    return s.kf[ik], s.indgk[ik], s.nvk, s.Aeig[ik], s.kf[jk], s.indgk[jk], s.nvk, np.conj(s.Aeig[jk])


def _oracle_bands(s, kl, indgk, A, iop):
    ls = dataclasses.replace(s.lapw, kl=np.asarray(kl), indgk=indgk, nvk=s.nvk)
    return OL.band_coefficients(A, *OL.gap2_set_lapwcoef(*ls.coef_args(iop, True)))


def _oracle_mmat(s, ik, nst, nend, mst, mend, core):
    """mmat (MT | IS) and the core part, as calc_minm builds them."""
    kil, igk, nvk, Ak, kjl, igq, nvq, Aq = _k(s, ik)
    bk = _oracle_bands(s, kil, igk, Ak, 1)
    bq = _oracle_bands(s, kjl, igq, Aq, 2)
    ls = s.lapw
    rest = (ls.mult, s.nmix, s.big_l, ls.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loct)
    mt = OM.calc_minm_mt(s.ql, nst, nend, mst, mend, *bk, *bq, s.s3r, ls.pos, *rest)
    umk = np.rint(kil - kjl - s.ql).astype(int)
    it = OC.minm_is(nst, nend, mst, mend, Ak, Aq, umk, s.mpwipw, nvk, nvq, igk, igq, s.indgq_inverse, ls.gindex,
                    s.i_g0, ls.vol)
    mc = None
    ncg = len(s.corind)
    if core == "polarization" and ncg:
        mc = OM.calc_minc2(kjl, 0, ncg, mst, mend, *bq, s.s3r, s.corind, ls.pos, *rest)
    elif core == "selfe" and ncg:
        mc = OM.calc_minc(kil, nst, nend, 0, ncg, *bk, s.s3r, s.corind, ls.pos, *rest)
    return np.concatenate((mt, it), axis=-1), mc


def _oracle_pol(s, ik, barcvm):
    nvbm, ncbm, nb = s.nomx + 1, s.numin, s.enk.shape[0]
    mmat, mc = _oracle_mmat(s, ik, 0, nvbm, ncbm, nb, "polarization")
    minm = OC.minm(mmat, barcvm)
    if mc is not None:
        minm = np.concatenate((OC.minm(mc, barcvm, s.loct), minm), axis=0)
    return minm


def _oracle_selfe(s, ik, barcvm):
    ib = s.nomx
    mmat, mc = _oracle_mmat(s, ik, ib, ib + NB1, 0, s.enk.shape[0], "selfe")
    minm = OC.minm(mmat, barcvm)
    if mc is not None:
        minm = np.concatenate((OC.minm(mc, barcvm, s.loct), minm), axis=1)
    return minm


def _gpu_pol(s, me, ik):
    return me.polarization_minm(*_k(s, ik), s.nomx + 1, s.numin, s.enk.shape[0], len(s.corind))


def _gpu_selfe(s, me, ik):
    return me.selfe_minm(*_k(s, ik), s.nomx, s.nomx + NB1, s.enk.shape[0], len(s.corind))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("cores", [True, False])
def test_matrix_elements_vs_oracle(cores):
    """minm of calc_eps (core rows first) and of calc_selfc (core columns
    first) at every k, in the form barcvm's phases leave alone."""
    s = synth.pipeline_system(cores=cores)
    print(f"case pipeline matrix elements cores={cores}", flush=True)
    me, cb = _gpu_builder(s)
    B = cb.barcvm.cpu().numpy()
    _, _, _, Bo, _ = OC.coul_setev(s.vmat, 1e-10)
    for ik in range(len(s.kf)):
        got = _gpu_pol(s, me, ik).cpu().numpy()
        want = _oracle_pol(s, ik, Bo)
        assert got.shape == want.shape == (len(s.corind) + s.nomx + 1, s.enk.shape[0] - s.numin, s.vmat.shape[0])
        e = _rel(got @ B.T, want @ Bo.T)
        assert e < 1e-12, (ik, e)
        # the minc part against calc_minm's separate minc (pygw.py:2684)
        if cores:
            kil, igk, nvk, Ak, kjl, igq, nvq, Aq = _k(s, ik)
            _, minc = me.calc_minm(kil, igk, nvk, Ak, kjl, igq, nvq, Aq,
                                   (0, s.nomx + 1, s.numin, s.enk.shape[0], 0, len(s.corind)), "polarization")
            assert np.abs(minc.cpu().numpy() - got[:len(s.corind)]).max() < 1e-13 * np.abs(got).max()
    for ik in KS:
        got = _gpu_selfe(s, me, ik).cpu().numpy()
        want = _oracle_selfe(s, ik, Bo)
        assert got.shape == want.shape == (NB1, len(s.corind) + s.enk.shape[0], s.vmat.shape[0])
        assert _rel(got @ B.T, want @ Bo.T) < 1e-12


def test_pipeline_through_the_q_loop_vs_oracle():
    """LAPW -> minm -> weights -> chi0 -> eps^-1 -> MWM -> Sigma_c (FP64)."""
    import torch
    from pygw_amd.qloop import SelfEnergyQLoop
    s = synth.pipeline_system()
    print("case pipeline q loop", flush=True)
    me, cb = _gpu_builder(s)
    B = cb.barcvm.cpu().numpy()
    _, _, _, Bo, _ = OC.coul_setev(s.vmat, 1e-10)
    ncg, nb, nk = len(s.corind), s.enk.shape[0], len(s.kf)
    enk = np.vstack([np.repeat(s.eig_core[:, None], nk, axis=1), s.enk])
    fm = mesh.FrequencyMesh(8, 0.05 / mesh.H2EV, 80.0 / mesh.H2EV, iop_multiple=5)
    tetc = mesh.tetra_corners(s.div)
    kq = s.kqid[:, s.iq]
    pb = cb.matsize
    L = SelfEnergyQLoop(enk, 0.0, ncg, s.nomx, s.numin, fm.omega, fm.womeg, tetc, pb, 1, len(KS), NB1, ncg + nb,
                        omega_p=fm.omega_precise, womeg_p=fm.womeg_precise, precision="fp64")
    M = torch.stack([_gpu_pol(s, me, ik).reshape(-1, pb) for ik in range(nk)])
    Ms = torch.stack([_gpu_selfe(s, me, ik).reshape(-1, pb) for ik in KS])
    enq = np.ascontiguousarray(enk[:, kq[list(KS)]].T)
    L.add_q(kq, M, Ms, enq, keep_chi0=True)
    kn = L.kcw.cpu().numpy()
    Mo = [_oracle_pol(s, ik, Bo).reshape(-1, pb) for ik in range(nk)]
    chi = chain.calc_eps_accumulate(((ik, Mo[ik]) for ik in range(nk)), lambda ik: kn[ik].T)
    W = chain.eps_inverse(chi)
    got_chi, got_w = L.chi.cpu().numpy(), L.eps.cpu().numpy()
    for w in range(fm.omega.size):
        a = B @ got_chi[:, :, w] @ B.conj().T
        b = Bo @ chi[:, :, w] @ Bo.conj().T
        assert _rel(a, b) < 1e-12, w
        a = B @ got_w[:, :, w] @ B.conj().T
        b = Bo @ W[:, :, w] @ Bo.conj().T
        assert _rel(a, b) < 1e-11, w
    sc = np.empty((len(KS), NB1, fm.omega.size), dtype=complex)
    for j, ik in enumerate(KS):
        ms = _oracle_selfe(s, ik, Bo).reshape(-1, pb)
        mw = chain.calc_mwm(ms, W, 1.0).reshape(fm.omega.size, NB1, -1)
        sc[j] = chain.selfc_spline(enq[j], mw, fm.omega, fm.omega_precise, fm.womeg_precise)
    got_sc = L.sigc.cpu().numpy()
    e = np.abs(got_sc - sc).max() / np.abs(sc).max()
    print(f"  Sigma_c from the LAPW matrix elements: max rel {e:.2e}", flush=True)
    assert e < 1e-11
    torch.cuda.synchronize()


def test_pipeline_split_precision_vs_oracle():
    """The same chain with the products with conj(barcvm) (calc_eps's and
    calc_selfc's minm) and the q loop in split precision (minm handed over in
    complex64) against the FP64 oracle
    composition: chi0 and eps^-1 within 1e-6 per frequency, Sigma_c within
    1e-5 (the split-precision loop's own bars)."""
    import torch
    from pygw_amd.qloop import SelfEnergyQLoop
    s = synth.pipeline_system(seed=67)
    print("case pipeline q loop, split precision", flush=True)
    me, cb = _gpu_builder(s)
    B = cb.barcvm.cpu().numpy()
    _, _, _, Bo, _ = OC.coul_setev(s.vmat, 1e-10)
    ncg, nb, nk = len(s.corind), s.enk.shape[0], len(s.kf)
    enk = np.vstack([np.repeat(s.eig_core[:, None], nk, axis=1), s.enk])
    fm = mesh.FrequencyMesh(8, 0.05 / mesh.H2EV, 80.0 / mesh.H2EV, iop_multiple=5)
    tetc = mesh.tetra_corners(s.div)
    kq = s.kqid[:, s.iq]
    pb = cb.matsize
    L = SelfEnergyQLoop(enk, 0.0, ncg, s.nomx, s.numin, fm.omega, fm.womeg, tetc, pb, 1, len(KS), NB1, ncg + nb,
                        omega_p=fm.omega_precise, womeg_p=fm.womeg_precise, precision="mixed")
    M = torch.stack([me.polarization_minm(*_k(s, ik), s.nomx + 1, s.numin, nb, ncg, complex64=True,
                                          precision="mixed").reshape(-1, pb) for ik in range(nk)])
    Ms = torch.stack([me.selfe_minm(*_k(s, ik), s.nomx, s.nomx + NB1, nb, ncg, complex64=True,
                                    precision="mixed").reshape(-1, pb) for ik in KS])
    enq = np.ascontiguousarray(enk[:, kq[list(KS)]].T)
    L.add_q(kq, M, Ms, enq, keep_chi0=True)
    kn = L.kcw.cpu().numpy()
    Mo = [_oracle_pol(s, ik, Bo).reshape(-1, pb) for ik in range(nk)]
    chi = chain.calc_eps_accumulate(((ik, Mo[ik]) for ik in range(nk)), lambda ik: kn[ik].T)
    W = chain.eps_inverse(chi)
    got_chi, got_w = L.chi.cpu().numpy(), L.eps.cpu().numpy()
    one = np.eye(pb)
    rc = ri = 0.0
    for w in range(fm.omega.size):
        rc = max(rc, _rel(B @ got_chi[:, :, w] @ B.conj().T, Bo @ chi[:, :, w] @ Bo.conj().T))
        # eps^-1 itself (W + 1) in the basis-free form
        ri = max(ri, _rel(B @ (got_w[:, :, w] + one) @ B.conj().T, Bo @ (W[:, :, w] + one) @ Bo.conj().T))
    sc = np.empty((len(KS), NB1, fm.omega.size), dtype=complex)
    for j, ik in enumerate(KS):
        ms = _oracle_selfe(s, ik, Bo).reshape(-1, pb)
        mw = chain.calc_mwm(ms, W, 1.0).reshape(fm.omega.size, NB1, -1)
        sc[j] = chain.selfc_spline(enq[j], mw, fm.omega, fm.omega_precise, fm.womeg_precise)
    got_sc = L.sigc.cpu().numpy()
    rs = max(np.linalg.norm(got_sc[:, :, w] - sc[:, :, w]) / np.linalg.norm(sc[:, :, w])
             for w in range(fm.omega.size))
    print(f"  chi0 {rc:.2e} eps^-1 {ri:.2e} Sigma_c {rs:.2e}", flush=True)
    assert rc < 1e-6 and ri < 1e-6 and rs < 1e-5
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_polarization_minm_many_equals_one_by_one(precision):
    """Several k-points through one product with conj(barcvm) give what
    each gives alone."""
    import torch
    s = synth.pipeline_system()
    me, _ = _gpu_builder(s)
    print(f"case polarization_minm_many {precision}", flush=True)
    nb, ncg = s.enk.shape[0], len(s.corind)
    ks = [_k(s, ik) for ik in range(len(s.kf))]
    many = me.polarization_minm_many(ks, s.nomx + 1, s.numin, nb, ncg, precision=precision).cpu().numpy()
    for ik, k in enumerate(ks):
        one = me.polarization_minm(*k, s.nomx + 1, s.numin, nb, ncg, precision=precision).cpu().numpy()
        assert np.abs(many[ik] - one).max() <= 1e-14 * np.abs(one).max()
    torch.cuda.synchronize()
