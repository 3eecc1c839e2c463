"""GPU parity of the muffin-tin and core matrix elements (minm_mt.hip through
pygw_amd.coulomb) against the oracle (oracle/minm_mt.py): calc_minm_MT with
and without local orbitals, band windows that start inside the arrays, a
result written into the concatenated mmat (col0 = 0, a wider row), a
PyGW-sized lmax 10; calc_minc and calc_minc2 over all and over part of the
core list."""
import numpy as np
import pytest
import torch

from oracle import minm_mt as O
from pygw_amd import coulomb as C
from pygw_amd import synth

pytestmark = pytest.mark.gpu


def _plan(s):
    return C.MinmMTPlan(s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.lmax)


def _oracle_mt(s, nst, nend, mst, mend):
    return O.calc_minm_mt(s.ql, nst, nend, mst, mend, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, s.s3r, s.vpos,
                          s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)


@pytest.mark.parametrize("los", [True, False])
@pytest.mark.parametrize("window", [(0, 7, 0, 9), (1, 6, 2, 9)])
def test_minm_mt_matches_oracle(los, window):
    s = synth.minm_mt_system(los=los)
    nst, nend, mst, mend = window
    print(f"minm_mt los={los} window={window}", flush=True)
    want = _oracle_mt(s, nst, nend, mst, mend)
    got = C.calc_minm_mt(_plan(s), s.ql, nst, nend, mst, mend, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp,
                         s.s3r, s.vpos, s.cgcoef).cpu().numpy()
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_minm_mt_into_concatenated_mmat():
    """Written into columns [0, loctmatsize) of rows of ld = loctmatsize + 23;
    the interstitial columns stay as they were."""
    s = synth.minm_mt_system()
    want = _oracle_mt(s, 1, 6, 2, 9)
    out = torch.full((5, 7, s.loctmatsize + 23), 7.0 + 1j, dtype=torch.complex128, device="cuda")
    C.calc_minm_mt(_plan(s), s.ql, 1, 6, 2, 9, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, s.s3r, s.vpos,
                   s.cgcoef, out=out)
    o = out.cpu().numpy()
    assert np.abs(o[:, :, :s.loctmatsize] - want).max() < 1e-13 * np.abs(want).max()
    assert np.all(o[:, :, s.loctmatsize:] == 7.0 + 1j)


def test_minm_mt_lmax10():
    """lmax 10 (121 (l, m) per channel kind) on the small atom set, 20 x 40 bands."""
    s = synth.minm_mt_system(lmax=10, nbmax=22, mbmax=44)
    print("minm_mt lmax 10", flush=True)
    want = _oracle_mt(s, 2, 22, 3, 43)
    got = C.calc_minm_mt(_plan(s), s.ql, 2, 22, 3, 43, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, s.s3r,
                         s.vpos, s.cgcoef).cpu().numpy()
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("crange", [None, (1, 5)])
def test_minc_matches_oracle(crange):
    s = synth.minm_mt_system()
    c0, c1 = crange or (0, len(s.corind))
    args = (s.s3r, s.corind, s.vpos, s.mult, s.nmix, s.big_l, s.nLO_at, s.ncore, s.cgcoef, s.lmax, s.loctmatsize)
    p = _plan(s)
    want = O.calc_minc(s.kl, 1, 6, c0, c1, s.alfa, s.beta, s.gama, *args)
    got = C.calc_minc(p, s.kl, 1, 6, c0, c1, s.alfa, s.beta, s.gama, s.s3r, s.corind, s.vpos, s.cgcoef)
    assert np.abs(got.cpu().numpy() - want).max() < 1e-13 * np.abs(want).max()
    want = O.calc_minc2(s.kl, c0, c1, 2, 9, s.alfp, s.betp, s.gamp, *args)
    got = C.calc_minc2(p, s.kl, c0, c1, 2, 9, s.alfp, s.betp, s.gamp, s.s3r, s.corind, s.vpos, s.cgcoef)
    assert np.abs(got.cpu().numpy() - want).max() < 1e-13 * np.abs(want).max()


def test_uploaded_constants_give_the_same_results():
    """s3r/vpos/cgcoef/corind uploaded once (mt_constants) in place of the
    host arrays: calc_minm_mt, calc_minc and calc_minc2 unchanged, bit for bit."""
    s = synth.minm_mt_system()
    p = _plan(s)
    k = C.mt_constants(p, s.s3r, s.vpos, s.cgcoef, s.corind)
    a = C.calc_minm_mt(p, s.ql, 1, 6, 2, 9, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, s.s3r, s.vpos, s.cgcoef)
    b = C.calc_minm_mt(p, s.ql, 1, 6, 2, 9, s.alfa, s.beta, s.gama, s.alfp, s.betp, s.gamp, k, None, None)
    assert torch.equal(a, b)
    a = C.calc_minc(p, s.kl, 1, 6, 1, 5, s.alfa, s.beta, s.gama, s.s3r, s.corind, s.vpos, s.cgcoef)
    b = C.calc_minc(p, s.kl, 1, 6, 1, 5, s.alfa, s.beta, s.gama, k, None, None, None)
    assert torch.equal(a, b)
    a = C.calc_minc2(p, s.kl, 0, len(s.corind), 2, 9, s.alfp, s.betp, s.gamp, s.s3r, s.corind, s.vpos, s.cgcoef)
    b = C.calc_minc2(p, s.kl, 0, len(s.corind), 2, 9, s.alfp, s.betp, s.gamp, k, None, None, None)
    assert torch.equal(a, b)
