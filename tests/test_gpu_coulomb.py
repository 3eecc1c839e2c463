"""GPU parity of the Coulomb v^(1/2) basis (pygw_coul_setev: Jacobi
eigensolver + Coul_setev selection) and of minm = mmat conj(barcvm)
(pygw_minm_vhalf) against the scipy/numpy oracle (oracle/coulomb.py).

Eigenvectors are compared where they are defined: eigenvalues, kept count
and order, barcvm barcvm^H (phase-free), |<gpu column, oracle column>| for
separated eigenvalues, and the residual and orthonormality of the GPU
eigenvectors themselves."""
import numpy as np
import pytest

from oracle import coulomb as O

pytestmark = pytest.mark.gpu


def _spectrum(n, kind, rng):
    if kind == "flat":
        return rng.uniform(0.5, 2.0, n)
    if kind == "graded":
        return np.concatenate([10 ** rng.uniform(-3, 3, n - n // 5), 10 ** rng.uniform(-13, -9, n // 5)])
    if kind == "coulomb":   # 4 pi / |q + G|^2 - like values and a tail of near-null vectors
        g = rng.uniform(0.25, 9.0, n - n // 10)
        return np.concatenate([4 * np.pi / g ** 2, 10 ** rng.uniform(-12, -8, n // 10)])
    raise ValueError(kind)


def _check_basis(A, got, evtol, lam, wi0=None, iop=0):
    ev, V, barcev, barcvm, immax = O.coul_setev(A, evtol, wi0, iop)
    nrm = np.abs(lam).max()
    ev_g = got.ev.cpu().numpy()
    assert np.abs(ev_g - ev).max() < 1e-12 * nrm
    assert got.matsize == len(barcev) and got.immax == immax
    bg = got.barcev.cpu().numpy()
    assert np.abs(bg - barcev).max() < 1e-12 * nrm
    B = got.barcvm.cpu().numpy()
    assert B.shape == barcvm.shape
    assert np.abs(B @ B.conj().T - barcvm @ barcvm.conj().T).max() < 1e-12 * nrm
    gap = np.minimum(np.abs(np.diff(np.r_[np.inf, barcev])), np.abs(np.diff(np.r_[barcev, -np.inf])))
    sep = gap > 1e-6 * nrm
    ov = np.abs(np.einsum("im,im->m", B.conj(), barcvm))
    assert np.all(np.abs(ov - np.abs(barcev))[sep] < 1e-9 * nrm)
    if got.vmat is not None:
        Vg = got.vmat.cpu().numpy()
        assert np.abs(A @ Vg - Vg * ev_g).max() < 1e-12 * nrm
        assert np.abs(Vg.conj().T @ Vg - np.eye(len(ev))).max() < 1e-12
    return B


@pytest.mark.parametrize("n,kind", [(1, "flat"), (2, "flat"), (7, "flat"), (64, "graded"), (301, "graded"),
                                    (1000, "coulomb"), (2000, "coulomb")])
def test_coul_setev_vs_oracle(n, kind):
    from pygw_amd import coulomb
    rng = np.random.default_rng(n)
    lam = _spectrum(n, kind, rng)
    A, _ = O.hermitian_with_spectrum(lam, rng)
    print(f"case coul_setev n={n} {kind}", flush=True)
    # one threshold at the largest size (the oracle's eigh there takes seconds)
    for evtol in ((-1e-10,) if n > 1000 else (-1e-10, 1e-7)):
        got = coulomb.coul_setev(A, evtol, eigenvectors=n <= 301)
        print(f"  evtol={evtol} matsize={got.matsize} sweeps={got.sweeps}", flush=True)
        _check_basis(A, got, evtol, lam)


def test_coul_setev_lower_triangle_and_q0():
    """Only the lower triangle is read (the upper one here is garbage); at
    q = 0 the vector with the largest overlap with wi0 is dropped
    (iop_coul_x = 0) or kept below evtol (iop_coul_x = 1)."""
    from pygw_amd import coulomb
    rng = np.random.default_rng(11)
    n = 90
    lam = _spectrum(n, "graded", rng)
    A, Q = O.hermitian_with_spectrum(lam, rng)
    Ag = np.tril(A) + np.triu(rng.normal(size=(n, n)), 1)
    j0 = int(np.argsort(lam)[40])
    wi0 = 3.0 * Q[:, j0] + 1e-3 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    print("case coul_setev q=0", flush=True)
    for iop, evtol in ((0, 1e-7), (1, 1.5 * lam[j0])):
        got = coulomb.coul_setev(Ag, evtol, wi0, iop)
        assert got.immax == 40
        _check_basis(A, got, evtol, lam, wi0, iop)


@pytest.mark.parametrize("shape,mb,ms,n_in", [((3, 37), 101, 90, None), ((1007,), 150, 130, 77),
                                              ((2, 64), 64, 64, None), ((5,), 16, 1, 3)])
def test_minm_vhalf_vs_oracle(shape, mb, ms, n_in):
    import torch
    from pygw_amd import coulomb
    rng = np.random.default_rng(mb + ms)
    ld = mb if n_in is None else n_in
    mm = rng.normal(size=shape + (ld,)) + 1j * rng.normal(size=shape + (ld,))
    bv = np.asfortranarray(rng.normal(size=(mb, ms)) + 1j * rng.normal(size=(mb, ms)))
    want = O.minm(mm, bv, n_in)
    print(f"case minm shape={shape} mb={mb} ms={ms} n_in={n_in}", flush=True)
    mt = torch.as_tensor(mm, device="cuda")
    bt = torch.as_tensor(bv.T.copy(), device="cuda").T    # unit stride along mbsize
    got = coulomb.calc_minm_vhalf(mt, bt, n_in).cpu().numpy()
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    g32 = coulomb.calc_minm_vhalf(mt, bt, n_in, complex64=True).cpu().numpy()
    assert g32.dtype == np.complex64
    assert np.abs(g32 - want).max() <= 2.0 ** -23 * np.abs(want).max()


def test_vhalf_chain_is_phase_free():
    """M from the GPU basis and M from the oracle's basis give the same
    M M^H (the rows' Gram matrix, what every contraction over the product
    basis sees up to a unitary change of basis)."""
    import torch
    from pygw_amd import coulomb
    rng = np.random.default_rng(5)
    n = 240
    lam = _spectrum(n, "coulomb", rng)
    A, _ = O.hermitian_with_spectrum(lam, rng)
    got = coulomb.coul_setev(A, 1e-7)
    _, _, _, bref, _ = O.coul_setev(A, 1e-7)
    mm = rng.normal(size=(200, n)) + 1j * rng.normal(size=(200, n))
    print("case vhalf chain", flush=True)
    Mg = coulomb.calc_minm_vhalf(torch.as_tensor(mm, device="cuda"), got.barcvm).cpu().numpy()
    Mr = O.minm(mm, bref)
    G, R = Mg @ Mg.conj().T, Mr @ Mr.conj().T
    assert np.abs(G - R).max() < 1e-11 * np.abs(R).max()


@pytest.mark.parametrize("kw,rng_", [(dict(seed=7, gmax=3, nvi=37, nvj=41, ngq=23, ngq_barc=29, nb=9), (1, 6, 2, 9)),
                                     (dict(seed=8, gmax=5, nvi=300, nvj=321, ngq=117, ngq_barc=150, nb=70),
                                      (3, 23, 10, 70)),
                                     (dict(seed=9, gmax=4, nvi=90, nvj=80, ngq=40, ngq_barc=60, nb=60), (0, 45, 5, 50)),
                                     (dict(seed=10, gmax=4, nvi=90, nvj=80, ngq=10, ngq_barc=60, nb=60),
                                      (0, 20, 5, 55))])
def test_minm_is_vs_oracle(kw, rng_):
    """The interstitial matrix elements against the restatement of
    calc_minm_IS in every association the library has (the pair density
    first where it costs fewer products, the first three cases; T first,
    dense and gathered), alone and written after loctmatsize MT columns of
    the concatenated mmat (the columns before them left alone); the
    preferred form twice gives the same bits."""
    import torch
    from pygw_amd import coulomb, synth
    s = synth.interstitial_system(**kw)
    want = O.minm_is(*s.args(*rng_))
    print(f"case minm_is {kw} bands {rng_}", flush=True)
    nd, md, ngq, nb = rng_[1] - rng_[0], rng_[3] - rng_[2], kw["ngq"], kw["ngq_barc"]
    rho = coulomb._rho_cheaper(ngq, nb, kw["nvj"], nd, md)
    assert rho == (kw["ngq"] != 10)
    got = coulomb.calc_minm_is(*s.args(*rng_)).cpu().numpy()
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    ws = torch.empty(coulomb.minm_is_workspace(ngq, nb, kw["nvi"], kw["nvj"], nd, md), dtype=torch.uint8,
                     device="cuda")
    again = coulomb.calc_minm_is(*s.args(*rng_), workspace=ws).cpu().numpy()
    assert np.array_equal(again, got)
    forms = ["gather"] + ([] if rho else ["dense"])
    if rho:
        try:
            forms.append(coulomb.calc_minm_is(*s.args(*rng_), form="dense") is not None and "dense")
        except ValueError:
            pass
    for f in forms:
        g = coulomb.calc_minm_is(*s.args(*rng_), form=f).cpu().numpy()
        assert np.abs(g - want).max() < 1e-13 * np.abs(want).max(), f
    loct = 13
    full = torch.full(want.shape[:2] + (loct + want.shape[2],), 7.0 + 0j, dtype=torch.complex128, device="cuda")
    coulomb.calc_minm_is(*s.args(*rng_), out=full, col0=loct)
    f = full.cpu().numpy()
    assert np.all(f[..., :loct] == 7.0) and np.abs(f[..., loct:] - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("nrows,n_in,dense_rows", [(300, None, False), (1111, 640, True), (5, None, True)])
def test_minm_vhalf_mixed_vs_fp64(nrows, n_in, dense_rows):
    """The split-precision product with conj(barcvm) against the FP64 one:
    rows of very different sizes (and a zero row), all of mbsize or the
    first n_in (minc's loctmatsize); every entry within 1e-6 of its row's
    largest, the whole within 3e-7 (Frobenius: the two 22-bit operand
    splits, ~1.2e-7 each, and the FP32 accumulation), complex64 and
    complex128 results alike."""
    import torch
    from pygw_amd import coulomb, synth
    vmat, _ = synth.coulomb_matrix(700, 60, seed=31)
    cb = coulomb.coul_setev(torch.as_tensor(vmat, device="cuda"), 1e-7)
    rng = np.random.default_rng(nrows)
    mb = vmat.shape[0]
    m = (rng.normal(size=(nrows, mb)) + 1j * rng.normal(size=(nrows, mb)))
    m *= 10.0 ** rng.uniform(-6, 3, size=(nrows, 1))
    if dense_rows:
        m[0] = 0.0
    print(f"case minm_vhalf_mixed rows {nrows} n_in {n_in}", flush=True)
    mm = torch.as_tensor(m, device="cuda")
    basis = coulomb.SplitBasis(cb.barcvm)
    want = coulomb.calc_minm_vhalf(mm, cb.barcvm, n_in).cpu().numpy()
    for c64 in (True, False):
        got = coulomb.calc_minm_vhalf_mixed(mm, basis, n_in, complex64=c64).cpu().numpy()
        assert got.dtype == (np.complex64 if c64 else np.complex128)
        scale = np.abs(want).max(axis=1, keepdims=True)
        err = np.abs(got - want) / np.where(scale > 0, scale, 1.0)
        fro = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"  complex64 {c64}: max {err.max():.2e} of the row's largest, Frobenius {fro:.2e}", flush=True)
        assert err.max() < 1e-6 and fro < 3e-7, (err.max(), fro)
        if dense_rows:
            assert np.all(got[0] == 0)


def test_minm_is_repeated_plane_wave():
    """A plane wave listed twice among k's first nvi: every association sums
    both terms, as the reference's sum over i does (the pair-density form
    marks such slots and sums them in the gather)."""
    from pygw_amd import coulomb, synth
    s = synth.interstitial_system(seed=12, gmax=4, nvi=90, nvj=80, ngq=40, ngq_barc=60, nb=60)
    s.indgk_i[7] = s.indgk_i[3]
    s.indgk_i[50] = s.indgk_i[3]
    rng_ = (0, 45, 5, 50)
    want = O.minm_is(*s.args(*rng_))
    print("case minm_is repeated plane wave", flush=True)
    for f in ("auto", "gather"):
        got = coulomb.calc_minm_is(*s.args(*rng_), form=f).cpu().numpy()
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max(), f


def test_coul_setev_nothing_kept_and_degenerate():
    """Every eigenvalue at or below evtol: PB = 0 and an empty minm, as
    numpy gives; a degenerate spectrum (a multiple eigenvalue) keeps its
    whole eigenspace: barcvm barcvm^H equals the oracle's."""
    import torch
    from pygw_amd import coulomb
    rng = np.random.default_rng(21)
    lam = np.concatenate([np.full(5, 1e-9), [-1e-12, 0.0]])
    A, _ = O.hermitian_with_spectrum(lam, rng)
    print("case coul_setev nothing kept", flush=True)
    got = coulomb.coul_setev(A, 1e-7)
    assert got.matsize == 0 and got.barcvm.shape == (7, 0)
    m = coulomb.calc_minm_vhalf(torch.zeros((3, 7), dtype=torch.complex128, device="cuda"), got.barcvm)
    assert m.shape == (3, 0)
    lam = np.concatenate([np.full(6, 2.0), np.full(4, 0.5), [3.0, 1e-3]])
    A, _ = O.hermitian_with_spectrum(lam, rng)
    print("case coul_setev degenerate", flush=True)
    got = coulomb.coul_setev(A, 1e-7, eigenvectors=True)
    _check_basis(A, got, 1e-7, lam)


@pytest.mark.parametrize("ngq,ngq_barc,q0", [(60, 90, False), (150, 210, True), (7, 7, False)])
def test_pw_basis_and_coulomb_block_vs_oracle(ngq, ngq_barc, q0):
    """mpwipw (olap^(-1/2) tmat, phase-free) and the plane-wave block of the
    Coulomb matrix against the oracle, the block written at row0 =
    loctmatsize of a larger Vmat (the rest left alone); then that Coulomb
    matrix through Coul_setev."""
    import torch
    from pygw_amd import coulomb, synth
    s = synth.plane_wave_system(ngq=ngq, ngq_barc=ngq_barc, q=(0, 0, 0) if q0 else (0.1, 0.0, 0.0))
    print(f"case pw_basis ngq={ngq} ngq_barc={ngq_barc} q0={q0}", flush=True)
    want = O.pw_basis(s.indgq, ngq, ngq_barc, s.ipwint, s.gindex, s.i_g0)
    got = coulomb.pw_basis(s.indgq, ngq, ngq_barc, s.ipwint, s.gindex, s.i_g0)
    assert np.abs(got.cpu().numpy() - want).max() < 1e-12 * np.abs(want).max()
    vw = O.coul_pw_block(want, s.gqlen, q0)
    loct = 11
    full = torch.full((loct + ngq, loct + ngq), 3.0 + 0j, dtype=torch.complex128, device="cuda")
    coulomb.coul_pw_block(got, s.gqlen, q0, vmat=full, row0=loct)
    f = full.cpu().numpy()
    assert np.all(f[:loct] == 3.0) and np.all(f[:, :loct] == 3.0)
    assert np.abs(f[loct:, loct:] - vw).max() < 1e-12 * np.abs(vw).max()
    cb = coulomb.coul_setev(f[loct:, loct:], 1e-7)
    ev = np.linalg.eigvalsh(vw)
    assert np.abs(cb.ev.cpu().numpy() - ev).max() < 1e-12 * ev.max()


@pytest.mark.parametrize("q", [(0.25, 0.0, 0.125), (0.0, 0.0, 0.0)])
def test_mixed_coulomb_and_block_vs_oracle(q):
    """The MT / plane-wave rows of Vmatit of both atoms (written at their row
    offsets) against the restatement of Mixed_Coulomb (itself pinned to
    PyGW's compiled routine), then the two mixed blocks of Vmat."""
    import torch
    from pygw_amd import coulomb, synth
    s = synth.mixed_system(q=q)
    print(f"case mixed_coulomb q={q}", flush=True)
    want = [O.mixed_coulomb(*s.args(iat)) for iat in (0, 1)]
    loct = sum(w.shape[0] for w in want)
    W = torch.zeros((loct, s.ngq_barc), dtype=torch.complex128, device="cuda")
    row = 0
    for iat in (0, 1):
        a = s.args(iat)
        coulomb.mixed_coulomb(*a, out=W, row0=row)
        row += want[iat].shape[0]
    Wn, Wr = W.cpu().numpy(), np.concatenate(want)
    assert np.abs(Wn - Wr).max() < 1e-12 * np.abs(Wr).max()
    rng = np.random.default_rng(2)
    ngq = 70
    mp = rng.normal(size=(ngq, s.ngq_barc)) + 1j * rng.normal(size=(ngq, s.ngq_barc))
    V = torch.full((loct + ngq, loct + ngq), 5.0 + 0j, dtype=torch.complex128, device="cuda")
    coulomb.coul_mixed_block(torch.as_tensor(mp, device="cuda"), W, V)
    mat2 = mp @ Wr.T
    Vn = V.cpu().numpy()
    tol = 1e-12 * np.abs(mat2).max()
    assert np.abs(Vn[loct:, :loct] - mat2).max() < tol and np.abs(Vn[:loct, loct:] - mat2.conj().T).max() < tol
    assert np.all(Vn[:loct, :loct] == 5.0) and np.all(Vn[loct:, loct:] == 5.0)


def test_pw_basis_singular_overlap_rejected():
    """A repeated plane wave in indgq makes olap singular: pygw_pw_basis
    reports PYGW_E_NOTPD instead of writing inf/NaN into mpwipw."""
    from pygw_amd import _lib, coulomb, synth
    s = synth.plane_wave_system(ngq=30, ngq_barc=45)
    indgq = s.indgq.copy()
    indgq[1] = indgq[0]
    print("case pw_basis repeated plane wave", flush=True)
    with pytest.raises(_lib.PygwError) as e:
        coulomb.pw_basis(indgq, 30, 45, s.ipwint, s.gindex, s.i_g0)
    assert e.value.code == -13
    good = coulomb.pw_basis(s.indgq, 30, 45, s.ipwint, s.gindex, s.i_g0)     # the call after the error is unaffected
    assert bool(good.isfinite().all())
