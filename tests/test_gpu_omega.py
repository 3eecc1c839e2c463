"""The contraction in a compressed frequency basis: PyGW's SVD mode (iopfreq
5: svd_functions, pygw.py:3478-3501; bz_calcqdepw_par3; calc_eps contracting
n_l basis functions and expanding eps = dot(eps, Ul), pygw.py:3206-3226) and
the same steps on a basis taken from the q's own weights (pygw_omega_basis).

Pins:
- the Jacobi eigen-decomposition on weights of known singular values
  (kcw = A diag(s) Q^T with A orthonormal columns and Q orthogonal: the
  singular values are s and the basis vectors the columns of Q), including an
  odd n_om (dummy index of the round-robin ordering) and n_om = 96 (the
  limit);
- projection and expansion against numpy;
- the SVD-mode chain (weights in PyGW's basis -> contraction -> dot(eps, Ul))
  against the oracle's calc_eps on the same weights, FP64 to 1e-13;
- the compressed contraction against the uncompressed FP64 one (itself at
  1e-13 of the oracle in test_gpu_chain): the truncation's change of chi0 is
  below 1e-8 per frequency at the default rank, and the split-precision
  compressed chi0 stays within 1e-6 per frequency.
"""
import numpy as np
import pytest

from oracle import chain
from pygw_amd import mesh, synth

pytestmark = pytest.mark.gpu


def _flat_to_native(flat, nk, npair):
    """[nk*npair, nom] (row k*npair + p) -> kcw native [nk, nom, npair]."""
    return np.ascontiguousarray(flat.reshape(nk, npair, -1).transpose(0, 2, 1))


@pytest.mark.parametrize("nom", [1, 7, 33, 64, 96])
def test_basis_of_known_spectrum(nom):
    import torch
    from pygw_amd import polarization as P
    print(f"case omega basis nom={nom}", flush=True)
    rng = np.random.default_rng(nom)
    nk, npair = 5, 61
    A, _ = np.linalg.qr(rng.normal(size=(nk * npair, nom)))
    Q, _ = np.linalg.qr(rng.normal(size=(nom, nom)))
    s = np.geomspace(1.0, 1e-7, nom) if nom > 1 else np.ones(1)
    flat = (A * s) @ Q.T
    U, Ue, sv = P.omega_basis(_flat_to_native(flat, nk, npair), balance=False)
    U, Ue, sv = U.cpu().numpy(), Ue.cpu().numpy(), sv.cpu().numpy()
    assert np.array_equal(U, Ue)
    # eigenvalues of the Gram matrix to n*eps of its norm
    assert np.all(np.abs(sv ** 2 - s ** 2) <= 1e-14 * nom)
    assert np.all(np.diff(sv) <= 0)
    np.testing.assert_allclose(U.T @ U, np.eye(nom), atol=1e-13)
    # vectors of well-separated eigenvalues (lambda >= 1e-6) up to sign
    good = s >= 1e-3
    dots = np.abs(np.sum(U[:, good] * Q[:, good], axis=0))
    assert np.all(dots >= 1 - 1e-10), dots.min()
    torch.cuda.synchronize()


def test_balanced_basis():
    """Columns scaled to unit norm first: sv are the singular values of the
    column-normalised matrix (numpy SVD), proj^T expand = 1, and at full
    rank projection then expansion returns the weights."""
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(11)
    nk, nom, npair = 4, 40, 53
    flat = rng.normal(size=(nk * npair, nom)) * np.geomspace(1.0, 1e-4, nom)   # columns of very different norm
    flat[:, :5] = flat[:, 5:10] * 3.0 + 1e-9 * flat[:, :5]                       # a nearly dependent block
    print("case omega balanced basis", flush=True)
    Pm, Em, sv = P.omega_basis(_flat_to_native(flat, nk, npair))
    Pm, Em, sv = Pm.cpu().numpy(), Em.cpu().numpy(), sv.cpu().numpy()
    s_ref = np.linalg.svd(flat / np.linalg.norm(flat, axis=0), compute_uv=False)
    assert np.all(np.abs(sv ** 2 - s_ref ** 2) <= 1e-14 * nom * s_ref[0] ** 2)
    np.testing.assert_allclose(Pm.T @ Em, np.eye(nom), atol=1e-11)
    np.testing.assert_allclose(flat @ Pm @ Em.T, flat, rtol=0, atol=1e-11 * np.abs(flat).max())
    torch.cuda.synchronize()


def test_project_and_expand_match_numpy():
    import torch
    from pygw_amd import polarization as P
    rng = np.random.default_rng(3)
    nk, nom, npair, rank = 5, 48, 77, 40          # rank > 32: two passes of the kernels
    kcw = rng.normal(size=(nk, nom, npair))
    proj = rng.normal(size=(nom, rank))
    want = np.einsum("wr,kwp->krp", proj, kcw)
    print("case omega project", flush=True)
    got = P.omega_project(kcw, proj, rank).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    # a Fortran-ordered projection (PyGW's dUl [n_om, n_l]) through the strides
    projF = torch.as_tensor(np.asfortranarray(proj), device="cuda")
    projF = torch.as_strided(projF, (nom, rank), (1, nom))
    projF.copy_(torch.as_tensor(proj, device="cuda"))
    got = P.omega_project(kcw, projF, 12).cpu().numpy()
    np.testing.assert_allclose(got, want[:, :12], rtol=1e-13, atol=1e-13 * np.abs(want).max())

    pb = 30
    X = rng.normal(size=(pb, pb, rank)) + 1j * rng.normal(size=(pb, pb, rank))
    ex = rng.normal(size=(nom, rank))
    Xd = torch.as_tensor(np.asfortranarray(X).reshape(-1, order="F"), device="cuda").as_strided(
        (pb, pb, rank), (1, pb, pb * pb))
    want = np.einsum("ijr,wr->ijw", X, ex)
    print("case omega expand", flush=True)
    got = P.omega_expand(Xd, ex, rank)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    got2 = P.omega_expand(Xd, ex, rank, eps=got)                # beta = 1 adds
    np.testing.assert_allclose(got2.cpu().numpy(), 2 * want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    # batched: [n_b][rank][nel] -> [n_b][n_om][nel] (the MWM rows of every k)
    import ctypes as C
    from pygw_amd import _lib
    nb, nel, r2 = 6, 333, 9
    xb = rng.normal(size=(nb, r2, nel)) + 1j * rng.normal(size=(nb, r2, nel))
    Ed = torch.as_tensor(ex, device="cuda")
    xd = torch.as_tensor(xb, device="cuda")
    out = torch.empty((nb, nom, nel), dtype=torch.complex128, device="cuda")
    print("case omega expand batched", flush=True)
    _lib.call("pygw_omega_expand_batched", C.c_void_p(xd.data_ptr()), C.c_void_p(Ed.data_ptr()), rank, 1,
              nb, nel, r2, nom, 0, C.c_void_p(out.data_ptr()), None)
    want = np.einsum("brn,wr->bwn", xb, ex[:, :r2])
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    torch.cuda.synchronize()


def _system(div, n_occ, n_emp, n_om, pb, seed):
    cfg = synth.small_config(div=div, n_occ=n_occ, n_emp=n_emp, n_om=n_om, pb=pb, seed=seed, vhalf="coulomb",
                             wmin_ev=0.05, wmax_ev=80.0)
    b = synth.band_energies(cfg)
    fm = mesh.FrequencyMesh(cfg.n_om, cfg.wmin_ev / mesh.H2EV, cfg.wmax_ev / mesh.H2EV, iop_multiple=5)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, synth.q_index(cfg)]
    M = synth.matrix_elements(cfg, b, kq)
    return cfg, b, fm, tetc, kq, M


def test_svd_mode_chain_vs_oracle():
    """PyGW's SVD mode: weights in the basis dUl (par3), chi0 contracted with
    the n_l basis functions, eps = dot(eps, Ul)."""
    import torch
    from pygw_amd import polarization as P
    cfg, b, fm, tetc, kq, M = _system((3, 3, 3), 3, 6, 16, 40, 43)
    Ul, dUl = mesh.svd_functions(fm.omega, fm.womeg, 0.005, 10.0, 48, 1e-6)
    nl = Ul.shape[0]
    print(f"case svd mode n_l={nl}", flush=True)
    kl = P.bz_calcqdepw_par3(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, dUl, fm.omega, tetc, layout="native")
    X = P.chi0_accumulate(M, kl, precision="fp64")
    ex = torch.as_tensor(np.ascontiguousarray(Ul), device="cuda").T      # E(w, l) = Ul[l, w]
    eps = P.omega_expand(X, ex, nl).cpu().numpy()
    kln = kl.cpu().numpy()
    chi_l = chain.calc_eps_accumulate(((ik, M[ik]) for ik in range(cfg.nk)), lambda ik: kln[ik].T)
    want = np.dot(chi_l, Ul)
    for w in range(cfg.n_om):
        assert np.linalg.norm(eps[:, :, w] - want[:, :, w]) <= 1e-13 * np.linalg.norm(want[:, :, w])
    torch.cuda.synchronize()


@pytest.mark.parametrize("mesh_kind", ["tan", "log"])
def test_compressed_chi0_vs_uncompressed(mesh_kind):
    """tan: PyGW's mesh of the chain workload; log: 64 points from 0.005 to
    200 eV, the mesh of the c2/c3 workloads, where the weights fall by
    orders of magnitude across the mesh (the balanced basis keeps every
    frequency's relative error)."""
    import torch
    from pygw_amd import polarization as P
    cfg, b, fm, tetc, kq, M = _system((4, 4, 4), 4, 12, 32, 72, 44)
    om = fm.omega if mesh_kind == "tan" else mesh.log_mesh(0.005, 200.0, cfg.n_om)
    kn = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
    print(f"case compressed chi0 {mesh_kind}", flush=True)
    ref = P.chi0_accumulate(M, kn, precision="fp64").cpu().numpy()
    basis = P.omega_basis(kn)
    rank = P.omega_rank(basis[2])
    assert rank < cfg.n_om
    nrm = np.linalg.norm(ref.reshape(-1, cfg.n_om), axis=0)

    def rel(x):
        return (np.linalg.norm((x - ref).reshape(-1, cfg.n_om), axis=0) / nrm).max()

    e64, r64, _ = P.chi0_compressed(M, kn, basis=basis, rank=rank, precision="fp64")
    emx, _, _ = P.chi0_compressed(M.astype(np.complex64), kn, basis=basis, rank=rank, precision="mixed")
    full_mx = P.chi0_accumulate(M.astype(np.complex64), kn, precision="mixed").cpu().numpy()
    d64, dmx, dfull = rel(e64.cpu().numpy()), rel(emx.cpu().numpy()), rel(full_mx)
    print(f"  rank {rank}/{cfg.n_om}: FP64 compressed {d64:.2e}, split compressed {dmx:.2e}, "
          f"split uncompressed {dfull:.2e}", flush=True)
    assert d64 < 1e-8
    assert dmx < 1e-6
    torch.cuda.synchronize()
