"""Batched split-precision chi0 over many q (uneven PB per q, ragged per-k
pair windows, metal weights) in one persistent launch, against the FP64
contraction of the same inputs and the numpy calc_eps restatement."""
import numpy as np
import pytest

from oracle import chain
from pygw_amd import mesh, synth

pytestmark = pytest.mark.gpu


def _batch_inputs(n_q=4, n_om=9):
    import torch
    from pygw_amd import polarization as P
    cfg = synth.small_config(div=(4, 4, 4), n_occ=4, n_emp=7, n_metal=3, n_om=n_om, seed=61)
    batch = synth.metal_batch(cfg, n_q=n_q, pb_range=(40, 300), window_range=(5, 40))
    b = batch.bands
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kqid = mesh.k_minus_q(cfg.div)
    Ms, kcws = [], []
    for i in range(batch.n_q):
        kq = kqid[:, batch.q_index[i]]
        kn = P.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc, layout="native")
        win = torch.as_tensor(batch.windows[i], device="cuda").long()
        kcws.append(kn[win[:, 0], :, win[:, 1]].T.contiguous())     # [n_om, K_q]
        Ms.append(synth.device_window_elements(batch, i))              # complex64 [K_q, PB_q]
    return batch, Ms, kcws


def test_batch_matches_fp64_and_oracle():
    import torch
    from pygw_amd import polarization as P
    batch, Ms, kcws = _batch_inputs()
    print(f"case batch pb={list(batch.pb)} K={[m.shape[0] for m in Ms]}", flush=True)
    B = P.Chi0Batch(Ms, kcws)
    eps = [e.cpu().numpy() for e in B.run()]
    for i in range(batch.n_q):
        ref = P.chi0_accumulate(Ms[i][None].to(torch.complex128), kcws[i][None]).cpu().numpy()
        for w in range(ref.shape[2]):
            err = np.linalg.norm(eps[i][:, :, w] - ref[:, :, w]) / np.linalg.norm(ref[:, :, w])
            assert err < 1e-6, (i, w, err)
        if i == 0:  # numpy restatement of calc_eps on the same pairs
            m = Ms[0].cpu().numpy().astype(complex)
            want = chain.calc_eps_accumulate([(0, m)], lambda ik: kcws[0].cpu().numpy().T)
            assert np.linalg.norm(ref - want) / np.linalg.norm(want) < 1e-13
        # Hermitian: the mirrored tiles are exact conjugates
        e = eps[i]
        assert np.abs(e - np.conj(e.transpose(1, 0, 2))).max() <= 1e-6 * np.abs(e).max()
    B.close()
    torch.cuda.synchronize()


def test_batch_graph_schedules_autotune():
    import torch
    from pygw_amd import polarization as P
    batch, Ms, kcws = _batch_inputs(n_q=3)
    print("case batch graph", flush=True)
    B = P.Chi0Batch(Ms, kcws)
    eager = [e.clone() for e in B.run(schedule=2)]
    torch.cuda.synchronize()
    for s in (1, 4):
        got = B.run(schedule=s)
        for a, b in zip(got, eager):
            assert (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item() < 1e-6
    g = B.capture()
    for rep in range(3):
        for e in B.eps:
            e.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(B.eps, eager):
            assert torch.equal(a, b), rep
        B.run()  # eager work between the replays
        torch.cuda.synchronize()
    best, times = B.autotune()
    print(f"  autotune best={best} times={times}", flush=True)
    assert best in (1, 2) and times[4] == -1.0   # s = 4 is never auto-selected (accuracy)
    B.close()
    torch.cuda.synchronize()


def test_batch_inverse_matches_eps_inverse():
    import torch
    from pygw_amd import polarization as P
    batch, Ms, kcws = _batch_inputs(n_q=2)
    print("case batch inverse", flush=True)
    B = P.Chi0Batch(Ms, kcws, inverse=True)
    chi = [e.clone() for e in B.run(inverse=False)]
    got = B.invert()
    for a, c in zip(got, chi):
        want = P.eps_inverse(c.clone())
        assert (torch.linalg.norm(a - want) / torch.linalg.norm(want)).item() < 1e-12
    again = B.run()              # contraction + inversion in one call
    for a, c in zip(again, got):
        assert (torch.linalg.norm(a - c) / torch.linalg.norm(c)).item() < 1e-12
    B.close()
    torch.cuda.synchronize()


def test_batch_inverse_streams_graph():
    """eps^-1 of every q dealt to side streams: eager and graph replays (with
    eager work between them) equal the one-stream inversion bit for bit."""
    import torch
    from pygw_amd import polarization as P
    batch, Ms, kcws = _batch_inputs(n_q=4)
    print("case batch inverse streams graph", flush=True)
    B = P.Chi0Batch(Ms, kcws, inverse=True)
    chi = [e.clone() for e in B.run(inverse=False)]
    one = [e.clone() for e in B.invert(streams=1)]
    for e, c in zip(B.eps, chi):
        e.copy_(c)
    four = [e.clone() for e in B.invert(streams=4)]
    for a, b in zip(four, one):
        assert torch.equal(a, b)
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        B.invert(streams=4)
    for rep in range(3):
        for e, c in zip(B.eps, chi):
            e.copy_(c)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(B.eps, one):
            assert torch.equal(a, b), rep
        B.run(inverse=False)   # eager work between the replays
        torch.cuda.synchronize()
    B.close()
    torch.cuda.synchronize()


def test_batch_in_frequency_bases():
    """Every q contracted in its own balanced frequency basis (all 32 basis
    functions: the projection, contraction and expansion of every q with
    nothing truncated), expanded back: within 1e-6 of the FP64 contraction
    per q and frequency; a captured graph replays the eager result bit for
    bit.  These metal weights on a log mesh need 28 of 32 functions, so
    "auto" contracts frequency by frequency (more than 3/4 of n_om)."""
    import torch
    from pygw_amd import polarization as P
    batch, Ms, kcws = _batch_inputs(n_q=4, n_om=32)
    A = P.Chi0Batch(Ms, kcws, omega_rank="auto")
    assert A.omega_rank is None and A.rank == 32
    A.close()
    B = P.Chi0Batch(Ms, kcws, omega_rank=32)
    print(f"case batch frequency bases rank {B.rank}/32", flush=True)
    eager = [e.clone() for e in B.run()]
    torch.cuda.synchronize()
    for i in range(batch.n_q):
        ref = P.chi0_accumulate(Ms[i][None].to(torch.complex128), kcws[i][None])
        d = torch.linalg.matrix_norm((eager[i] - ref).permute(2, 0, 1))
        err = (d / torch.linalg.matrix_norm(ref.permute(2, 0, 1))).max().item()
        print(f"  q {i}: max rel {err:.2e}", flush=True)
        assert err < 1e-6
    g = B.capture()
    for rep in range(2):
        for e in B.eps:
            e.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(B.eps, eager):
            assert torch.equal(a, b), rep
        B.run()
        torch.cuda.synchronize()
    B.close()
