"""pygw_eps_inverse_noalloc (the inversion on the library's fixed device
array, no allocation) against pygw_eps_inverse with a caller workspace:
two calls in a row, the second on a smaller input, equal bit for bit (the
same kernels on the same inputs); sizes beyond the array are rejected
before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eps(pb, nom, seed):
    """chi0-like Fortran [pb, pb, nom] complex128 with 1 + eps positive definite."""
    import torch
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(nom, pb, 2 * pb)) + 1j * rng.normal(size=(nom, pb, 2 * pb))
    a = np.einsum("wik,wjk->ijw", x, x.conj()) / (2 * pb)        # Hermitian PSD per frequency
    t = torch.as_tensor(np.asfortranarray(a).reshape(-1, order="F"), device="cuda")
    return t


def _workspace_version(eps, pb, nom):
    import torch
    from pygw_amd import _lib
    sz = C.c_size_t(0)
    _lib.call("pygw_eps_inverse_workspace", pb, nom, C.byref(sz))
    ws = torch.empty(sz.value, dtype=torch.uint8, device="cuda")
    out = eps.clone()
    info = (C.c_int * 1)()
    _lib.call("pygw_eps_inverse", C.c_void_p(out.data_ptr()), pb, nom, C.c_void_p(ws.data_ptr()), sz.value,
              C.cast(info, C.c_void_p), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return out, info[0]


def _noalloc_version(eps, pb, nom, check=True):
    import torch
    from pygw_amd import _lib
    out = eps.clone()
    info = (C.c_int * 1)()
    _lib.call("pygw_eps_inverse_noalloc", C.c_void_p(out.data_ptr()), pb, nom,
              C.cast(info, C.c_void_p) if check else None, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return out, info[0]


def test_noalloc_matches_workspace_entry_point():
    import torch
    for pb, nom, seed in ((300, 12, 1), (130, 5, 2)):        # the second call is the smaller one
        print(f"case eps_inverse_noalloc pb={pb} nom={nom}", flush=True)
        eps = _eps(pb, nom, seed)
        want, i0 = _workspace_version(eps, pb, nom)
        got, i1 = _noalloc_version(eps, pb, nom)
        torch.cuda.synchronize()
        assert i0 == 0 and i1 == 0
        assert torch.equal(got, want)
        # and against numpy's inverse (hermitize, +1, inverse, -1)
        a = eps.cpu().numpy().reshape(pb, pb, nom, order="F")
        g = got.cpu().numpy().reshape(pb, pb, nom, order="F")
        for w in range(nom):
            h = (a[:, :, w] + a[:, :, w].conj().T) / 2 + np.eye(pb)
            ref = np.linalg.inv(h) - np.eye(pb)
            assert np.linalg.norm(g[:, :, w] - ref) / np.linalg.norm(ref + np.eye(pb)) < 1e-12


def test_noalloc_async_and_limits():
    import torch
    from pygw_amd import _lib
    pb, nom = 200, 4
    eps = _eps(pb, nom, 3)
    want, _ = _workspace_version(eps, pb, nom)
    got, _ = _noalloc_version(eps, pb, nom, check=False)   # no info: stays asynchronous
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    # beyond the fixed array: -15 before any launch (the pointer is never touched)
    sz = C.c_size_t(0)
    _lib.call("pygw_eps_inverse_workspace", 4000, 64, C.byref(sz))
    assert sz.value > (128 << 20)
    rc = _lib.lib().pygw_eps_inverse_noalloc(C.c_void_p(16), 4000, 64, None, None)
    assert rc == -15
    assert _lib.lib().pygw_eps_inverse_noalloc(None, 100, 4, None, None) == -10
    # an indefinite 1 + eps is reported as with the workspace entry point
    a = np.zeros((40, 40, 2), dtype=complex)
    a[:, :, 0] = a[:, :, 1] = np.diag(np.linspace(0.5, 2.0, 40))
    a[7, 7, 1] = -3.0
    t = torch.as_tensor(a.reshape(-1, order="F"), device="cuda")
    with pytest.raises(_lib.PygwError) as e:
        _noalloc_version(t, 40, 2)
    assert e.value.code == -13
