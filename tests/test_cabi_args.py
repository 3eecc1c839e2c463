"""C-ABI argument checks and workspace queries: every entry point rejects
null arrays and non-positive sizes with PYGW_E_ARG (-10) before touching a
device, and the workspace queries give the documented sizes.  Runs on the
host (no device calls)."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pygw_amd", "libpygw_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
E_ARG = -10


def _fn(name):
    from pygw_amd import _lib
    return getattr(_lib.lib(), name)


def _zeros(fn):
    """Arguments of the signature all null / zero."""
    out = []
    for t in fn.argtypes:
        if t in (C.c_double,):
            out.append(0.0)
        elif t is C.c_void_p or (isinstance(t, type) and issubclass(t, C._Pointer)):
            out.append(None)
        else:
            out.append(0)
    return out


@pytest.mark.parametrize("name", [
    "pygw_tetra_weights", "pygw_tetra_weights_svd", "pygw_chi0_accumulate", "pygw_chi0_autotune",
    "pygw_eps_inverse", "pygw_eps_inverse_noalloc", "pygw_mwm", "pygw_padecof", "pygw_padeg", "pygw_selfc_spline", "pygw_selfc_convolve", "pygw_pade_qp",
    "pygw_pade_qp_gw0", "pygw_selfc_qsum", "pygw_gw0_enk", "pygw_tetra_idos", "pygw_chi0_wings",
    "pygw_eps_inverse_head", "pygw_mwm_head", "pygw_chi0_batch_run", "pygw_chi0_batch_autotune",
    "pygw_chi0_batch_free", "pygw_chi0_workspace", "pygw_eps_inverse_workspace", "pygw_mwm_workspace",
    "pygw_tetra_weights_workspace", "pygw_chi0_wings_workspace", "pygw_eps_inverse_head_workspace",
    "pygw_coul_setev", "pygw_coul_setev_workspace", "pygw_minm_vhalf", "pygw_minm_is",
    "pygw_minm_is_workspace", "pygw_pw_basis", "pygw_pw_basis_workspace", "pygw_coul_pw_block",
    "pygw_mixed_coulomb", "pygw_mixed_coulomb_workspace", "pygw_coul_mixed_block",
    "pygw_omega_basis_workspace", "pygw_omega_basis", "pygw_omega_basis_many_workspace", "pygw_omega_basis_many",
    "pygw_omega_project", "pygw_omega_expand", "pygw_omega_expand_batched",
    "pygw_ewald_workspace", "pygw_ewald_summation", "pygw_mt_coulomb", "pygw_radial_olap_workspace", "pygw_radial_olap",
    "pygw_zero_momentum_workspace", "pygw_zero_momentum_coulomb",
    "pygw_minm_mt_plan", "pygw_minm_mt_workspace", "pygw_minm_mt", "pygw_minc", "pygw_minc2",
    "pygw_lapw_plan", "pygw_lapw_coef", "pygw_lapw_band",
    "pygw_mmat_vv_workspace", "pygw_mmat_vv", "pygw_mmat_cv", "pygw_calc_head_workspace", "pygw_calc_head",
    "pygw_wing_pm", "pygw_minm_vhalf_split_basis_bytes", "pygw_minm_vhalf_split_basis",
    "pygw_minm_vhalf_mixed_workspace", "pygw_minm_vhalf_mixed",
])
def test_null_and_zero_arguments_rejected(name):
    fn = _fn(name)
    assert fn(*_zeros(fn)) == E_ARG


def test_chi0_workspace_sizes():
    q = _fn("pygw_chi0_workspace")
    s = [C.c_size_t(0) for _ in range(4)]
    assert q(900, 2400, 64, 1, 1, C.byref(s[0])) == 0
    assert q(900, 2400, 64, 2, 1, C.byref(s[1])) == 0
    assert q(900, 2400, 64, 512, 1, C.byref(s[2])) == 0
    per_k = s[1].value - s[0].value
    # planes (16 h8 of 16 B per 32-pair chunk and padded column) + FP32 weights
    assert per_k == 75 * 16 * 1024 * 16 + 64 * 75 * 32 * 4
    # the slab is capped at 1 GiB of k-points; the part slots are 64 MiB
    assert s[2].value - s[0].value == ((1 << 30) // per_k - 1) * per_k
    fixed = s[0].value - per_k
    assert fixed >= 256 * 128 * 128 * 16
    assert q(900, 2400, 64, 1, 0, C.byref(s[3])) == 0 and s[3].value == 0   # FP64 needs none
    assert q(900, 2400, 64, 1, 2, C.byref(s[3])) == E_ARG


def test_other_workspace_queries():
    b = C.c_size_t(0)
    assert _fn("pygw_eps_inverse_workspace")(900, 64, C.byref(b)) == 0 and b.value > 0
    assert _fn("pygw_mwm_workspace")(64, 1120, 900, 64, 0, C.byref(b)) == 0 and b.value == 0
    assert _fn("pygw_mwm_workspace")(64, 1120, 900, 64, 1, C.byref(b)) == 0 and b.value > 0
    assert _fn("pygw_tetra_weights_workspace")(512, 2400, 64, C.byref(b)) == 0 and b.value > 0
    assert _fn("pygw_chi0_wings_workspace")(900, 64, 512, C.byref(b)) == 0 and b.value > 0
    a = C.c_size_t(0)
    assert _fn("pygw_eps_inverse_head_workspace")(900, 64, C.byref(a)) == 0
    _fn("pygw_eps_inverse_workspace")(900, 64, C.byref(b))
    assert a.value > b.value
    pb = (C.c_int * 3)(302, 700, 1099)
    npair = (C.c_int * 3)(60000, 77000, 90000)
    assert _fn("pygw_chi0_batch_workspace")(3, pb, npair, 48, C.byref(b)) == 0 and b.value > 0
    bad = (C.c_int * 3)(302, 0, 1099)
    assert _fn("pygw_chi0_batch_workspace")(3, bad, npair, 48, C.byref(b)) == E_ARG


def test_coulomb_workspace_and_bounds():
    q = _fn("pygw_coul_setev_workspace")
    s = C.c_size_t(0)
    # the matrix padded to whole 32 x 32 pivot blocks and the eigenvectors (complex128), the U's
    assert q(999, C.byref(s)) == 0 and s.value >= 2 * 1024 * 1024 * 16 + 1024 * 32 * 16
    assert q(16385, C.byref(s)) == E_ARG
    mv = _fn("pygw_minm_vhalf")
    p = C.c_void_p(16)     # never dereferenced: the checks come first
    # n_in above a leading dimension, and a bad output kind
    assert mv(p, 10, 8, 7, p, 8, 4, 0, p, None) == E_ARG
    assert mv(p, 10, 8, 8, p, 7, 4, 0, p, None) == E_ARG
    assert mv(p, 10, 8, 8, p, 8, 4, 2, p, None) == E_ARG


def test_small_workspaces_rejected_before_any_device_call():
    """A caller workspace below the size the work needs gives PYGW_E_ARG
    (checked before anything is launched; the pointers are never touched)."""
    p = C.c_void_p(16)
    ms, im, sw = C.c_int(0), C.c_int(0), C.c_int(0)
    f = _fn("pygw_coul_setev")
    assert f(p, 100, 1e-8, None, 0, p, p, p, None, C.byref(ms), C.byref(im), C.byref(sw), p, 1024, None) == E_ARG
    f = _fn("pygw_pw_basis")
    assert f(p, 50, 80, p, 1000, p, p, 3, 3, 3, p, p, 1024, None) == E_ARG
    f = _fn("pygw_minm_is")
    umk = (C.c_int * 3)(0, 0, 0)
    assert f(0, 4, 0, 8, p, 10, 50, p, 10, 50, umk, p, 30, 40, 50, 50, p, p, p, 100, p, 1000, p, 3, 3, 3, 100.0, p,
             30, 0, p, 64, None) == E_ARG


def test_muffin_tin_coulomb_workspaces_and_bounds():
    s = C.c_size_t(0)
    q = _fn("pygw_radial_olap_workspace")
    # the table of j_L(k r) r, L <= lmax (mode 0), or of sin(k r) (mode 1, lmax 0), at 256 bytes, then the weights
    up = lambda b: (b + 255) // 256 * 256  # noqa: E731
    assert q(0, 3, 781, 100, C.byref(s)) == 0 and s.value == up(4 * 781 * 100 * 8) + 781 * 8
    assert q(1, 0, 640, 7, C.byref(s)) == 0 and s.value == up(640 * 7 * 8) + 640 * 8
    assert q(1, 1, 640, 7, C.byref(s)) == E_ARG
    assert q(0, 2, 3, 7, C.byref(s)) == E_ARG      # RINT13g needs 4 points
    q = _fn("pygw_zero_momentum_workspace")
    assert q(3, 40, C.byref(s)) == 0 and s.value == 3 * 3 * 40 * 16
    p = C.c_void_p(16)     # never dereferenced: the checks come first
    # the Ewald lattice arguments are host arrays: a cubic cell of side 10 and 3 atoms
    dbl = lambda *v: (C.c_double * len(v))(*v)  # noqa: E731
    pq, pos, alat = dbl(0.1, 0.0, 0.0), dbl(*([0.0] * 9)), dbl(10.0, 10.0, 10.0)
    rbas = dbl(10.0, 0, 0, 0, 10.0, 0, 0, 0, 10.0)
    br2 = dbl(*(2 * 3.141592653589793 / 10.0 * x for x in (1, 0, 0, 0, 1, 0, 0, 0, 1)))
    ew = _fn("pygw_ewald_summation")
    ewq = _fn("pygw_ewald_workspace")
    assert ewq(pq, 8, pos, br2, rbas, alat, 0, 10.0, 10.0, 3, C.byref(s)) == 0 and s.value > 0
    # lam_max above the table of l, a bad ortho flag, eta 0, a cutoff of 64 cells, a small workspace
    assert ew(pq, 17, pos, br2, rbas, alat, 1000.0, 0, 10.0, 10.0, 1.0, 3, p, None, 0, None) == E_ARG
    assert ew(pq, 8, pos, br2, rbas, alat, 1000.0, 2, 10.0, 10.0, 1.0, 3, p, None, 0, None) == E_ARG
    assert ew(pq, 8, pos, br2, rbas, alat, 1000.0, 0, 10.0, 10.0, 0.0, 3, p, None, 0, None) == E_ARG
    assert ew(pq, 8, pos, br2, rbas, alat, 1000.0, 0, 640.0, 10.0, 1.0, 3, p, None, 0, None) == E_ARG
    assert ew(pq, 8, pos, br2, rbas, alat, 1000.0, 0, 10.0, 10.0, 1.0, 3, p, p, s.value - 1, None) == E_ARG
    mt = _fn("pygw_mt_coulomb")
    ims = (C.c_int * 3)(0, 20, 64)
    # iat outside [-1, nat), rows past loctmatsize
    assert mt(p, 2, 64, p, p, p, ims, p, 2, 6, p, 2145, p, p, 50, p, p, 6, p, None) == E_ARG
    assert mt(p, 1, 60, p, p, p, ims, p, 2, 6, p, 2145, p, p, 50, p, p, 6, p, None) == E_ARG
    ro = _fn("pygw_radial_olap")
    assert ro(0, 1, p, p, p, 4, 3, 781, p, 0.01, p, 100, p, p, 1024, None) == E_ARG   # workspace too small


def test_minm_mt_plan_layout():
    """The plan of two atom types (mult 1, 2), lmax 3: loctmatsize, channel
    counts (u and udot of 16 (l, m) plus the local orbitals), tiles of 64
    product functions, and the size query; bad arguments rejected."""
    f = _fn("pygw_minm_mt_plan")
    ints = lambda *v: (C.c_int * len(v))(*v)  # noqa: E731
    mult, nmix, ncore = ints(1, 2), ints(6, 5), ints(2, 1)
    big_l = ints(0, 0, 1, 1, 2, 3, 0, 1, 1, 2, 0, 0)          # (max_nmix 6, nat 2) Fortran
    nlo = ints(1, 0, 2, 0, 0, 1, 1, 0)                         # (lomax+1 4, nat 2) Fortran
    n = C.c_int(0)
    assert f(2, mult, nmix, big_l, 6, nlo, 3, ncore, 3, None, C.byref(n)) == 0 and n.value > 12
    plan = (C.c_int * n.value)()
    assert f(2, mult, nmix, big_l, 6, nlo, 3, ncore, 3, plan, C.byref(n)) == 0
    h = list(plan)
    nat, ndf, lmax, nlm, cmax, loct, ntile = h[:7]
    assert (nat, ndf, lmax, nlm) == (2, 3, 3, 16)
    assert loct == (1 + 1 + 3 + 3 + 5 + 7) + 2 * (1 + 3 + 3 + 5 + 1)
    # channels: type 0: 32 + 1 (l=0) + 2*5 (l=2) = 43; type 1: 32 + 3 + 5 = 40
    assert cmax == 43 and ntile == 3
    q = _fn("pygw_minm_mt_workspace")
    s = C.c_size_t(0)
    assert q(plan, 5, 7, C.byref(s)) == 0 and s.value == (5 * 43 * loct * 16 + 255) // 256 * 256 + 3 * 43 * 7 * 16
    small = C.c_int(n.value - 1)
    assert f(2, mult, nmix, big_l, 6, nlo, 3, ncore, 3, plan, C.byref(small)) == E_ARG
    assert f(2, ints(1, 0), nmix, big_l, 6, nlo, 3, ncore, 3, None, C.byref(n)) == E_ARG   # mult 0
    p = C.c_void_p(16)
    mt = _fn("pygw_minm_mt")
    # a window past nbmax, and nlomax below the plan's largest nLO_at (2)
    assert mt(plan, p, p, 0, 8, 0, 5, p, p, p, 7, 7, p, p, p, 9, 9, 2, p, 6, 13, p, p, 100, p, loct, 0, None, 0,
              None) == E_ARG
    assert mt(plan, p, p, 0, 5, 0, 5, p, p, p, 7, 7, p, p, p, 9, 9, 1, p, 6, 13, p, p, 100, p, loct, 0, None, 0,
              None) == E_ARG


def test_lapw_plan_layout_and_checks():
    """The plan of lapw_system's two atom types (mult 1, 2; nt 4, lomax 3,
    nloat 3): header and size; inconsistent local orbitals, a zero APW p and
    a short buffer rejected."""
    import numpy as np
    from pygw_amd import synth
    s = synth.lapw_system()
    f = _fn("pygw_lapw_plan")
    I = lambda a: np.ascontiguousarray(np.asarray(a).ravel(order="F"), dtype=np.int32)  # noqa: E731
    D = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))  # noqa: E731
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    nt, lomax, nloat = s.umt.shape[1], s.nLO_at.shape[0] - 1, s.abcelo.shape[2]

    def call(nlo_at=s.nLO_at, umt=s.umt, plan=None, n=None):
        arrs = [I(s.mult), I(s.lapw), I(s.nlo), I(nlo_at), D(umt), D(s.abcelo), D(s.rmt), D(s.pos), D(s.rotloc),
                D(s.trotij), D(s.k2cartes)]
        n = n or C.c_int(0)
        rc = f(2, P(arrs[0]), nt, lomax, nloat, P(arrs[1]), P(arrs[2]), P(arrs[3]), P(arrs[4]), P(arrs[5]),
               P(arrs[6]), P(arrs[7]), P(arrs[8]), P(arrs[9]), s.vol, P(arrs[10]), plan, C.byref(n))
        return rc, n.value

    rc, n = call()
    atw = 3 + nt + 2 * (lomax + 1) + 4 * nt + 3 * (lomax + 1) * nloat + 9
    assert rc == 0 and n == 16 + 9 + 2 * atw + 13 * 3
    plan = np.zeros(n)
    assert call(plan=P(plan), n=C.c_int(n))[0] == 0
    assert list(plan[:6]) == [2, nt, lomax, nloat, 3, 1] and plan[6] == s.vol
    assert call(plan=P(plan), n=C.c_int(n - 1))[0] == E_ARG
    bad = s.nLO_at.copy()
    bad[1, 1] += 1          # more local orbitals than the lohns block holds
    assert call(nlo_at=bad)[0] == E_ARG
    u = s.umt.copy()
    u[1, 0, 0] = 0.0        # APW l = 0 of type 1 with p = 0
    assert call(umt=u)[0] == E_ARG


def test_minm_is_workspace_follows_the_cheaper_association():
    """The interstitial workspace query gives the pair-density buffers (map,
    inverse map, rho, the slice sums) where that association costs fewer
    products, and the dense T-first buffers (map, T, P) otherwise."""
    q = _fn("pygw_minm_is_workspace")
    b = C.c_size_t(0)
    al = lambda x: (x + 255) // 256 * 256  # noqa: E731
    # 600 plane waves, 400 of 700 basis plane waves, 20 x 120 bands: the pair density
    assert q(400, 700, 600, 600, 20, 120, C.byref(b)) == 0
    s1 = 3 * 120 * 20 * 700          # three slices of j (600) over 2 x 219 tiles
    s2 = 4 * 20 * 120 * 400          # four slices of K (700) over 38 x 7 tiles
    assert b.value == al(600 * 600 * 4) + al(700 * 600 * 4) + al(20 * 120 * 700 * 16) + max(s1, s2) * 16
    # 10 basis plane waves against 60, 20 x 50 bands: T first (dense)
    assert q(10, 60, 90, 80, 20, 50, C.byref(b)) == 0
    assert b.value == al(90 * 80 * 4) + 10 * 80 * 20 * 16 + 60 * 80 * 20 * 16


def test_vhalf_split_sizes():
    b = C.c_size_t(0)
    assert _fn("pygw_minm_vhalf_split_basis_bytes")(1000, 900, C.byref(b)) == 0
    assert b.value == 3840 + 32 * 16 * 1024 * 16        # column scales (900 floats, 256-B aligned), planes
    assert _fn("pygw_minm_vhalf_mixed_workspace")(6720, 1000, C.byref(b)) == 0
    assert b.value == 26880 + 32 * 16 * 6784 * 16
