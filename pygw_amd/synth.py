"""Seeded synthetic inputs for the polarization path.

Band model: e_n(k) = c_n + t_n * S(k), S(k) = (1/3) * sum_i cos(2 pi k_i) in
[-1, 1].  Occupied c_n ~ U[-10,-1] eV, empty c_n ~ U[1,20] eV, t_n ~ U[0.2,1]
eV; metals add bands with c_n ~ U[-2,2] eV that cross E_F = 0.  (The unnormalised
sum has amplitude 3 t_n, which would push roughly half of all seeds through
E_F and turn the stated "insulator" into a metal; the 1/3 keeps every stated
range and keeps the gap open.)  Energies are sorted per k and converted to
Hartree, E_F = 0 (pygw.py:785-794).

Matrix elements: M[k, p, I] with p = i_occ * n_e + i_emp (the C-order ravel of
pygw.py:3170), complex Gaussian with Re, Im ~ N(0, 1/2), scaled by
exp(-|e_n(k) - e_m(k-q)| / 20 eV) / sqrt(PB) and by v_half[I].  The Gaussian
is counter-based (SplitMix64 of (seed, linear index), Box-Muller), so the
device generator and `gaussian_pairs` below produce the same draws up to the
last-ulp differences of log/cos/sin.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh

BASE_SEED = 20261203
MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


@dataclass
class PolConfig:
    name: str
    div: tuple
    n_occ: int
    n_emp: int
    n_om: int
    wmin_ev: float
    wmax_ev: float
    pb: int
    q: tuple = (1, 0, 0)          # integer q on the k-mesh
    n_metal: int = 0              # bands crossing E_F (c_n ~ U[-2,2] eV)
    vhalf: str = "none"           # "none" | "coulomb" (4 pi / |q+G|^2, literal)
    seed: int = BASE_SEED
    mesh_kind: str = "log"        # "log" | "tan"
    extra: dict = field(default_factory=dict)

    @property
    def nk(self):
        return int(np.prod(self.div))

    @property
    def ntet(self):
        return 6 * self.nk


# The reference's five workloads (stated in the benchmark description).
CONFIGS = {
    "c1": PolConfig("c1", (4, 4, 4), 8, 24, 32, 0.01, 100.0, 200, q=(1, 0, 0), seed=BASE_SEED + 1),
    "c2": PolConfig("c2", (8, 8, 8), 16, 48, 64, 0.005, 200.0, 600, q=(1, 0, 0), vhalf="coulomb",
                    seed=BASE_SEED + 2),
    "c3": PolConfig("c3", (10, 10, 10), 32, 200, 64, 0.005, 200.0, 1200, q=(1, 0, 0), seed=BASE_SEED + 3),
    "c4": PolConfig("c4", (8, 8, 8), 16, 48, 48, 0.005, 200.0, 700, n_metal=3, seed=BASE_SEED + 4),
    "c5": PolConfig("c5", (8, 8, 8), 20, 120, 64, 0.005, 200.0, 900, q=(1, 0, 0), vhalf="coulomb",
                    seed=BASE_SEED + 5, mesh_kind="tan"),
}


def small_config(name="tiny", div=(2, 2, 2), n_occ=2, n_emp=3, n_om=6, pb=8, n_metal=0, seed=7,
                 wmin_ev=0.01, wmax_ev=100.0, vhalf="none", q=(1, 0, 0)):
    return PolConfig(name, tuple(div), n_occ, n_emp, n_om, wmin_ev, wmax_ev, pb, q=tuple(q),
                     n_metal=n_metal, vhalf=vhalf, seed=seed)


@dataclass
class Bands:
    enk: np.ndarray        # Fortran (nbnd, N_k), Hartree, E_F = 0
    ncg: int
    nomx: int
    numin: int
    ef: float = 0.0

    @property
    def n_o(self):
        return self.ncg + self.nomx + 1

    @property
    def n_e(self):
        return self.enk.shape[0] - self.ncg - self.numin


def band_energies(cfg: PolConfig) -> Bands:
    rng = np.random.default_rng(cfg.seed)
    c_occ = rng.uniform(-10.0, -1.0, cfg.n_occ)
    c_emp = rng.uniform(1.0, 20.0, cfg.n_emp)
    c_met = rng.uniform(-2.0, 2.0, cfg.n_metal)
    c = np.concatenate([c_occ, c_met, c_emp])
    t = rng.uniform(0.2, 1.0, c.size)
    kf = mesh.kpoint_coords(cfg.div)
    s = np.cos(2 * np.pi * kf).sum(axis=1) / 3.0
    e = (c[:, None] + t[:, None] * s[None, :]) / mesh.H2EV   # [nbnd, nk]
    e = np.sort(e, axis=0)
    nocc_k = (e < 0.0).sum(axis=0)
    nomx = int(nocc_k.max()) - 1      # pygw.py:1002-1009
    numin = int(nocc_k.min())
    return Bands(np.asfortranarray(e), 0, nomx, numin, 0.0)


def omega_mesh(cfg: PolConfig):
    """(omega, womeg) in Hartree.  Log mesh: womeg is the trapezoid rule."""
    if cfg.mesh_kind == "tan":
        fm = mesh.FrequencyMesh(cfg.n_om, cfg.wmin_ev / mesh.H2EV, cfg.wmax_ev / mesh.H2EV)
        return fm.omega, fm.womeg
    om = mesh.log_mesh(cfg.wmin_ev, cfg.wmax_ev, cfg.n_om)
    w = np.zeros_like(om)
    d = np.diff(om)
    w[:-1] += 0.5 * d
    w[1:] += 0.5 * d
    return om, w


def g_vectors(cfg: PolConfig) -> np.ndarray:
    """PB distinct integer G with |G| <= 6 (more if PB exceeds their count),
    seeded; dropping G = 0 at q = 0."""
    r = 6
    while True:
        rng_ = np.arange(-r, r + 1)
        g = np.array(np.meshgrid(rng_, rng_, rng_, indexing="ij")).reshape(3, -1).T
        g = g[(g ** 2).sum(1) <= r * r]
        if all(x == 0 for x in cfg.q):
            g = g[(g ** 2).sum(1) > 0]
        if len(g) >= cfg.pb:
            break
        r += 1
    rng = np.random.default_rng(cfg.seed + 77)
    sel = rng.choice(len(g), cfg.pb, replace=False)
    return g[np.sort(sel)]


def vhalf(cfg: PolConfig) -> np.ndarray:
    """Diagonal v_half[I]: 4 pi / |q+G_I|^2 taken literally (q, G in units of
    the reciprocal-lattice vector of a unit cube) or ones."""
    if cfg.vhalf == "none":
        return np.ones(cfg.pb)
    g = g_vectors(cfg).astype(np.float64)
    qf = np.array(cfg.q, dtype=np.float64) / np.array(cfg.div, dtype=np.float64)
    qg2 = ((qf[None, :] + g) ** 2).sum(1)
    return 4.0 * np.pi / qg2


def q_index(cfg: PolConfig) -> int:
    return int(mesh.k2indx(np.array(cfg.q) % np.array(cfg.div), np.array(cfg.div)))


# ---------------------------------------------------------------------------
# counter-based Gaussian (host twin of the device generator)
# ---------------------------------------------------------------------------

def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15)) & MASK64
        z = x
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK64
        return z ^ (z >> np.uint64(31))


def gaussian_pairs(seed: int, idx: np.ndarray):
    """(re, im) with Re, Im ~ N(0, 1/2) for linear indices idx (uint64)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)) & MASK64
        a = _splitmix64((base ^ (idx * np.uint64(2))) & MASK64)
        b = _splitmix64((base ^ (idx * np.uint64(2) + np.uint64(1))) & MASK64)
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)  # (0, 1]
    u2 = (b >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)          # [0, 1)
    r = np.sqrt(-np.log(u1))
    return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)


def pair_scale(bands: Bands, kqid_col: np.ndarray, ik: int) -> np.ndarray:
    """exp(-|e_n(k) - e_m(k-q)| / 20 eV) for p = i_o * n_e + i_e."""
    no, ne = bands.n_o, bands.n_e
    eo = bands.enk[:no, ik]
    ee = bands.enk[bands.ncg + bands.numin:, kqid_col[ik]]
    d = np.abs(eo[:, None] - ee[None, :]) * mesh.H2EV
    return np.exp(-d / 20.0).reshape(no * ne)


def matrix_elements(cfg: PolConfig, bands: Bands, kqid_col: np.ndarray, ks=None,
                    dtype=np.complex128) -> np.ndarray:
    """M[k, p, I] for the k in `ks` (default: all), host generator."""
    ks = range(cfg.nk) if ks is None else ks
    no, ne, pb = bands.n_o, bands.n_e, cfg.pb
    np_ = no * ne
    vh = vhalf(cfg)
    out = np.empty((len(ks), np_, pb), dtype=dtype)
    for i, ik in enumerate(ks):
        idx = (np.uint64(ik) * np.uint64(np_ * pb)) + np.arange(np_ * pb, dtype=np.uint64)
        re, im = gaussian_pairs(cfg.seed, idx)
        m = (re + 1j * im).reshape(np_, pb)
        m *= (pair_scale(bands, kqid_col, ik) / np.sqrt(pb))[:, None]
        m *= vh[None, :]
        out[i] = m
    return out


def device_matrix_elements(cfg: PolConfig, bands: Bands, kqid_col, k0=0, nkc=None, complex128=True,
                           device="cuda", out=None, rows=None, seed=None):
    """M[k0:k0+nkc] generated on the device (same draws as matrix_elements
    up to last-ulp differences of log/sincos).  Returns a torch tensor
    [nkc, n_rows*n_cols, PB].  rows = (row_band0, n_rows, col_band0, n_cols)
    selects other band windows (default: the polarization pairs)."""
    import ctypes as C

    import torch

    from . import _lib

    nkc = cfg.nk - k0 if nkc is None else nkc
    pb = cfg.pb
    if rows is None:
        rows = (0, bands.n_o, bands.ncg + bands.numin, bands.n_e)
    r0, no, c0, ne = rows
    dt = torch.complex128 if complex128 else torch.complex64
    enk = torch.as_tensor(np.ascontiguousarray(bands.enk.T), dtype=torch.float64, device=device)
    kq = torch.as_tensor(np.asarray(kqid_col, dtype=np.int32), device=device)
    vh = torch.as_tensor(vhalf(cfg), dtype=torch.float64, device=device)
    if out is None:
        out = torch.empty((nkc, no * ne, pb), dtype=dt, device=device)
    _lib.call("pygw_synth_minm", C.c_void_p(enk.data_ptr()), bands.enk.shape[0], C.c_void_p(kq.data_ptr()),
              r0, c0, no, ne, pb, C.c_void_p(vh.data_ptr()), C.c_ulonglong(cfg.seed if seed is None else seed), k0, nkc,
              1 if complex128 else 0, C.c_void_p(out.data_ptr()),
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.current_stream().synchronize()
    return out


# ---------------------------------------------------------------------------
# Batched uneven metal workload: many q with their own PB and ragged per-k
# pair windows
# ---------------------------------------------------------------------------

@dataclass
class QBatch:
    cfg: PolConfig              # band model, mesh, frequencies (pb unused)
    bands: Bands
    q_index: np.ndarray         # [n_q] linear k-index of each q
    pb: np.ndarray              # [n_q] product-basis size per q
    windows: list               # per q: int32 [K_q, 2] (k, p) pairs, p = i_o*n_e + i_e

    @property
    def n_q(self):
        return len(self.pb)

    def q_vector(self, i):
        d = np.array(self.cfg.div)
        j = int(self.q_index[i])
        return (int(j // (d[1] * d[2])), int((j // d[2]) % d[1]), int(j % d[2]))

    def q_config(self, i):
        return PolConfig(**{**self.cfg.__dict__, "q": self.q_vector(i), "pb": int(self.pb[i])})


def metal_batch(cfg: PolConfig, n_q=32, pb_range=(300, 1100), window_range=(40, 260)) -> QBatch:
    """Seeded batch: n_q distinct q != 0, PB_q ~ U{pb_range}, and for every
    (q, k) a window of n_k ~ U{window_range} pairs -- the n_k pairs with the
    smallest transition energy |e_m(k-q) - e_n(k)| (lowest-energy
    transitions first)."""
    bands = band_energies(cfg)
    rng = np.random.default_rng(cfg.seed + 404)
    nk = cfg.nk
    qidx = rng.permutation(np.arange(1, nk))[:n_q]
    pbs = rng.integers(pb_range[0], pb_range[1] + 1, size=n_q)
    kqid = mesh.k_minus_q(cfg.div)
    no, ne = bands.n_o, bands.n_e
    c0 = bands.ncg + bands.numin
    wins = []
    for iq in range(n_q):
        kq = kqid[:, qidx[iq]]
        nwin = rng.integers(window_range[0], min(window_range[1], no * ne) + 1, size=nk)
        rows = []
        for k in range(nk):
            d = np.abs(bands.enk[c0:, kq[k]][None, :] - bands.enk[:no, k][:, None]).reshape(-1)
            sel = np.sort(np.argsort(d, kind="stable")[:nwin[k]])
            rows.append(np.stack([np.full(sel.size, k), sel], axis=1))
        wins.append(np.ascontiguousarray(np.concatenate(rows).astype(np.int32)))
    return QBatch(cfg, bands, qidx.astype(np.int64), pbs.astype(np.int64), wins)


def device_window_elements(batch: QBatch, i, complex128=False, device="cuda"):
    """M rows of q i for its window pairs, [K_q, PB_q] on the device: the same
    draws as matrix_elements / device_matrix_elements at (k, p)."""
    import ctypes as C

    import torch

    from . import _lib

    cfg = batch.q_config(i)
    b = batch.bands
    kq = mesh.k_minus_q(cfg.div)[:, batch.q_index[i]]
    win = batch.windows[i]
    dt = torch.complex128 if complex128 else torch.complex64
    enk = torch.as_tensor(np.ascontiguousarray(b.enk.T), dtype=torch.float64, device=device)
    kqd = torch.as_tensor(np.asarray(kq, dtype=np.int32), device=device)
    vh = torch.as_tensor(vhalf(cfg), dtype=torch.float64, device=device)
    pr = torch.as_tensor(win, device=device)
    out = torch.empty((win.shape[0], cfg.pb), dtype=dt, device=device)
    _lib.call("pygw_synth_minm_pairs", C.c_void_p(enk.data_ptr()), b.enk.shape[0], C.c_void_p(kqd.data_ptr()),
              0, b.ncg + b.numin, b.n_o, b.n_e, cfg.pb, C.c_void_p(vh.data_ptr()), C.c_ulonglong(cfg.seed),
              C.c_void_p(pr.data_ptr()), win.shape[0], 1 if complex128 else 0, C.c_void_p(out.data_ptr()),
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.current_stream().synchronize()
    return out


@dataclass
class GW0System:
    """Inputs of the GW0 iteration on stored mwm (postprocess.py:206-321) for
    a mesh without symmetry (every k irreducible).  ebnd [n_k, n_band] Kohn-
    Sham energies (Hartree, E_F = 0); the GW bands are ibgw .. ibgw+nb1-1;
    core [n_core]; jrk [n_q, n_k] the index of k - q; mwm [n_q, n_k, n_om,
    nb1, nb2] (the mwm.{iq}.{irk}.npy stack); sigc0 [n_k, nb1, n_om] the
    G0W0 Sigma_c (the sum over q at the Kohn-Sham bands); sigx, vxc
    [n_k, nb1]; atet [n_tet, 4], wtet [n_tet]; nval the valence electrons
    (spin 2) counted from band 0."""
    div: tuple
    ebnd: np.ndarray
    core: np.ndarray
    jrk: np.ndarray
    mwm: np.ndarray
    sigx: np.ndarray
    vxc: np.ndarray
    omega: np.ndarray
    womeg: np.ndarray
    omega_p: np.ndarray
    womeg_p: np.ndarray
    atet: np.ndarray
    wtet: np.ndarray
    nval: int
    ibgw: int


def gw0_system(div=(2, 2, 2), n_occ=4, n_emp=6, nbgw=4, n_core=1, n_om=16, iop_multiple=5, q_list=None, seed=11,
               mwm_array=True):
    """A seeded insulating GW0 system.  Bands as band_energies with a gap of
    at least 3 eV (occupied c_n ~ U[-10,-1.5] eV, empty ~ U[1.5,20] eV);
    nb2 = n_core + n_occ + n_emp.  mwm(i w) of each (q, k, ie1, ie2) is a
    pole form -a 2W/(w^2 + W^2) with a small imaginary part ib 2W/(w^2 +
    W^2), W ~ U[0.3, 1.5] Ha, a ~ U[0, 0.15]/(nb2 n_q) (Sigma_c of order
    0.1 Ha), b = a/10 for ie1 != ie2.  sigx = -U[0.3,0.6] Ha (occupied)
    / -U[0.05,0.15] (empty); vxc = sigx + U[0, 0.03] (occupied) / - U[0, 0.03]
    (empty), so that sigx - vxc opens the gap by up to 1.6 eV."""
    rng = np.random.default_rng(seed)
    nk = int(np.prod(div))
    c = np.concatenate([rng.uniform(-10.0, -1.5, n_occ), rng.uniform(1.5, 20.0, n_emp)])
    t = rng.uniform(0.2, 1.0, c.size)
    s = np.cos(2 * np.pi * mesh.kpoint_coords(div)).sum(axis=1) / 3.0
    e = np.sort((c[None, :] + t[None, :] * s[:, None]) / mesh.H2EV, axis=1)   # [nk, nband]
    core = np.sort(rng.uniform(-8.0, -4.0, n_core))
    ibgw = n_occ - nbgw // 2
    nb1, nb2 = nbgw, n_core + n_occ + n_emp
    kqid = mesh.k_minus_q(div)
    q_list = np.arange(nk) if q_list is None else np.asarray(q_list)
    jrk = np.ascontiguousarray(kqid[:, q_list].T.astype(np.int32))      # [n_q, n_k]
    fm = mesh.FrequencyMesh(n_om, 0.02, 4.0, iop_multiple=iop_multiple)
    nq = len(q_list)
    mwm = None
    if mwm_array:
        a = rng.uniform(0.0, 0.15, (nq, nk, nb1, nb2)) / (nb2 * nq)
        W = rng.uniform(0.3, 1.5, (nq, nk, nb1, nb2))
        b = a / 10.0 * rng.uniform(-1.0, 1.0, a.shape)
        diag = np.zeros((nb1, nb2), bool)   # (ie1, ie2) of the same state: real
        for i in range(nb1):
            diag[i, n_core + ibgw + i] = True
        b[:, :, diag] = 0.0
        om = fm.omega[None, None, :, None, None]
        f = 2 * W[:, :, None] / (om ** 2 + W[:, :, None] ** 2)          # [nq, nk, nom, nb1, nb2]
        mwm = (-a[:, :, None] * f + 1j * b[:, :, None] * f).astype(np.complex128)
    occ = np.arange(ibgw, ibgw + nb1)[None, :] < n_occ
    sigx = np.where(occ, -rng.uniform(0.3, 0.6, (nk, nb1)), -rng.uniform(0.05, 0.15, (nk, nb1)))
    vxc = sigx + np.where(occ, rng.uniform(0.0, 0.03, (nk, nb1)), -rng.uniform(0.0, 0.03, (nk, nb1)))
    tetc = mesh.tetra_corners(div)
    atet = np.ascontiguousarray(tetc.T.astype(np.int32))
    wtet = np.full(atet.shape[0], 1.0 / atet.shape[0])
    return GW0System(tuple(div), e, core, jrk, mwm, sigx, vxc, fm.omega, fm.womeg, fm.omega_precise,
                     fm.womeg_precise, atet, wtet, 2 * n_occ, ibgw)


def coulomb_matrix(mbsize=1000, n_null=100, seed=20261211):
    """Synthetic Hermitian Coulomb matrix [mbsize, mbsize] (complex128, C
    order) of known spectrum: mbsize - n_null eigenvalues 4 pi / g^2 with
    g ~ U[0.25, 9] Bohr^-1 (the plane-wave part's range), and n_null
    near-null eigenvalues 10^U[-12, -8] (the linear dependences that
    barcevtol removes), in the eigenbasis of a random unitary.  Returns
    (vmat, eigenvalues)."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.25, 9.0, mbsize - n_null)
    lam = np.concatenate([4 * np.pi / g ** 2, 10 ** rng.uniform(-12, -8, n_null)])
    Q, R = np.linalg.qr(rng.normal(size=(mbsize, mbsize)) + 1j * rng.normal(size=(mbsize, mbsize)))
    Q = Q * (np.diag(R) / np.abs(np.diag(R)))
    A = (Q * lam) @ Q.conj().T
    return (A + A.conj().T) / 2, np.sort(lam)


@dataclass
class InterstitialSystem:
    """Inputs of calc_minm_IS (for_vxcnn.f90:1030) with PyGW's conventions:
    gindex [ng, 3] integer G vectors (shortest first), i_g0 the inverse map
    [2m+1, 2m+1, 2m+1] (-1 = none), indgk_i / indgk_j the plane waves of k and
    k - q (indices into gindex, ngi >= nvi entries), indgq the product-basis
    plane waves and indgq_inverse its inverse, mpwipw [ngq, ngq_barc],
    Aeigk / Aeigq [nbmax, ngi / ngj] coefficients; all indices 0-based."""
    gindex: np.ndarray
    i_g0: np.ndarray
    indgk_i: np.ndarray
    indgk_j: np.ndarray
    indgq: np.ndarray
    indgq_inverse: np.ndarray
    mpwipw: np.ndarray
    Aeigk: np.ndarray
    Aeigq: np.ndarray
    iumklap: np.ndarray
    nvi: int
    nvj: int
    vol: float

    def args(self, nstart, nend, mstart, mend):
        return (nstart, nend, mstart, mend, self.Aeigk, self.Aeigq, self.iumklap, self.mpwipw, self.nvi, self.nvj,
                self.indgk_i, self.indgk_j, self.indgq_inverse, self.gindex, self.i_g0, self.vol)


def interstitial_system(seed=7, gmax=3, nvi=37, nvj=41, ngq=23, ngq_barc=29, nb=9, vol=270.0,
                        iumklap=(1, 0, -1)) -> InterstitialSystem:
    rng = np.random.default_rng(seed)
    r = np.arange(-gmax, gmax + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    g = g[np.lexsort((g[:, 2], g[:, 1], g[:, 0], (g ** 2).sum(1)))]      # shortest first
    i_g0 = -np.ones((2 * gmax + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gmax, g[:, 1] + gmax, g[:, 2] + gmax] = np.arange(len(g))
    short = np.arange(min(len(g), 3 * max(nvi, nvj)))
    indgk_i = np.concatenate([rng.permutation(short)[:nvi], rng.integers(0, len(g), 3)])
    indgk_j = np.concatenate([rng.permutation(short)[:nvj], rng.integers(0, len(g), 2)])
    indgq = rng.permutation(np.arange(2 * ngq_barc))[:ngq_barc]
    inv = -np.ones(indgq.max() + 1, dtype=np.int64)
    inv[indgq] = np.arange(ngq_barc)
    cplx = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    return InterstitialSystem(gindex=g.astype(np.int64), i_g0=i_g0, indgk_i=indgk_i, indgk_j=indgk_j, indgq=indgq,
                              indgq_inverse=inv, mpwipw=cplx(ngq, ngq_barc), Aeigk=cplx(nb, len(indgk_i)),
                              Aeigq=cplx(nb, len(indgk_j)), iumklap=np.array(iumklap), nvi=nvi, nvj=nvj, vol=vol)


@dataclass
class PlaneWaveSystem:
    """Inputs of OrthogonalizedBasisInterstitials and of the plane-wave block
    of Coulomb for a cubic cell of side a with two muffin-tin spheres:
    gindex, i_g0 as in InterstitialSystem; indgq the product-basis plane
    waves (the ngq_barc shortest, their differences inside the i_g0 box);
    ipwint[G] = (1/V) integral over the interstitial of e^{iG.r}; gqlen =
    |q + G| of the basis."""
    gindex: np.ndarray
    i_g0: np.ndarray
    indgq: np.ndarray
    ipwint: np.ndarray
    gqlen: np.ndarray
    ngq: int
    ngq_barc: int


def plane_wave_system(ngq=60, ngq_barc=90, gmax=8, a=10.0, rmt=2.0, q=(0.1, 0.0, 0.0), seed=5) -> PlaneWaveSystem:
    rng = np.random.default_rng(seed)
    r = np.arange(-gmax, gmax + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    g = g[np.lexsort((g[:, 2], g[:, 1], g[:, 0], (g ** 2).sum(1)))]
    i_g0 = -np.ones((2 * gmax + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gmax, g[:, 1] + gmax, g[:, 2] + gmax] = np.arange(len(g))
    half = np.flatnonzero(np.abs(g).max(1) <= gmax // 2)
    indgq = np.concatenate([half[:ngq], rng.permutation(half[ngq:])[:ngq_barc - ngq]])
    vol = a ** 3
    gc = 2 * np.pi / a * g
    x = np.linalg.norm(gc, axis=1) * rmt
    with np.errstate(invalid="ignore", divide="ignore"):
        j1x = np.where(x > 1e-8, (np.sin(x) / x ** 2 - np.cos(x) / x) / x, 1.0 / 3.0)
    tau = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    ph = np.exp(2j * np.pi * g @ tau.T).sum(1)
    ipwint = (np.all(g == 0, axis=1).astype(float) - ph * 4 * np.pi * rmt ** 3 / vol * j1x).astype(complex)
    gqlen = np.linalg.norm(gc[indgq] + 2 * np.pi / a * np.asarray(q), axis=1)
    return PlaneWaveSystem(gindex=g.astype(np.int64), i_g0=i_g0, indgq=indgq, ipwint=ipwint, gqlen=gqlen, ngq=ngq,
                           ngq_barc=ngq_barc)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@dataclass
class MixedSystem:
    """Inputs of Mixed_Coulomb (for_Coulomb.f90:327) for a cubic cell: two
    inequivalent atoms (mult 1 and 2), product functions of L <= 3 on each,
    the product-basis plane waves sorted by |q + G| with the starts of the
    equal-length groups (G_unique) and those lengths (gqlen0), and
    synthetic Bessel integrals jlam[atom] [nmix, ngs]."""
    vq: np.ndarray
    gindex: np.ndarray
    indgq: np.ndarray
    gqlen: np.ndarray
    gqlen0: np.ndarray
    G_unique: np.ndarray
    jlam: list
    big_l: list
    k2cartes: np.ndarray
    rotloc: np.ndarray
    trotij: np.ndarray
    mult: np.ndarray
    pos: np.ndarray
    vol: float
    ngq_barc: int

    def args(self, iat):
        return (self.vq, iat, True, self.jlam[iat], self.gindex, self.indgq, self.gqlen0, self.G_unique,
                self.ngq_barc, self.big_l[iat], self.k2cartes, self.rotloc, self.trotij, self.mult, self.pos, self.vol)


def mixed_system(ngq_barc=120, q=(0.25, 0.0, 0.125), a=10.0, gmax=6, seed=3) -> MixedSystem:
    rng = np.random.default_rng(seed)
    r = np.arange(-gmax, gmax + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    vq = np.asarray(q, dtype=float)
    k2c = 2 * np.pi / a * np.eye(3)
    ln = np.linalg.norm((vq[None, :] + g) @ k2c.T, axis=1)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.round(ln, 10)))
    indgq = order[:ngq_barc]
    gql = ln[indgq]
    starts = np.flatnonzero(np.r_[True, np.abs(np.diff(gql)) > 1e-10])
    big_l = [np.array([0, 0, 1, 1, 2, 3]), np.array([0, 1, 2, 2, 3])]
    mult = np.array([1, 2])
    jl = [rng.normal(size=(len(b), len(starts))) / (1 + gql[starts][None, :]) for b in big_l]
    return MixedSystem(vq=vq, gindex=g.astype(np.int64), indgq=indgq, gqlen=gql, gqlen0=gql[starts],
                       G_unique=starts, jlam=jl, big_l=big_l, k2cartes=k2c,
                       rotloc=np.stack([_rotation(rng) for _ in range(2)]),
                       trotij=np.stack([_rotation(rng) for _ in range(3)]), mult=mult,
                       pos=rng.uniform(0, 1, (3, 3)), vol=a ** 3, ngq_barc=ngq_barc)


@dataclass
class MTSystem:
    """Inputs of Ewald_Summation, MT_Coulomb, Calc_Olap_PW_Product and
    Zero_Momentum_Coulomb (for_Coulomb.f90:1-283, pygw.py:2055-2124) for a
    cell of two inequivalent atoms (mult 1 and 2): the lattice (rbas rows,
    br2 columns), positions [3, ndf], product functions of L <= lmax_b per
    atom type with their radial functions on each type's logarithmic mesh
    (r0, dh, npt), the packed tables (tilg, djmm), rrint, rtlij, Ewald
    parameters and the length-sorted plane waves of the q = 0 term."""
    rbas: np.ndarray
    br2: np.ndarray
    alat: np.ndarray
    vol: float
    ortho: bool
    pos: np.ndarray
    mult: np.ndarray
    atm: list
    pq: np.ndarray
    eta: float
    r_cf: float
    g_cf: float
    lam_max: int
    big_l: list
    big_l_f: np.ndarray
    nmix: np.ndarray
    im_start: np.ndarray
    loctmatsize: int
    iprod: dict
    tilg: np.ndarray
    rrint: np.ndarray
    djmm: np.ndarray
    rtlij: np.ndarray
    meshes: list
    ul: list
    us: list
    rel: bool
    glen: np.ndarray
    gindex: np.ndarray


def _mt_build(rbas, alat, ortho, pos, mult, big_l, mesh_spec, q, eta, rcut, gcut, lam_max, gmax,
              identity_rotations, rng) -> MTSystem:
    from .coulomb import calc_tildeg
    vol = abs(np.linalg.det(rbas))
    br2 = 2 * np.pi * np.linalg.inv(rbas)
    mult = np.asarray(mult)
    atm = [iat for iat in range(len(mult)) for _ in range(mult[iat])]
    nat = len(mult)
    lmax_b = int(max(int(np.max(b)) for b in big_l))
    nmix = np.array([len(b) for b in big_l])
    max_nmix = int(nmix.max())
    big_l_f = np.zeros((max_nmix, nat), dtype=np.int64)
    im_start = np.zeros(nat + 1, dtype=np.int64)
    iprod = {}
    im, idf = 0, 0
    for iat in range(nat):
        im_start[iat] = im
        big_l_f[: nmix[iat], iat] = big_l[iat]
        for _ in range(mult[iat]):
            for irm, L in enumerate(big_l[iat]):
                if L == 0:
                    iprod[(idf, irm)] = im
                im += 2 * L + 1
            idf += 1
    im_start[nat] = im
    tilg = calc_tildeg(2 * (lmax_b + 1))
    dimdj = sum((2 * l + 1) * (l + 1) for l in range(lmax_b + 1))
    ndf = int(mult.sum())
    if identity_rotations:
        djmm = np.zeros((dimdj, ndf), dtype=complex)
        for l in range(lmax_b + 1):
            for m in range(0, l + 1):
                djmm[l * (l + 1) * (4 * l - 1) // 6 + (l + 1) * (m + l) + m, :] = 1.0
    else:
        djmm = rng.normal(size=(dimdj, ndf)) + 1j * rng.normal(size=(dimdj, ndf))
    nij = max_nmix * (max_nmix + 1) // 2
    rrint = rng.uniform(0.1, 1.0, (nat, nij))
    rtl = rng.uniform(0.2, 1.5, (nat, max_nmix))
    rtlij = np.zeros((max_nmix, max_nmix, nat, nat))
    for iat in range(nat):
        for jat in range(nat):
            rtlij[: nmix[jat], : nmix[iat], jat, iat] = np.outer(rtl[jat, : nmix[jat]], rtl[iat, : nmix[iat]])
    # logarithmic meshes and smooth radial functions r u, r us
    meshes, ul, us = [], [], []
    for iat, (r0, rmt, npt) in enumerate(mesh_spec):
        dh = np.log(rmt / r0) / (npt - 1)
        meshes.append((r0, dh, npt))
        rx = r0 * np.exp(dh) ** np.arange(npt)
        c = rng.uniform(0.5, 2.0, (len(big_l[iat]), 3))
        ul.append(np.array([rx ** (L + 1) * (c[i, 0] + c[i, 1] * rx) * np.exp(-c[i, 2] * rx)
                            for i, L in enumerate(big_l[iat])]))
        us.append(0.3 * ul[-1] * np.cos(rx)[None, :])
    n = int(gmax / min(np.linalg.norm(br2, axis=0))) + 1
    r = np.arange(-n, n + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    ln = np.linalg.norm(g @ br2.T, axis=1)
    keep = ln <= gmax
    g, ln = g[keep], ln[keep]
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.round(ln, 10)))
    return MTSystem(rbas=rbas, br2=br2, alat=alat, vol=vol, ortho=ortho, pos=pos, mult=mult, atm=atm,
                    pq=np.asarray(q, dtype=float), eta=eta, r_cf=rcut * eta, g_cf=gcut / eta,
                    lam_max=lam_max, big_l=big_l, big_l_f=big_l_f, nmix=nmix, im_start=im_start,
                    loctmatsize=int(im), iprod=iprod, tilg=tilg, rrint=rrint, djmm=djmm, rtlij=rtlij,
                    meshes=meshes, ul=ul, us=us, rel=True, glen=ln[order], gindex=g[order])


def mt_system(q=(0.11, -0.07, 0.05), ortho=False, eta=1.6, rcut=7.0, gcut=14.0, lmax_b=3, gmax=2.5 * np.pi,
              identity_rotations=False, seed=13) -> MTSystem:
    """r_cf = rcut eta and g_cf = gcut / eta.  identity_rotations: djmm of the
    identity (every local frame aligned); else random complex entries."""
    rng = np.random.default_rng(seed)
    if ortho:
        alat = np.array([5.2, 5.9, 7.1])
        rbas = np.diag(alat)
    else:
        a, c = 5.4, 8.3
        alat = np.array([a, a, c])
        rbas = np.array([[a, 0.0, 0.0], [-a / 2, a * np.sqrt(3) / 2, 0.0], [0.0, 0.0, c]])
    pos = np.array([[0.0, 0.0, 0.0], [1 / 3, 2 / 3, 0.27], [2 / 3, 1 / 3, 0.77]]).T
    big_l = [np.array([0, 0, 1, 1, 2, 3]), np.array([0, 1, 2, 2, 3, 0])]
    big_l = [b[b <= lmax_b] for b in big_l]
    return _mt_build(rbas, alat, ortho, pos, [1, 2], big_l, ((1e-4, 2.1, 781), (5e-5, 1.9, 640)), q, eta, rcut, gcut,
                     2 * lmax_b, gmax, identity_rotations, rng)


def mt_bench_system(q=(0.07, 0.05, -0.04), seed=29) -> MTSystem:
    """A PyGW-sized muffin-tin block: 8 atoms (4 types of mult 2) in a
    monoclinic cell of about 700 bohr^3, lmbmax = 3 with 26-30 product
    functions per atom (L = 0..3; 112 (L, M) functions per atom, loctmatsize
    904), structure constants to lam_max = 4 (lmbmax + 1) = 16, 781-point
    radial meshes, plane waves to |G| = 7 bohr^-1 for the q = 0 term."""
    rng = np.random.default_rng(seed)
    rbas = np.array([[7.9, 0.0, 0.0], [1.3, 8.4, 0.0], [0.0, 0.0, 10.6]])
    alat = np.array([7.9, 8.5, 10.6])
    pos = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.25, 0.1, 0.3], [0.75, 0.9, 0.7],
                    [0.1, 0.6, 0.2], [0.9, 0.4, 0.8], [0.6, 0.2, 0.1], [0.4, 0.8, 0.9]]).T
    per_l = [(6, 8, 8, 6), (7, 8, 8, 6), (6, 7, 8, 6), (6, 8, 7, 7)]
    big_l = [np.repeat(np.arange(4), n) for n in per_l]
    meshes = [(1e-4, 2.3, 781), (5e-5, 2.1, 781), (1e-4, 1.9, 781), (1e-4, 1.8, 781)]
    return _mt_build(rbas, alat, False, pos, [2, 2, 2, 2], big_l, meshes, q, 2.2, 7.0, 14.0, 16, 7.0, False, rng)


@dataclass
class MinmMTSystem:
    """Inputs of calc_minm_MT / calc_minc / calc_minc2 (for_vxcnn.f90:839-1397):
    band coefficients alfa/beta/gama at k and alfp/betp/gamp at k - q in
    PyGW's index order, the radial integrals s3r [n_mix, n_fnc, n_fnc, nat]
    (symmetric in the two radial channels; the first ncore channels of an
    atom type are its core states), the product basis (big_l [max_nmix,
    nat]), local orbitals nLO_at [lomax+1, nat], the core list corind [ncg, 5]
    and the Gaunt table."""
    mult: np.ndarray
    nmix: np.ndarray
    big_l: np.ndarray
    nLO_at: np.ndarray
    ncore: np.ndarray
    lmax: int
    loctmatsize: int
    alfa: np.ndarray
    beta: np.ndarray
    gama: np.ndarray
    alfp: np.ndarray
    betp: np.ndarray
    gamp: np.ndarray
    s3r: np.ndarray
    corind: np.ndarray
    core_l: list
    vpos: np.ndarray
    ql: np.ndarray
    kl: np.ndarray
    cgcoef: np.ndarray


def minm_mt_system(lmax=3, lmbmax=2, nbmax=7, mbmax=9, los=True, big=False, seed=17) -> MinmMTSystem:
    """Two atom types (mult 1 and 2).  los: local orbitals on l <= 2 of
    both types (else none).  big: 4 types of mult 2, lmax 10, lmbmax 3 and
    product bases of 26-30 functions (a PyGW-sized atom set)."""
    from .coulomb import cmp_all_gaunt
    rng = np.random.default_rng(seed)
    if big:
        mult = np.array([2, 2, 2, 2])
        big_l_list = [np.repeat(np.arange(4), n) for n in ((6, 8, 8, 6), (7, 8, 8, 6), (6, 7, 8, 6), (6, 8, 7, 7))]
        lomax = lmax
        nlo = np.zeros((lmax + 1, 4), dtype=int)
        nlo[:4] = np.array([[1, 1, 1, 0], [1, 0, 1, 0], [0, 1, 1, 0], [1, 1, 0, 0]]).T[: min(4, lmax + 1)]
        core_l = [[0, 0, 1], [0, 1], [0], [0, 0, 1, 1]]
    else:
        mult = np.array([1, 2])
        base = [np.array([0, 0, 1, 1, 2, 3]), np.array([0, 1, 1, 2, 0])]
        big_l_list = [b[b <= lmbmax] for b in base]
        # nLO_at over every l <= lmax (PyGW runs with lomax >= lmax), local orbitals on l <= 2
        lomax = lmax
        nlo = np.zeros((lmax + 1, 2), dtype=int)
        if los:
            nlo[:3] = np.array([[1, 0, 2], [0, 1, 1]]).T[: min(3, lmax + 1)]
        core_l = [[0, 1], [0]]
    nat = len(mult)
    ndf = int(mult.sum())
    nmix = np.array([len(b) for b in big_l_list])
    max_nmix = int(nmix.max())
    big_l = np.zeros((max_nmix, nat), dtype=np.int64)
    for a in range(nat):
        big_l[: nmix[a], a] = big_l_list[a]
    ncore = np.array([len(c) for c in core_l])
    nlo_tot = [int(nlo[: min(lomax, lmax) + 1, a].sum()) for a in range(nat)]
    n_fnc = max(int(ncore[a]) + 2 * (lmax + 1) + nlo_tot[a] for a in range(nat))
    nlomax = int(nlo.max()) if nlo.size else 0
    nlm = (lmax + 1) ** 2
    cplx = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)  # noqa: E731
    alfa, beta = cplx(nbmax, nlm, ndf), 0.3 * cplx(nbmax, nlm, ndf)
    alfp, betp = cplx(mbmax, nlm, ndf), 0.3 * cplx(mbmax, nlm, ndf)
    gama = 0.5 * cplx(nbmax, max(nlomax, 1), nlm, ndf)
    gamp = 0.5 * cplx(mbmax, max(nlomax, 1), nlm, ndf)
    s3r = rng.normal(size=(max_nmix, n_fnc, n_fnc, nat))
    s3r = 0.5 * (s3r + s3r.transpose(0, 2, 1, 3))
    loct = int(sum(mult[a] * np.sum(2 * big_l_list[a] + 1) for a in range(nat)))
    corind = []
    idf = 0
    for a in range(nat):
        for _ in range(mult[a]):
            for ic, lc in enumerate(core_l[a]):
                for mc in range(-lc, lc + 1):
                    corind.append([a, idf, ic, lc, mc])
            idf += 1
    maxnt = max(lmax + 1, 2 * (lmbmax + 1) if not big else 8)
    return MinmMTSystem(mult=mult, nmix=nmix, big_l=big_l, nLO_at=nlo, ncore=ncore, lmax=lmax, loctmatsize=loct,
                        alfa=alfa, beta=beta, gama=gama, alfp=alfp, betp=betp, gamp=gamp, s3r=s3r,
                        corind=np.array(corind, dtype=np.int64), core_l=core_l,
                        vpos=rng.uniform(0, 1, (3, ndf)), ql=np.array([0.25, -0.125, 0.5]),
                        kl=np.array([0.375, 0.25, -0.125]), cgcoef=cmp_all_gaunt(maxnt))


@dataclass
class LapwSystem:
    """Inputs of gap2_set_lapwcoef (lapwcoeff.f90:3-222) for one k-point:
    the per-run arrays in f2py's shapes (abcelo [nat, lomax+1, nloat, 3],
    umt [nat, nt, 5], rotloc [nat, 3, 3], trotij [ndf, 3, 3], pos [3, ndf],
    nLO_at/nlo [lomax+1, nat], lapw [nt, nat]), the G list gindex [ng, 3],
    the k-point's rows indgk (nvk plane waves by |k+G|, then the
    local-orbital rows) and eigenvectors Aeig [nbmax, ngk]."""
    abcelo: np.ndarray
    rmt: np.ndarray
    pos: np.ndarray
    mult: np.ndarray
    umt: np.ndarray
    rotloc: np.ndarray
    trotij: np.ndarray
    vol: float
    k2cartes: np.ndarray
    nLO_at: np.ndarray
    nlo: np.ndarray
    lapw: np.ndarray
    nlomax: int
    gindex: np.ndarray
    kl: np.ndarray
    indgk: np.ndarray
    nvk: int
    Aeig: np.ndarray
    alat: float

    def plan_args(self):
        return (self.abcelo, self.rmt, self.pos, self.mult, self.umt, self.rotloc, self.trotij, self.vol,
                self.k2cartes, self.nLO_at, self.nlo, self.lapw)

    def coef_args(self, iop=1, ifrot=True):
        return (self.kl, self.indgk, iop, ifrot, self.nvk, self.gindex, self.abcelo, self.rmt, self.pos, self.mult,
                self.umt, self.rotloc, self.trotij, self.vol, self.k2cartes, self.nLO_at, self.nlo, self.lapw,
                self.nlomax)


def lapw_system(lmax=3, lomax=3, nvk=60, nbmax=12, mult=(1, 2), apw=(False, True), los=True, cubic=False,
                wronskian=False, kl=(0.25, -0.125, 0.375), gmax=4, alat=8.0, seed=23) -> LapwSystem:
    """A cell of len(mult) atom types.  apw[a]: atom type a is APW+lo for
    l <= 2 and LAPW above (all LAPW without local orbitals); local orbitals (los) on l = 0 and 2
    of the first type and l = 1 of every type.  cubic: a simple cubic cell
    (k2cartes = 2 pi/alat I, so 2 pi (G+k).pos is (k+G).r in Cartesian
    units).  wronskian: umt with dp pe - dpe p = 1/R^2, the normalisation
    under which the coefficients match a plane wave in value and slope on the
    sphere."""
    rng = np.random.default_rng(seed)
    mult = np.asarray(mult, dtype=np.int64)
    nat, ndf, nt = len(mult), int(np.sum(mult)), lmax + 1
    if cubic:
        rbas = alat * np.eye(3)
    else:
        rbas = alat * (np.eye(3) + 0.15 * rng.uniform(-1, 1, (3, 3)))
    vol = abs(np.linalg.det(rbas))
    k2cartes = 2 * np.pi * np.linalg.inv(rbas)      # columns: reciprocal vectors
    lapw = np.ones((nt, nat), dtype=np.int64)
    nlo = np.zeros((lomax + 1, nat), dtype=np.int64)
    for a in range(nat):
        if apw[a] and los:                          # APW l always carries its lo
            lapw[: min(3, nt), a] = 0
            if los:
                nlo[: min(3, lomax + 1), a] = 1     # the APW+lo orbitals
        if los:
            if lomax >= 1 and lmax >= 1:
                nlo[1, a] += 1
            if a == 0:
                nlo[0, a] += 1
                if lomax >= 2 and lmax >= 2:
                    nlo[2, a] += 1
    nLO_at = np.zeros_like(nlo)
    for a in range(nat):
        for l in range(lomax + 1):
            nLO_at[l, a] = nlo[l, a] - (1 - lapw[l, a]) if l < nt else nlo[l, a]
    nLO_at = np.maximum(nLO_at, 0)
    nloat = int(nlo.max()) + 1 if nlo.size else 1
    rmt = rng.uniform(1.6, 2.2, nat)
    umt = np.zeros((nat, nt, 5))
    for a in range(nat):
        for l in range(nt):
            p, pe, dp = rng.uniform(0.5, 1.5) * rng.choice([-1, 1]), rng.uniform(-1, 1), rng.uniform(-1, 1)
            dpe = (dp * pe - 1.0 / rmt[a] ** 2) / p if wronskian else rng.uniform(-1, 1)
            umt[a, l] = (p, pe, dp, dpe, rng.uniform(0.1, 1))
    abcelo = rng.normal(size=(nat, lomax + 1, nloat, 3))
    rotloc = np.stack([_rotation(rng) for _ in range(nat)])
    trotij = np.stack([_rotation(rng) for _ in range(ndf)])
    pos = rng.uniform(0, 1, (3, ndf))
    r = np.arange(-gmax, gmax + 1)
    G = np.array([(i, j, k) for i in r for j in r for k in r], dtype=np.int64)
    kl = np.asarray(kl, dtype=np.float64)
    kpg = np.linalg.norm((G + kl) @ k2cartes.T, axis=1)
    order = np.argsort(kpg, kind="stable")
    gindex = G[order]
    kpg = kpg[order]
    nlo_tot = int(sum(nlo[l, a] * (2 * l + 1) * mult[a] for a in range(nat) for l in range(lomax + 1)))
    # plane waves: the nvk shortest k+G; local-orbital rows: further G vectors
    lo_rows = rng.choice(np.arange(nvk, min(len(G), nvk + 4 * nlo_tot + 8)), size=nlo_tot, replace=False) \
        if nlo_tot else np.zeros(0, dtype=np.int64)
    indgk = np.concatenate([np.arange(nvk), lo_rows]).astype(np.int64)
    ngk = indgk.size
    Aeig = (rng.normal(size=(nbmax, ngk)) + 1j * rng.normal(size=(nbmax, ngk))) / np.sqrt(ngk)
    return LapwSystem(abcelo=abcelo, rmt=rmt, pos=pos, mult=mult, umt=umt, rotloc=rotloc, trotij=trotij, vol=vol,
                      k2cartes=k2cartes, nLO_at=nLO_at, nlo=nlo, lapw=lapw, nlomax=int(nLO_at.max()),
                      gindex=gindex, kl=kl, indgk=indgk, nvk=nvk, Aeig=Aeig, alat=alat)


@dataclass
class HeadSystem:
    """Inputs of calc_head (pygw.py:2808-3037) for nirk irreducible k-points
    of a full mesh (k_ind = identity, weight 1): per irk the band
    coefficients (alfa, beta, gama), eigenvectors Aeig [nbmax, ngk], kl and
    indgk/nvk; the radial integrals of the momentum operator (vv_tables:
    iul_ul, iul_udl, iudl_ul, iudl_udl [2, lmax, nat]; iulol_ul, iulol_udl,
    iul_ulol, iudl_ulol [2, nlo, lmax+1, nat]; iulol_ulol [2, nlo, nlo,
    lmax+1, nat]) and of the core states (cv_tables: iul_ucl, iudl_ucl,
    iucl_ul, iucl_udl [2, ncore_total]; iulol_ucl, iucl_ulol [2, nloat,
    ncore_total]); the interstitial step function ipwint over gindex with
    its index map i_g0; energies Ebnd [nirk, nbnd] (valence first), core
    energies per core state, weights kcw [n_o, n_e, N_k, n_om] (Fortran
    semantics, core rows first)."""
    mult: np.ndarray
    nLO_at: np.ndarray
    lmax: int
    vv_tables: tuple
    corind: np.ndarray
    core_off: np.ndarray
    cv_tables: tuple
    eig_core: np.ndarray
    gindex: np.ndarray
    i_g0: np.ndarray
    ipwint: np.ndarray
    k2cartes: np.ndarray
    vol: float
    kl: np.ndarray
    Aeig: list
    indgk: list
    nvk: list
    abc: list
    Ebnd: np.ndarray
    nomx: int
    numin: int
    nbmaxpol: int
    ncg_p: int
    k_ind: np.ndarray
    weight: np.ndarray
    kwfer: np.ndarray
    omega: np.ndarray
    kcw: np.ndarray


def head_system(nirk=3, lmax=3, nbmax=9, nomx=3, numin=4, nvk=30, gk=2, n_om=6, mult=(1, 2), los=True,
                metal=False, seed=31) -> HeadSystem:
    """A small cell: band coefficients, plane-wave coefficients and radial
    integrals drawn at random (the momentum elements are bilinear in them),
    the plane waves of each k inside |G|_inf <= gk and gindex covering their
    differences; local orbitals on l <= 2 (nLO_at has lmax + 1 rows); cores
    (l = 0, 1) on every atom type.  metal: nomx >= numin (bands in both
    windows), Fermi-surface weights kwfer and Eg < 0."""
    rng = np.random.default_rng(seed)
    mult = np.asarray(mult, dtype=np.int64)
    nat, ndf, nt = len(mult), int(mult.sum()), lmax + 1
    nLO_at = np.zeros((lmax + 1, nat), dtype=np.int64)
    if los:
        nLO_at[:3, 0] = [1, 0, 2][:min(3, lmax + 1)]
        nLO_at[:3, -1] = np.maximum(nLO_at[:3, -1], [0, 1, 1][:min(3, lmax + 1)])
    nlo = max(int(nLO_at.max()), 1)
    R = lambda *s: rng.normal(size=s)  # noqa: E731
    vv = tuple(R(2, lmax, nat) for _ in range(4)) + tuple(R(2, nlo, lmax + 1, nat) for _ in range(4)) + \
        (R(2, nlo, nlo, lmax + 1, nat),)
    core_l = [[0, 1], [0]] + [[0]] * max(0, nat - 2)
    corind, core_off, off = [], [], 0
    idf = 0
    for a in range(nat):
        core_off.append(off)
        for _ in range(mult[a]):
            for ic, lc in enumerate(core_l[a]):
                for mc in range(-lc, lc + 1):
                    corind.append([a, idf, ic, lc, mc])
            idf += 1
        off += len(core_l[a])
    nct = off
    cv = tuple(R(2, nct) for _ in range(4)) + tuple(R(2, nlo, nct) for _ in range(2))
    corind = np.array(corind, dtype=np.int64)
    eig_core = np.array([-20.0 - 5.0 * c[2] - 3.0 * c[0] for c in corind])
    alat = 9.0
    rbas = alat * (np.eye(3) + 0.1 * rng.uniform(-1, 1, (3, 3)))
    vol = abs(np.linalg.det(rbas))
    k2cartes = 2 * np.pi * np.linalg.inv(rbas)
    gm = 2 * gk
    r = np.arange(-gm, gm + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    g = g[np.lexsort((g[:, 2], g[:, 1], g[:, 0], (g ** 2).sum(1)))]
    i_g0 = -np.ones((2 * gm + 1,) * 3, dtype=np.int64)
    i_g0[g[:, 0] + gm, g[:, 1] + gm, g[:, 2] + gm] = np.arange(len(g))
    # the interstitial step function: ipwint(-G) = conj(ipwint(G)), ipwint(0) real
    ipw = (R(len(g)) + 1j * R(len(g))) * np.exp(-0.3 * (g ** 2).sum(1))
    neg = i_g0[-g[:, 0] + gm, -g[:, 1] + gm, -g[:, 2] + gm]
    ipwint = 0.5 * (ipw + np.conj(ipw[neg]))
    ipwint[0] = 0.7
    inner = np.where(np.abs(g).max(axis=1) <= gk)[0]
    kls, Aeig, indgk, nvks, abc = [], [], [], [], []
    nlm = (lmax + 1) ** 2
    nlomax = int(nLO_at.max())
    for irk in range(nirk):
        kls.append(rng.uniform(-0.5, 0.5, 3))
        ig = np.concatenate([rng.permutation(inner)[:nvk], rng.choice(inner, 3)])
        indgk.append(ig)
        nvks.append(nvk)
        Aeig.append((R(nbmax, len(ig)) + 1j * R(nbmax, len(ig))) / np.sqrt(nvk))
        cplx = lambda *s: R(*s) + 1j * R(*s)  # noqa: E731
        abc.append((cplx(nbmax, nlm, ndf), 0.3 * cplx(nbmax, nlm, ndf),
                    0.5 * cplx(nbmax, max(nlomax, 1), nlm, ndf) if nlomax else np.zeros((1, 1, 1, 1), complex)))
    nbmaxpol = nbmax
    Ebnd = np.zeros((nirk, nbmax))
    for irk in range(nirk):
        e = np.sort(np.concatenate([rng.uniform(-0.8, -0.05, nomx + 1), rng.uniform(0.05, 1.5, nbmax - nomx - 1)]))
        if metal:
            e[numin:nomx + 1] = rng.uniform(-0.05, 0.05, nomx + 1 - numin)
        Ebnd[irk] = np.sort(e)
    ncg_p = len(corind)
    n_o, n_e = ncg_p + nomx + 1, nbmaxpol - numin
    omega = np.geomspace(0.01, 5.0, n_om)
    kcw = -np.abs(R(n_o, n_e, nirk, n_om)) * 0.01
    kwfer = np.abs(R(nirk, nbmax)) * (0.1 if metal else 0.0)
    return HeadSystem(mult=mult, nLO_at=nLO_at, lmax=lmax, vv_tables=vv, corind=corind,
                      core_off=np.array(core_off), cv_tables=cv, eig_core=eig_core, gindex=g, i_g0=i_g0,
                      ipwint=ipwint, k2cartes=k2cartes, vol=vol, kl=np.array(kls), Aeig=Aeig, indgk=indgk, nvk=nvks,
                      abc=abc, Ebnd=Ebnd, nomx=nomx, numin=numin, nbmaxpol=nbmaxpol, ncg_p=ncg_p,
                      k_ind=np.arange(nirk), weight=np.ones(nirk), kwfer=kwfer, omega=omega, kcw=kcw)


@dataclass
class PipelineSystem:
    """Everything calc_minm needs for the k-points of one q on a full k-mesh
    (no symmetry: every k irreducible), and the band energies calc_eps and
    calc_selfc need after it.  The atoms and the G list come from a
    ``LapwSystem`` (lapw, at its kl); per k: kl (fractional), indgk (the nvk
    shortest |k+G| in gindex, then the local-orbital rows) and Aeig [nbmax,
    ngk].  The product basis: the muffin-tin functions of a
    ``minm_mt_system`` on the LAPW atoms (nmix, big_l, s3r, the core list
    corind with core_l per atom type) and ngq plane waves against ngq_barc
    (indgq, its inverse, mpwipw [ngq, ngq_barc]); vmat a Hermitian positive
    definite [loct + ngq]^2 Coulomb matrix.  enk [nbmax, N_k] band energies
    (Hartree, E_F = 0) with windows nomx/numin, eig_core per corind row."""
    lapw: LapwSystem
    nmix: np.ndarray
    big_l: np.ndarray
    ncore: np.ndarray
    lmax: int
    loct: int
    s3r: np.ndarray
    corind: np.ndarray
    cgcoef: np.ndarray
    div: tuple
    kf: np.ndarray
    kqid: np.ndarray
    iq: int
    ql: np.ndarray
    i_g0: np.ndarray
    indgk: list
    nvk: int
    Aeig: list
    indgq: np.ndarray
    indgq_inverse: np.ndarray
    mpwipw: np.ndarray
    vmat: np.ndarray
    enk: np.ndarray
    nomx: int
    numin: int
    eig_core: np.ndarray


def pipeline_system(div=(2, 2, 2), q=(1, 0, 0), n_occ=2, n_emp=5, nvk=40, ngq=20, ngq_barc=40, cores=True,
                    seed=61) -> PipelineSystem:
    """q in mesh units (q / div is the fractional q-point)."""
    from .coulomb import cmp_all_gaunt
    rng = np.random.default_rng(seed)
    cfg = small_config(div=div, n_occ=n_occ, n_emp=n_emp, q=q, seed=seed)
    b = band_energies(cfg)
    nbmax = b.enk.shape[0]
    ls = lapw_system(lmax=3, lomax=3, nvk=nvk, nbmax=nbmax, gmax=4, seed=seed + 1)
    nat = len(ls.mult)
    base = minm_mt_system(lmax=3, nbmax=nbmax, mbmax=nbmax, seed=seed + 2)
    core_l = base.core_l if cores else [[] for _ in range(nat)]
    ncore = np.array([len(c) for c in core_l])
    nfnc = max(int(ncore[a]) + 2 * (ls.lapw.shape[0]) + int(ls.nLO_at[:, a].sum()) for a in range(nat))
    s3r = rng.normal(size=(base.s3r.shape[0], nfnc, nfnc, nat))
    s3r = 0.5 * (s3r + s3r.transpose(0, 2, 1, 3))
    corind = []
    idf = 0
    for a in range(nat):
        for _ in range(ls.mult[a]):
            for ic, lc in enumerate(core_l[a]):
                for mc in range(-lc, lc + 1):
                    corind.append([a, idf, ic, lc, mc])
            idf += 1
    corind = np.array(corind, dtype=np.int64).reshape(-1, 5)
    eig_core = np.array([-20.0 - 5.0 * c[2] - 3.0 * c[0] for c in corind])
    # k-mesh, k - q, the G list's inverse map
    kf = mesh.kpoint_coords(div)
    kqid = mesh.k_minus_q(div)
    iq = q_index(cfg)
    ql = np.asarray(q, dtype=np.float64) / np.asarray(div)
    gmax = int(np.abs(ls.gindex).max())
    i_g0 = -np.ones((2 * gmax + 1,) * 3, dtype=np.int64)
    g = ls.gindex
    i_g0[g[:, 0] + gmax, g[:, 1] + gmax, g[:, 2] + gmax] = np.arange(len(g))
    nlo_tot = ls.indgk.size - ls.nvk
    indgk, Aeig = [], []
    for k in kf:
        kpg = np.linalg.norm((g + k) @ ls.k2cartes.T, axis=1)
        order = np.argsort(kpg, kind="stable")
        lo = rng.choice(order[nvk:nvk + 4 * nlo_tot + 8], size=nlo_tot, replace=False) if nlo_tot else order[:0]
        ig = np.concatenate([order[:nvk], lo]).astype(np.int64)
        indgk.append(ig)
        Aeig.append((rng.normal(size=(nbmax, ig.size)) + 1j * rng.normal(size=(nbmax, ig.size))) / np.sqrt(ig.size))
    # the interstitial product basis of q
    indgq = np.argsort(np.linalg.norm((g + ql) @ ls.k2cartes.T, axis=1), kind="stable")[:ngq_barc]
    inv = -np.ones(indgq.max() + 1, dtype=np.int64)
    inv[indgq] = np.arange(ngq_barc)
    mpwipw = (rng.normal(size=(ngq, ngq_barc)) + 1j * rng.normal(size=(ngq, ngq_barc))) / np.sqrt(ngq_barc)
    mb = base.loctmatsize + ngq
    qm, _ = np.linalg.qr(rng.normal(size=(mb, mb)) + 1j * rng.normal(size=(mb, mb)))
    lam = np.geomspace(0.05, 20.0, mb) * (1 + 0.01 * rng.uniform(size=mb))
    vmat = (qm * lam) @ qm.conj().T
    vmat = 0.5 * (vmat + vmat.conj().T)
    return PipelineSystem(lapw=ls, nmix=base.nmix, big_l=base.big_l, ncore=ncore, lmax=base.lmax,
                          loct=base.loctmatsize, s3r=s3r, corind=corind, cgcoef=cmp_all_gaunt(6), div=tuple(div),
                          kf=kf, kqid=kqid, iq=iq, ql=ql, i_g0=i_g0, indgk=indgk, nvk=nvk, Aeig=Aeig, indgq=indgq,
                          indgq_inverse=inv, mpwipw=mpwipw, vmat=vmat, enk=b.enk, nomx=b.nomx, numin=b.numin,
                          eig_core=eig_core)
