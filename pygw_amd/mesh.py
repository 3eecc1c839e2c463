"""k/q meshes, linked tetrahedra and Matsubara meshes for the polarization path.

These are the input contracts of the weight kernel: `tetc`, `kqid` and the
frequency mesh.  Layout and numbering follow PyGW exactly (0-based indices,
Fortran-ordered `tetc[4, 6*N_k]`, k enumerated in `itertools.product` order).
"""
from __future__ import annotations

import itertools
import math

import numpy as np

# Units (inout.py:5-7).  Everything on the device path is in Hartree.
H2EV = 27.2113961
RY2H = 0.5

# The six tetrahedra of a cube sharing the main diagonal (0,0,0)-(1,1,1)
# (kqmesh.py:220-225).
_TET0 = np.array([[[0, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1]],
                  [[0, 0, 0], [0, 1, 1], [0, 1, 0], [1, 1, 1]],
                  [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 1, 1]],
                  [[0, 0, 0], [1, 1, 0], [1, 0, 0], [1, 1, 1]],
                  [[0, 0, 0], [1, 0, 0], [1, 0, 1], [1, 1, 1]],
                  [[0, 0, 0], [1, 0, 1], [0, 0, 1], [1, 1, 1]]], dtype=np.int64)


def k2indx(k, div):
    """Linear index of integer k-point (for_kpts.f90:2-7)."""
    return (k[..., 0] * div[1] + k[..., 1]) * div[2] + k[..., 2]


def cube_tetrahedra(mnd: int = 0) -> np.ndarray:
    """Corner offsets of the six tetrahedra for main diagonal `mnd`
    (kqmesh.py:226-240).  Returns int array [6, 4, 3]."""
    t0 = _TET0
    tet = np.zeros_like(t0)
    if mnd == 0:
        tet[:] = t0
    elif mnd == 1:
        tet[:, :, 0] = t0[:, :, 0]
        tet[:, :, 1] = t0[:, :, 2]
        tet[:, :, 2] = 1 - t0[:, :, 1]
    elif mnd == 2:
        tet[:, :, 0] = t0[:, :, 0]
        tet[:, :, 1] = 1 - t0[:, :, 2]
        tet[:, :, 2] = t0[:, :, 1]
    elif mnd == 3:
        tet[:, :, 0] = t0[:, :, 0]
        tet[:, :, 1] = 1 - t0[:, :, 1]
        tet[:, :, 2] = 1 - t0[:, :, 2]
    else:
        raise ValueError("mnd must be 0..3")
    return tet


def main_diagonal(div, gbas=None) -> int:
    """Shortest cube diagonal (kqmesh.py:203-218).  With gbas=None the
    reciprocal lattice is taken as the unit cube, for which all four diagonals
    tie and argmin returns 0."""
    gbas = np.eye(3) if gbas is None else np.asarray(gbas, dtype=float)
    p0 = np.array([[0, 0, 0, 0, 1, 1, 1, 1], [0, 0, 1, 1, 0, 0, 1, 1], [0, 1, 0, 1, 0, 1, 0, 1]])
    diag = np.zeros(4)
    for i in range(4):
        sm = np.dot([(p0[k, i] - p0[k, 7 - i]) / float(div[k]) for k in range(3)], gbas)
        diag[i] = np.sum(sm ** 2)
    return int(np.argmin(diag))


def tetra_corners(div, mnd: int = 0) -> np.ndarray:
    """`tetc[4, 6*N_k]` (Fortran order, 0-based k indices) as built by
    KQmesh.tetra (kqmesh.py:242-257)."""
    div = np.asarray(div, dtype=np.int64)
    nk = int(np.prod(div))
    tet = cube_tetrahedra(mnd)
    orig = np.array(list(itertools.product(range(div[0]), range(div[1]), range(div[2]))), dtype=np.int64)
    corn = (orig[:, None, None, :] + tet[None, :, :, :]) % div  # [nk, 6, 4, 3]
    ids = k2indx(corn, div)  # [nk, 6, 4]
    tetc = np.asfortranarray(ids.reshape(nk * 6, 4).T.astype(np.int32))
    return tetc


def k_minus_q(div) -> np.ndarray:
    """kqid[ik, iq] = index of k-q (for_kpts.f90:21-45), int32 [N_k, N_k]."""
    div = np.asarray(div, dtype=np.int64)
    pts = np.array(list(itertools.product(range(div[0]), range(div[1]), range(div[2]))), dtype=np.int64)
    kmq = (pts[:, None, :] - pts[None, :, :]) % div
    return k2indx(kmq, div).astype(np.int32)


def kpoint_coords(div) -> np.ndarray:
    """Fractional coordinates k_i = j_i/N_i in k-index order, float [N_k, 3]."""
    div = np.asarray(div, dtype=np.int64)
    pts = np.array(list(itertools.product(range(div[0]), range(div[1]), range(div[2]))), dtype=np.float64)
    return pts / div[None, :]


def log_mesh(wmin_ev: float, wmax_ev: float, n: int) -> np.ndarray:
    """Log-spaced Matsubara mesh, returned in Hartree."""
    return np.geomspace(wmin_ev, wmax_ev, n) / H2EV


def give2tan_mesh(x0: float, L: float, Nw: int):
    """Tangent mesh and quadrature weights on (-L, L) with 2*Nw points
    (mcommon.py:13-31).  The 2x2 system L = w/tan(d),
    x0 = w*tan(pi/(2Nw) - d/Nw) is solved by Newton's method from the same
    starting point the reference passes to scipy.optimize.root."""
    tnw = math.tan(math.pi / (2 * Nw))
    if x0 > L * 0.25 * Nw * tnw ** 2:
        x0 = L * 0.25 * Nw * tnw ** 2 - 1e-15
    xi = x0 / L
    d = Nw / 2.0 * (tnw - math.sqrt(tnw ** 2 - 4 * xi / Nw))
    w = L * d
    for _ in range(100):
        a = math.pi / (2 * Nw) - d / Nw
        f1 = L - w / math.tan(d)
        f2 = x0 - w * math.tan(a)
        # Jacobian
        j11 = w / math.sin(d) ** 2
        j12 = -1.0 / math.tan(d)
        j21 = w / (Nw * math.cos(a) ** 2)
        j22 = -math.tan(a)
        det = j11 * j22 - j12 * j21
        dd = (f1 * j22 - j12 * f2) / det
        dw = (j11 * f2 - j21 * f1) / det
        d -= dd
        w -= dw
        if abs(dd) < 1e-16 * max(1.0, abs(d)) and abs(dw) < 1e-16 * max(1.0, abs(w)):
            break
    xt = np.linspace(0.0, 1.0, 2 * Nw) * (math.pi - 2 * d) - math.pi / 2 + d
    om = w * np.tan(xt)
    dom = (w * (math.pi - 2 * d) / (2.0 * Nw)) / np.cos(xt) ** 2
    return om, dom


class FrequencyMesh:
    """Matsubara mesh with convolution weights (pygw.py:1643-1687, iopfreq=4:
    tangent mesh plus an `iop_multiple` times denser mesh for the Sigma_c
    convolution).  Arguments in Hartree."""

    def __init__(self, nomeg: int, omeg_min: float, omeg_max: float, iop_multiple: int = 20):
        om0, dom0 = give2tan_mesh(omeg_min, omeg_max, nomeg)
        n = len(om0) // 2
        self.omega, self.womeg = om0[n:].copy(), dom0[n:].copy()
        minx = omeg_min / (1 + 0.2 * (iop_multiple - 1.0))
        om1, dom1 = give2tan_mesh(minx, omeg_max, nomeg * iop_multiple)
        n = len(om1) // 2
        self.omega_precise, self.womeg_precise = om1[n:].copy(), dom1[n:].copy()
        self.iopfreq = 4


def svd_functions(iomega, wiomeg, om_min, om_max, om_nom, svd_cutoff):
    """SVD basis of the bosonic Matsubara kernel (pygw.py:3478-3501).
    Returns (Ul [n_l, nom], dUl [nom, n_l])."""
    om, dom = give2tan_mesh(om_min, om_max, om_nom)
    iom = np.hstack((-iomega[::-1], iomega[:]))
    diom = np.hstack((wiomeg[::-1], wiomeg[:]))
    num = om * dom / np.pi
    om2 = om ** 2
    ker = num[None, :] * np.sqrt(diom)[:, None] / (om2[None, :] + iom[:, None] ** 2)
    u, s, vh = np.linalg.svd(ker, full_matrices=True)
    n = int(np.where(s < svd_cutoff)[0][0])
    ul = np.transpose(u[:, :n]) / np.sqrt(diom)[None, :]
    n2 = len(iom) // 2
    dul = (ul[:, n2:] * diom[n2:] * 2.0).T
    return ul[:, n2:], dul
