"""ORACLE (test infrastructure only): independent pins of the tetrahedron
weights that share no code or formula with PyGW's closed forms.

  unit_tet_weights  w_i = int_T lambda_i(x) g(Delta(x)) dV on the unit
                    tetrahedron T (volume 1/6), g(D) = -2 D / (w^2 + D^2),
                    Delta linear with corner values D_1..D_4, lambda =
                    (1-x-y-z, x, y, z).  Conical-product Gauss-Legendre rule
                    (Stroud's collapsed-cube map), numpy.polynomial.legendre.
  region_weights    the same integrand over {e(x) <= E_F <= f(x)} within T:
                    the polyhedron is built by half-space clipping of T and
                    tetrahedralised with scipy.spatial.Delaunay, so it checks
                    the reference's node geometry (surfnodes..sortnodes).
  flat_band_kcw     closed form kcw = g(D) / N_k when Delta is the same
                    constant D on every tetrahedron.

PyGW normalises one tetrahedron's corner weights as int_T lambda_i g dV with
|T| = 1/6 (for_tetrahedra3.f90:392-403 with vol = |det| = 1 for the unit
tetrahedron), and kcw = sum_t 6/N_tet * w (for_tetrahedra.f90:455-458).
"""
from __future__ import annotations

import itertools

import numpy as np


def _collapsed_rule(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    u, v, t = np.meshgrid(x, x, x, indexing="ij")
    wu, wv, wt = np.meshgrid(w, w, w, indexing="ij")
    # (u,v,t) in unit cube -> tetrahedron x+y+z <= 1
    px = u
    py = (1.0 - u) * v
    pz = (1.0 - u) * (1.0 - v) * t
    jac = (1.0 - u) ** 2 * (1.0 - v)
    pts = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1)
    wts = (wu * wv * wt * jac).ravel()
    return pts, wts


_RULES = {}


def rule(n=48):
    if n not in _RULES:
        _RULES[n] = _collapsed_rule(n)
    return _RULES[n]


def _g(d, om):
    return -2.0 * d / (om * om + d * d)


def unit_tet_weights(delta, omega, n=48):
    """[4, nom] exact corner weights on the unit tetrahedron."""
    delta = np.asarray(delta, dtype=np.float64)
    omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
    pts, wts = rule(n)
    lam = np.stack([1.0 - pts.sum(1), pts[:, 0], pts[:, 1], pts[:, 2]], axis=1)  # [np,4]
    d = lam @ delta
    out = np.zeros((4, omega.size))
    for j, om in enumerate(omega):
        gv = _g(d, om) * wts
        out[:, j] = lam.T @ gv
    return out


def _clip(poly_pts, a, b):
    """Vertices of {x in hull(poly_pts): a.x <= b}, for a convex point set."""
    s = poly_pts @ a - b
    keep = list(poly_pts[s <= 0])
    n = len(poly_pts)
    for i in range(n):
        for j in range(i + 1, n):
            if (s[i] < 0 < s[j]) or (s[j] < 0 < s[i]):
                t = s[i] / (s[i] - s[j])
                keep.append(poly_pts[i] + t * (poly_pts[j] - poly_pts[i]))
    return np.array(keep) if keep else np.zeros((0, 3))


def region_polytope(e, f, ef=0.0):
    """Vertex set of {x in T : e(x) <= ef, f(x) >= ef} (convex)."""
    T = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    e = np.asarray(e, float)
    f = np.asarray(f, float)
    ae = np.array([e[1] - e[0], e[2] - e[0], e[3] - e[0]])
    af = np.array([f[1] - f[0], f[2] - f[0], f[3] - f[0]])
    # e(x) = e1 + ae.x <= ef ; f(x) = f1 + af.x >= ef  <=> -af.x <= f1 - ef
    pts = _clip(T, ae, ef - e[0])
    if len(pts) == 0:
        return pts
    # note: edges between vertices of a convex set include non-edges; the
    # extra points lie inside the hull, so the hull is unchanged.
    pts = _clip(pts, -af, f[0] - ef)
    return pts


def region_weights(e, f, omega, ef=0.0, n=32):
    """[4, nom] corner weights over the occupied/empty region of T."""
    from scipy.spatial import ConvexHull, Delaunay, QhullError

    omega = np.atleast_1d(np.asarray(omega, dtype=np.float64))
    out = np.zeros((4, omega.size))
    pts = region_polytope(e, f, ef)
    if len(pts) < 4:
        return out
    try:
        hull = ConvexHull(pts)
    except QhullError:
        return out  # flat region
    if hull.volume < 1e-14:
        return out
    verts = pts[hull.vertices]
    tri = Delaunay(verts)
    e = np.asarray(e, float)
    f = np.asarray(f, float)
    dbig = f - e
    qp, qw = rule(n)
    for simplex in tri.simplices:
        c = verts[simplex]
        J = np.stack([c[1] - c[0], c[2] - c[0], c[3] - c[0]], axis=1)
        vol6 = abs(np.linalg.det(J))
        if vol6 < 1e-14:
            continue
        x = c[0][None, :] + qp @ J.T
        lam = np.stack([1.0 - x.sum(1), x[:, 0], x[:, 1], x[:, 2]], axis=1)
        d = lam @ dbig
        for j, om in enumerate(omega):
            gv = _g(d, om) * qw * vol6
            out[:, j] += lam.T @ gv
    return out


def flat_band_kcw(delta, omega, nk):
    omega = np.asarray(omega, dtype=np.float64)
    return _g(delta, omega) / nk
