"""Write the golden fixtures of this directory from PyGW's own Fortran
(compiled in place from the reference sources by `make -C oracle ref`).

    python tests/golden/make_golden.py

Each fixture is an .npz holding inputs and the reference's outputs for a
small case.  Inputs come from pygw_amd.synth (seeded), so the files can be
regenerated bit for bit.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pygw_amd import mesh, synth  # noqa: E402

CASES = {
    # name: (div, n_occ, n_emp, n_metal, n_om, seed, q)
    "tetra_insulator": ((3, 3, 3), 3, 5, 0, 8, 101, (1, 0, 0)),
    "tetra_metal": ((3, 3, 3), 3, 4, 3, 6, 202, (1, 1, 0)),
    "tetra_metal_b": ((4, 2, 3), 2, 3, 3, 5, 303, (0, 1, 2)),
}


def tetra_case(name):
    div, no, ne, nm, nom, seed, q = CASES[name]
    cfg = synth.small_config(name, div=div, n_occ=no, n_emp=ne, n_metal=nm, n_om=nom, seed=seed, q=q)
    b = synth.band_energies(cfg)
    om, _ = synth.omega_mesh(cfg)
    tetc = mesh.tetra_corners(cfg.div)
    kq = mesh.k_minus_q(cfg.div)[:, synth.q_index(cfg)]
    return cfg, b, om, tetc, kq


def main():
    if not ref.available():
        raise SystemExit("build the reference first: make -C oracle ref")
    for name in CASES:
        cfg, b, om, tetc, kq = tetra_case(name)
        kcw = ref.bz_calcqdepw_par2(b.enk, 0.0, kq, b.ncg, b.nomx, b.numin, om, tetc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), enk=b.enk, kqid=kq, tetc=tetc, omega=om,
                            win=np.array([b.ncg, b.nomx, b.numin]), kcw=kcw)
        print(name, kcw.shape)
    # Pade (for_pade.f90:3-51) on a five-pole function sampled at i*omega_n
    z = 1j * np.geomspace(0.02, 5.0, 10)
    poles = np.array([0.2 - 0.05j, -0.5 + 0.02j, 0.9 - 0.3j, -1.3 + 0.4j, 0.05 - 0.6j])
    res = np.array([0.3, 0.7, 0.25, 0.4, 0.15])
    y = (res[None, :] / (z[:, None] - poles[None, :])).sum(1)
    apar = ref.padecof(y, z)
    x = np.linspace(-1, 1, 9) + 0.0j
    g = ref.padeg(x, z, apar)
    np.savez_compressed(os.path.join(HERE, "pade.npz"), z=z, y=y, apar=apar, x=x, g=g)
    # Sigma_c convolution (fnc.f90:139-159) on a small synthetic mwm
    rng = np.random.default_rng(404)
    nom, nb1, nb2 = 10, 2, 5
    om = np.geomspace(0.01, 4.0, nom)
    wo = np.gradient(om)
    enk = rng.uniform(-0.5, 0.5, nb2)
    mwm = (rng.normal(size=(nom, nb1, nb2)) + 1j * rng.normal(size=(nom, nb1, nb2))) / (1 + om[:, None, None] ** 2)
    sc = ref.all_fr_convolutions(enk, mwm, om, wo)
    np.savez_compressed(os.path.join(HERE, "conv.npz"), enk=enk, mwm=mwm, omega=om, womeg=wo, sc=sc)
    # the spline branch of Compute_selfc_inside (mcommon.py:126-144): mwm on
    # a 12-point tan mesh splined (scipy, as PyGW) onto the x5 precise mesh,
    # then PyGW's compiled few_fr_convolutions (fnc.f90:161-176) per (ie1, iom)
    from oracle import chain
    fm = mesh.FrequencyMesh(12, 0.02, 15.0, iop_multiple=5)
    nom, nb1, nb2 = fm.omega.size, 2, 6
    a = rng.uniform(0.2, 1.4, (nb1, nb2))
    mwm = (1.0 / (fm.omega[:, None, None] ** 2 + a[None] ** 2)) * (1 + 0.2j * rng.normal(size=(1, nb1, nb2)))
    enk = rng.uniform(-0.6, 0.6, nb2)
    enk[0] = 0.0                           # E = 0: atan(Omega_max/(-i w)) on its branch cut
    mwmp = np.zeros((nb1, nb2, fm.omega_precise.size), dtype=complex)
    sc = np.zeros((nb1, nom), dtype=complex)
    for ie1 in range(nb1):
        for ie2 in range(nb2):
            mwmp[ie1, ie2] = chain.spline_precise(mwm[:, ie1, ie2], fm.omega, fm.omega_precise)
        for iom in range(nom):
            sc[ie1, iom] = ref.few_fr_convolutions(enk, mwmp[ie1], mwm[iom, ie1, :], fm.omega_precise, fm.omega[iom],
                                                   fm.womeg_precise)
    np.savez_compressed(os.path.join(HERE, "conv_spline.npz"), enk=enk, mwm=mwm, omega=fm.omega, womeg=fm.womeg,
                        omega_p=fm.omega_precise, womeg_p=fm.womeg_precise, mwm_p=mwmp, sc=sc)
    print("pade, conv, conv_spline written")


if __name__ == "__main__":
    main()
