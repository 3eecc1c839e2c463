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
    print("pade, conv written")


if __name__ == "__main__":
    main()
