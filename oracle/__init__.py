"""ORACLE -- test infrastructure only.

CPU restatements of PyGW's polarization -> dielectric-inversion -> Sigma_c
path, used by tests/, the smoke check and bench.py's CPU-baseline leg as the
checker; never by the product path.

  tetra_oracle.c / tetra.py   tetrahedron weights (C, branch-for-branch)
  chain.py                    chi0, eps^-1, MWM, Sigma_c convolution, Pade (numpy)
  cubature.py                 independent pins (cubature, half-space clipping)
  ref.py                      PyGW's own Fortran compiled into oracle/_ref
"""
