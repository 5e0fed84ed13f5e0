"""Shared helpers of the GPU parity tests."""
import numpy as np

from oracle import oracle as O


def basis_dict(b):
    return dict(n=b.n, l=b.l, m=b.m, zeta=b.zeta, norm=b.norm, centers=b.O)


def oracle_on_grid(basis, pts, grad):
    return O.collocation(basis_dict(basis), pts, grad=grad)


def orbitals_on_grid(w, pts, wts):
    """S on the given grid (oracle), then the seeded Loewdin orbitals."""
    from nukexc_amd import configs
    phi = O.collocation(basis_dict(w["basis"]), pts, grad=False)
    S = O.overlap(phi, wts)
    C, P = configs.loewdin_orbitals(S, w["nocc"], w["rng"])
    return S, C, P
