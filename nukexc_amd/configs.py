"""Synthetic workloads of the XC path (geometries, STO bases, orbitals).

All inputs are seeded (NumPy PCG64, seed = workload index) so the GPU forms,
the tests and the CPU baseline read identical numbers.  Coordinates in bohr.

  water      O (0,0,0), H (0, +-1.4305, 1.1073); TA 75 x Lebedev 302; LDA_X;
             5 occupied orbitals
  benzene    6 C on a ring of radius 2.64 bohr, 6 H at 4.68 bohr;
             TA 99 x Lebedev 590; PBE; 21 occupied orbitals
  water100   100 waters, random orientations, O uniformly in a cube of
             1 g/cm^3 with O-O >= 2.5 bohr; TA 75 x 302; PBE; 500 occupied
  molset64   64 molecules of 3..60 atoms (H/C/O with p = 0.5/0.3/0.2),
             random-walk coordinates (steps 2.0-2.9 bohr, >= 1.8 bohr apart)
  mol30      one 30-atom molecule of the molset recipe

Basis recipe (per atom, shells expanded over m = -l..l): O and C get 1s x2,
2s x2, 2p x2, 3d x1 with zeta log-uniform on [0.5, 12]; H gets 1s x2, 2p x1
with zeta log-uniform on [0.3, 3].  Orbitals: C0 ~ N(0, 1) of shape
(N_bf, N_occ), Loewdin-orthonormalised in the numerically integrated overlap
S of the same grid, C = C0 (C0^T S C0)^(-1/2), P = 2 C C^T.
"""
import numpy as np

from .api import make_manual_basis

SHELLS = {
    1: ((1, 0, 2), (2, 1, 1)),
    6: ((1, 0, 2), (2, 0, 2), (2, 1, 2), (3, 2, 1)),
    8: ((1, 0, 2), (2, 0, 2), (2, 1, 2), (3, 2, 1)),
}
ZETA_RANGE = {1: (0.3, 3.0), 6: (0.5, 12.0), 8: (0.5, 12.0)}
WATER = np.array([[0.0, 0.0, 0.0], [0.0, 1.4305, 1.1073], [0.0, -1.4305, 1.1073]])
BOHR_PER_ANG = 1.8897261254578281
AVOGADRO = 6.02214076e23


def sto_functions(Z, xyz, rng):
    """(n, l, m, zeta, x, y, z) tuples in atom, shell, copy, m order."""
    out = []
    for z, c in zip(Z, xyz):
        lo, hi = ZETA_RANGE[int(z)]
        for n, l, count in SHELLS[int(z)]:
            for _ in range(count):
                zeta = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
                for m in range(-l, l + 1):
                    out.append((n, l, m, zeta, float(c[0]), float(c[1]), float(c[2])))
    return out


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(nmol, rng, min_oo=2.5):
    vol_ang3 = nmol * 18.015 / AVOGADRO * 1e24  # cm^3 -> A^3 at 1 g/cm^3
    L = vol_ang3 ** (1.0 / 3.0) * BOHR_PER_ANG
    O = []
    while len(O) < nmol:
        p = rng.uniform(0.0, L, size=3)
        if all(np.linalg.norm(p - q) >= min_oo for q in O):
            O.append(p)
    xyz, Z = [], []
    for o in O:
        R = _rotation(rng)
        for k, zz in enumerate((8, 1, 1)):
            xyz.append(o + R @ WATER[k])
            Z.append(zz)
    return np.array(xyz), np.array(Z, dtype=np.int32), L


def random_walk_molecule(natoms, rng, step=(2.0, 2.9), min_sep=1.8):
    Z = rng.choice([1, 6, 8], size=natoms, p=[0.5, 0.3, 0.2]).astype(np.int32)
    xyz = [np.zeros(3)]
    while len(xyz) < natoms:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = xyz[-1] + d * rng.uniform(*step)
        if all(np.linalg.norm(p - q) >= min_sep for q in xyz):
            xyz.append(p)
    return np.array(xyz), Z


def workload(name):
    """Geometry, basis, grid and functional of a named workload.  The
    returned rng continues after the basis draws and yields C0."""
    if name == "water":
        rng = np.random.Generator(np.random.PCG64(0))
        xyz, Z = WATER.copy(), np.array([8, 1, 1], dtype=np.int32)
        grid = dict(nrad=75, nang_order=29)
        funcs, nocc = (1,), 5
    elif name == "benzene":
        rng = np.random.Generator(np.random.PCG64(1))
        ang = np.arange(6) * np.pi / 3
        ring = lambda r: np.column_stack([r * np.cos(ang), r * np.sin(ang), np.zeros(6)])
        xyz = np.vstack([ring(2.64), ring(4.68)])
        Z = np.array([6] * 6 + [1] * 6, dtype=np.int32)
        grid = dict(nrad=99, nang_order=41)
        funcs, nocc = (101, 130), 21
    elif name == "water100":
        rng = np.random.Generator(np.random.PCG64(2))
        xyz, Z, _ = water_cluster(100, rng)
        grid = dict(nrad=75, nang_order=29)
        funcs, nocc = (101, 130), 500
    elif name == "mol30":
        rng = np.random.Generator(np.random.PCG64(4))
        xyz, Z = random_walk_molecule(30, rng)
        grid = dict(nrad=75, nang_order=29)
        funcs, nocc = (101, 130), None
    else:
        raise KeyError(name)
    basis = make_manual_basis(sto_functions(Z, xyz, rng))
    if nocc is None:
        nocc = int(Z.sum()) // 2
    return dict(name=name, xyz=xyz, Z=Z, basis=basis, grid=grid, funcs=funcs, nocc=nocc, rng=rng)


def molset(nmol=64, seed=3):
    """The 64-molecule batching workload: list of (xyz, Z, basis, nocc)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mols = []
    for _ in range(nmol):
        na = int(rng.integers(3, 61))
        xyz, Z = random_walk_molecule(na, rng)
        basis = make_manual_basis(sto_functions(Z, xyz, rng))
        mols.append(dict(xyz=xyz, Z=Z, basis=basis, nocc=int(Z.sum()) // 2))
    return mols, rng


def loewdin_orbitals(S, nocc, rng):
    """C = C0 (C0^T S C0)^(-1/2) with C0 ~ N(0, 1); returns C, P = 2 C C^T."""
    nbf = S.shape[0]
    C0 = rng.standard_normal((nbf, nocc))
    M = C0.T @ S @ C0
    e, U = np.linalg.eigh(0.5 * (M + M.T))
    C = C0 @ (U * e ** -0.5) @ U.T
    return C, 2.0 * C @ C.T
