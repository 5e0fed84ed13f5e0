"""Open-shell nuclear gradient (nkx_xc_grad_polarized, FP64 and the int8
mixed form) against fourth-order central differences of the oracle's
spin-polarised energy, the grid rebuilt at each displaced geometry and the
basis centres moved with their atoms; the spin-symmetric case against the
closed-shell gradient.  Bars: |dE/dR - FD| <= 1e-6 Eh/bohr, sum_A dE/dR_A = 0."""
import ctypes
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from conftest import lebedev  # noqa: E402
from gpu_util import basis_dict, orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import _capi, configs  # noqa: E402
from nukexc_amd.autograd import function_atoms  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


def _setup(ctx, name, nrad, order):
    """Doublet cation occupation of the seeded orbitals: alpha holds nocc,
    beta nocc - 1 orbitals."""
    w = configs.workload(name)
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], nrad=nrad, nang_order=order)
    ctx.set_basis(w["basis"])
    S, C, P = orbitals_on_grid(w, grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy())
    Pa, Pb = C @ C.T, C[:, :-1] @ C[:, :-1].T
    fa = function_atoms(w["basis"], w["xyz"])
    return w, P, Pa, Pb, fa


def _oracle_energy(w, xyz, fa, Pa, Pb, funcs, nrad, order):
    pts, wts, _ = O.make_flat_grid(xyz, w["Z"], nrad, lebedev(order), scheme=0, alpha=0.6)
    b = basis_dict(dataclasses.replace(w["basis"], O=np.ascontiguousarray(xyz[fa])))
    phi, gx, gy, gz = O.collocation(b, pts, grad=True)
    g = dict(gx=gx, gy=gy, gz=gz) if any(f > 100 for f in funcs) else {}
    return O.vxc_polarized(phi, wts, funcs, Pa, Pb, **g)[0]


def dev(a):
    return torch.tensor(a, device="cuda")


@pytest.mark.parametrize("funcs", [(1, 12), (101, 130)])
def test_xc_grad_polarized_water_vs_oracle_fd(ctx, funcs):
    nrad, order = 30, 17
    w, P, Pa, Pb, fa = _setup(ctx, "water", nrad, order)
    E, g = ctx.xc_grad_polarized(dev(Pa), dev(Pb), funcs=funcs)
    Em, gm = ctx.xc_grad_polarized(dev(Pa), dev(Pb), funcs=funcs, precision="mixed")
    E0 = _oracle_energy(w, w["xyz"], fa, Pa, Pb, funcs, nrad, order)
    # the energy of the gradient call is the polarised build's
    Eb = ctx.vxc_polarized(dev(Pa), dev(Pb), funcs=funcs).energy
    assert abs(E - E0) < 1e-8 and abs(E - Eb) < 1e-10
    h = 5e-4
    fd = np.zeros_like(g)
    for a in range(len(w["Z"])):
        for k in range(3):
            e = []
            for t in (2, 1, -1, -2):
                xyz = w["xyz"].copy()
                xyz[a, k] += t * h
                e.append(_oracle_energy(w, xyz, fa, Pa, Pb, funcs, nrad, order))
            fd[a, k] = (-e[0] + 8 * e[1] - 8 * e[2] + e[3]) / (12 * h)
    print(f"{funcs}: max|g - fd| = {np.abs(g - fd).max():.2e}  mixed: |E - E0| = "
          f"{abs(Em - E0):.1e}  max|g - fd| = {np.abs(gm - fd).max():.2e}")
    assert np.abs(g - fd).max() <= 1e-6
    assert np.abs(g.sum(0)).max() <= 1e-10
    assert abs(Em - E0) <= 1e-6
    assert np.abs(gm - fd).max() <= 1e-6
    assert np.abs(gm.sum(0)).max() <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_xc_grad_polarized_spin_symmetric_equals_closed_shell(ctx, prec):
    """Pa = Pb = P/2: the polarised functional reduces to the unpolarised
    one and the two spins' basis-centre terms add up to the closed-shell
    one (in the int8 form the slices of P/2 are those of P one exponent
    down; the two functional forms round differently, by about 1e-9 of a
    Y slice in the int8 form)."""
    w, P, Pa, Pb, fa = _setup(ctx, "benzene", 20, 11)
    E1, g1 = ctx.xc_grad(dev(P), funcs=(101, 130), precision=prec)
    E2, g2 = ctx.xc_grad_polarized(dev(0.5 * P), dev(0.5 * P), funcs=(101, 130),
                                   precision=prec)
    print(f"{prec}: dE = {E2 - E1:.1e}, max|dg| = {np.abs(g2 - g1).max():.1e}")
    assert abs(E2 - E1) < 1e-10 * abs(E1)
    assert np.abs(g2 - g1).max() < (1e-10 if prec == "fp64" else 1e-9)


def test_xc_grad_polarized_mixed_matches_fp64_benzene(ctx):
    w, P, Pa, Pb, fa = _setup(ctx, "benzene", 20, 11)
    E, g = ctx.xc_grad_polarized(dev(Pa), dev(Pb), funcs=(101, 130))
    Em, gm = ctx.xc_grad_polarized(dev(Pa), dev(Pb), funcs=(101, 130), precision="mixed")
    print(f"|dE| = {abs(Em - E):.1e}, max|dg| = {np.abs(gm - g).max():.1e}")
    assert abs(Em - E) <= 1e-7
    assert np.abs(gm - g).max() <= 1e-7
    E2, g2 = ctx.xc_grad_polarized(dev(Pa), dev(Pb), funcs=(101, 130), precision="mixed")
    assert E2 == Em and np.array_equal(g2, gm)  # fixed summation order


def test_xc_grad_polarized_rejects_bf16(ctx):
    """The gradient has FP64 and int8 forms only; the C ABI refuses the
    split-bf16 precision code."""
    w, P, Pa, Pb, fa = _setup(ctx, "water", 20, 11)
    A, B = dev(Pa), dev(Pb)
    ids = (ctypes.c_int * 1)(1)
    E = ctypes.c_double()
    g = np.zeros(3 * len(w["Z"]))
    rc = _capi.lib().nkx_xc_grad_polarized(
        ctx.h, A.data_ptr(), B.data_ptr(), 1, ids, 1, _capi.PREC_MIXED_BF16, ctypes.byref(E),
        g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert _capi.STATUS_NAMES[rc] == "NKX_ERR_INVALID_ARGUMENT"
