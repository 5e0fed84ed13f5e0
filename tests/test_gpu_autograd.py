"""dE_xc/dP and dE_xc/dR through torch.autograd: gradcheck of the FP64 build
(numerical Jacobian by central differences in every element of P against
V_xc) and of the nuclear gradient (central differences in every coordinate,
grid and basis rebuilt at each displaced geometry, against nkx_xc_grad)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpu_util import orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402
from nukexc_amd.autograd import (function_atoms, xc_energy,  # noqa: E402
                                 xc_energy_of_geometry, xc_energy_polarized)


@pytest.fixture(scope="module")
def setup():
    c = nk.Context(0)
    w = configs.workload("water")
    grid = c.make_flat_grid(w["xyz"], w["Z"], nrad=30, nang_order=17)
    c.set_basis(w["basis"])
    S, C, P = orbitals_on_grid(w, grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy())
    yield c, w, C, P
    c.close()


@pytest.mark.parametrize("funcs", [(1, 12), (101, 130)])
def test_gradcheck_vxc(setup, funcs):
    c, w, C, P = setup
    Pt = torch.tensor(P, dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: xc_energy(c, x, funcs=funcs), (Pt,), eps=1e-6,
                                    atol=1e-7, rtol=1e-5)


def test_gradcheck_vxc_polarized(setup):
    c, w, C, P = setup
    Pa = torch.tensor(C @ C.T, dtype=torch.float64, device="cuda", requires_grad=True)
    Pb = torch.tensor(C[:, :-1] @ C[:, :-1].T, dtype=torch.float64, device="cuda",
                      requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: xc_energy_polarized(c, a, b, funcs=(101, 130)),
                                    (Pa, Pb), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_backward_is_vxc(setup):
    c, w, C, P = setup
    Pt = torch.tensor(P, dtype=torch.float64, device="cuda", requires_grad=True)
    E = xc_energy(c, Pt, funcs=(101, 130))
    E.backward()
    r = c.vxc(torch.tensor(P, dtype=torch.float64, device="cuda"), funcs=(101, 130))
    assert abs(E.item() - r.energy) == 0.0
    assert torch.equal(Pt.grad, r.potential)


@pytest.mark.parametrize("funcs", [(1, 12), (101, 130)])
def test_gradcheck_nuclear_gradient(setup, funcs):
    c, w, C, P = setup
    grid = dict(nrad=30, nang_order=17)
    Pt = torch.tensor(P, dtype=torch.float64, device="cuda")
    xyz = torch.tensor(w["xyz"], dtype=torch.float64, requires_grad=True)
    fatom = function_atoms(w["basis"], w["xyz"])
    try:
        assert torch.autograd.gradcheck(
            lambda x: xc_energy_of_geometry(c, x, w["Z"], w["basis"], fatom, Pt, funcs=funcs,
                                            grid=grid),
            (xyz,), eps=1e-5, atol=1e-6, rtol=1e-5)
    finally:  # the module's other tests use the grid and basis at xyz
        c.make_flat_grid(w["xyz"], w["Z"], **grid)
        c.set_basis(w["basis"])
