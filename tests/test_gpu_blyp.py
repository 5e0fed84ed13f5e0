"""B88 exchange, LYP and VWN correlation (ids 106, 131, 7) and PW91
exchange and correlation (109, 134) on the GPU
against the oracle: pointwise (unpolarised and polarised, including one-spin points),
in the batched V_xc build (FP64 and int8 mixed), and the helium atom built
end to end from the Thakkar table (reader, grid, build) against the radial
integrals the oracle tests pin to Becke's and Lee-Yang-Parr's values."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpu_util import oracle_on_grid, orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs, readers  # noqa: E402
from oracle import oracle as O  # noqa: E402
from test_oracle_blyp import HE, _he_density, _radial  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.tensor(a, device="cuda")


@pytest.mark.parametrize("ids", [[106], [131], [106, 131], [7], [1, 7], [109], [134],
                                 [109, 134]])
def test_functional_eval_matches_oracle(ctx, ids):
    rng = np.random.default_rng(11)
    n = 4000
    rho = 10 ** rng.uniform(-8, 2, n)
    sig = rho ** (8 / 3) * 10 ** rng.uniform(-4, 2, n)
    e0, v0, s0 = O.xc_eval(ids, rho, sig)
    e1, v1, s1 = (x.cpu().numpy() for x in nk.functional_eval(ctx, ids, dev(rho), dev(sig)))
    de = np.abs(e1 - e0) / np.abs(e0).max(initial=1e-300)
    rel = lambda a, b: np.abs(a - b) / (np.abs(b) + 1e-30)  # noqa: E731
    print(f"{ids}: e {rel(e1, e0).max():.1e}, vrho {rel(v1, v0).max():.1e}, "
          f"vsigma {rel(s1, s0).max():.1e}")
    assert rel(e1, e0).max() < 1e-11 and rel(v1, v0).max() < 1e-11 and de.max() < 1e-12
    assert rel(s1, s0).max() < 1e-10
    # polarised, a tenth of the points with one empty spin channel
    r2 = np.stack([rho * rng.uniform(0.2, 0.8, n), rho * rng.uniform(0.2, 0.8, n)], 1)
    r2[: n // 10, 1] = 0.0
    s3 = np.stack([sig / 4, sig / 8, sig / 4], 1)
    e0, v0, s0 = O.xc_eval_polarized(ids, r2, s3)
    e1, v1, s1 = (x.cpu().numpy() for x in
                  nk.functional_eval_polarized(ctx, ids, dev(r2), dev(s3)))
    sc = np.abs(e0) + 1e-3 * np.abs(O.xc_eval_polarized([106], r2, s3)[0])
    print(f"{ids} polarised: e {(np.abs(e1 - e0) / sc).max():.1e}")
    assert (np.abs(e1 - e0) / sc).max() < 1e-11
    assert np.allclose(v1, v0, rtol=1e-10, atol=1e-14 * np.abs(v0).max())
    if any(i > 100 for i in ids):  # an LDA-only list leaves vsigma unwritten
        assert np.allclose(s1, s0, rtol=1e-9, atol=1e-12 * np.abs(s0).max())


@pytest.mark.parametrize("funcs", [(106, 131), (109, 134)], ids=["BLYP", "PW91"])
def test_blyp_build_water_matches_oracle(ctx, funcs):
    w = configs.workload("water")
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], nrad=40, nang_order=23)
    ctx.set_basis(w["basis"])
    pts, wts = grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy()
    S, C, P = orbitals_on_grid(w, pts, wts)
    phi, gx, gy, gz = oracle_on_grid(w["basis"], pts, grad=True)
    E0, N0, V0 = O.vxc_from_collocation(phi, wts, funcs, P=P, gx=gx, gy=gy, gz=gz)
    r = ctx.vxc(dev(P), funcs=funcs)
    rm = ctx.vxc(dev(P), funcs=funcs, precision="mixed")
    dV = np.abs(r.potential.cpu().numpy() - V0).max()
    dVm = np.abs(rm.potential.cpu().numpy() - V0).max()
    print(f"{funcs} water: E = {r.energy:.10f}, dE = {r.energy - E0:.1e}, dV = {dV:.1e}; "
          f"mixed dE = {rm.energy - E0:.1e}, dV = {dVm:.1e}")
    assert abs(r.energy - E0) < 1e-8 and dV < 1e-9
    assert abs(rm.energy - E0) < 1e-6 and dVm < 1e-7


def test_helium_from_thakkar_table(ctx):
    """Reader -> grid -> build: B88 and LYP energies of the Hartree-Fock He
    density on the molecular grid agree with the radial integrals."""
    xyz, Z = np.zeros((1, 3)), np.array([2])
    ctx.make_flat_grid(xyz, Z, nrad=100, nang_order=17)
    b = readers.load_thakkar_basis(xyz, Z, HE.rsplit("/", 1)[0])
    ctx.set_basis(b)
    c = np.array([float(l.split()[2]) for l in open(HE)
                  if len(l.split()) == 3 and l.split()[0][0].isdigit()])
    P = 2.0 * np.outer(c, c)
    r, wr = _radial()
    rho, drho = _he_density(r)
    for ids in ((106,), (131,)):
        E = ctx.vxc(dev(P), funcs=ids).energy
        E0 = float(wr @ (rho * O.xc_eval(list(ids), rho, drho * drho)[0]))
        print(f"He {ids}: grid {E:.8f}, radial {E0:.8f}")
        assert abs(E - E0) < 1e-6
