"""The XC-path workloads at their real sizes on the GPU against the oracle's
dense build (oracle.vxc_dense, the reference's GEMM route in FP64):

* benzene, PBE, TA 99 x Lebedev 590 (the whole 700 k-point grid), FP64 and
  the int8 mixed form;
* the 100-water cluster (2500 functions): the density and sigma of the
  whole 6.3 M-point grid on a seeded sample of points, and the V_xc build on
  the grid points of one interior water handed over with nkx_set_grid (so
  the 2500-function screening of compact point batches is exercised);
* the smallest, the median and the largest molecule of the 64-molecule set
  at TA 75 x Lebedev 302 in one set build, each against the oracle;
* all 90 components of the 30-atom gradient against central differences.

Bars: FP64 |dE_xc| <= 1e-8 Eh and max|dV_xc| <= 1e-9; mixed 1e-6 and 1e-7;
gradients 1e-6 Eh/bohr."""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpu_util import basis_dict  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402
from nukexc_amd.api import _density_points  # noqa: E402
from oracle import oracle as O  # noqa: E402

PBE = (101, 130)


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), device="cuda")


def _check(tag, r, E0, V0, tolE, tolV):
    dE = r.energy - E0
    dV = float(np.abs(r.potential.cpu().numpy() - V0).max())
    print(f"{tag}: E = {r.energy:.10f}, dE = {dE:.2e}, max|dV| = {dV:.2e}", flush=True)
    assert abs(dE) <= tolE and dV <= tolV


def test_benzene_full_grid(ctx):
    w = configs.workload("benzene")
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
    ctx.set_basis(w["basis"])
    C, P = configs.loewdin_orbitals(ctx.overlap().cpu().numpy(), w["nocc"], w["rng"])
    pts, wts = grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy()
    print(f"benzene: {len(wts)} points, {w['basis'].nbf()} functions", flush=True)
    t = time.time()
    E0, N0, V0 = O.vxc_dense(basis_dict(w["basis"]), pts, wts, PBE, P)
    print(f"oracle {time.time() - t:.1f} s; N_el = {N0:.10f}", flush=True)
    assert abs(N0 - 2 * w["nocc"]) < 1e-8
    Pd = dev(P)
    r = ctx.vxc(Pd, funcs=PBE, precision="fp64")
    _check("benzene FP64", r, E0, V0, 1e-8, 1e-9)
    assert abs(r.n_el - N0) < 1e-9
    _check("benzene mixed", ctx.vxc(Pd, funcs=PBE, precision="mixed"), E0, V0, 1e-6, 1e-7)


def test_water100_density_and_one_water(ctx):
    w = configs.workload("water100")
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
    ctx.set_basis(w["basis"])
    C, P = configs.loewdin_orbitals(ctx.overlap().cpu().numpy(), w["nocc"], w["rng"])
    Pd = dev(P)
    pts, wts = grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy()
    own = grid.point_owner.cpu().numpy()
    b = basis_dict(w["basis"])
    # density and sigma of the whole grid, on a seeded sample
    rho, gx, gy, gz, sig = (x.cpu().numpy() for x in _density_points(ctx, Pd, True))
    sel = np.sort(np.random.default_rng(7).choice(len(wts), 20000, replace=False))
    _, _, _, rho0, sig0 = O.vxc_dense(b, pts[sel], wts[sel], PBE, P, want_density=True)
    drho = np.abs(rho[sel] - rho0) / (np.abs(rho0) + 1e-12)
    dsig = np.abs(sig[sel] - sig0) / (np.abs(sig0) + 1e-12)
    print(f"water100 density on {len(sel)} of {len(wts)} points: max rel drho {drho.max():.1e}, "
          f"dsigma {dsig.max():.1e}", flush=True)
    assert drho.max() < 1e-11 and dsig.max() < 1e-10
    # V_xc on the points of the water nearest the cluster's centre
    ctr = w["xyz"].mean(0)
    io = min((i for i in range(len(w["Z"])) if w["Z"][i] == 8),
             key=lambda i: np.linalg.norm(w["xyz"][i] - ctr))
    keep = np.isin(own, [io, io + 1, io + 2])
    sub = nk.FlatGrid(dev(pts[keep]), dev(wts[keep]), grid.atom_centers, grid.Z,
                      torch.tensor(own[keep], dtype=torch.int32, device="cuda"))
    ctx.set_grid(sub)
    st = ctx.plan_stats()
    print(f"one water: {st['npts']} points, {st['pairs_kept']} of {st['pairs_total']} "
          f"(block, batch) pairs live", flush=True)
    t = time.time()
    E0, N0, V0 = O.vxc_dense(b, pts[keep], wts[keep], PBE, P)
    print(f"oracle {time.time() - t:.1f} s", flush=True)
    _check("water100 one water FP64", ctx.vxc(Pd, funcs=PBE, precision="fp64"), E0, V0,
           1e-8, 1e-9)
    _check("water100 one water mixed", ctx.vxc(Pd, funcs=PBE, precision="mixed"), E0, V0,
           1e-6, 1e-7)


def test_molset_three_molecules_in_one_set(ctx):
    mols, rng = configs.molset(64, 3)
    order = np.argsort([len(m["Z"]) for m in mols], kind="stable")
    pick = [mols[i] for i in (order[0], order[len(order) // 2], order[-1])]
    grid_kw = dict(nrad=75, nang_order=29)
    Ps, grids = [], []
    for m in pick:
        g = ctx.make_flat_grid(m["xyz"], m["Z"], **grid_kw)
        ctx.set_basis(m["basis"])
        C, P = configs.loewdin_orbitals(ctx.overlap().cpu().numpy(), m["nocc"], rng)
        Ps.append(P)
        grids.append((g.quad_points.cpu().numpy(), g.weights.cpu().numpy()))
    npts = ctx.molset_create(pick, **grid_kw)
    assert [len(g[1]) for g in grids] == list(npts)
    res = {p: ctx.molset_vxc([dev(P) for P in Ps], funcs=PBE, precision=p)
           for p in ("fp64", "mixed")}
    for k, (m, P, (pts, wts)) in enumerate(zip(pick, Ps, grids)):
        t = time.time()
        E0, N0, V0 = O.vxc_dense(basis_dict(m["basis"]), pts, wts, PBE, P)
        tag = f"set molecule {k}: {len(m['Z'])} atoms, {m['basis'].nbf()} functions, {len(wts)} points"
        print(f"{tag}; oracle {time.time() - t:.1f} s", flush=True)
        _check(tag + " FP64", res["fp64"][k], E0, V0, 1e-8, 1e-9)
        _check(tag + " mixed", res["mixed"][k], E0, V0, 1e-6, 1e-7)


def test_mol30_gradient_all_components_vs_fd():
    """The 30-atom gradient (FP64 and mixed) on all 90 components against
    central differences (h = 1e-4 bohr) of the FP64 build's E_xc, whose value
    at the reference geometry is checked against the oracle first."""
    from test_gpu_grad import _fn_atom, _moved
    c = nk.Context(0)
    try:
        w = configs.workload("mol30")
        grid = c.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
        c.set_basis(w["basis"])
        C, P = configs.loewdin_orbitals(c.overlap().cpu().numpy(), w["nocc"], w["rng"])
        Pt = dev(P)
        E0, _, _ = O.vxc_dense(basis_dict(w["basis"]), grid.quad_points.cpu().numpy(),
                               grid.weights.cpu().numpy(), PBE, P)
        E, g = c.xc_grad(Pt, funcs=PBE)
        Em, gm = c.xc_grad(Pt, funcs=PBE, precision="mixed")
        print(f"mol30: E FP64 - oracle {E - E0:.1e}", flush=True)
        assert abs(E - E0) < 1e-8
        fa = _fn_atom(w["basis"], w["xyz"])
        h = 1e-4
        fd = np.zeros_like(g)
        for a in range(len(w["Z"])):
            for k in range(3):
                e = []
                for t in (1, -1):
                    xyz = w["xyz"].copy()
                    xyz[a, k] += t * h
                    c.make_flat_grid(xyz, w["Z"], **w["grid"])
                    c.set_basis(_moved(w["basis"], fa, xyz))
                    e.append(c.vxc(Pt, funcs=PBE, precision="fp64").energy)
                fd[a, k] = (e[0] - e[1]) / (2 * h)
            print(f"atom {a}: max |g - fd| FP64 {np.abs(g[a] - fd[a]).max():.1e}, "
                  f"mixed {np.abs(gm[a] - fd[a]).max():.1e}", flush=True)
        assert np.abs(g - fd).max() <= 1e-6
        assert np.abs(gm - fd).max() <= 1e-6
    finally:
        c.close()
