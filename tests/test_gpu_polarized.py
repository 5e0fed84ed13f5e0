"""GPU parity of the spin-polarised path (compute_lsda / compute_gga_lsda,
xc_integrals.hpp:169-262, :543-718, and the batched open-shell build) against
the FP64 CPU oracle, through the C ABI.  Bars as for the closed-shell path:
|dE_xc| <= 1e-8 Eh, max|dV_s| <= 1e-9; functional values to ~1e-10 relative
of the exchange scale."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpu_util import oracle_on_grid, orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-8, 2, (n, 2))
    rho[: n // 10, 1] = 0.0                 # fully polarised points
    rho[n // 10: n // 5, 0] = 1e-14          # a channel below threshold
    ga = rng.normal(size=(n, 3)) * np.sqrt(rho[:, :1]) * 10 ** rng.uniform(-2, 1, (n, 1))
    gb = rng.normal(size=(n, 3)) * np.sqrt(rho[:, 1:]) * 10 ** rng.uniform(-2, 1, (n, 1))
    sig = np.column_stack([(ga * ga).sum(1), (ga * gb).sum(1), (gb * gb).sum(1)])
    return rho, sig


@pytest.mark.parametrize("ids", [[1], [12], [13], [101], [130], [1, 12], [101, 130]])
def test_functional_eval_polarized_matches_oracle(ctx, ids):
    rho, sig = _inputs(20000, 7)
    gga = any(i > 100 for i in ids)
    e0, v0, s0 = O.xc_eval_polarized(ids, rho, sig if gga else None)
    e1, v1, s1 = nk.functional_eval_polarized(ctx, ids, dev(rho), dev(sig) if gga else None)
    e1, v1 = e1.cpu().numpy(), v1.cpu().numpy()
    # error scale: the exchange values at the same point (the correlation
    # energy of a saturated PBE enhancement cancels to rounding level)
    ex, vx, _ = O.xc_eval_polarized([1], rho, sig)
    _, _, sx = O.xc_eval_polarized([101], rho, sig)
    rt = rho.sum(1)
    de = np.abs(e1 - e0) * rt / (np.abs(e0) * rt + 1e-3 * np.abs(ex) * rt + 1e-18)
    dv = np.abs(v1 - v0) / (np.abs(v0) + 1e-3 * np.abs(vx).max(1, keepdims=True) + 1e-18)
    print(f"{ids}: max scaled err e {de.max():.1e} vrho {dv.max():.1e}")
    assert de.max() < 1e-10 and dv.max() < 1e-9
    if gga:
        s1 = s1.cpu().numpy()
        # where the PBE enhancement saturates (rho ~ 1e-8 with large
        # gradients) vsigma is ~1e-19 and rounding noise in both codes; the
        # last term measures a vsigma error by what it contributes to the
        # energy density, |dv| sigma against 1e-9 |e_x|
        st = (sig[:, 0] + sig[:, 2] + np.abs(sig[:, 1]))[:, None]
        ds = np.abs(s1 - s0) / (np.abs(s0) + 1e-3 * np.abs(sx[:, [0, 2]]).max(1, keepdims=True)
                                + 1e-9 * (np.abs(ex) * rt)[:, None] / (st + 1e-300) + 1e-300)
        print(f"{ids}: vsigma {ds.max():.1e}")
        assert ds.max() < 1e-9


def _open_shell(ctx, name, **grid_kw):
    """Grid and basis on the ctx; spin density matrices of an open-shell
    occupation built from the seeded Loewdin orbitals: alpha holds nocc,
    beta nocc - 1 orbitals (a doublet cation of the closed-shell workload)."""
    w = configs.workload(name)
    gkw = dict(w["grid"])
    gkw.update(grid_kw)
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], **gkw)
    ctx.set_basis(w["basis"])
    pts = grid.quad_points.cpu().numpy()
    wts = grid.weights.cpu().numpy()
    S, C, P = orbitals_on_grid(w, pts, wts)
    Ca, Cb = C, C[:, :-1]
    return w, grid, pts, wts, Ca, Cb, Ca @ Ca.T, Cb @ Cb.T


@pytest.mark.parametrize("name,grid_kw", [("water", {}), ("benzene", dict(nrad=40, nang_order=23))])
def test_vxc_polarized_matches_oracle(ctx, name, grid_kw):
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, name, **grid_kw)
    gga = any(f > 100 for f in w["funcs"])
    phi, gx, gy, gz = oracle_on_grid(w["basis"], pts, grad=True)
    g = dict(gx=gx, gy=gy, gz=gz) if gga else {}
    E0, N0, Va0, Vb0 = O.vxc_polarized(phi, wts, w["funcs"], Pa, Pb, **g)
    r = ctx.vxc_polarized(dev(Pa), dev(Pb), funcs=w["funcs"])
    Va, Vb = r.potential_alpha.cpu().numpy(), r.potential_beta.cpu().numpy()
    print(f"{name}: E={r.energy:.12f} dE={r.energy - E0:.2e} max|dVa|={np.abs(Va - Va0).max():.2e} "
          f"max|dVb|={np.abs(Vb - Vb0).max():.2e} N={r.n_el}")
    assert abs(r.energy - E0) <= 1e-8
    assert np.abs(Va - Va0).max() <= 1e-9 and np.abs(Vb - Vb0).max() <= 1e-9
    assert np.array_equal(Va, Va.T) and np.array_equal(Vb, Vb.T)
    nocc = w["nocc"]
    assert abs(r.n_el[0] - nocc) < 1e-9 and abs(r.n_el[1] - (nocc - 1)) < 1e-9


@pytest.mark.parametrize("family", ["lda", "gga"])
def test_reference_shaped_lsda_matches_oracle(ctx, family):
    """compute_lsda / compute_gga_lsda on caller collocation arrays (MO route)
    against the oracle (P route)."""
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, "water", nrad=30, nang_order=17)
    funcs = [1, 12] if family == "lda" else [101, 130]
    phi = nk.fill_collocation(ctx, w["basis"], grid.quad_points)
    if family == "lda":
        r = nk.compute_lsda(ctx, phi, grid.weights, Ca, np.ones(Ca.shape[1]), Cb,
                            np.ones(Cb.shape[1]), funcs)
    else:
        gx, gy, gz = nk.fill_grad_collocation(ctx, w["basis"], grid.quad_points)
        r = nk.compute_gga_lsda(ctx, phi, gx, gy, gz, grid.weights, Ca, np.ones(Ca.shape[1]),
                                Cb, np.ones(Cb.shape[1]), funcs)
    ophi, ogx, ogy, ogz = oracle_on_grid(w["basis"], pts, grad=True)
    g = dict(gx=ogx, gy=ogy, gz=ogz) if family == "gga" else {}
    E0, _, Va0, Vb0 = O.vxc_polarized(ophi, wts, funcs, Pa, Pb, **g)
    Va, Vb = r.potential_alpha.cpu().numpy(), r.potential_beta.cpu().numpy()
    assert abs(r.energy - E0) <= 1e-8
    assert np.abs(Va - Va0).max() <= 1e-9 and np.abs(Vb - Vb0).max() <= 1e-9


def test_vxc_polarized_spin_symmetric_equals_closed_shell(ctx):
    """Pa = Pb = P/2: the polarised build reproduces nkx_vxc(P) (two GPU
    pipelines, one identity)."""
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, "water")
    r1 = ctx.vxc(dev(2 * Pa), funcs=w["funcs"])
    r2 = ctx.vxc_polarized(dev(Pa), dev(Pa), funcs=w["funcs"])
    V = r1.potential.cpu().numpy()
    assert abs(r2.energy - r1.energy) < 1e-11 * abs(r1.energy)
    assert np.abs(r2.potential_alpha.cpu().numpy() - V).max() < 1e-11
    assert np.abs(r2.potential_beta.cpu().numpy() - V).max() < 1e-11


@pytest.mark.parametrize("name", ["water", "benzene"])
def test_vxc_polarized_mixed_matches_fp64(ctx, name):
    """The int8 form of the open-shell build (one density and one V
    contraction per spin, polarised functional in FP64) against the FP64
    form: the mixed-form bars on E_xc and both potentials, and N per spin."""
    kw = dict(nrad=40, nang_order=23) if name == "benzene" else {}
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, name, **kw)
    r64 = ctx.vxc_polarized(dev(Pa), dev(Pb), funcs=(101, 130))
    rmx = ctx.vxc_polarized(dev(Pa), dev(Pb), funcs=(101, 130), precision="mixed")
    dVa = (rmx.potential_alpha - r64.potential_alpha).abs().max().item()
    dVb = (rmx.potential_beta - r64.potential_beta).abs().max().item()
    print(f"{name}: dE = {rmx.energy - r64.energy:.1e}, dV_a = {dVa:.1e}, dV_b = {dVb:.1e}, "
          f"dN = {rmx.n_el[0] - r64.n_el[0]:.1e}, {rmx.n_el[1] - r64.n_el[1]:.1e}")
    assert abs(rmx.energy - r64.energy) <= 1e-6
    assert dVa <= 1e-7 and dVb <= 1e-7
    assert abs(rmx.n_el[0] - r64.n_el[0]) <= 1e-6 and abs(rmx.n_el[1] - r64.n_el[1]) <= 1e-6


def test_vxc_polarized_mixed_spin_symmetric_equals_closed_shell(ctx):
    """Pa = Pb = P/2 in the int8 form: the slices of P/2 are those of P one
    exponent down, so the polarised build repeats the closed-shell int8 build
    up to the rounding of the two functional forms."""
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, "water")
    r1 = ctx.vxc(dev(2 * Pa), funcs=(101, 130), precision="mixed")
    r2 = ctx.vxc_polarized(dev(Pa), dev(Pa), funcs=(101, 130), precision="mixed")
    V = r1.potential.cpu().numpy()
    print(f"dE = {r2.energy - r1.energy:.1e}, "
          f"dV = {np.abs(r2.potential_alpha.cpu().numpy() - V).max():.1e}")
    assert abs(r2.energy - r1.energy) < 1e-10 * abs(r1.energy)
    assert np.abs(r2.potential_alpha.cpu().numpy() - V).max() < 1e-9
    assert np.abs(r2.potential_beta.cpu().numpy() - V).max() < 1e-9


def test_polarized_errors(ctx):
    w, grid, pts, wts, Ca, Cb, Pa, Pb = _open_shell(ctx, "water", nrad=20, nang_order=11)
    with pytest.raises(nk.NukexcError) as ei:
        ctx.vxc_polarized(dev(Pa), dev(Pb), funcs=w["funcs"], precision="mixed_bf16")
    assert ei.value.status == "NKX_ERR_UNSUPPORTED"
    phi = nk.fill_collocation(ctx, w["basis"], grid.quad_points)
    with pytest.raises(nk.NukexcError) as ei:
        nk.compute_lsda(ctx, phi, grid.weights, Ca, np.ones(Ca.shape[1]), Cb,
                        np.ones(Cb.shape[1]), [101, 130])
    assert ei.value.status == "NKX_ERR_FUNCTIONAL_FAMILY"
