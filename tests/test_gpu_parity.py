"""GPU parity: every kernel of libnukexc_hip.so against the FP64 CPU oracle on
the same inputs, through the C ABI.  FP64 bars from the XC-path contract:
|dE_xc| <= 1e-8 Eh, max|dV_xc| <= 1e-9; element-wise kernels to ~1e-13."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from conftest import ROOT, lebedev, next_order  # noqa: E402
from gpu_util import basis_dict, oracle_on_grid, orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


# ------------------------------------------------------------ MFMA layout ----
def test_mfma_f64_lane_layout():
    """v_mfma_f64_16x16x4f64: lane l feeds A[l%16][l/16] and B[l/16][l%16];
    result register r of lane l is D[4r + l/16][l%16] (rows interleaved
    across lane groups, unlike the FP32-output forms)."""
    probe = ctypes.CDLL(os.path.join(ROOT, "tests", "gpu_probe", "libmfma_probe.so"))
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 8, size=(16, 4)).astype(np.float64)
    B = rng.integers(-8, 8, size=(4, 16)).astype(np.float64)
    a = np.array([A[l % 16, l // 16] for l in range(64)])
    b = np.array([B[l // 16, l % 16] for l in range(64)])
    d = np.zeros(256)
    pd = ctypes.POINTER(ctypes.c_double)
    assert probe.mfma_probe_f64(a.ctypes.data_as(pd), b.ctypes.data_as(pd), d.ctypes.data_as(pd)) == 0
    D = A @ B
    got = np.array([[d[(j + 16 * (i % 4)) * 4 + i // 4] for j in range(16)] for i in range(16)])
    assert np.array_equal(got, D)


def _bf16_bits(x):
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_mfma_bf16_lane_layout():
    """v_mfma_f32_16x16x32_bf16: lane l feeds A[l%16][8(l/16)..+7] and
    B[8(l/16)..+7][l%16]; result as for the f64 form."""
    probe = ctypes.CDLL(os.path.join(ROOT, "tests", "gpu_probe", "libmfma_probe.so"))
    rng = np.random.default_rng(1)
    A = rng.integers(-4, 4, size=(16, 32)).astype(np.float32)
    B = rng.integers(-4, 4, size=(32, 16)).astype(np.float32)
    a = np.array([A[l % 16, 8 * (l // 16) + i] for l in range(64) for i in range(8)], np.float32)
    b = np.array([B[8 * (l // 16) + i, l % 16] for l in range(64) for i in range(8)], np.float32)
    ab, bb = _bf16_bits(a), _bf16_bits(b)
    d = np.zeros(256, np.float32)
    pu = ctypes.POINTER(ctypes.c_ushort)
    assert probe.mfma_probe_bf16(ab.ctypes.data_as(pu), bb.ctypes.data_as(pu),
                                 d.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    D = A @ B
    got = np.array([[d[(j + 16 * (i // 4)) * 4 + i % 4] for j in range(16)] for i in range(16)])
    assert np.array_equal(got, D)


# ------------------------------------------------------------------ Becke ----
def _raw_points(xyz, Z, nrad, order):
    ang = nk.lebedev_rule(order)
    P, W, Ow = [], [], []
    for a, (c, z) in enumerate(zip(xyz, Z)):
        r, w = nk.radial_rule(0, nrad, nk.radial_scale(0, z), 0.6)
        P.append((r[:, None, None] * ang[None, :, :3]).reshape(-1, 3) + c)
        W.append((w[:, None] * ang[None, :, 3]).reshape(-1))
        Ow.append(np.full(len(r) * len(ang), a, np.int32))
    return np.concatenate(P), np.concatenate(W), np.concatenate(Ow)


@pytest.mark.parametrize("capped", [True, False])
def test_becke_matches_oracle(ctx, capped):
    w = configs.workload("benzene")
    P, W, Ow = _raw_points(w["xyz"], w["Z"], 30, 17)
    ref = O.becke(w["xyz"], P, Ow, W, capped=capped)
    wd = dev(W)
    nk.partition_becke(ctx, dev(w["xyz"]), dev(P), dev(Ow, torch.int32), wd, capped=capped)
    got = wd.cpu().numpy()
    # Partition fractions p = w / w_raw in [0, 1]: absolute agreement.  Tiny
    # fractions come from 1 - p(p(p(mu))) cancellation and carry ~1e-16
    # absolute noise from the last bit of the arithmetic.
    err = np.abs(got - ref) / W
    print(f"becke capped={capped}: max |dp| = {err.max():.2e}")
    assert err.max() < 1e-14


@pytest.mark.parametrize("name,nrad,order,stride", [("water", 75, 29, 1), ("benzene", 99, 41, 1),
                                                    ("water100", 75, 29, 7)])
def test_becke_exact_screening_is_bitwise(ctx, name, nrad, order, stride):
    """The capped partition's exact screening (skipped atoms with P_i = 0,
    skipped factors s = 1) against every factor evaluated, bit for bit, on
    the raw product points of the water (75 x 302), benzene (99 x 590) and
    100-water (75 x 302, every 7th point) grids; plus the oracle's
    unscreened partition on a sample."""
    w = configs.workload(name)
    P, W, Ow = _raw_points(w["xyz"], w["Z"], nrad, order)
    P, W, Ow = P[::stride], W[::stride], Ow[::stride]
    print(f"{name}: {len(W)} points, {len(w['Z'])} atoms", flush=True)
    a = dev(W)
    b = dev(W)
    C, Pd, Od = dev(w["xyz"]), dev(P), dev(Ow, torch.int32)
    nk.partition_becke(ctx, C, Pd, Od, a, capped=True, screen=True)
    nk.partition_becke(ctx, C, Pd, Od, b, capped=True, screen=False)
    ga, gb = a.cpu().numpy(), b.cpu().numpy()
    assert np.array_equal(ga.view(np.int64), gb.view(np.int64))
    sel = np.random.default_rng(0).choice(len(W), size=min(len(W), 20000), replace=False)
    ref = O.becke(w["xyz"], P[sel], Ow[sel], W[sel], capped=True)
    assert (np.abs(ga[sel] - ref) / W[sel]).max() < 1e-14


def test_becke_partition_of_unity_gpu(ctx):
    """test_partition.cxx:197-258 on the GPU kernel."""
    centers = np.array([[i, i % 2, i % 3] for i in range(10)], float)
    jk = np.array([(0.0, j / 10, k / 10) for j in range(10) for k in range(10)])
    pts = np.tile(jk, (10, 1))
    owner = np.repeat(np.arange(10), 100).astype(np.int32)
    wd = dev(np.ones(1000))
    nk.partition_becke_team(ctx, dev(centers), dev(pts), dev(owner, torch.int32), wd)
    assert np.abs(wd.cpu().numpy().reshape(10, 100).sum(0) - 1).max() < 1e-10


# ------------------------------------------------------------------- grid ----
@pytest.mark.parametrize("nrad,s_ref", [(10, 0.859122134805326), (20, 0.858383667211934),
                                        (30, 0.858385391682404)])
def test_h2plus_grid_and_overlap_gpu(ctx, nrad, s_ref):
    """latex/Data/convergence_h2plus.csv through the GPU grid build and the
    GPU overlap build."""
    H2 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    grid = ctx.make_flat_grid(H2, [1, 1], nrad=nrad, nang_order=59)
    ctx.set_basis(nk.make_manual_basis([(1, 0, 0, 1.0, 0, 0, 0), (1, 0, 0, 1.0, 1.0, 0, 0)]))
    S = ctx.overlap().cpu().numpy()
    assert abs(S[0, 1] - s_ref) < 5e-13
    P, W, Ow = O.make_flat_grid(H2, [1, 1], nrad, nk.lebedev_rule(59), scheme=0)
    _same_grid_but_band(grid, H2, [1, 1], nrad, 59, f"H2+ {nrad} x 59")
    assert abs(grid.weights.sum().item() - W.sum()) < 1e-10 * W.sum()


def _same_grid_but_band(grid, xyz, Z, nrad, order, tag):
    """The device grid against the oracle's on the library's Lebedev table
    (same rule, same point order).  The product points are formed in one
    rounding order on both sides, so they agree bit for bit; the partition
    agrees to rounding, and the two compacted sets agree except for points
    whose partition fraction lies in the band where 1 - p(p(p(mu))) has no
    digits left (the last bit of a distance or quotient decides between an
    exact zero and a weight ~1e-16 of the raw one).  Every point kept by
    only one side must carry less than 1e-14 of its raw weight."""
    ang = nk.lebedev_rule(order)
    xyz = np.asarray(xyz, float)
    raw_p, raw_w = [], []
    for a, z in enumerate(Z):
        r, wr = O.radial(0, nrad, O.radial_scale(0, int(z)))
        raw_p.append(xyz[a] + r[:, None, None] * ang[None, :, :3])
        raw_w.append(wr[:, None] * ang[None, :, 3])
    raw_p = np.concatenate([x.reshape(-1, 3) for x in raw_p])
    raw_w = np.concatenate([x.reshape(-1) for x in raw_w])
    Pa, Wa, Oa = O.make_flat_grid(xyz, Z, nrad, ang, scheme=0, w_thr=-1.0)
    assert np.array_equal(Pa, raw_p)
    keep_o = Wa > 1e-30
    gp = grid.quad_points.cpu().numpy()
    gw = grid.weights.cpu().numpy()
    go = grid.point_owner.cpu().numpy()
    # the device points are a subsequence of the raw points, bit for bit
    pos = {row.tobytes(): i for i, row in enumerate(raw_p)}
    idx = np.array([pos[row.tobytes()] for row in gp], np.int64)
    assert np.all(np.diff(idx) > 0)
    keep_d = np.zeros(len(raw_p), bool)
    keep_d[idx] = True
    only_o = keep_o & ~keep_d
    only_d = keep_d & ~keep_o
    print(f"{tag}: device {len(gp)} points, oracle {keep_o.sum()}; kept by the oracle only "
          f"{only_o.sum()}, by the device only {only_d.sum()}")
    assert np.array_equal(go, Oa[idx])
    assert np.all(Wa[only_o] < 1e-14 * raw_w[only_o])
    dev_w = np.zeros(len(raw_p))
    dev_w[idx] = gw
    assert np.all(dev_w[only_d] < 1e-14 * raw_w[only_d])
    both = keep_o & keep_d
    assert np.all(np.abs(dev_w[both] - Wa[both]) <= 1e-12 * Wa[both] + 1e-15 * raw_w[both])


@pytest.mark.parametrize("name,nrad,order", [("water", 40, 29), ("water", 75, 29),
                                             ("benzene", 99, 41)])
def test_make_flat_grid_matches_oracle(ctx, name, nrad, order):
    """make_flat_grid on the device against the oracle's host loop."""
    w = configs.workload(name)
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], nrad=nrad, nang_order=order)
    _same_grid_but_band(grid, w["xyz"], w["Z"], nrad, order, f"{name} {nrad} x {order}")


# ------------------------------------------------------------ collocation ----
def test_collocation_matches_oracle_high_l(ctx):
    rng = np.random.default_rng(5)
    funcs = []
    for l in range(8):
        for m in range(-l, l + 1):
            n = l + 1 + (m % 2)
            funcs.append((n, l, m, float(rng.uniform(0.4, 3)), *rng.normal(size=3)))
    basis = nk.make_manual_basis(funcs)
    pts = rng.normal(scale=2.0, size=(3000, 3))
    phi = nk.fill_collocation(ctx, basis, dev(pts)).cpu().numpy()
    gx, gy, gz = (g.cpu().numpy() for g in nk.fill_grad_collocation(ctx, basis, dev(pts)))
    r = oracle_on_grid(basis, pts, grad=True)
    for got, ref in zip((phi, gx, gy, gz), r):
        scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300
        assert np.abs(got - ref).max() / scale.min() < 1e-11 or np.all(np.abs(got - ref) <= 1e-12 * scale)


# --------------------------------------------------------------- functional ----
@pytest.mark.parametrize("ids", [[1], [12], [13], [101], [130], [101, 130], [1, 12]])
def test_functional_matches_oracle(ctx, ids):
    rng = np.random.default_rng(sum(ids))
    rho = 10 ** rng.uniform(-16, 3, 4000)
    sig = 10 ** rng.uniform(-20, 4, 4000) * np.maximum(rho, 1e-10) ** (8 / 3)
    exc, vr, vs = nk.functional_eval(ctx, ids, dev(rho), dev(sig))
    e0, v0, s0 = O.xc_eval(ids, rho, sig)
    # Scale of each output: the PBE exchange values at the same point.  PBE
    # correlation cancels against its PW92 part at low density (eps_c + H ~ 0)
    # and PBE-C v_sigma cancels PBE-X v_sigma at small s by construction, so
    # agreement is measured relative to |ref| plus that scale.
    ex, vx, sx = O.xc_eval([101], rho, sig)
    scales = (np.abs(ex), np.abs(vx), np.abs(sx))
    for name, got, ref, sc in (("exc", exc, e0, scales[0]), ("vrho", vr, v0, scales[1]),
                               ("vsigma", vs, s0, scales[2])):
        got = got.cpu().numpy()
        err = np.abs(got - ref) / (np.abs(ref) + 1e-3 * sc + 1e-300)
        print(f"{ids} {name}: max scaled err {err.max():.2e}")
        assert err.max() < 1e-11


# ------------------------------------------- reference-shaped XC on GPU ----
def _single_atom(ctx, key):
    g = GOLD[key]
    Z = 1 if g["zeta"] == 1.0 else 2
    order = next_order(g["grid"]["nang_order"]) if g["grid"]["nrad"] < 1000 else 3
    grid = ctx.make_flat_grid([[0.0, 0.0, 0.0]], [Z], nrad=g["grid"]["nrad"], nang_order=order,
                              radial="becke")
    basis = nk.make_manual_basis([(1, 0, 0, g["zeta"], 0.0, 0.0, 0.0)])
    return g, grid, basis


def test_gpu_lda_x_hydrogen_golden(ctx):
    g, grid, basis = _single_atom(ctx, "h_lda_x")
    phi = nk.fill_collocation(ctx, basis, grid.quad_points)
    r = nk.compute_lda(ctx, phi, grid.weights, [[1.0]], [g["occ"]], g["func"])
    assert abs(r.energy - g["energy"]) < g["rel"] * abs(g["energy"])


@pytest.mark.parametrize("key", ["h_pbe_x", "he_pbe_x"])
def test_gpu_pbe_x_goldens(ctx, key):
    g, grid, basis = _single_atom(ctx, key)
    phi = nk.fill_collocation(ctx, basis, grid.quad_points)
    gx, gy, gz = nk.fill_grad_collocation(ctx, basis, grid.quad_points)
    r = nk.compute_gga(ctx, phi, gx, gy, gz, grid.weights, [[1.0]], [g["occ"]], g["func"])
    assert abs(r.energy - g["energy"]) < g["rel"] * abs(g["energy"])
    assert abs(r.potential[0, 0].item() - g["v00"]) < g["rel"] * abs(g["v00"])
    # the batched build gives the same numbers from P = f C C^T
    r2 = ctx.vxc(torch.tensor([[g["occ"]]], dtype=torch.float64), funcs=g["func"])
    assert abs(r2.energy - g["energy"]) < g["rel"] * abs(g["energy"])
    assert abs(r2.potential[0, 0].item() - g["v00"]) < g["rel"] * abs(g["v00"])


@pytest.mark.parametrize("key", ["h_density", "he_density"])
def test_gpu_density_integrals(ctx, key):
    g, grid, basis = _single_atom(ctx, key)
    phi = nk.fill_collocation(ctx, basis, grid.quad_points)
    gx, gy, gz = nk.fill_grad_collocation(ctx, basis, grid.quad_points)
    rho, grad, sig = nk.compute_density_and_sigma(ctx, phi, gx, gy, gz, [[1.0]], [g["occ"]])
    w = grid.weights
    z = g["zeta"]
    ref_sig = 1 / (2 * math.pi) if key == "h_density" else 2 * z ** 5 / math.pi
    assert abs((w * rho).sum().item() - g["int_rho"]) < g["rel"] * g["int_rho"]
    assert abs((w * sig).sum().item() - ref_sig) < g["rel"] * ref_sig


def test_gga_rejects_lda_functional(ctx):
    g, grid, basis = _single_atom(ctx, "h_lda_x")
    phi = nk.fill_collocation(ctx, basis, grid.quad_points)
    gx, gy, gz = nk.fill_grad_collocation(ctx, basis, grid.quad_points)
    with pytest.raises(nk.NukexcError) as e:
        nk.compute_gga(ctx, phi, gx, gy, gz, grid.weights, [[1.0]], [1.0], [1])
    assert e.value.status == "NKX_ERR_FUNCTIONAL_FAMILY"
    with pytest.raises(nk.NukexcError) as e:
        nk.compute_lda(ctx, phi, grid.weights, [[1.0]], [1.0], [101])
    assert e.value.status == "NKX_ERR_FUNCTIONAL_FAMILY"


def test_vxc_needs_basis():
    c = nk.Context(0)
    c.make_flat_grid([[0.0, 0, 0]], [1], nrad=10, nang_order=5)
    with pytest.raises(nk.NukexcError) as e:
        c.vxc(torch.ones((1, 1), dtype=torch.float64))
    assert e.value.status == "NKX_ERR_STATE"
    c.close()


# ----------------------------------------------------- batched V_xc (FP64) ----
def _workload_on_gpu_grid(ctx, name, **grid_kw):
    w = configs.workload(name)
    gkw = dict(w["grid"])
    gkw.update(grid_kw)
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], **gkw)
    ctx.set_basis(w["basis"])
    pts = grid.quad_points.cpu().numpy()
    wts = grid.weights.cpu().numpy()
    S, C, P = orbitals_on_grid(w, pts, wts)
    return w, pts, wts, S, C, P


@pytest.mark.parametrize("name,grid_kw", [("water", {}), ("benzene", dict(nrad=50, nang_order=29))])
def test_vxc_fp64_matches_oracle(ctx, name, grid_kw):
    w, pts, wts, S, C, P = _workload_on_gpu_grid(ctx, name, **grid_kw)
    gga = any(f > 100 for f in w["funcs"])
    phi, gx, gy, gz = oracle_on_grid(w["basis"], pts, grad=True)
    E0, N0, V0 = O.vxc_from_collocation(phi, wts, w["funcs"], P=P,
                                        gx=gx if gga else None, gy=gy if gga else None,
                                        gz=gz if gga else None)
    r = ctx.vxc(dev(P), funcs=w["funcs"], precision="fp64")
    V = r.potential.cpu().numpy()
    print(f"{name}: npts={len(wts)} E={r.energy:.12f} dE={r.energy - E0:.2e} "
          f"max|dV|={np.abs(V - V0).max():.2e} N_el={r.n_el:.12f}")
    assert abs(r.energy - E0) <= 1e-8
    assert np.abs(V - V0).max() <= 1e-9
    assert np.array_equal(V, V.T)
    # N_el = tr(P S_num) = 2 N_occ exactly for Loewdin orbitals in the same-grid S
    assert abs(r.n_el - 2 * w["nocc"]) < 1e-9
    assert abs(N0 - 2 * w["nocc"]) < 1e-9


def test_vxc_fp64_matches_reference_shaped_path(ctx):
    """nkx_vxc (batched, P route) == nkx_xc_from_collocation (compute_gga, MO
    route) on the same grid: two independent GPU pipelines."""
    w, pts, wts, S, C, P = _workload_on_gpu_grid(ctx, "benzene", nrad=30, nang_order=17)
    grid = ctx.grid
    phi = nk.fill_collocation(ctx, w["basis"], grid.quad_points)
    gx, gy, gz = nk.fill_grad_collocation(ctx, w["basis"], grid.quad_points)
    r1 = nk.compute_gga(ctx, phi, gx, gy, gz, grid.weights, C, 2.0 * np.ones(C.shape[1]),
                        w["funcs"])
    r2 = ctx.vxc(dev(P), funcs=w["funcs"])
    assert abs(r1.energy - r2.energy) < 1e-9
    assert (r1.potential - r2.potential).abs().max().item() < 1e-10


def test_overlap_matches_oracle(ctx):
    w, pts, wts, S, C, P = _workload_on_gpu_grid(ctx, "water", nrad=40, nang_order=17)
    Sg = ctx.overlap().cpu().numpy()
    assert np.abs(Sg - S).max() < 1e-12


# ------------------------------------------------------ batched V_xc (mixed) ----
@pytest.mark.parametrize("name,grid_kw", [("water", {}), ("benzene", dict(nrad=50, nang_order=29))])
def test_vxc_mixed_matches_oracle(ctx, name, grid_kw):
    """Mixed form (int8 slices with per-row exponents, exact integer MFMA,
    ten slice products, FP64 functional and accumulation) vs the FP64 oracle:
    |dE_xc| <= 1e-6 Eh, max|dV_xc| <= 1e-7."""
    w, pts, wts, S, C, P = _workload_on_gpu_grid(ctx, name, **grid_kw)
    gga = any(f > 100 for f in w["funcs"])
    phi, gx, gy, gz = oracle_on_grid(w["basis"], pts, grad=True)
    E0, N0, V0 = O.vxc_from_collocation(phi, wts, w["funcs"], P=P,
                                        gx=gx if gga else None, gy=gy if gga else None,
                                        gz=gz if gga else None)
    r = ctx.vxc(dev(P), funcs=w["funcs"], precision="mixed")
    V = r.potential.cpu().numpy()
    dE, dV, dN = r.energy - E0, np.abs(V - V0).max(), r.n_el - N0
    k = np.unravel_index(np.abs(V - V0).argmax(), V.shape)
    b = w["basis"]
    print(f"mixed {name}: dE={dE:.2e} max|dV|={dV:.2e} dN={dN:.2e} at V{k}={V0[k]:.4f} "
          f"zeta=({b.zeta[k[0]]:.2f},{b.zeta[k[1]]:.2f}) l=({b.l[k[0]]},{b.l[k[1]]})")
    assert abs(dE) <= 1e-6
    assert dV <= 1e-7
    assert abs(dN) <= 1e-6
    assert np.array_equal(V, V.T)


def test_vxc_mixed_no_density_bias(ctx):
    """Mixed vs FP64 on the GPU for an 8-water cluster (PBE, ~290 K points):
    the integer accumulation leaves no systematic density error, so N_el
    agrees to 1e-9 relative (the bf16-MFMA form was off by ~1e-8 relative,
    growing with the grid) and E_xc to 1e-7 Eh."""
    rng = np.random.Generator(np.random.PCG64(8))
    xyz, Z, _ = configs.water_cluster(8, rng)
    basis = nk.make_manual_basis(configs.sto_functions(Z, xyz, rng))
    grid = ctx.make_flat_grid(xyz, Z, nrad=50, nang_order=29)
    ctx.set_basis(basis)
    S = ctx.overlap().cpu().numpy()
    C, P = configs.loewdin_orbitals(S, 40, rng)
    r64 = ctx.vxc(dev(P), funcs=(101, 130), precision="fp64")
    rm = ctx.vxc(dev(P), funcs=(101, 130), precision="mixed")
    dV = (rm.potential - r64.potential).abs().max().item()
    print(f"8 waters, {grid.weights.shape[0]} points: mixed dE={rm.energy - r64.energy:.2e} "
          f"dN={rm.n_el - r64.n_el:.2e} max|dV|={dV:.2e}")
    assert abs(rm.n_el - r64.n_el) <= 1e-9 * r64.n_el
    assert abs(rm.energy - r64.energy) <= 1e-7
    assert dV <= 1e-7


def test_vxc_mixed_density_shapes(ctx):
    """The int8 density contraction's workgroup shapes on an 8-water cluster
    (several function blocks per batch): one or two batches per workgroup
    (65, 66) and the XCD-tiled absolute-pair forms (71, 72) give bitwise the
    same build (each batch's blocks in the same ascending order, the same
    column sums), within the mixed bar of the FP64 build.  FP64 group codes
    run the int8 form as 65, and out-of-range shapes are rejected by
    nkx_set_launch."""
    rng = np.random.Generator(np.random.PCG64(8))
    xyz, Z, _ = configs.water_cluster(8, rng)
    basis = nk.make_manual_basis(configs.sto_functions(Z, xyz, rng))
    ctx.make_flat_grid(xyz, Z, nrad=40, nang_order=23)
    ctx.set_basis(basis)
    S = ctx.overlap().cpu().numpy()
    C, P = configs.loewdin_orbitals(S, 40, rng)
    r64 = ctx.vxc(dev(P), funcs=(101, 130), precision="fp64")
    res = {}
    try:
        for shape in (65, 66, 71, 72, 2):
            ctx.set_launch(0, shape)
            assert ctx.get_launch() == (0, shape)
            r = ctx.vxc(dev(P), funcs=(101, 130), precision="mixed")
            res[shape] = (r.energy, r.n_el, r.potential.cpu().numpy())
            print(f"shape {shape}: E={r.energy:.12f} N={r.n_el:.12f}")
        for bad in (3, 17, 33, 67, 129):
            with pytest.raises(RuntimeError):
                ctx.set_launch(0, bad)
        with pytest.raises(RuntimeError):
            ctx.set_launch(-1, 0)
    finally:
        ctx.set_launch(0, 0)
    E, N, V = res[65]
    for shape in (66, 71, 72, 2):
        assert res[shape][0] == E and res[shape][1] == N, shape
        assert np.array_equal(res[shape][2], V), shape
    assert abs(E - r64.energy) <= 1e-6
    assert np.abs(V - r64.potential.cpu().numpy()).max() <= 1e-7


def test_mfma_i8_lane_layout():
    """v_mfma_i32_16x16x64_i8: lane l feeds A[l%16][16(l/16) + e] and
    B[16(l/16) + e][l%16] (e = 0..15, byte e of the lane's 16 bytes); result
    register r of lane l is D[4(l/16) + r][l%16].  Integer products and sums
    are exact."""
    probe = ctypes.CDLL(os.path.join(ROOT, "tests", "gpu_probe", "libmfma_probe.so"))
    i8p = ctypes.POINTER(ctypes.c_int8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    probe.mfma_probe_i8.argtypes = [i8p, i8p, i32p]
    rng = np.random.default_rng(3)
    A = rng.integers(-128, 128, size=(16, 64)).astype(np.int64)
    B = rng.integers(-128, 128, size=(64, 16)).astype(np.int64)
    lane = np.arange(64)
    a = np.zeros((64, 16), np.int8)
    b = np.zeros((64, 16), np.int8)
    for e in range(16):
        a[:, e] = A[lane % 16, 16 * (lane // 16) + e]
        b[:, e] = B[16 * (lane // 16) + e, lane % 16]
    d = np.zeros(256, np.int32)
    assert probe.mfma_probe_i8(a.ctypes.data_as(i8p), b.ctypes.data_as(i8p),
                               d.ctypes.data_as(i32p)) == 0
    D = A @ B
    got = np.zeros((16, 16), np.int64)
    for r in range(4):
        got[4 * (lane // 16) + r, lane % 16] = d.reshape(64, 4)[:, r]
    assert np.array_equal(got, D)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chunked_build_matches_single_chunk(precision):
    """A workspace cap small enough to split the build into many chunks of
    batches (collocation, density and V partials per chunk) gives the same
    E_xc, N_el and V as one chunk, up to the order of the per-chunk partial
    sums of V; the nuclear gradient likewise."""
    rng = np.random.Generator(np.random.PCG64(8))
    xyz, Z, _ = configs.water_cluster(8, rng)
    basis = nk.make_manual_basis(configs.sto_functions(Z, xyz, rng))
    c = nk.Context(0)
    try:
        c.make_flat_grid(xyz, Z, nrad=40, nang_order=23)
        c.set_basis(basis)
        S = c.overlap().cpu().numpy()
        C, P = configs.loewdin_orbitals(S, 40, rng)
        Pd = dev(P)
        one = c.vxc(Pd, funcs=(101, 130), precision=precision)
        E1, g1 = c.xc_grad(Pd, funcs=(101, 130), precision=precision)
        c.set_memory_budget(64 << 20)  # the minimum: tens of chunks here
        many = c.vxc(Pd, funcs=(101, 130), precision=precision)
        E2, g2 = c.xc_grad(Pd, funcs=(101, 130), precision=precision)
    finally:
        c.close()
    dV = (many.potential - one.potential).abs().max().item()
    print(f"{precision}: dE = {many.energy - one.energy:.1e}, dV = {dV:.1e}, "
          f"dg = {np.abs(g2 - g1).max():.1e}")
    assert abs(many.energy - one.energy) <= 1e-12 * abs(one.energy)
    assert abs(many.n_el - one.n_el) <= 1e-12 * one.n_el
    assert dV <= 1e-13
    assert abs(E2 - E1) <= 1e-12 * abs(E1)
    assert np.abs(g2 - g1).max() <= 1e-12
