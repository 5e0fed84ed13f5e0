"""Pruned grids on the GPU (nkx_make_flat_grid_pruned): the point counts of
the reference's pruning study on its water geometry, and the points,
weights and owners of the oracle's pruned grid."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from conftest import ROOT  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd.readers import read_xyz  # noqa: E402
from oracle import oracle as O  # noqa: E402
from test_oracle_pruning import ROWS  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = nk.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("row", [r for r in ROWS if int(r[1]) <= 60], ids=lambda r: "-".join(r[:3]))
def test_pruned_grid_matches_study_and_oracle(ctx, row):
    scheme, nrad, order, npts = row[0], int(row[1]), int(row[2]), int(row[3])
    xyz, Z = read_xyz(os.path.join(ROOT, "tests", "golden", "readers", "water.xyz"))
    g = ctx.make_flat_grid(xyz, Z, nrad=nrad, nang_order=order, pruning=scheme.lower()
                           if scheme != "Unpruned" else "none")
    print(f"{scheme} {nrad} x {order}: {g.weights.shape[0]} points (reference {npts})")
    assert g.weights.shape[0] == npts
    # the library's Lebedev tables (same rules as the oracle's, in the
    # library's point order), so the grids compare point by point
    full, low, med = (nk.lebedev_rule(o) for o in nk.pruning_orders(order, scheme))
    pr = {"Unpruned": 0, "Treutler": 1, "Robust": 2}[scheme]
    P0, W0, O0 = O.make_flat_grid(xyz, Z, nrad, full, pruning=pr, ang_low=low, ang_med=med)
    # same points in the same order (the oracle's host compiler may contract
    # centre + r u into an FMA, so to rounding rather than bit for bit)
    assert np.allclose(g.quad_points.cpu().numpy(), P0, rtol=0, atol=1e-12)
    assert np.array_equal(g.point_owner.cpu().numpy(), O0)
    # Becke products in another order on the device: weights to rounding
    assert np.allclose(g.weights.cpu().numpy(), W0, rtol=1e-10, atol=1e-14 * W0.max())


@pytest.mark.parametrize("order,npts_med", [(32, 302), (36, 434)])
def test_robust_middle_shells_use_base_order(ctx, order, npts_med):
    """nang_order 32 / 36 (tabulated bases 35 / 41): Robust's middle shells
    take order 29 / 35 rules; the grid equals the oracle's built from those
    rules, point for point."""
    xyz, Z = read_xyz(os.path.join(ROOT, "tests", "golden", "readers", "water.xyz"))
    g = ctx.make_flat_grid(xyz, Z, nrad=30, nang_order=order, pruning="robust")
    full, low, med = (nk.lebedev_rule(o) for o in nk.pruning_orders(order, "robust"))
    assert len(med) == npts_med
    P0, W0, O0 = O.make_flat_grid(xyz, Z, 30, full, pruning=2, ang_low=low, ang_med=med)
    own = g.point_owner.cpu().numpy()
    print("per-atom points: library", np.bincount(own, minlength=3), "oracle",
          np.bincount(O0, minlength=3))
    assert g.weights.shape[0] == len(W0)
    assert np.array_equal(own, O0)


def test_robust_negative_weight_order_raises(ctx):
    """nang_order 30 (base 31): Robust's middle rule is order 25, which has
    negative weights; the build stops with NKX_ERR_NEGATIVE_WEIGHT as the
    reference throws "Detected negative weights aborting" (grid.hpp:284-286)."""
    xyz, Z = read_xyz(os.path.join(ROOT, "tests", "golden", "readers", "water.xyz"))
    with pytest.raises(nk.NukexcError) as e:
        ctx.make_flat_grid(xyz, Z, nrad=30, nang_order=30, pruning="robust")
    assert e.value.code == 6 and "negative weights" in str(e.value)
