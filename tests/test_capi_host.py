"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/nukexc_hip.h declares, and its host-side tables (Lebedev
rules, radial rules, STO normalisation / cutoff, solid harmonics, functional
families, error codes) agree with independent sources."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT, lebedev
import nukexc_amd as nk
from nukexc_amd import _capi
from oracle import oracle as O

HEADER = os.path.join(ROOT, "include", "nukexc_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(nkx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = nk.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
        assert n in _capi.SIGNATURES, f"{n} has no ctypes signature"


def test_lebedev_tables_match_scipy():
    for order in (3, 17, 29, 41, 59, 101, 131):
        mine = nk.lebedev_rule(order)
        ref = lebedev(order)
        assert mine.shape == ref.shape
        a = {tuple(np.round(r[:3], 15)): r[3] for r in mine}
        b = {tuple(np.round(r[:3], 15)): r[3] for r in ref}
        assert a.keys() == b.keys()
        assert max(abs(a[k] - b[k]) for k in a) < 1e-15
        assert abs(mine[:, 3].sum() - 4 * math.pi) < 1e-12


def test_lebedev_next_order():
    n, act = ctypes.c_int(), ctypes.c_int()
    for order, npts in [(20, 170), (30, 350), (40, 590), (100, 3470), (29, 302)]:
        _capi.check(nk.lib().nkx_lebedev_npts(order, ctypes.byref(n), ctypes.byref(act)))
        assert n.value == npts
    assert nk.lib().nkx_lebedev_npts(132, ctypes.byref(n), ctypes.byref(act)) == 1


@pytest.mark.parametrize("scheme,Z", [(0, 1), (0, 6), (0, 8), (1, 1), (1, 2), (1, 8)])
def test_radial_rules_match_oracle(scheme, Z):
    R = nk.radial_scale(scheme, Z)
    assert R == pytest.approx(O.radial_scale(scheme, Z), rel=1e-15)
    r, w = nk.radial_rule(scheme, 75, R, 0.6)
    r0, w0 = O.radial(scheme, 75, R, 0.6)
    assert np.allclose(r, r0, rtol=1e-13, atol=0)
    assert np.allclose(w, w0, rtol=1e-12, atol=0)


def test_becke_slater_table_length():
    # grid.hpp:71-96 lists Z = 0..118; beyond that the scale defaults to 1.0
    assert nk.radial_scale(1, 118) == pytest.approx(0.5 * 1.75 * 1.8897261254578281)
    assert nk.radial_scale(1, 119) == 1.0


def test_sto_norm_cutoff_matches_oracle():
    for n in range(1, 6):
        for zeta in (0.3, 1.0, 2.7, 11.0):
            for tol in (1e-10, 1e-12):
                a = nk.sto_norm_cutoff(n, zeta, tol)
                b = O.sto_norm_cutoff(n, zeta, tol)
                assert a[0] == pytest.approx(b[0], rel=1e-14)
                assert a[1] == pytest.approx(b[1], rel=1e-9)
    with pytest.raises(nk.NukexcError):
        nk.sto_norm_cutoff(0, 1.0, 1e-10)


def test_solid_harmonics_recurrence_matches_oracle():
    """The library builds S_lm by recurrence; the oracle from the explicit
    Legendre-sum definition.  l <= 7, values and gradients."""
    rng = np.random.default_rng(11)
    for x, y, z in rng.normal(size=(6, 3)):
        for l in range(8):
            for m in range(-l, l + 1):
                a = nk.solid_harmonic(l, m, x, y, z)
                b = O.solid_harmonic(l, m, x, y, z)
                assert np.allclose(a, b, rtol=1e-12, atol=1e-13), (l, m)
    with pytest.raises(nk.NukexcError):
        nk.solid_harmonic(8, 0, 1, 1, 1)


def test_functional_info_and_errors():
    assert nk.functional_info(1) == (1, 0)
    assert nk.functional_info(101) == (2, 0)
    assert nk.functional_info(130) == (2, 1)
    with pytest.raises(nk.NukexcError) as e:
        nk.functional_info(999)
    assert e.value.status == "NKX_ERR_UNKNOWN_FUNCTIONAL"


def test_null_context_is_rejected():
    L = nk.lib()
    assert L.nkx_sync(None) == 1
    assert b"null" in L.nkx_last_error(None)


def test_function_atoms_map_and_rejection():
    """The function -> atom map of the geometry gradcheck binding: water's 25
    functions sit on O (15) and the two H (5 each); a basis moved off the
    atoms is rejected."""
    import dataclasses
    torch = pytest.importorskip("torch")  # noqa: F841 (the module imports torch)
    from nukexc_amd import configs
    from nukexc_amd.autograd import function_atoms
    w = configs.workload("water")
    fa = function_atoms(w["basis"], w["xyz"])
    assert list(np.bincount(fa, minlength=3)) == [15, 5, 5]
    moved = dataclasses.replace(w["basis"], O=w["basis"].O + 1e-6)
    with pytest.raises(ValueError):
        function_atoms(moved, w["xyz"])


def test_launch_points_reports_the_batch_size():
    """nkx_launch_points: grid points per density workgroup for a precision
    and density shape (the point-batch size the autotuner picks): FP64 1, 2,
    4 batches of 64 points (default 2), int8 65 / 71 / 72 one batch, 66 two."""
    from nukexc_amd.api import Context
    assert [Context.launch_points("fp64", s) for s in (0, 1, 2, 4)] == [128, 64, 128, 256]
    assert [Context.launch_points("mixed", s) for s in (0, 65, 66, 71, 72)] == [64, 64, 128, 64, 64]
