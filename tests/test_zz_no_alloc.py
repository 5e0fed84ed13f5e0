"""nkx_becke_noalloc (the partition with its distance table in a fixed
per-GPU device array) against nkx_becke through the C ABI: bit for bit, on
two calls in a row, the second with fewer atoms and points, capped and
plain; and the size limit."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402
from nukexc_amd._capi import lib  # noqa: E402


def _raw(xyz, Z, nrad, order, stride):
    ang = nk.lebedev_rule(order)
    P, W, O = [], [], []
    for a, (c, z) in enumerate(zip(xyz, Z)):
        r, w = nk.radial_rule(0, nrad, nk.radial_scale(0, z), 0.6)
        P.append((r[:, None, None] * ang[None, :, :3]).reshape(-1, 3) + c)
        W.append((w[:, None] * ang[None, :, 3]).reshape(-1))
        O.append(np.full(len(r) * len(ang), a, np.int32))
    P, W, O = np.concatenate(P)[::stride], np.concatenate(W)[::stride], np.concatenate(O)[::stride]
    return P, W, O


def _call(fn, ctx, xyz, P, W, O, capped):
    d = lambda a, t=torch.float64: torch.tensor(np.ascontiguousarray(a), dtype=t, device="cuda")  # noqa: E731
    C, Pd, Od, Wd = d(xyz), d(P), d(O, torch.int32), d(W)
    rc = fn(ctx.h, len(xyz), ctypes.c_void_p(C.data_ptr()), len(W),
            ctypes.c_void_p(Pd.data_ptr()), ctypes.c_void_p(Od.data_ptr()),
            ctypes.c_void_p(Wd.data_ptr()), capped)
    assert rc == 0, lib().nkx_last_error(ctx.h)
    ctx.sync()
    return Wd.cpu().numpy()


@pytest.mark.parametrize("capped", [1, 0])
def test_becke_noalloc_matches_becke(capped):
    ctx = nk.Context(0)
    try:
        big = configs.workload("water100")      # 300 atoms
        small = configs.workload("water")       # 3 atoms
        for tag, w, nrad, order, stride in (("300 atoms", big, 40, 17, 5),
                                            ("3 atoms", small, 30, 17, 1)):
            P, W, O = _raw(w["xyz"], w["Z"], nrad, order, stride)
            print(f"{tag}: {len(W)} points, capped={capped}", flush=True)
            a = _call(lib().nkx_becke, ctx, w["xyz"], P, W, O, capped)
            b = _call(lib().nkx_becke_noalloc, ctx, w["xyz"], P, W, O, capped)
            assert np.array_equal(a.view(np.int64), b.view(np.int64))
            assert np.all(b >= 0) and np.any(b > 0)
    finally:
        ctx.close()


def test_becke_noalloc_rejects_too_many_atoms():
    ctx = nk.Context(0)
    try:
        xyz = torch.zeros((513, 3), dtype=torch.float64, device="cuda")
        rc = lib().nkx_becke_noalloc(ctx.h, 513, ctypes.c_void_p(xyz.data_ptr()), 0, None, None,
                                     None, 1)
        assert rc == 1 and b"at most 512 atoms" in lib().nkx_last_error(ctx.h)
    finally:
        ctx.close()
