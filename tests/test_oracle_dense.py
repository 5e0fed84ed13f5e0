"""The oracle's chunked GEMM route (oracle.vxc_dense, used for the large
parity cases) against its per-point loops (orc_vxc_from_collocation) on the
water workload: LDA and PBE, several chunk sizes."""
import numpy as np
import pytest

from conftest import lebedev
from gpu_util import basis_dict
from nukexc_amd import configs
from oracle import oracle as O


@pytest.mark.parametrize("ids,chunk", [((1,), 997), ((101, 130), 4096), ((109, 134), 50000)])
def test_vxc_dense_matches_loops(ids, chunk):
    w = configs.workload("water")
    pts, wts, _ = O.make_flat_grid(w["xyz"], w["Z"], 30, lebedev(17))
    b = basis_dict(w["basis"])
    phi, gx, gy, gz = O.collocation(b, pts, grad=True)
    C, P = configs.loewdin_orbitals(O.overlap(phi, wts), w["nocc"], w["rng"])
    E0, N0, V0 = O.vxc_from_collocation(phi, wts, ids, P=P, gx=gx, gy=gy, gz=gz)
    E1, N1, V1, rho, sig = O.vxc_dense(b, pts, wts, ids, P, chunk=chunk, want_density=True)
    r0, g0, s0 = O.density_P(phi, P, gx, gy, gz)
    assert abs(E1 - E0) < 1e-12 and abs(N1 - N0) < 1e-12
    assert np.abs(V1 - V0).max() < 1e-12
    assert np.allclose(rho, r0, rtol=1e-12, atol=1e-14)
    if len(ids) > 1 or ids[0] > 100:
        assert np.allclose(sig, s0, rtol=1e-10, atol=1e-14)
