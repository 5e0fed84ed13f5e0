"""HIP graph of one V_xc build (nkx_graph_capture / nkx_graph_launch):
every replay, with eager builds and density-matrix changes in between,
matches the eager build of the same P."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from gpu_util import orbitals_on_grid  # noqa: E402
import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402


@pytest.fixture(scope="module")
def setup():
    c = nk.Context(0)
    w = configs.workload("benzene")
    grid = c.make_flat_grid(w["xyz"], w["Z"], nrad=30, nang_order=17)
    c.set_basis(w["basis"])
    S, C, P1 = orbitals_on_grid(w, grid.quad_points.cpu().numpy(), grid.weights.cpu().numpy())
    C2, P2 = configs.loewdin_orbitals(S, w["nocc"], np.random.default_rng(9))
    yield c, w, P1, P2
    c.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(setup, precision):
    c, w, P1, P2 = setup
    nbf = P1.shape[0]
    Pd = torch.tensor(P1, device="cuda")
    V = torch.empty((nbf, nbf), dtype=torch.float64, device="cuda")
    c.graph_capture(Pd, V, funcs=w["funcs"], precision=precision)
    for it in range(5):
        Pn = P1 if it % 2 == 0 else P2
        Pd.copy_(torch.tensor(Pn, device="cuda"))
        torch.cuda.synchronize()
        E, N = c.graph_launch()
        Vg = V.clone()
        r = c.vxc(torch.tensor(Pn, device="cuda"), funcs=w["funcs"], precision=precision)
        dV = (Vg - r.potential).abs().max().item()
        print(f"{precision} replay {it}: dE={E - r.energy:.1e} dN={N - r.n_el:.1e} dV={dV:.1e}")
        assert abs(E - r.energy) <= 1e-12 * abs(r.energy)
        assert abs(N - r.n_el) <= 1e-12 * r.n_el
        assert dV <= 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_open_shell_graph_replays_match_eager(setup, precision):
    """nkx_graph_capture_polarized: every replay, with the spin density
    matrices changed in place and eager builds in between, matches the eager
    open-shell build bit for bit."""
    c, w, P1, P2 = setup
    nbf = P1.shape[0]
    Pa = torch.tensor(P1 / 2, device="cuda")
    Pb = torch.tensor(P2 / 2, device="cuda")
    Va = torch.empty((nbf, nbf), dtype=torch.float64, device="cuda")
    Vb = torch.empty_like(Va)
    c.graph_capture_polarized(Pa, Pb, Va, Vb, funcs=w["funcs"], precision=precision)
    for it in range(4):
        pa, pb = (P1 / 2, P2 / 2) if it % 2 == 0 else (P2 / 2, P1 / 2)
        Pa.copy_(torch.tensor(pa, device="cuda"))
        Pb.copy_(torch.tensor(pb, device="cuda"))
        torch.cuda.synchronize()
        E, (Na, Nb) = c.graph_launch()
        Vga, Vgb = Va.clone(), Vb.clone()
        r = c.vxc_polarized(torch.tensor(pa, device="cuda"), torch.tensor(pb, device="cuda"),
                            funcs=w["funcs"], precision=precision)
        print(f"{precision} open-shell replay {it}: dE={E - r.energy:.1e}")
        assert E == r.energy and Na == r.n_el[0] and Nb == r.n_el[1]
        assert torch.equal(Vga, r.potential_alpha) and torch.equal(Vgb, r.potential_beta)


def test_graph_invalidated_by_new_grid(setup):
    c, w, P1, P2 = setup
    nbf = P1.shape[0]
    Pd = torch.tensor(P1, device="cuda")
    V = torch.empty((nbf, nbf), dtype=torch.float64, device="cuda")
    c.graph_capture(Pd, V, funcs=w["funcs"])
    c.graph_launch()
    c.make_flat_grid(w["xyz"], w["Z"], nrad=30, nang_order=17)
    with pytest.raises(nk.NukexcError) as ei:
        c.graph_launch()
    assert ei.value.status == "NKX_ERR_STATE"


def test_autotune_keeps_results(setup):
    """The autotuner picks a split count among its candidates; builds after it
    agree with builds before it to summation-order rounding."""
    c, w, P1, P2 = setup
    c.make_flat_grid(w["xyz"], w["Z"], nrad=30, nang_order=17)
    r0 = c.vxc(torch.tensor(P1, device="cuda"), funcs=w["funcs"])
    k, g, ms = c.autotune(torch.tensor(P1, device="cuda"), funcs=w["funcs"])
    r1 = c.vxc(torch.tensor(P1, device="cuda"), funcs=w["funcs"])
    print(f"autotune: split {k}, group {g}, {ms:.2f} ms per build")
    assert k >= 1 and g in (1, 2, 4) and ms > 0
    assert abs(r1.energy - r0.energy) <= 1e-12 * abs(r0.energy)
    assert (r1.potential - r0.potential).abs().max().item() <= 1e-13
