"""Molecule sets: one launch sequence over many molecules of different sizes
(nkx_molset_create / nkx_molset_vxc) against the same molecules built one at
a time on the ctx grid and basis (two paths through the plan builder, the
per-molecule tables of every kernel and the per-molecule reductions)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import nukexc_amd as nk  # noqa: E402
from nukexc_amd import configs  # noqa: E402

GRID = dict(nrad=30, nang_order=17)


@pytest.fixture(scope="module")
def setup():
    c = nk.Context(0)
    mols, rng = configs.molset(64, 3)
    mols = mols[:10]
    Ps, singles = [], []
    for m in mols:
        grid = c.make_flat_grid(m["xyz"], m["Z"], **GRID)
        c.set_basis(m["basis"])
        S = c.overlap().cpu().numpy()
        C, P = configs.loewdin_orbitals(S, m["nocc"], rng)
        Ps.append(P)
        singles.append({p: c.vxc(torch.tensor(P, device="cuda"), funcs=(101, 130), precision=p)
                        for p in ("fp64", "mixed")})
    yield c, mols, Ps, singles
    c.close()


@pytest.mark.parametrize("precision,shape", [("fp64", 0), ("mixed", 0), ("mixed", 66),
                                             ("mixed", 71), ("mixed", 72)])
def test_molset_matches_single_builds(setup, precision, shape):
    """shape: the density workgroup shape of the set build (the single builds
    use the default); 66 puts two batches (possibly of two molecules) in one
    workgroup, 71 and 72 take absolute block pairs of each batch's
    molecule."""
    c, mols, Ps, singles = setup
    npts = c.molset_create(mols, **GRID)
    assert np.all(npts > 0)
    c.set_launch(0, shape)
    try:
        res = c.molset_vxc([torch.tensor(P, device="cuda") for P in Ps], funcs=(101, 130),
                           precision=precision)
    finally:
        c.set_launch(0, 0)
    worst = 0.0
    for k, (r, one) in enumerate(zip(res, singles)):
        ref = one[precision]
        dV = (r.potential - ref.potential).abs().max().item()
        worst = max(worst, dV)
        assert abs(r.energy - ref.energy) <= 1e-11 * max(1.0, abs(ref.energy)), k
        assert abs(r.n_el - ref.n_el) <= 1e-11 * ref.n_el, k
        assert dV <= 1e-12, (k, dV)
        assert abs(r.n_el - 2 * mols[k]["nocc"]) < 1e-6
    print(f"{precision} shape {shape}: {len(mols)} molecules, {npts.sum()} points, "
          f"worst |dV| {worst:.1e}")


def test_molset_rejects_split_bf16(setup):
    """The split-bf16 form builds one molecule at a time: NKX_ERR_UNSUPPORTED."""
    c, mols, Ps, singles = setup
    c.molset_create(mols[:2], **GRID)
    Pd = [torch.tensor(P, device="cuda") for P in Ps[:2]]
    Vs = [torch.empty_like(P) for P in Pd]
    Pp = (ctypes.c_void_p * 2)(*[x.data_ptr() for x in Pd])
    Vp = (ctypes.c_void_p * 2)(*[x.data_ptr() for x in Vs])
    E = (ctypes.c_double * 2)()
    ids = (ctypes.c_int * 1)(101)
    rc = nk.lib().nkx_molset_vxc(c.h, Pp, 1, ids, nk._capi.PREC_MIXED_BF16, E, None, Vp)
    assert nk._capi.STATUS_NAMES[rc] == "NKX_ERR_UNSUPPORTED"


def test_molset_graph_replays_match_eager(setup):
    """One int8 build of the whole set captured as a HIP graph: replays after
    P changes in place give the eager set build's E, N and V for every
    molecule, and a later molset_create invalidates the graph."""
    c, mols, Ps, singles = setup
    c.molset_create(mols, **GRID)
    c.set_launch(0, 0)
    Pd = [torch.tensor(P, device="cuda") for P in Ps]
    Vs = [torch.empty_like(P) for P in Pd]
    c.molset_graph_capture(Pd, Vs, funcs=(101, 130), precision="mixed")
    for rep in range(3):
        scale = 1.0 + 0.01 * rep
        for P, P0 in zip(Pd, Ps):
            P.copy_(torch.tensor(P0 * scale, device="cuda"))
        torch.cuda.synchronize()
        E, N = c.molset_graph_launch()
        eager = c.molset_vxc([P.clone() for P in Pd], funcs=(101, 130), precision="mixed")
        for k in range(len(mols)):
            assert E[k] == eager[k].energy and N[k] == eager[k].n_el, (rep, k)
            assert torch.equal(Vs[k], eager[k].potential), (rep, k)
        print(f"replay {rep}: {len(mols)} molecules, sum E = {sum(E):.10f}")
    with pytest.raises(RuntimeError):
        c.graph_launch()
    c.molset_create(mols[:2], **GRID)
    with pytest.raises(RuntimeError):
        c.molset_graph_launch()


def test_molset_autotune(setup):
    """nkx_molset_autotune times the candidates on set builds, keeps a valid
    pair, and the set build with it matches the single builds."""
    c, mols, Ps, singles = setup
    c.molset_create(mols, **GRID)
    Pd = [torch.tensor(P, device="cuda") for P in Ps]
    try:
        split, shape, ms = c.molset_autotune(Pd, funcs=(101, 130), precision="mixed")
        print(f"molset autotune: split {split}, shape {shape}, {ms:.2f} ms")
        assert split >= 1 and shape in (65, 71) and ms > 0
        assert c.get_launch() == (split, shape)
        res = c.molset_vxc(Pd, funcs=(101, 130), precision="mixed")
    finally:
        c.set_launch(0, 0)
    for r, one in zip(res, singles):
        ref = one["mixed"]
        assert abs(r.energy - ref.energy) <= 1e-11 * max(1.0, abs(ref.energy))
        assert (r.potential - ref.potential).abs().max().item() <= 1e-12
