#!/usr/bin/env python3
"""Throughput of the PBE V_xc build on one MI355X.

A step is one V_xc build -- collocation, density and density gradient,
on-device PBE, Z, V_xc, E_xc and N_el -- over the whole grid of the workload,
from a density matrix already resident in HBM.  Default workload: the
100-water cluster (TA 75 x Lebedev 302 per atom, 2500 STO functions, PBE).

Prints one JSON line.  The roofline object is built from HIP events around
each kernel launch inside the timed region (ctx stream) and the plan's
algorithmic flop count; cpu_baseline times the FP64 CPU restatement of the
reference's dense algorithm on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "i8": 5000.0, "bf16": 2500.0}
# The mixed form runs each tile product as ten int8 slice products on MFMA.
MIXED_OPS_PER_FLOP = 10
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs scaled by a plain copy kernel (written by hand from the
# profiles in profiles/; see the .md files there for the commands)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def kernel_name(precision, stage, shape):
    """Name of the kernel behind a timing stage for a density shape code."""
    if precision == "fp64":
        return {"density": "density_fused_f64", "vgemm": "vgemm_f64",
                "collocation": "colloc_tiles_f64"}[stage]
    if precision == "mixed_bf16":
        return {"density": "density_fused_mixed", "vgemm": "vgemm_mixed",
                "collocation": "colloc_tiles_mixed"}[stage]
    if stage == "density":
        return "density_x2_i8"
    return {"vgemm": "vgemm_i8", "collocation": "colloc_tiles_i8"}[stage]


def measured_traffic(workload, precision, split, shape, kernel):
    """(bytes per launch, source) of a profiled configuration, or (None, None).
    A density kernel must match the density shape, a V kernel the split count."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    with open(TRAFFIC_FILE) as f:
        entries = json.load(f)
    for e in entries:
        if e["workload"] != workload or e["precision"] != precision or kernel not in e["kernels"]:
            continue
        if kernel.startswith("density") and e["shape"] != shape:
            continue
        if kernel.startswith("vgemm") and e["split"] != split:
            continue
        return e["kernels"][kernel]["bytes_per_launch"], e["source"]
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup(name):
    import numpy as np
    import torch

    import nukexc_amd as nk
    from nukexc_amd import configs

    w = configs.workload(name)
    ctx = nk.Context(0)
    t0 = time.perf_counter()
    grid = ctx.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
    ctx.set_basis(w["basis"])
    log(f"{name}: {len(w['Z'])} atoms, {w['basis'].nbf()} functions, "
        f"{grid.weights.shape[0]} points (grid {time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    S = ctx.overlap().cpu().numpy()
    C, P = configs.loewdin_orbitals(S, w["nocc"], w["rng"])
    log(f"overlap + Loewdin orbitals {time.perf_counter() - t0:.1f} s; plan {ctx.plan_stats()}")
    Pd = torch.as_tensor(P, device="cuda")
    return w, ctx, grid, C, Pd


def cpu_baseline(w, grid, C, budget_s=15.0, chunk=16384):
    """FP64 CPU restatement of the reference dense algorithm (collocation,
    MO-route density with GEMMs, functional, Z, Phi Z^T), on chunks of a
    seeded random sample of the grid until about budget_s of work."""
    import numpy as np

    from oracle import oracle as O

    pts = grid.quad_points.cpu().numpy()
    wts = grid.weights.cpu().numpy()
    rng = np.random.default_rng(12345)
    order = rng.permutation(len(wts))
    b = w["basis"]
    basis = dict(n=b.n, l=b.l, m=b.m, zeta=b.zeta, norm=b.norm, centers=b.O)
    gga = any(f > 100 for f in w["funcs"])
    done, t_total = 0, 0.0
    while t_total < budget_s and done < len(order):
        idx = order[done:done + chunk]
        p, ww = pts[idx], wts[idx]
        t0 = time.perf_counter()
        phi, gx, gy, gz = O.collocation(basis, p, grad=True)
        psi = C.T @ phi
        rho = 2.0 * np.einsum("ig,ig->g", psi, psi)
        if gga:
            gr = np.stack([4.0 * np.einsum("ig,ig->g", psi, C.T @ g) for g in (gx, gy, gz)])
            sig = (gr * gr).sum(0)
        else:
            sig = None
        exc, vr, vs = O.xc_eval(list(w["funcs"]), rho, sig)
        Z = 0.5 * phi * vr
        if gga:
            Z += 2.0 * vs * (gr[0] * gx + gr[1] * gy + gr[2] * gz)
        Z *= ww
        Vb = phi @ Z.T
        _ = Vb + Vb.T
        _ = float(np.sum(ww * rho * exc))
        t_total += time.perf_counter() - t0
        done += len(idx)
    return done, t_total, O.num_threads()


def molset_main(args):
    """The 64-molecule batching workload: one build of the whole set (eager
    and as a replayed HIP graph) against 64 sequential single-molecule builds
    (default launch parameters, then each molecule autotuned; the set build
    autotuned once), per-molecule errors of the int8 form against the FP64
    set build."""
    import numpy as np
    import torch

    import nukexc_amd as nk
    from nukexc_amd import configs

    mols, rng = configs.molset(64, 3)
    grid_kw = dict(nrad=75, nang_order=29)
    funcs = (101, 130)
    order = np.argsort([len(m["Z"]) for m in mols], kind="stable")
    checked = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    grids = {}
    ctxs, Ps, npts = [], [], 0
    t0 = time.perf_counter()
    for k, m in enumerate(mols):
        c = nk.Context(0)
        c.set_memory_budget(1 << 30)  # 64 resident contexts
        g = c.make_flat_grid(m["xyz"], m["Z"], **grid_kw)
        c.set_basis(m["basis"])
        S = c.overlap().cpu().numpy()
        _, P = configs.loewdin_orbitals(S, m["nocc"], rng)
        Pd = torch.as_tensor(P, device="cuda")
        c.vxc(Pd, funcs, precision=args.precision)  # plan and workspace
        if k in checked:
            grids[k] = (g.quad_points.cpu().numpy(), g.weights.cpu().numpy(), P)
        ctxs.append(c)
        Ps.append(Pd)
        npts += g.weights.shape[0]
    log(f"molset64: {len(mols)} molecules, {sum(m['basis'].nbf() for m in mols)} functions, "
        f"{npts} points (set-up {time.perf_counter() - t0:.1f} s)")

    def timed(fn, reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t) / reps

    def sequential():
        for _ in range(args.warmup):
            for c, P in zip(ctxs, Ps):
                c.vxc(P, funcs, precision=args.precision)
        return timed(lambda: [c.vxc(P, funcs, precision=args.precision)
                              for c, P in zip(ctxs, Ps)], args.steps)

    seq_default_ms = sequential()
    for c, P in zip(ctxs, Ps):  # each molecule's own launch parameters
        c.autotune(P, funcs, precision=args.precision)
    seq_ms = sequential()
    for c in ctxs:  # the set build gets the device memory to itself
        c.close()
    torch.cuda.synchronize()
    cs = nk.Context(0)
    cs.molset_create(mols, **grid_kw)
    ref = cs.molset_vxc(Ps, funcs, precision="fp64")
    tuned = cs.molset_autotune(Ps, funcs, precision=args.precision)
    log(f"molset autotune: V split {tuned[0]}, density shape {tuned[1]}, {tuned[2]:.1f} ms")
    for _ in range(args.warmup):
        res = cs.molset_vxc(Ps, funcs, precision=args.precision)
    set_ms = timed(lambda: cs.molset_vxc(Ps, funcs, precision=args.precision), args.steps)
    res = cs.molset_vxc(Ps, funcs, precision=args.precision)
    Vs = [torch.empty_like(P) for P in Ps]
    cs.molset_graph_capture(Ps, Vs, funcs, precision=args.precision)
    for _ in range(args.warmup):
        cs.molset_graph_launch()
    graph_ms = timed(cs.molset_graph_launch, args.steps)
    E, N = cs.molset_graph_launch()
    dE = max(abs(E[k] - ref[k].energy) for k in range(len(mols)))
    dN = max(abs(N[k] - ref[k].n_el) for k in range(len(mols)))
    dV = max(float((Vs[k] - ref[k].potential).abs().max()) for k in range(len(mols)))
    log(f"64 sequential builds {seq_default_ms:.1f} ms (default launch), {seq_ms:.1f} ms "
        f"(each autotuned), one set build {set_ms:.1f} ms, "
        f"graph replay {graph_ms:.1f} ms; max |dE| {dE:.2e}, |dN| {dN:.2e}, |dV| {dV:.2e}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = molset_cpu_baseline(mols, grids, funcs, ref, res)
    out = {
        "metric": "V_xc builds/s for a 64-molecule PBE set on 1 GPU (one launch sequence vs 64)",
        "value": len(mols) / (graph_ms * 1e-3),
        "unit": "molecule builds/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": graph_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "i8 slices x4 (exact i32 MFMA) + f64"}[args.precision],
        "data": "synthetic (seeded geometries, STO exponents and orbitals)",
        "config": {"workload": f"molset64: 64 molecules of 3-60 atoms, {npts} points, PBE",
                   "precision": args.precision, "sequential_ms": seq_ms,
                   "sequential_default_launch_ms": seq_default_ms,
                   "autotuned": {"vgemm_split": tuned[0], "density_group": tuned[1],
                                 "points_per_density_workgroup":
                                     cs.launch_points(args.precision, tuned[1])},
                   "set_build_ms": set_ms, "graph_replay_ms": graph_ms,
                   "speedup_vs_sequential": seq_ms / graph_ms,
                   "max_dE_vs_fp64": dE, "max_dN_vs_fp64": dN, "max_dV_vs_fp64": dV},
    }
    if cpu is not None:
        out["config"]["vs_oracle"] = cpu["molecules"]
        out["cpu_baseline"] = cpu["baseline"]
    print(json.dumps(out), flush=True)
    return 0


def molset_cpu_baseline(mols, grids, funcs, ref, res):
    """CPU leg of the molecule-set run: the FP64 CPU restatement of the
    reference's dense build (oracle.vxc_dense) on the smallest, the median and
    the largest molecule of the set, timed as the CPU baseline, and the set
    build's per-molecule errors (FP64 and the run's precision) against it."""
    import numpy as np

    from oracle import oracle as O

    rows, pts_done, t_total = [], 0, 0.0
    for k, (pts, wts, P) in sorted(grids.items()):
        b = mols[k]["basis"]
        basis = dict(n=b.n, l=b.l, m=b.m, zeta=b.zeta, norm=b.norm, centers=b.O)
        t = time.perf_counter()
        E0, N0, V0 = O.vxc_dense(basis, pts, wts, funcs, P)
        t_total += time.perf_counter() - t
        pts_done += len(wts)
        row = {"molecule": k, "atoms": int(len(mols[k]["Z"])), "functions": int(b.nbf()),
               "points": int(len(wts))}
        for tag, r in (("fp64", ref[k]), ("run", res[k])):
            row[f"dE_{tag}"] = r.energy - E0
            row[f"max_dV_{tag}"] = float(np.abs(r.potential.cpu().numpy() - V0).max())
        log(f"molecule {k} ({row['atoms']} atoms, {row['functions']} functions, "
            f"{row['points']} points) vs oracle: FP64 dE {row['dE_fp64']:.1e}, dV "
            f"{row['max_dV_fp64']:.1e}; this run dE {row['dE_run']:.1e}, dV {row['max_dV_run']:.1e}")
        rows.append(row)
    return {"molecules": rows,
            "baseline": {"value": pts_done / t_total, "unit": "grid pts/s", "cores": O.num_threads(),
                         "kind": "port",
                         "sample": "the smallest, median and largest molecule of the set, whole "
                                   "grids, FP64 dense build (collocation, GEMM-route density, "
                                   "functional, Z, Phi Z^T)"}}


def grad_main(args):
    """The differentiation workload (mol30: 30 atoms, TA 75 x Lebedev 302,
    PBE, 90 components): XC nuclear gradients per second in FP64 and in the
    mixed form, each against central differences (h = 1e-4 bohr) of the
    FP64 E_xc on six coordinates (grid and basis rebuilt per displacement)."""
    import dataclasses

    import numpy as np
    import torch

    import nukexc_amd as nk
    from nukexc_amd import configs
    from nukexc_amd.autograd import function_atoms

    w = configs.workload("mol30")
    c = nk.Context(0)
    grid = c.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
    c.set_basis(w["basis"])
    S = c.overlap().cpu().numpy()
    _, P = configs.loewdin_orbitals(S, w["nocc"], w["rng"])
    Pt = torch.as_tensor(P, device="cuda")
    funcs = (101, 130)
    npts = grid.weights.shape[0]
    res = {}
    for prec in ("fp64", "mixed"):
        for _ in range(max(args.warmup, 1)):
            c.xc_grad(Pt, funcs=funcs, precision=prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            E, g = c.xc_grad(Pt, funcs=funcs, precision=prec)
        res[prec] = (1e3 * (time.perf_counter() - t) / args.steps, E, g)
    fa = function_atoms(w["basis"], w["xyz"])
    h = 1e-4
    coords = [(0, 0), (7, 1), (13, 2), (21, 0), (25, 1), (29, 2)]
    fd = []
    for a, k in coords:
        e = []
        for sgn in (1, -1):
            xyz = w["xyz"].copy()
            xyz[a, k] += sgn * h
            c.make_flat_grid(xyz, w["Z"], **w["grid"])
            c.set_basis(dataclasses.replace(w["basis"], O=np.ascontiguousarray(xyz[fa])))
            e.append(c.vxc(Pt, funcs=funcs, precision="fp64").energy)
        fd.append((e[0] - e[1]) / (2 * h))
    err = {p: max(abs(res[p][2][a, k] - f) for (a, k), f in zip(coords, fd)) for p in res}
    ms64, E64, g64 = res["fp64"]
    msm, Em, gm = res["mixed"]
    log(f"mol30: {len(w['Z'])} atoms, {w['basis'].nbf()} functions, {npts} points; gradient "
        f"FP64 {ms64:.1f} ms (vs FD {err['fp64']:.1e}), mixed {msm:.1f} ms (vs FD "
        f"{err['mixed']:.1e}, vs FP64 {np.abs(gm - g64).max():.1e})")
    out = {
        "metric": "XC nuclear gradients/s (30-atom PBE, 90 components) on 1 GPU",
        "value": 1e3 / msm, "unit": "gradients/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": msm,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i8 slices x4 (exact i32 MFMA) + f64",
        "data": "synthetic (seeded geometry, STO exponents and orbitals)",
        "config": {"workload": f"mol30: {len(w['Z'])} atoms, {w['basis'].nbf()} STO functions, "
                               f"{npts} points, PBE gradient",
                   "fp64_ms": ms64, "mixed_ms": msm, "speedup_vs_fp64": ms64 / msm,
                   "grid_points_per_s_mixed": npts / (msm * 1e-3),
                   "max_abs_g_minus_fd_fp64": err["fp64"],
                   "max_abs_g_minus_fd_mixed": err["mixed"],
                   "max_abs_g_mixed_minus_fp64": float(np.abs(gm - g64).max()),
                   "dE_mixed_vs_fp64": Em - E64,
                   "sum_over_atoms_mixed": float(np.abs(gm.sum(0)).max()),
                   "fd_step_bohr": h, "fd_coordinates": coords},
    }
    print(json.dumps(out), flush=True)
    c.close()
    return 0


def sparse_main(args):
    """The screened-grid set-up on the 100-water grid: Morton order of the
    points with the box of every 64 consecutive points (create_bounding_boxes)
    and the per-box list of basis functions whose cutoff sphere meets the box
    (build_neighbor_list), then the per-point density of the sorted grid
    (compute_density_and_sigma_sparse).  Grid points per second of the
    set-up; the oracle's numpy restatement on a bounded sample beside it."""
    import numpy as np
    import torch

    import nukexc_amd as nk
    from nukexc_amd import configs

    w = configs.workload("water100")
    c = nk.Context(0)
    c.make_flat_grid(w["xyz"], w["Z"], **w["grid"])  # warm-up
    t = time.perf_counter()
    grid = c.make_flat_grid(w["xyz"], w["Z"], **w["grid"])
    ms_grid = 1e3 * (time.perf_counter() - t)
    c.set_basis(w["basis"])
    npts, per = grid.weights.shape[0], 64
    S = c.overlap().cpu().numpy()
    C, P = configs.loewdin_orbitals(S, w["nocc"], w["rng"])
    Ct = torch.as_tensor(C, device="cuda")
    f = torch.full((C.shape[1],), 2.0, dtype=torch.float64, device="cuda")

    # a copy of the grid that the ctx does not hold, so the timed set-up is
    # the two C-ABI calls alone
    g2 = nk.FlatGrid(grid.quad_points.clone(), grid.weights.clone(), grid.atom_centers,
                     grid.Z, grid.point_owner.clone())

    def setup():
        boxes = nk.create_bounding_boxes(c, g2, per)
        return boxes, nk.build_neighbor_list(c, w["basis"], boxes, per, npts)

    for _ in range(max(args.warmup, 1)):
        boxes, nl = setup()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        boxes, nl = setup()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / args.steps
    nk.create_bounding_boxes(c, grid, per)  # the ctx's grid, now in the same order as g2
    nk.compute_density_and_sigma_sparse(c, w["basis"], grid, nl, Ct, f)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        rho, _, _ = nk.compute_density_and_sigma_sparse(c, w["basis"], grid, nl, Ct, f)
    torch.cuda.synchronize()
    ms_rho = 1e3 * (time.perf_counter() - t) / args.steps
    nel = float(grid.weights @ rho)
    pairs = int(nl.offsets[-1])
    nbox = boxes.shape[0]
    log(f"screened set-up: {npts} points, {nbox} boxes, {pairs} box-function pairs "
        f"({pairs / (nbox * w['basis'].nbf()):.3f} of all); set-up {ms:.2f} ms, sparse density "
        f"{ms_rho:.1f} ms, N_el {nel:.9f}; make_flat_grid {ms_grid:.1f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        pts = grid.quad_points.cpu().numpy()
        m = min(npts, 1 << 20)
        t = time.perf_counter()
        _, b0 = O.bounding_boxes(pts[:m], per)
        O.neighbor_list(b0, w["basis"].O, w["basis"].cutoff_radii)
        sec = time.perf_counter() - t
        cpu = {"value": m / sec, "unit": "grid points/s", "cores": 1, "kind": "port",
               "sample": f"oracle numpy set-up (Morton sort, boxes, neighbor list) of the "
                         f"first {m} points, {sec:.1f} s"}
    # the permutation is the minimal traffic: points, weights and owners read
    # and written once (36 + 36 bytes per point)
    gbs = 72.0 * npts / (ms * 1e-3) / 1e9
    out = {
        "metric": "screened-grid set-up, grid points/s (Morton order, boxes, neighbor list) "
                  "on 1 GPU",
        "value": npts / (ms * 1e-3), "unit": "grid points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 / u64 keys", "data": "synthetic (seeded geometry and STO exponents)",
        "config": {"workload": f"water100: {npts} points, {w['basis'].nbf()} functions, "
                               f"{per} points per box",
                   "boxes": nbox, "pairs": pairs, "sparse_density_ms": ms_rho, "n_el": nel,
                   "make_flat_grid_ms": ms_grid},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                     "frac": gbs / 8000.0, "traffic": None,
                     "note": "whole set-up (several kernels and two host round trips) "
                             "against the permutation's 72 B per point"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    c.close()
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water100")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed", "mixed_bf16"])
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the FP64 reference build of a mixed run (profiling)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (density batch group, V split count)")
    ap.add_argument("--launch", default=None, metavar="SPLIT,SHAPE",
                    help="fixed launch parameters (V split count, density shape) instead "
                         "of the autotuner")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="timed FP64 builds for the mixed-vs-FP64 speedup (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path "
              "(the per-GPU share of a sharded build is the single-GPU workload).",
              file=sys.stderr)
        return 2

    if args.workload == "molset64":
        return molset_main(args)
    if args.workload == "mol30grad":
        return grad_main(args)
    if args.workload == "sparse":
        return sparse_main(args)

    import numpy as np
    import torch

    if args.copy_calibration:
        # one plain 16-byte vector copy kernel (nkx_copy_d2d) over a buffer
        # twice the size of the MI355X last-level cache, after a fill of the
        # source, so the counters see MIB MiB read from and written to HBM
        import nukexc_amd as nk
        from nukexc_amd import _capi
        c = nk.Context(0)
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        _capi.check(_capi.lib().nkx_copy_d2d(c.h, ctypes.c_void_p(dst.data_ptr()),
                                             ctypes.c_void_p(src.data_ptr()),
                                             args.copy_calibration << 20), c.h)
        c.sync()
        log(f"copy calibration: nkx_copy_d2d of {args.copy_calibration} MiB")
        del src, dst
        c.close()

    w, ctx, grid, C, P = setup(args.workload)
    funcs = w["funcs"]
    npts = grid.weights.shape[0]

    ref, fp64_ms, fp64_tuned = None, None, None
    if args.precision != "fp64" and not args.no_reference:
        ref = ctx.vxc(P, funcs, precision="fp64")
        log(f"FP64 reference build: E_xc = {ref.energy:.10f}")
        if args.fp64_steps > 0:
            fp64_tuned = None
            if not args.no_autotune:  # the FP64 form at its own best launch parameters
                k, g, ms = ctx.autotune(P, funcs, precision="fp64")
                fp64_tuned = {"vgemm_split": k, "density_group": g,
                              "points_per_density_workgroup": ctx.launch_points("fp64", g)}
                log(f"FP64 autotune: V split {k}, density group {g}, {ms:.1f} ms/build")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.fp64_steps):
                ctx.vxc(P, funcs, precision="fp64")
            torch.cuda.synchronize()
            fp64_ms = 1e3 * (time.perf_counter() - t0) / args.fp64_steps
            log(f"FP64 form: {fp64_ms:.2f} ms/build over {args.fp64_steps} builds")
            ctx.set_launch(0, 0)

    tuned = None
    if args.launch:
        k, g = (int(x) for x in args.launch.split(","))
        ctx.set_launch(k, g)
        tuned = {"vgemm_split": k, "density_group": g, "ms": None}
        log(f"launch parameters: V split {k}, density shape {g}")
    elif not args.no_autotune:
        k, g, ms = ctx.autotune(P, funcs, precision=args.precision)
        tuned = {"vgemm_split": k, "density_group": g, "ms": ms,
                 "points_per_density_workgroup": ctx.launch_points(args.precision, g)}
        log(f"autotune: V split {k}, density group {g} "
            f"({tuned['points_per_density_workgroup']} points per workgroup), {ms:.1f} ms/build")

    for i in range(args.warmup):
        r = ctx.vxc(P, funcs, precision=args.precision)
        log(f"warmup {i}: E_xc = {r.energy:.10f} N_el = {r.n_el:.10f}")

    ctx.set_timing(True)
    ctx.reset_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ms = []
    for i in range(args.steps):
        ts = time.perf_counter()
        r = ctx.vxc(P, funcs, precision=args.precision)  # returns after E, N reach the host
        step_ms.append(1e3 * (time.perf_counter() - ts))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(False)
    tm = ctx.timing()
    work = ctx.plan_work()
    ms_step = 1e3 * elapsed / args.steps
    log(f"timed {args.steps} builds: {ms_step:.2f} ms/build; stages {tm}")

    # Dominant kernel and its roofline (MFMA-bound contractions).
    stage_flops = {"density": work["flops_density"], "vgemm": work["flops_v"]}
    dom = max(("density", "vgemm", "collocation"), key=lambda k: tm[k][0])
    dom_ms, dom_launches = tm[dom]
    split_used, shape_used = ctx.get_launch()
    dom_kernel = kernel_name(args.precision, dom, shape_used)
    if dom == "collocation":
        elem_bytes = (8 if args.precision == "fp64" else 4) * (4 if len(funcs) and max(funcs) > 100 else 1)
        traffic_alg = work["colloc_elems"] * elem_bytes
        achieved = traffic_alg * args.steps / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}
    else:
        peak = PEAK_TFLOPS[{"fp64": "f64", "mixed": "i8", "mixed_bf16": "bf16"}[args.precision]]
        achieved = stage_flops[dom] * args.steps / (dom_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
        if args.precision != "fp64":
            # algorithmic flops above; the MFMA executes ten int8 slice
            # products (six bf16 products) per tile product
            ops = MIXED_OPS_PER_FLOP if args.precision == "mixed" else 6
            roof["mfma_issue_frac"] = achieved * ops / peak
    traffic, source = measured_traffic(args.workload, args.precision, split_used, shape_used,
                                       dom_kernel)
    roof["traffic"] = traffic
    roof.update({"kernel": dom_kernel,
                 "avg_launch_ms": dom_ms / max(dom_launches, 1),
                 "launches": dom_launches,
                 "stages_ms_per_build": {k: v[0] / args.steps for k, v in tm.items()
                                         if k != "build" and v[1]}})
    if source:
        roof["traffic_source"] = source

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        n_cpu, t_cpu, cores = cpu_baseline(w, grid, C, budget_s=args.cpu_budget)
        cpu = {"value": n_cpu / t_cpu, "unit": "grid pts/s", "cores": cores, "kind": "port",
               "sample": f"{n_cpu} of {npts} grid points (seeded random sample, 16384-point "
                         f"chunks), full basis, FP64 dense algorithm: collocation, "
                         f"MO-route density by OpenBLAS GEMMs, functional, Z, Phi Z^T"}
        log(f"cpu baseline: {n_cpu} pts in {t_cpu:.1f} s on {cores} threads")

    out = {
        "metric": "grid pts/s for PBE Vxc build, mixed vs FP64 on 1 GPU at |dE_xc|<=1e-6 Eh",
        "value": npts / (ms_step * 1e-3),
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "i8 slices x4 (exact i32 MFMA) + f64",
                  "mixed_bf16": "bf16x3 split (f32 MFMA) + f64"}[args.precision],
        "data": "synthetic (seeded geometry, STO exponents and orbitals)",
        "config": {"workload": f"{w['name']}: {len(w['Z'])} atoms, {w['basis'].nbf()} STO "
                               f"functions, TA {w['grid']['nrad']} x Lebedev order "
                               f"{w['grid']['nang_order']}, {npts} points, "
                               f"{'PBE' if max(funcs) > 100 else 'LDA_X'}",
                   "precision": args.precision, "builds_per_s": 1e3 / ms_step,
                   "median_ms_per_build": float(np.median(step_ms)),
                   "autotuned": tuned,
                   "E_xc": r.energy, "N_el": r.n_el},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if ref is not None:
        out["config"]["dE_xc_vs_fp64"] = r.energy - ref.energy
        out["config"]["max_dV_vs_fp64"] = float((r.potential - ref.potential).abs().max())
        if fp64_ms is not None:
            out["config"]["fp64_ms_per_build"] = fp64_ms
            out["config"]["fp64_launch"] = fp64_tuned
            out["config"]["speedup_vs_fp64"] = fp64_ms / ms_step
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
