"""ctypes bindings of the FP64 CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Nothing in nukexc_amd/ imports
this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_radial.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _d, _d]
        L.orc_radial_scale.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_radial_scale.restype = ctypes.c_double
        L.orc_becke.argtypes = [ctypes.c_int, _d, ctypes.c_long, _d, _i, _d, ctypes.c_int]
        L.orc_make_flat_grid.argtypes = [ctypes.c_int, _d, _i, ctypes.c_int, ctypes.c_int, _d,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, _d, _d, _i]
        L.orc_make_flat_grid.restype = ctypes.c_long
        L.orc_make_flat_grid_pruned.argtypes = [
            ctypes.c_int, _d, _i, ctypes.c_int, _i, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _d, _d, _i]
        L.orc_make_flat_grid_pruned.restype = ctypes.c_long
        L.orc_sto_norm_cutoff.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _d, _d]
        L.orc_solid_harmonic.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _d]
        L.orc_collocation.argtypes = [ctypes.c_int, _i, _i, _i, _d, _d, _d, ctypes.c_long, _d,
                                      _d, _d, _d, _d]
        L.orc_density_P.argtypes = [ctypes.c_int, ctypes.c_long, _d, _d, _d, _d, _d, _d, _d, _d]
        L.orc_density_C.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _d, _d, _d, _d, _d,
                                    _d, _d, _d, _d]
        L.orc_xc_eval.argtypes = [ctypes.c_int, _i, ctypes.c_double, ctypes.c_long, _d, _d, _d, _d, _d]
        L.orc_xc_eval_polarized.argtypes = [ctypes.c_int, _i, ctypes.c_double, ctypes.c_long, _d,
                                             _d, _d, _d, _d]
        L.orc_vxc_polarized.argtypes = [ctypes.c_int, ctypes.c_long, _d, _d, _d, _d, _d, _d, _d,
                                        ctypes.c_int, _i, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_double), _d, _d, _d]
        L.orc_vxc_from_collocation.argtypes = [ctypes.c_int, ctypes.c_long, _d, _d, _d, _d, _d, _d,
                                               ctypes.c_int, _d, _d, ctypes.c_int, _i,
                                               ctypes.c_double, _d, _d, _d]
        L.orc_overlap.argtypes = [ctypes.c_int, ctypes.c_long, _d, _d, _d]
        _lib = L
    return _lib


def _p(a, t=_d):
    if a is None:
        return ctypes.cast(None, t)
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def num_threads():
    return lib().orc_num_threads()


def radial(scheme, n, R, alpha=0.6):
    r = np.empty(n)
    w = np.empty(n)
    if lib().orc_radial(scheme, n, R, alpha, _p(r), _p(w)):
        raise ValueError("bad radial scheme")
    return r, w


def radial_scale(scheme, Z):
    return lib().orc_radial_scale(scheme, Z)


def becke(centers, pts, owner, w, capped=True):
    centers = _f64(centers)
    pts = _f64(pts)
    owner = _i32(owner)
    w = _f64(w).copy()
    lib().orc_becke(len(centers), _p(centers), len(w), _p(pts), _p(owner, _i), _p(w), int(capped))
    return w


def make_flat_grid(xyz, Z, nrad, ang, scheme=0, alpha=0.6, w_thr=1e-30, capped=True,
                   pyscf_sizing=False, pruning=0, ang_low=None, ang_med=None):
    """ang: (nang, 4) array of x, y, z, weight (sum 4 pi).  pruning 1
    (Treutler) or 2 (Robust) with the low and medium rules ang_low, ang_med."""
    xyz = _f64(xyz)
    Z = _i32(Z)
    rules = [_f64(a) for a in (ang, ang if ang_low is None else ang_low,
                               ang if ang_med is None else ang_med)]
    nmax = len(Z) * int(nrad * (2.1 if pyscf_sizing else 1) + 1) * len(rules[0])
    P = np.empty((nmax, 3))
    W = np.empty(nmax)
    O = np.empty(nmax, dtype=np.int32)
    nangs = (ctypes.c_int * 3)(*[len(a) for a in rules])
    angs = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in rules])
    n = lib().orc_make_flat_grid_pruned(len(Z), _p(xyz), _p(Z, _i), nrad, nangs, angs,
                                        int(pruning), scheme, alpha, w_thr, int(capped),
                                        int(pyscf_sizing), _p(P), _p(W), _p(O, _i))
    if n < 0:
        raise RuntimeError("Detected negative weights aborting")
    return P[:n].copy(), W[:n].copy(), O[:n].copy()


def sto_norm_cutoff(n, zeta, tol=1e-10):
    a = ctypes.c_double()
    b = ctypes.c_double()
    lib().orc_sto_norm_cutoff(n, zeta, tol, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def solid_harmonic(l, m, x, y, z):
    out = np.empty(4)
    if lib().orc_solid_harmonic(l, m, x, y, z, _p(out)):
        raise ValueError("bad (l, m)")
    return out


def collocation(basis, pts, grad=True):
    """basis: dict with n, l, m, zeta, norm, centers (nbf, 3).  Returns
    phi (nbf, npts) and, if grad, gx, gy, gz of the same shape."""
    n, l, m = _i32(basis["n"]), _i32(basis["l"]), _i32(basis["m"])
    zeta, norm, O = _f64(basis["zeta"]), _f64(basis["norm"]), _f64(basis["centers"])
    pts = _f64(pts)
    nbf, npts = len(n), len(pts)
    phi = np.empty((nbf, npts))
    gx = np.empty((nbf, npts)) if grad else None
    gy = np.empty((nbf, npts)) if grad else None
    gz = np.empty((nbf, npts)) if grad else None
    lib().orc_collocation(nbf, _p(n, _i), _p(l, _i), _p(m, _i), _p(zeta), _p(norm), _p(O), npts,
                          _p(pts), _p(phi), _p(gx), _p(gy), _p(gz))
    return (phi, gx, gy, gz) if grad else phi


def density_P(phi, P, gx=None, gy=None, gz=None):
    nbf, npts = phi.shape
    rho = np.empty(npts)
    grad = np.empty((npts, 3)) if gx is not None else None
    sig = np.empty(npts) if gx is not None else None
    lib().orc_density_P(nbf, npts, _p(_f64(phi)), _p(gx), _p(gy), _p(gz), _p(_f64(P)), _p(rho),
                        _p(grad), _p(sig))
    return rho, grad, sig


def density_C(phi, C, f, gx=None, gy=None, gz=None):
    nbf, npts = phi.shape
    C = _f64(C)
    f = _f64(f)
    rho = np.empty(npts)
    grad = np.empty((npts, 3)) if gx is not None else None
    sig = np.empty(npts) if gx is not None else None
    lib().orc_density_C(nbf, C.shape[1], npts, _p(_f64(phi)), _p(gx), _p(gy), _p(gz), _p(C),
                        _p(f), _p(rho), _p(grad), _p(sig))
    return rho, grad, sig


def xc_eval(ids, rho, sigma=None, dens_thr=0.0):
    ids = _i32(ids)
    rho = _f64(rho)
    sigma = None if sigma is None else _f64(sigma)
    n = len(rho)
    exc, vr, vs = np.empty(n), np.empty(n), np.empty(n)
    if lib().orc_xc_eval(len(ids), _p(ids, _i), dens_thr, n, _p(rho), _p(sigma), _p(exc), _p(vr),
                         _p(vs)):
        raise ValueError(f"unknown functional id in {list(ids)}")
    return exc, vr, vs


def xc_eval_polarized(ids, rho2, sigma3=None, dens_thr=0.0):
    """rho2 (n, 2), sigma3 (n, 3) -> exc (n,), vrho (n, 2), vsigma (n, 3)."""
    ids = _i32(ids)
    rho2 = _f64(rho2)
    sigma3 = None if sigma3 is None else _f64(sigma3)
    n = rho2.shape[0]
    exc, vr, vs = np.empty(n), np.empty((n, 2)), np.empty((n, 3))
    if lib().orc_xc_eval_polarized(len(ids), _p(ids, _i), dens_thr, n, _p(rho2), _p(sigma3),
                                   _p(exc), _p(vr), _p(vs)):
        raise ValueError(f"unknown functional id in {list(ids)}")
    return exc, vr, vs


def vxc_from_collocation(phi, w, ids, P=None, C=None, f=None, gx=None, gy=None, gz=None,
                         dens_thr=0.0, chunk=None):
    """E_xc, N_el, V_xc from collocation data (compute_lda / compute_gga).
    Either P (nbf, nbf) or (C (nbf, nocc), f (nocc,)).  Columns (points) may be
    processed in chunks; results are accumulated in FP64."""
    nbf, npts = phi.shape
    ids = _i32(ids)
    E = ctypes.c_double(0.0)
    N = ctypes.c_double(0.0)
    V = np.zeros((nbf, nbf))
    C = None if C is None else _f64(C)
    f = None if f is None else _f64(f)
    P = None if P is None else _f64(P)
    nocc = 0 if C is None else C.shape[1]
    step = chunk or npts
    for s in range(0, npts, step):
        e = slice(s, min(npts, s + step))
        a = [_f64(x[:, e]) if x is not None else None for x in (phi, gx, gy, gz)]
        ww = _f64(w[e])
        rc = lib().orc_vxc_from_collocation(nbf, ww.shape[0], _p(a[0]), _p(a[1]), _p(a[2]),
                                            _p(a[3]), _p(ww), _p(P), nocc, _p(C), _p(f), len(ids),
                                            _p(ids, _i), dens_thr, ctypes.byref(E),
                                            ctypes.byref(N), _p(V))
        if rc:
            raise ValueError(f"oracle vxc failed (code {rc})")
    return E.value, N.value, V


def vxc_polarized(phi, w, ids, Pa, Pb, gx=None, gy=None, gz=None, dens_thr=0.0, chunk=None):
    """E_xc, (N_alpha, N_beta), V_alpha, V_beta from collocation data and spin
    density matrices (compute_lsda / compute_gga_lsda)."""
    nbf, npts = phi.shape
    ids = _i32(ids)
    E = ctypes.c_double(0.0)
    N = np.zeros(2)
    Va, Vb = np.zeros((nbf, nbf)), np.zeros((nbf, nbf))
    Pa, Pb = _f64(Pa), _f64(Pb)
    step = chunk or npts
    for s in range(0, npts, step):
        e = slice(s, min(npts, s + step))
        a = [_f64(x[:, e]) if x is not None else None for x in (phi, gx, gy, gz)]
        ww = _f64(w[e])
        rc = lib().orc_vxc_polarized(nbf, ww.shape[0], _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]),
                                     _p(ww), _p(Pa), _p(Pb), len(ids), _p(ids, _i), dens_thr,
                                     ctypes.byref(E), _p(N), _p(Va), _p(Vb))
        if rc:
            raise ValueError(f"oracle polarised vxc failed (code {rc})")
    return E.value, N, Va, Vb


def overlap(phi, w):
    nbf, npts = phi.shape
    S = np.zeros((nbf, nbf))
    lib().orc_overlap(nbf, npts, _p(_f64(phi)), _p(_f64(w)), _p(S))
    return S


# ---- screened-grid set-up (numpy restatement) ------------------------------
def _spread21(v):
    """Interleave the 21 low bits of v with two zero bits after each."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    out = np.zeros_like(v)
    for b in range(21):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def morton_codes(pts):
    """63-bit Morton code of each point (npts x 3): 21 bits per axis of the
    position in the points' bounding box, x in the highest bit of each
    triple.  Stands where ArborX's Morton64 curve orders the points in
    create_bounding_boxes (octree.hpp:88-106); ArborX is absent, so this
    curve is this build's own (same order within a 2^21 lattice up to the
    axis convention)."""
    pts = np.asarray(pts, dtype=np.float64)
    lo, hi = pts.min(0), pts.max(0)
    code = np.zeros(len(pts), dtype=np.uint64)
    for d in range(3):
        ext = max(hi[d] - lo[d], 1e-12)
        t = np.clip((pts[:, d] - lo[d]) / ext, 0.0, 1.0)
        code |= _spread21((t * 2097151.0).astype(np.uint64)) << np.uint64(2 - d)
    return code


def bounding_boxes(pts, per):
    """create_bounding_boxes (octree.hpp:71-138): stable Morton order of the
    points (returned as the permutation) and the (nboxes, 6) boxes (min xyz,
    max xyz) of each run of per consecutive sorted points."""
    pts = np.asarray(pts, dtype=np.float64)
    perm = np.argsort(morton_codes(pts), kind="stable")
    sp = pts[perm]
    nbox = (len(sp) + per - 1) // per
    boxes = np.empty((nbox, 6))
    for b in range(nbox):
        blk = sp[b * per:(b + 1) * per]
        boxes[b, :3] = blk.min(0)
        boxes[b, 3:] = blk.max(0)
    return perm, boxes


def neighbor_list(boxes, centers, radii):
    """build_neighbor_list (octree.hpp:317-389): for each box the ascending
    indices of the spheres (centre, radius) whose distance from the centre
    to the box is <= the radius; CSR (offsets, neighbors)."""
    boxes = np.asarray(boxes, dtype=np.float64).reshape(-1, 6)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(radii, dtype=np.float64)
    off, nbr = [0], []
    for b in boxes:
        e = np.maximum(np.maximum(b[:3] - c, c - b[3:]), 0.0)
        d = np.sqrt((e * e).sum(1))
        hit = np.nonzero(d <= r)[0]
        nbr.extend(hit.tolist())
        off.append(len(nbr))
    return np.array(off, dtype=np.int32), np.array(nbr, dtype=np.int32)


# ---- chunked dense build over BLAS (large parity cases) --------------------
def vxc_dense(basis, pts, w, ids, P, chunk=16384, want_density=False):
    """E_xc, N_el, V_xc of compute_gga / compute_lda (xc_integrals.hpp:424-541,
    :73-166) from a density matrix, evaluated chunk by chunk so that large
    grids fit in host memory: collocation and the functional from the C
    restatement above, the contractions as FP64 GEMMs as the reference's
    KokkosBlas calls (density.hpp:205-212, xc_integrals.hpp:520):
      X = P Phi,  rho = sum_mu Phi X,  grad rho = 2 sum_mu grad Phi X,
      Z = w (1/2 v_rho Phi + 2 v_sigma grad rho . grad Phi),  V = Phi Z^T + Z Phi^T.
    With want_density, also returns rho and sigma per point."""
    ids = _i32(ids)
    gga = any(lib().orc_is_gga(int(i)) for i in ids)
    P = _f64(P)
    pts = _f64(pts)
    w = _f64(w)
    nbf = P.shape[0]
    E = 0.0
    N = 0.0
    V = np.zeros((nbf, nbf))
    rho_all = np.empty(len(w)) if want_density else None
    sig_all = np.empty(len(w)) if want_density else None
    for s in range(0, len(w), chunk):
        e = slice(s, min(len(w), s + chunk))
        if gga:
            phi, gx, gy, gz = collocation(basis, pts[e], grad=True)
        else:
            phi = collocation(basis, pts[e], grad=False)
        X = P @ phi
        rho = np.einsum("ug,ug->g", phi, X)
        sig = None
        if gga:
            gr = 2.0 * np.stack([np.einsum("ug,ug->g", g, X) for g in (gx, gy, gz)])
            sig = np.einsum("kg,kg->g", gr, gr)
        exc, vr, vs = xc_eval(ids, rho, sig)
        ww = w[e]
        E += float(ww @ (rho * exc))
        N += float(ww @ rho)
        Z = (0.5 * ww * vr) * phi
        if gga:
            b = 2.0 * ww * vs
            Z += (b * gr[0]) * gx + (b * gr[1]) * gy + (b * gr[2]) * gz
        Vb = phi @ Z.T
        V += Vb + Vb.T
        if want_density:
            rho_all[e] = rho
            sig_all[e] = sig if gga else 0.0
    if want_density:
        return E, N, V, rho_all, sig_all
    return E, N, V
