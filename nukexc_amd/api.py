"""Host-side mirror of the reference's nukexc/ API over the C ABI.

Names, argument meaning and errors follow the reference headers
(grid.hpp, partitioning.hpp, stobasis.hpp, density.hpp, xc_integrals.hpp);
device arrays are torch tensors on the ctx's GPU (torch is only the
allocator and stream provider here).  Reference errors raise NukexcError
whose status names the failure (e.g. NKX_ERR_FUNCTIONAL_FAMILY where the
reference throws std::runtime_error at xc_integrals.hpp:81-84).

Layout: the reference's LayoutLeft (N_bf, N_pts) collocation view is a torch
tensor of shape (N_bf, N_pts) with strides (1, N_bf) -- the transpose of a
contiguous (N_pts, N_bf) tensor -- so phi[mu, g] indexes like the reference.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import NukexcError, check, lib

try:  # torch provides device memory and the stream
    import torch
except ImportError:  # pragma: no cover
    torch = None

FUNCTIONALS = {"LDA_X": 1, "LDA_C_PW": 12, "LDA_C_PW_MOD": 13, "GGA_X_PBE": 101,
               "GGA_C_PBE": 130, "GGA_X_B88": 106, "GGA_C_LYP": 131, "LDA_C_VWN": 7,
               "GGA_X_PW91": 109, "GGA_C_PW91": 134}
PBE = (101, 130)
TA_M4 = 0.6
TA_M3 = 0.0


def _ids(funcs):
    if isinstance(funcs, (int, np.integer, str)):
        funcs = [funcs]
    out = [FUNCTIONALS[f] if isinstance(f, str) else int(f) for f in funcs]
    return (ctypes.c_int * len(out))(*out), len(out)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _dev_f64(x, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)


# ------------------------------------------------------------------ types ----
@dataclass
class STOBasisSet:
    """stobasis.hpp:54-64 (host arrays; uploaded by Context.set_basis)."""
    n: np.ndarray
    l: np.ndarray
    m: np.ndarray
    norm: np.ndarray
    zeta: np.ndarray
    O: np.ndarray  # (nbf, 3) centres
    cutoff_radii: np.ndarray

    def nbf(self):
        return len(self.n)


@dataclass
class FlatGrid:
    """grid.hpp:150-160: device arrays."""
    quad_points: "torch.Tensor"  # (npts, 3)
    weights: "torch.Tensor"
    atom_centers: "torch.Tensor"  # (natoms, 3)
    Z: np.ndarray
    point_owner: "torch.Tensor"


@dataclass
class XC_result:
    """xc_integrals.hpp:62-65."""
    energy: float
    potential: "torch.Tensor"  # (N_bf, N_bf)
    n_el: float = float("nan")


@dataclass
class XC_result_polarized:
    """xc_integrals.hpp:67-71; n_el holds (N_alpha, N_beta) from the batched build."""
    energy: float
    potential_alpha: "torch.Tensor"  # (N_bf, N_bf)
    potential_beta: "torch.Tensor"
    n_el: tuple = (float("nan"), float("nan"))


def sto_norm_cutoff(n, zeta, tol=1e-10):
    a, b = ctypes.c_double(), ctypes.c_double()
    check(lib().nkx_sto_norm_cutoff(int(n), float(zeta), float(tol), ctypes.byref(a),
                                    ctypes.byref(b)))
    return a.value, b.value


def make_manual_basis(funcs, cutoff_tol=1e-10):
    """make_manual_basis (stobasis.hpp:175-225).  funcs: iterable of
    (n, l, m, zeta, ox, oy, oz)."""
    funcs = list(funcs)
    n = np.array([f[0] for f in funcs], dtype=np.int32)
    l = np.array([f[1] for f in funcs], dtype=np.int32)
    m = np.array([f[2] for f in funcs], dtype=np.int32)
    zeta = np.array([f[3] for f in funcs], dtype=np.float64)
    O = np.array([[f[4], f[5], f[6]] for f in funcs], dtype=np.float64).reshape(-1, 3)
    nc = [sto_norm_cutoff(a, z, cutoff_tol) for a, z in zip(n, zeta)]
    norm = np.array([x[0] for x in nc])
    cut = np.array([x[1] for x in nc])
    return STOBasisSet(n, l, m, norm, zeta, O, cut)


def lebedev_rule(order):
    npts, actual = ctypes.c_int(), ctypes.c_int()
    check(lib().nkx_lebedev_npts(int(order), ctypes.byref(npts), ctypes.byref(actual)))
    out = np.empty((npts.value, 4))
    check(lib().nkx_lebedev_rule(int(order), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def pruning_orders(nang_order, pruning="none"):
    """Lebedev orders (full, inner shells, middle shells) that a grid of this
    nang_order uses under the pruning scheme ("none", "treutler", "robust")."""
    pr = {"none": 0, "unpruned": 0, "treutler": 1, "robust": 2}[str(pruning).lower()]
    out = (ctypes.c_int * 3)()
    check(lib().nkx_pruning_orders(int(nang_order), pr, out))
    return tuple(out)


def radial_rule(scheme, n, R, alpha=TA_M4):
    r, w = np.empty(n), np.empty(n)
    pd = ctypes.POINTER(ctypes.c_double)
    check(lib().nkx_radial_rule(int(scheme), int(n), float(R), float(alpha),
                                r.ctypes.data_as(pd), w.ctypes.data_as(pd)))
    return r, w


def radial_scale(scheme, Z):
    R = ctypes.c_double()
    check(lib().nkx_radial_scale(int(scheme), int(Z), ctypes.byref(R)))
    return R.value


def solid_harmonic(l, m, x, y, z):
    """S_lm = r^l Y_lm and gradient (spherical_harmonics.hpp:750-1608)."""
    out = np.empty(4)
    check(lib().nkx_solid_harmonic(int(l), int(m), float(x), float(y), float(z),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def functional_info(fid):
    fam, kind = ctypes.c_int(), ctypes.c_int()
    check(lib().nkx_functional_info(int(fid), ctypes.byref(fam), ctypes.byref(kind)))
    return fam.value, kind.value


# ---------------------------------------------------------------- context ----
class Context:
    """An nkx_ctx bound to one GPU and torch's current stream on it."""

    def __init__(self, device=0):
        if torch is None or not torch.cuda.is_available():
            raise RuntimeError("nukexc_amd needs a ROCm GPU (torch.cuda is unavailable)")
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        h = ctypes.c_void_p()
        # stream 0 (torch's legacy default stream): the ctx makes a blocking
        # stream of its own, which is ordered with it.
        check(lib().nkx_create(device, ctypes.c_void_p(stream) if stream else None,
                               ctypes.byref(h)))
        self.h = h
        self.basis = None
        self.grid = None

    def close(self):
        if getattr(self, "h", None):
            lib().nkx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return check(rc, self.h)

    def _done(self, rc):
        """Check a call that leaves device outputs and wait for them."""
        check(rc, self.h)
        self.sync()

    def sync(self):
        self._chk(lib().nkx_sync(self.h))

    # -- argument checks: the C entry points take device pointers without
    # sizes, so shapes, dtypes and devices are checked here before any call
    def _nbf(self):
        n = ctypes.c_int()
        self._chk(lib().nkx_basis_nbf(self.h, ctypes.byref(n)))
        if n.value == 0:
            raise NukexcError(8, "no basis set on this context (set_basis first)")
        return n.value

    def _in_matrix(self, X, nbf, name):
        """A density matrix argument as a contiguous FP64 tensor on the ctx
        device, which must be nbf x nbf."""
        X = _dev_f64(X, self.device).contiguous()
        if tuple(X.shape) != (nbf, nbf):
            raise NukexcError(1, f"{name} must be {nbf} x {nbf} (the basis size), "
                                 f"got {tuple(X.shape)}")
        return X

    def _check_out(self, V, nbf, name):
        """A caller-supplied output matrix: nbf x nbf, float64, contiguous, on
        the ctx device (it is written through its device pointer)."""
        if not isinstance(V, torch.Tensor):
            raise NukexcError(1, f"{name} must be a torch tensor on {self.device}")
        if tuple(V.shape) != (nbf, nbf) or V.dtype != torch.float64:
            raise NukexcError(1, f"{name} must be a float64 {nbf} x {nbf} tensor, got "
                                 f"{V.dtype} {tuple(V.shape)}")
        if V.device != self.device or not V.is_contiguous():
            raise NukexcError(1, f"{name} must be contiguous on {self.device}")
        return V

    # -- grid -----------------------------------------------------------------
    def make_flat_grid(self, xyz, Z, nrad=50, nang_order=30, weight_threshold=1e-30,
                       ta_alpha=TA_M4, radial="ta", sizing="uniform", pruning="none"):
        """make_flat_grid<TA|Becke, LebedevLaikov> (grid.hpp:172-348); pruning
        "none", "treutler" or "robust" (grid.hpp:164-166)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        Zc = np.ascontiguousarray(Z, dtype=np.int32)
        n = ctypes.c_int64()
        scheme = {"ta": _capi.RADIAL_TA, "becke": _capi.RADIAL_BECKE}[radial]
        self._chk(lib().nkx_make_flat_grid_pruned(
            self.h, len(Zc), xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            Zc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(nrad), int(nang_order), scheme,
            float(ta_alpha), float(weight_threshold), {"uniform": 0, "pyscf": 1}[sizing],
            {"none": 0, "treutler": 1, "robust": 2}[pruning], ctypes.byref(n)))
        npts = n.value
        pts = torch.empty((npts, 3), dtype=torch.float64, device=self.device)
        w = torch.empty(npts, dtype=torch.float64, device=self.device)
        own = torch.empty(npts, dtype=torch.int32, device=self.device)
        self._chk(lib().nkx_grid_export(self.h, _ptr(pts), _ptr(w), _ptr(own)))
        self.sync()
        self.grid = FlatGrid(pts, w, torch.as_tensor(xyz, device=self.device), Zc, own)
        return self.grid

    def set_grid(self, grid):
        pts = grid.quad_points.contiguous()
        w = grid.weights.contiguous()
        own = grid.point_owner.to(torch.int32).contiguous()
        ac = grid.atom_centers.to(torch.float64).contiguous()
        self._done(lib().nkx_set_grid(self.h, ac.shape[0], _ptr(ac), pts.shape[0], _ptr(pts),
                                     _ptr(w), _ptr(own)))
        self.grid = grid

    # -- basis ----------------------------------------------------------------
    def set_basis(self, basis):
        pi = ctypes.POINTER(ctypes.c_int)
        pd = ctypes.POINTER(ctypes.c_double)
        n = np.ascontiguousarray(basis.n, np.int32)
        l = np.ascontiguousarray(basis.l, np.int32)
        m = np.ascontiguousarray(basis.m, np.int32)
        z = np.ascontiguousarray(basis.zeta, np.float64)
        nr = np.ascontiguousarray(basis.norm, np.float64)
        O = np.ascontiguousarray(basis.O, np.float64)
        cut = np.ascontiguousarray(basis.cutoff_radii, np.float64)
        self._chk(lib().nkx_set_basis(self.h, len(n), n.ctypes.data_as(pi), l.ctypes.data_as(pi),
                                      m.ctypes.data_as(pi), z.ctypes.data_as(pd),
                                      nr.ctypes.data_as(pd), O.ctypes.data_as(pd),
                                      cut.ctypes.data_as(pd)))
        self.basis = basis

    def set_dens_threshold(self, thr):
        self._chk(lib().nkx_set_dens_threshold(self.h, float(thr)))

    def set_screening(self, tol, batch_points=0):
        self._chk(lib().nkx_set_screening(self.h, float(tol), int(batch_points)))

    def plan_stats(self):
        v = [ctypes.c_int64() for _ in range(5)]
        self._chk(lib().nkx_plan_stats(self.h, *[ctypes.byref(x) for x in v]))
        keys = ("npts", "npts_padded", "nbatches", "pairs_kept", "pairs_total")
        return dict(zip(keys, [x.value for x in v]))

    # -- hot path -------------------------------------------------------------
    def vxc(self, P, funcs=PBE, precision="fp64", out=None):
        """One V_xc build on the ctx grid and basis from a density matrix P.
        Returns XC_result(energy, potential, n_el)."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED,
                "mixed_bf16": _capi.PREC_MIXED_BF16}[precision]
        nbf = self._nbf()
        P = self._in_matrix(P, nbf, "P")
        V = (self._check_out(out, nbf, "out") if out is not None else
             torch.empty((nbf, nbf), dtype=torch.float64, device=self.device))
        ids, nf = _ids(funcs)
        E, N = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().nkx_vxc(self.h, _ptr(P), nf, ids, prec, ctypes.byref(E), ctypes.byref(N),
                                _ptr(V)))
        return XC_result(E.value, V, N.value)

    def vxc_polarized(self, Pa, Pb, funcs=PBE, precision="fp64"):
        """Spin-polarised build from the spin density matrices (FP64 or the
        int8 mixed form).  Returns XC_result_polarized(energy, V_alpha,
        V_beta, (N_alpha, N_beta))."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED,
                "mixed_bf16": _capi.PREC_MIXED_BF16}[precision]
        nbf = self._nbf()
        Pa, Pb = self._in_matrix(Pa, nbf, "Pa"), self._in_matrix(Pb, nbf, "Pb")
        Va, Vb = (torch.empty((nbf, nbf), dtype=torch.float64, device=self.device)
                  for _ in range(2))
        ids, nf = _ids(funcs)
        E = ctypes.c_double()
        N = (ctypes.c_double * 2)()
        self._chk(lib().nkx_vxc_polarized(self.h, _ptr(Pa), _ptr(Pb), nf, ids, prec,
                                          ctypes.byref(E), N, _ptr(Va), _ptr(Vb)))
        return XC_result_polarized(E.value, Va, Vb, (N[0], N[1]))

    PARTITION = {"fixed": 0, "becke_capped": 1, "becke": 2}

    def xc_grad(self, P, funcs=PBE, partition="becke_capped", precision="fp64"):
        """Nuclear gradient of E_xc at fixed P on the ctx grid (points move
        with their owner atom, weights from the named partition), in FP64 or
        in the mixed form (int8-slice contractions, precision="mixed").
        Returns (E_xc, grad ndarray (natoms, 3))."""
        fn = {"fp64": "nkx_xc_grad", "mixed": "nkx_xc_grad_mixed"}[precision]
        P = self._in_matrix(P, self._nbf(), "P")
        natoms = self.grid.atom_centers.shape[0]
        g = np.zeros(3 * natoms)
        ids, nf = _ids(funcs)
        E = ctypes.c_double()
        self._chk(getattr(lib(), fn)(self.h, _ptr(P), nf, ids, self.PARTITION[partition],
                                     ctypes.byref(E),
                                     g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return E.value, g.reshape(natoms, 3)

    def xc_grad_polarized(self, Pa, Pb, funcs=PBE, partition="becke_capped", precision="fp64"):
        """Open-shell nuclear gradient of E_xc[P_alpha, P_beta] at fixed
        spin density matrices, in FP64 or the mixed form.  Returns (E_xc,
        grad ndarray (natoms, 3))."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED}[precision]
        nbf = self._nbf()
        Pa, Pb = self._in_matrix(Pa, nbf, "Pa"), self._in_matrix(Pb, nbf, "Pb")
        natoms = self.grid.atom_centers.shape[0]
        g = np.zeros(3 * natoms)
        ids, nf = _ids(funcs)
        E = ctypes.c_double()
        self._chk(lib().nkx_xc_grad_polarized(
            self.h, _ptr(Pa), _ptr(Pb), nf, ids, self.PARTITION[partition], prec,
            ctypes.byref(E), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return E.value, g.reshape(natoms, 3)

    def autotune(self, P, funcs=PBE, precision="fp64"):
        """Pick the fastest density workgroup shape and V-contraction split
        count for this workload; returns (split, shape, ms per build)."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED,
                "mixed_bf16": _capi.PREC_MIXED_BF16}[precision]
        P = self._in_matrix(P, self._nbf(), "P")
        V = torch.empty_like(P)
        ids, nf = _ids(funcs)
        k, g, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._chk(lib().nkx_autotune(self.h, _ptr(P), nf, ids, prec, _ptr(V), ctypes.byref(k),
                                     ctypes.byref(g), ctypes.byref(ms)))
        return k.value, g.value, ms.value

    def set_launch(self, vgemm_split=0, density_shape=0):
        """Set the launch parameters the autotuner picks (0 = automatic /
        default); density shapes 1, 2, 4 (FP64) or 65, 66, 71, 72 (int8
        form)."""
        self._chk(lib().nkx_set_launch(self.h, int(vgemm_split), int(density_shape)))

    def set_memory_budget(self, nbytes=0):
        """Tile-store bytes per build (0: automatic); larger builds run in
        chunks of batches."""
        self._chk(lib().nkx_set_memory_budget(self.h, int(nbytes)))

    def get_launch(self):
        """Current (vgemm_split, density_shape)."""
        k, g = ctypes.c_int(), ctypes.c_int()
        self._chk(lib().nkx_get_launch(self.h, ctypes.byref(k), ctypes.byref(g)))
        return k.value, g.value

    @staticmethod
    def launch_points(precision, density_shape):
        """Grid points per density workgroup (the point-batch size) of a
        build at this precision and density shape."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED,
                "mixed_bf16": _capi.PREC_MIXED_BF16}[precision]
        n = ctypes.c_int()
        check(lib().nkx_launch_points(prec, int(density_shape), ctypes.byref(n)))
        return n.value

    # -- HIP graph of one build ----------------------------------------------
    def graph_capture(self, P, V, funcs=PBE, precision="fp64"):
        """Capture one build for the device tensors P (input) and V (output);
        replay with graph_launch() after changing P in place."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED,
                "mixed_bf16": _capi.PREC_MIXED_BF16}[precision]
        nbf = self._nbf()
        self._check_out(P, nbf, "P")
        self._check_out(V, nbf, "V")
        ids, nf = _ids(funcs)
        self._graph_keep = (P, V)
        self._chk(lib().nkx_graph_capture(self.h, _ptr(P), nf, ids, prec, _ptr(V)))
        self._graph_nmol = 1
        self._graph_nspin = 1

    def graph_capture_polarized(self, Pa, Pb, Va, Vb, funcs=PBE, precision="fp64"):
        """Capture one open-shell build for the device tensors Pa, Pb (inputs)
        and Va, Vb (outputs); graph_launch() then returns (E_xc, (N_a, N_b))."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED}[precision]
        nbf = self._nbf()
        for t, name in ((Pa, "Pa"), (Pb, "Pb"), (Va, "Va"), (Vb, "Vb")):
            self._check_out(t, nbf, name)
        ids, nf = _ids(funcs)
        self._graph_keep = (Pa, Pb, Va, Vb)
        self._chk(lib().nkx_graph_capture_polarized(self.h, _ptr(Pa), _ptr(Pb), nf, ids, prec,
                                                    _ptr(Va), _ptr(Vb)))
        self._graph_nmol = 1
        self._graph_nspin = 2

    def graph_launch(self):
        """Replay the captured build; returns (E_xc, N_el), or (E_xc,
        (N_alpha, N_beta)) for an open-shell capture."""
        if getattr(self, "_graph_nmol", 1) != 1:
            raise RuntimeError("the captured build is a molecule set: use molset_graph_launch")
        E = ctypes.c_double()
        N = (ctypes.c_double * 2)()
        self._chk(lib().nkx_graph_launch(self.h, ctypes.byref(E), N))
        if getattr(self, "_graph_nspin", 1) == 2:
            return E.value, (N[0], N[1])
        return E.value, N[0]

    # -- molecule sets ---------------------------------------------------------
    def molset_create(self, mols, nrad=75, nang_order=29, weight_threshold=1e-30,
                      ta_alpha=TA_M4, radial="ta", sizing="uniform"):
        """Grids and bases of many molecules for one-launch builds.  mols: list
        of dicts with xyz (natoms, 3), Z, basis (STOBasisSet).  Returns the
        grid size of each molecule."""
        nmol = len(mols)
        natoms = np.array([len(m["Z"]) for m in mols], np.int32)
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(m["xyz"], float).reshape(-1, 3)
                                                   for m in mols]))
        Z = np.ascontiguousarray(np.concatenate([np.asarray(m["Z"], np.int32) for m in mols]))
        bs = [m["basis"] for m in mols]
        nbf = np.array([b.nbf() for b in bs], np.int32)
        cat = lambda f, t: np.ascontiguousarray(np.concatenate([np.asarray(f(b), t) for b in bs]))
        n, l, mm = cat(lambda b: b.n, np.int32), cat(lambda b: b.l, np.int32), cat(lambda b: b.m, np.int32)
        zeta, norm = cat(lambda b: b.zeta, np.float64), cat(lambda b: b.norm, np.float64)
        ctr = cat(lambda b: np.asarray(b.O).reshape(-1, 3), np.float64)
        npts = np.zeros(nmol, np.int64)
        scheme = {"ta": _capi.RADIAL_TA, "becke": _capi.RADIAL_BECKE}[radial]
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._chk(lib().nkx_molset_create(
            self.h, nmol, ip(natoms), dp(xyz), ip(Z), int(nrad), int(nang_order), scheme,
            float(ta_alpha), float(weight_threshold), {"uniform": 0, "pyscf": 1}[sizing],
            ip(nbf), ip(n), ip(l), ip(mm), dp(zeta), dp(norm), dp(ctr),
            npts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self._molset_nbf = [int(x) for x in nbf]
        return npts

    def molset_vxc(self, Ps, funcs=PBE, precision="fp64"):
        """One build over all molecules of the set.  Ps: list of density
        matrices.  Returns a list of XC_result."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED}[precision]
        nmol = len(self._molset_nbf)
        if len(Ps) != nmol:
            raise NukexcError(1, f"need {nmol} density matrices, got {len(Ps)}")
        Pd = [self._in_matrix(P, k, f"P[{i}]") for i, (P, k) in enumerate(zip(Ps, self._molset_nbf))]
        Vs = [torch.empty((k, k), dtype=torch.float64, device=self.device)
              for k in self._molset_nbf]
        Pp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Pd])
        Vp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Vs])
        E = (ctypes.c_double * nmol)()
        N = (ctypes.c_double * nmol)()
        ids, nf = _ids(funcs)
        self._chk(lib().nkx_molset_vxc(self.h, Pp, nf, ids, prec, E, N, Vp))
        return [XC_result(E[k], Vs[k], N[k]) for k in range(nmol)]

    def molset_autotune(self, Ps, funcs=PBE, precision="mixed"):
        """nkx_molset_autotune: fastest density shape and V split count for
        the molecule set (kept for later set builds and captures); returns
        (split, shape, ms per set build)."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED}[precision]
        nmol = len(self._molset_nbf)
        if len(Ps) != nmol:
            raise NukexcError(1, f"need {nmol} density matrices, got {len(Ps)}")
        Pd = [self._in_matrix(P, k, f"P[{i}]") for i, (P, k) in enumerate(zip(Ps, self._molset_nbf))]
        Vs = [torch.empty((k, k), dtype=torch.float64, device=self.device)
              for k in self._molset_nbf]
        Pp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Pd])
        Vp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Vs])
        split, shape, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        ids, nf = _ids(funcs)
        self._chk(lib().nkx_molset_autotune(self.h, Pp, nf, ids, prec, Vp, ctypes.byref(split),
                                            ctypes.byref(shape), ctypes.byref(ms)))
        return split.value, shape.value, ms.value

    def molset_graph_capture(self, Ps, Vs, funcs=PBE, precision="mixed"):
        """Capture one build of the molecule set for the device tensors Ps
        (inputs) and Vs (outputs); replay with molset_graph_launch()."""
        prec = {"fp64": _capi.PREC_FP64, "mixed": _capi.PREC_MIXED}[precision]
        nmol = len(self._molset_nbf)
        if len(Ps) != nmol or len(Vs) != nmol:
            raise NukexcError(1, f"need {nmol} density and potential matrices")
        for i, (P, V, k) in enumerate(zip(Ps, Vs, self._molset_nbf)):
            self._check_out(P, k, f"P[{i}]")
            self._check_out(V, k, f"V[{i}]")
        self._graph_keep = (list(Ps), list(Vs))
        Pp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Ps])
        Vp = (ctypes.c_void_p * nmol)(*[x.data_ptr() for x in Vs])
        ids, nf = _ids(funcs)
        self._chk(lib().nkx_molset_graph_capture(self.h, Pp, nf, ids, prec, Vp))
        self._graph_nmol = nmol
        self._graph_nspin = 1

    def molset_graph_launch(self):
        """Replay the captured molecule-set build; returns (E list, N list)."""
        n = self._graph_nmol
        E = (ctypes.c_double * n)()
        N = (ctypes.c_double * n)()
        self._chk(lib().nkx_graph_launch(self.h, E, N))
        return list(E), list(N)

    STAGES = ("collocation", "density", "vgemm", "final", "build", "functional_z")

    def set_timing(self, on=True):
        self._chk(lib().nkx_set_timing(self.h, int(bool(on))))

    def reset_timing(self):
        self._chk(lib().nkx_reset_timing(self.h))

    def timing(self):
        """{stage: (ms summed since reset, launches)} from HIP events."""
        out = {}
        for k, name in enumerate(self.STAGES):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            self._chk(lib().nkx_get_timing(self.h, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def plan_work(self):
        v = [ctypes.c_double() for _ in range(3)]
        self._chk(lib().nkx_plan_work(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("flops_density", "flops_v", "colloc_elems"), [x.value for x in v]))

    def overlap(self):
        """Numerical overlap on the ctx grid (core_hamiltonian.hpp:62-119)."""
        nbf = self.basis.nbf()
        S = torch.empty((nbf, nbf), dtype=torch.float64, device=self.device)
        self._chk(lib().nkx_overlap(self.h, _ptr(S)))
        self.sync()
        return S


# ------------------------------------------------- reference-shaped calls ----
def partition_becke(ctx, atom_centers, quad_points, point_owner, weights, capped=False,
                    screen=True):
    """partition_becke (partitioning.hpp:63-120); in place on weights.
    capped=True is partition_becke_team's R_ij cap; screen=False evaluates
    every factor of the capped form (the screened form must agree bit for
    bit)."""
    ac = atom_centers.contiguous()
    qp = quad_points.contiguous()
    own = point_owner.to(torch.int32).contiguous()
    assert weights.is_contiguous() and weights.dtype == torch.float64
    mode = (_capi.BECKE_CAPPED if screen else _capi.BECKE_CAPPED_UNSCREENED) if capped else \
        _capi.BECKE_PLAIN
    ctx._done(lib().nkx_becke(ctx.h, ac.shape[0], _ptr(ac), qp.shape[0], _ptr(qp), _ptr(own),
                             _ptr(weights), mode))
    return weights


def partition_becke_team(ctx, atom_centers, quad_points, point_owner, weights):
    """partition_becke_team (partitioning.hpp:126-213): R_ij capped at 5."""
    return partition_becke(ctx, atom_centers, quad_points, point_owner, weights, capped=True)


def _colloc_out(nbf, npts, device):
    # LayoutLeft (N_bf, N_pts): transpose view of a contiguous (N_pts, N_bf)
    return torch.empty((npts, nbf), dtype=torch.float64, device=device).t()


def fill_collocation(ctx, basis, points):
    """fill_collocation (stobasis.hpp:494-532)."""
    if ctx.basis is not basis:
        ctx.set_basis(basis)
    pts = points.contiguous()
    phi = _colloc_out(basis.nbf(), pts.shape[0], ctx.device)
    ctx._done(lib().nkx_collocation(ctx.h, pts.shape[0], _ptr(pts), _ptr(phi), None, None, None))
    return phi


def fill_grad_collocation(ctx, basis, points):
    """fill_grad_collocation (stobasis.hpp:534-584): returns gx, gy, gz."""
    if ctx.basis is not basis:
        ctx.set_basis(basis)
    pts = points.contiguous()
    g = [_colloc_out(basis.nbf(), pts.shape[0], ctx.device) for _ in range(3)]
    ctx._done(lib().nkx_collocation(ctx.h, pts.shape[0], _ptr(pts), None, _ptr(g[0]), _ptr(g[1]),
                                   _ptr(g[2])))
    return tuple(g)


def _left(t):
    """Check a (rows, cols) LayoutLeft view: the transpose of a contiguous tensor."""
    if not t.t().is_contiguous():
        t = t.t().contiguous().t()
    return t


def compute_density(ctx, phi, mo_orbitals, mo_coeff):
    """compute_density (density.hpp:150-186)."""
    phi = _left(phi)
    C = _left(_dev_f64(mo_orbitals, ctx.device).reshape(phi.shape[0], -1))
    f = _dev_f64(mo_coeff, ctx.device).contiguous()
    rho = torch.empty(phi.shape[1], dtype=torch.float64, device=ctx.device)
    ctx._done(lib().nkx_density(ctx.h, phi.shape[0], phi.shape[1], _ptr(phi), None, None, None,
                               C.shape[1], _ptr(C), _ptr(f), _ptr(rho), None, None, None, None))
    return rho


def compute_density_and_sigma(ctx, phi, gx, gy, gz, mo_orbitals, mo_coeff):
    """compute_density_and_sigma (density.hpp:188-255): rho, grad rho (3), sigma."""
    phi, gx, gy, gz = (_left(x) for x in (phi, gx, gy, gz))
    C = _left(_dev_f64(mo_orbitals, ctx.device).reshape(phi.shape[0], -1))
    f = _dev_f64(mo_coeff, ctx.device).contiguous()
    n = phi.shape[1]
    out = [torch.empty(n, dtype=torch.float64, device=ctx.device) for _ in range(5)]
    ctx._done(lib().nkx_density(ctx.h, phi.shape[0], n, _ptr(phi), _ptr(gx), _ptr(gy), _ptr(gz),
                               C.shape[1], _ptr(C), _ptr(f), *[_ptr(o) for o in out]))
    return out[0], (out[1], out[2], out[3]), out[4]


def functional_eval(ctx, funcs, rho, sigma=None):
    """exc, vrho, vsigma on device (the libxc calls of xc_integrals.hpp)."""
    ids, nf = _ids(funcs)
    n = rho.shape[0]
    exc, vr = (torch.empty(n, dtype=torch.float64, device=ctx.device) for _ in range(2))
    vs = torch.empty(n, dtype=torch.float64, device=ctx.device) if sigma is not None else None
    ctx._done(lib().nkx_functional_eval(ctx.h, nf, ids, n, _ptr(rho.contiguous()),
                                       _ptr(None if sigma is None else sigma.contiguous()),
                                       _ptr(exc), _ptr(vr), _ptr(vs)))
    return exc, vr, vs


def functional_eval_polarized(ctx, funcs, rho, sigma=None):
    """Spin-polarised exc, vrho, vsigma on device; rho (N, 2), sigma (N, 3)
    in libxc's interleaved XC_POLARIZED layout."""
    ids, nf = _ids(funcs)
    n = rho.shape[0]
    exc = torch.empty(n, dtype=torch.float64, device=ctx.device)
    vr = torch.empty((n, 2), dtype=torch.float64, device=ctx.device)
    vs = torch.empty((n, 3), dtype=torch.float64, device=ctx.device) if sigma is not None else None
    ctx._done(lib().nkx_functional_eval_polarized(
        ctx.h, nf, ids, n, _ptr(rho.contiguous()),
        _ptr(None if sigma is None else sigma.contiguous()), _ptr(exc), _ptr(vr), _ptr(vs)))
    return exc, vr, vs


def _xc_pol(ctx, phi, grads, weights, mo_alpha, occ_alpha, mo_beta, occ_beta, funcs, family):
    phi = _left(phi)
    nbf, n = phi.shape
    g = [None, None, None] if grads is None else [_left(x) for x in grads]
    Ca = _left(_dev_f64(mo_alpha, ctx.device).reshape(nbf, -1))
    Cb = _left(_dev_f64(mo_beta, ctx.device).reshape(nbf, -1))
    fa = _dev_f64(occ_alpha, ctx.device).contiguous()
    fb = _dev_f64(occ_beta, ctx.device).contiguous()
    w = weights.contiguous()
    Va, Vb = (torch.empty((nbf, nbf), dtype=torch.float64, device=ctx.device) for _ in range(2))
    E = ctypes.c_double()
    ids, nf = _ids(funcs)
    ctx._chk(lib().nkx_xc_from_collocation_polarized(
        ctx.h, nbf, n, _ptr(phi), _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(w), Ca.shape[1],
        _ptr(Ca), _ptr(fa), Cb.shape[1], _ptr(Cb), _ptr(fb), nf, ids, family, ctypes.byref(E),
        _ptr(Va), _ptr(Vb)))
    return XC_result_polarized(E.value, Va, Vb)


def compute_lsda(ctx, collocation_values, weights, mo_alpha, occ_alpha, mo_beta, occ_beta,
                 funcs):
    """compute_lsda (xc_integrals.hpp:169-262)."""
    return _xc_pol(ctx, collocation_values, None, weights, mo_alpha, occ_alpha, mo_beta,
                   occ_beta, funcs, _capi.FAMILY_LDA)


def compute_gga_lsda(ctx, collocation_values, gx, gy, gz, weights, mo_alpha, occ_alpha,
                     mo_beta, occ_beta, funcs):
    """compute_gga_lsda (xc_integrals.hpp:543-718)."""
    return _xc_pol(ctx, collocation_values, (gx, gy, gz), weights, mo_alpha, occ_alpha,
                   mo_beta, occ_beta, funcs, _capi.FAMILY_GGA)


def _xc(ctx, phi, grads, weights, mo_orbitals, mo_coeff, funcs, family):
    phi = _left(phi)
    nbf, n = phi.shape
    g = [None, None, None] if grads is None else [_left(x) for x in grads]
    C = _left(_dev_f64(mo_orbitals, ctx.device).reshape(nbf, -1))
    f = _dev_f64(mo_coeff, ctx.device).contiguous()
    w = weights.contiguous()
    V = torch.empty((nbf, nbf), dtype=torch.float64, device=ctx.device)
    E = ctypes.c_double()
    ids, nf = _ids(funcs)
    ctx._chk(lib().nkx_xc_from_collocation(ctx.h, nbf, n, _ptr(phi), _ptr(g[0]), _ptr(g[1]),
                                           _ptr(g[2]), _ptr(w), C.shape[1], _ptr(C), _ptr(f), nf,
                                           ids, family, ctypes.byref(E), _ptr(V)))
    return XC_result(E.value, V)


def compute_lda(ctx, collocation_values, weights, mo_orbitals, mo_coeff, funcs):
    """compute_lda (xc_integrals.hpp:73-166)."""
    return _xc(ctx, collocation_values, None, weights, mo_orbitals, mo_coeff, funcs,
               _capi.FAMILY_LDA)


def compute_gga(ctx, collocation_values, gx, gy, gz, weights, mo_orbitals, mo_coeff, funcs):
    """compute_gga (xc_integrals.hpp:424-541)."""
    return _xc(ctx, collocation_values, (gx, gy, gz), weights, mo_orbitals, mo_coeff, funcs,
               _capi.FAMILY_GGA)


# ------------------------------------------------------ screened grid API ----
@dataclass
class NeighborList:
    """octree.hpp:279-284: for box b, the basis functions
    neighbors[offsets[b]:offsets[b + 1]] (ascending) whose cutoff sphere
    meets the box."""
    neighbors: "torch.Tensor"  # int32
    offsets: "torch.Tensor"  # int32, nboxes + 1
    max_points_per_box: int
    total_points: int


def create_bounding_boxes(ctx, grid, max_points_per_bb):
    """create_bounding_boxes (octree.hpp:71-138): permutes grid.quad_points,
    grid.weights and grid.point_owner in place into Morton order and returns
    the (nboxes, 6) boxes (min x, y, z, max x, y, z) of consecutive runs of
    max_points_per_bb points.  Anything derived from the old point order is
    invalid afterwards; a ctx holding this grid is handed the new order."""
    pts, w, own = grid.quad_points, grid.weights, grid.point_owner
    if not (pts.is_contiguous() and pts.dtype == torch.float64 and w.is_contiguous()
            and w.dtype == torch.float64 and own.is_contiguous() and own.dtype == torch.int32):
        raise ValueError("the grid arrays must be contiguous (float64 points and weights, "
                         "int32 owners) to be permuted in place")
    npts, m = pts.shape[0], int(max_points_per_bb)
    if m <= 0:
        raise ValueError("max_points_per_bb must be positive")
    boxes = torch.empty(((npts + m - 1) // m, 6), dtype=torch.float64, device=pts.device)
    ctx._chk(lib().nkx_bounding_boxes(ctx.h, npts, _ptr(pts), _ptr(w), _ptr(own), m,
                                      _ptr(boxes)))
    if getattr(ctx, "grid", None) is grid:
        ctx.set_grid(grid)
    return boxes


def build_neighbor_list(ctx, basis, bounding_boxes, max_points_per_box, total_points):
    """build_neighbor_list (octree.hpp:317-389): the functions of basis whose
    sphere (centre, cutoff radius) meets each box, as a NeighborList."""
    boxes = bounding_boxes.contiguous()
    nbox = boxes.shape[0]
    O = np.ascontiguousarray(basis.O, np.float64)
    r = np.ascontiguousarray(basis.cutoff_radii, np.float64)
    pd = ctypes.POINTER(ctypes.c_double)
    off = torch.empty(nbox + 1, dtype=torch.int32, device=boxes.device)
    tot = ctypes.c_int64()
    args = (ctx.h, nbox, _ptr(boxes), len(r), O.ctypes.data_as(pd), r.ctypes.data_as(pd),
            _ptr(off))
    ctx._chk(lib().nkx_neighbor_list(*args, None, ctypes.byref(tot)))
    nbr = torch.empty(max(tot.value, 1), dtype=torch.int32, device=boxes.device)
    ctx._chk(lib().nkx_neighbor_list(*args, _ptr(nbr), ctypes.byref(tot)))
    return NeighborList(nbr[:tot.value], off, int(max_points_per_box), int(total_points))


def _sparse_setup(ctx, basis, grid, nl):
    if nl.total_points != grid.weights.shape[0]:
        raise ValueError("the neighbor list was built for another grid")
    if getattr(ctx, "grid", None) is not grid:
        ctx.set_grid(grid)
    if getattr(ctx, "basis", None) is not basis:
        ctx.set_basis(basis)


def density_matrix(ctx, mo_orbitals, mo_coeff):
    """P = C diag(f) C^T on the device (C LayoutLeft N_bf x N_occ)."""
    C = _dev_f64(mo_orbitals, ctx.device)
    C = _left(C.reshape(C.shape[0], -1))
    f = _dev_f64(mo_coeff, ctx.device).contiguous()
    P = torch.empty((C.shape[0], C.shape[0]), dtype=torch.float64, device=ctx.device)
    ctx._done(lib().nkx_density_matrix(ctx.h, C.shape[0], C.shape[1], _ptr(C), _ptr(f), _ptr(P)))
    return P


def _density_points(ctx, P, grads):
    n = ctx.grid.weights.shape[0]
    out = [torch.empty(n, dtype=torch.float64, device=ctx.device) for _ in range(5 if grads else 1)]
    ptrs = [_ptr(o) for o in out] + [None] * (5 - len(out))
    ctx._done(lib().nkx_density_points(ctx.h, _ptr(P.contiguous()), *ptrs))
    return out


def compute_density_sparse(ctx, basis, grid, nl, mo_orbitals, mo_coeff):
    """compute_density_sparse (density.hpp:97-148): rho at every grid point.
    The batched path screens on its own 64-point batches; the neighbor list
    only has to belong to this grid."""
    _sparse_setup(ctx, basis, grid, nl)
    return _density_points(ctx, density_matrix(ctx, mo_orbitals, mo_coeff), False)[0]


def compute_density_and_sigma_sparse(ctx, basis, grid, nl, mo_orbitals, mo_coeff):
    """compute_density_and_sigma_sparse (density.hpp:261-328): rho, grad rho
    (3), sigma."""
    _sparse_setup(ctx, basis, grid, nl)
    o = _density_points(ctx, density_matrix(ctx, mo_orbitals, mo_coeff), True)
    return o[0], (o[1], o[2], o[3]), o[4]


def _family_check(funcs, gga):
    fams = [functional_info(f)[0] for f in funcs]
    if gga and not all(f == _capi.FAMILY_GGA for f in fams):
        raise NukexcError(4, "Provided functional is not a part of the GGA family")
    if not gga and any(f == _capi.FAMILY_GGA for f in fams):
        raise NukexcError(4, "Provided functional is not part of the LDA family")


def _lsda_sparse(ctx, basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, funcs, gga):
    _family_check(funcs, gga)
    _sparse_setup(ctx, basis, grid, nl)
    Pa = density_matrix(ctx, mo_alpha, occ_alpha)
    Pb = density_matrix(ctx, mo_beta, occ_beta)
    return ctx.vxc_polarized(Pa, Pb, funcs=funcs)


def compute_lsda_sparse(ctx, basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, funcs):
    """compute_lsda_sparse (xc_integrals.hpp:263-422) on the batched FP64
    build from P_sigma = C_sigma diag(occ_sigma) C_sigma^T."""
    return _lsda_sparse(ctx, basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, funcs,
                        False)


def compute_gga_lsda_sparse(ctx, basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta,
                            funcs):
    """compute_gga_lsda_sparse (xc_integrals.hpp:728-977), as
    compute_lsda_sparse."""
    return _lsda_sparse(ctx, basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, funcs,
                        True)
