"""ctypes binding of libnukexc_hip.so (include/nukexc_hip.h).

The library is the product: every compute call here goes to the HIP kernels.
There is no CPU fallback; a missing or unbuildable library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnukexc_hip.so")

NKX_OK = 0
STATUS_NAMES = {
    0: "NKX_OK", 1: "NKX_ERR_INVALID_ARGUMENT", 2: "NKX_ERR_HIP", 3: "NKX_ERR_OUT_OF_MEMORY",
    4: "NKX_ERR_FUNCTIONAL_FAMILY", 5: "NKX_ERR_NSPIN", 6: "NKX_ERR_NEGATIVE_WEIGHT",
    7: "NKX_ERR_UNKNOWN_FUNCTIONAL", 8: "NKX_ERR_STATE", 9: "NKX_ERR_UNSUPPORTED",
}
BECKE_PLAIN, BECKE_CAPPED, BECKE_CAPPED_UNSCREENED = 0, 1, 3
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_BF16 = 2
RADIAL_TA = 0
RADIAL_BECKE = 1
FAMILY_LDA = 1
FAMILY_GGA = 2

_c_int, _c_double, _c_i64, _c_size = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_size_t
_vp = ctypes.c_void_p
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "nkx_create": [_c_int, _vp, ctypes.POINTER(_vp)],
    "nkx_destroy": [_vp],
    "nkx_sync": [_vp],
    "nkx_last_error": [_vp],
    "nkx_version": [_pi, _pi, _pi],
    "nkx_malloc": [_vp, _c_size, ctypes.POINTER(_vp)],
    "nkx_free": [_vp, _vp],
    "nkx_memcpy_h2d": [_vp, _vp, _vp, _c_size],
    "nkx_memcpy_d2h": [_vp, _vp, _vp, _c_size],
    "nkx_memset": [_vp, _vp, _c_int, _c_size],
    "nkx_copy_d2d": [_vp, _vp, _vp, _c_size],
    "nkx_lebedev_npts": [_c_int, _pi, _pi],
    "nkx_lebedev_rule": [_c_int, _pd],
    "nkx_pruning_orders": [_c_int, _c_int, _pi],
    "nkx_radial_rule": [_c_int, _c_int, _c_double, _c_double, _pd, _pd],
    "nkx_radial_scale": [_c_int, _c_int, _pd],
    "nkx_make_flat_grid": [_vp, _c_int, _pd, _pi, _c_int, _c_int, _c_int, _c_double, _c_double,
                           _c_int, _pi64],
    "nkx_make_flat_grid_pruned": [_vp, _c_int, _pd, _pi, _c_int, _c_int, _c_int, _c_double,
                                  _c_double, _c_int, _c_int, _pi64],
    "nkx_set_grid": [_vp, _c_int, _vp, _c_i64, _vp, _vp, _vp],
    "nkx_grid_npts": [_vp, _pi64],
    "nkx_grid_export": [_vp, _vp, _vp, _vp],
    "nkx_becke": [_vp, _c_int, _vp, _c_i64, _vp, _vp, _vp, _c_int],
    "nkx_becke_noalloc": [_vp, _c_int, _vp, _c_i64, _vp, _vp, _vp, _c_int],
    "nkx_sto_norm_cutoff": [_c_int, _c_double, _c_double, _pd, _pd],
    "nkx_set_basis": [_vp, _c_int, _pi, _pi, _pi, _pd, _pd, _pd, _pd],
    "nkx_basis_nbf": [_vp, _pi],
    "nkx_solid_harmonic": [_c_int, _c_int, _c_double, _c_double, _c_double, _pd],
    "nkx_collocation": [_vp, _c_i64, _vp, _vp, _vp, _vp, _vp],
    "nkx_density": [_vp, _c_int, _c_i64, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp,
                    _vp, _vp],
    "nkx_functional_info": [_c_int, _pi, _pi],
    "nkx_set_dens_threshold": [_vp, _c_double],
    "nkx_functional_eval": [_vp, _c_int, _pi, _c_i64, _vp, _vp, _vp, _vp, _vp],
    "nkx_xc_from_collocation": [_vp, _c_int, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp,
                                _c_int, _pi, _c_int, _pd, _vp],
    "nkx_vxc": [_vp, _vp, _c_int, _pi, _c_int, _pd, _pd, _vp],
    "nkx_functional_eval_polarized": [_vp, _c_int, _pi, _c_i64, _vp, _vp, _vp, _vp, _vp],
    "nkx_xc_from_collocation_polarized": [_vp, _c_int, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_int,
                                          _vp, _vp, _c_int, _vp, _vp, _c_int, _pi, _c_int, _pd,
                                          _vp, _vp],
    "nkx_vxc_polarized": [_vp, _vp, _vp, _c_int, _pi, _c_int, _pd, _pd, _vp, _vp],
    "nkx_xc_grad": [_vp, _vp, _c_int, _pi, _c_int, _pd, _pd],
    "nkx_xc_grad_mixed": [_vp, _vp, _c_int, _pi, _c_int, _pd, _pd],
    "nkx_xc_grad_polarized": [_vp, _vp, _vp, _c_int, _pi, _c_int, _c_int, _pd, _pd],
    "nkx_autotune": [_vp, _vp, _c_int, _pi, _c_int, _vp, _pi, _pi, _pd],
    "nkx_molset_autotune": [_vp, ctypes.POINTER(_vp), _c_int, _pi, _c_int, ctypes.POINTER(_vp),
                            _pi, _pi, _pd],
    "nkx_set_launch": [_vp, _c_int, _c_int],
    "nkx_set_memory_budget": [_vp, _c_i64],
    "nkx_get_launch": [_vp, _pi, _pi],
    "nkx_launch_points": [_c_int, _c_int, _pi],
    "nkx_graph_capture": [_vp, _vp, _c_int, _pi, _c_int, _vp],
    "nkx_graph_capture_polarized": [_vp, _vp, _vp, _c_int, _pi, _c_int, _vp, _vp],
    "nkx_graph_launch": [_vp, _pd, _pd],
    "nkx_molset_graph_capture": [_vp, ctypes.POINTER(_vp), _c_int, _pi, _c_int,
                                 ctypes.POINTER(_vp)],
    "nkx_molset_create": [_vp, _c_int, _pi, _pd, _pi, _c_int, _c_int, _c_int, _c_double,
                          _c_double, _c_int, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pi64],
    "nkx_molset_vxc": [_vp, ctypes.POINTER(_vp), _c_int, _pi, _c_int, _pd, _pd,
                       ctypes.POINTER(_vp)],
    "nkx_set_screening": [_vp, _c_double, _c_int],
    "nkx_plan_stats": [_vp, _pi64, _pi64, _pi64, _pi64, _pi64],
    "nkx_overlap": [_vp, _vp],
    "nkx_set_timing": [_vp, _c_int],
    "nkx_reset_timing": [_vp],
    "nkx_get_timing": [_vp, _c_int, _pd, _pi64],
    "nkx_plan_work": [_vp, _pd, _pd, _pd],
    "nkx_bounding_boxes": [_vp, _c_i64, _vp, _vp, _vp, _c_int, _vp],
    "nkx_neighbor_list": [_vp, _c_i64, _vp, _c_int, _pd, _pd, _vp, _vp, _pi64],
    "nkx_density_matrix": [_vp, _c_int, _c_int, _vp, _vp, _vp],
    "nkx_density_points": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
}
_RESTYPE = {"nkx_last_error": ctypes.c_char_p}

_lib = None


class NukexcError(RuntimeError):
    """A non-OK nkx_status, with the library's message."""

    def __init__(self, code, msg):
        self.code = code
        self.status = STATUS_NAMES.get(code, str(code))
        super().__init__(f"{self.status}: {msg}")


def lib():
    """Load libnukexc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C nukexc_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != NKX_OK:
        msg = lib().nkx_last_error(ctx)
        raise NukexcError(rc, msg.decode() if msg else "")
    return rc
