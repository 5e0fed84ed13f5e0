"""nukexc_amd: MI355X-native exchange-correlation integration for STO bases.

The compute path is libnukexc_hip.so (hand-written HIP kernels for gfx950)
behind the C ABI in include/nukexc_hip.h; this package is its Python binding
and a mirror of the reference's nukexc/ interface (see api.py).
"""
from ._capi import LIB_PATH, NukexcError, lib  # noqa: F401
from .api import (PBE, TA_M3, TA_M4, Context, FlatGrid, STOBasisSet, XC_result,  # noqa: F401
                  XC_result_polarized, compute_density, compute_density_and_sigma, compute_gga,
                  compute_gga_lsda, compute_lda, compute_lsda, fill_collocation,
                  fill_grad_collocation, functional_eval, functional_eval_polarized,
                  functional_info,
                  lebedev_rule, make_manual_basis, partition_becke, partition_becke_team,
                  pruning_orders,
                  radial_rule, radial_scale, solid_harmonic, sto_norm_cutoff)
from .api import (NeighborList, build_neighbor_list, compute_density_and_sigma_sparse,  # noqa: F401
                  compute_density_sparse, compute_gga_lsda_sparse, compute_lsda_sparse,
                  create_bounding_boxes, density_matrix)
from . import readers  # noqa: F401  (read_xyz, load_adf_basis, load_thakkar_basis)

__version__ = "0.1.0"
