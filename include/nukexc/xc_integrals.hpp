// nukexc/xc_integrals.hpp -- E_xc and V_xc (xc_integrals.hpp:62-262, :424-718
// of the reference: compute_lda, compute_lsda, compute_gga, compute_gga_lsda).  The reference passes a libxc xc_func_type and evaluates
// the functional on the host; here a small Functional handle carries libxc ids
// and the functional runs on the GPU.
#pragma once

#include <stdexcept>
#include <vector>

#include "density.hpp"
#include "grid.hpp"
#include "octree.hpp"
#include "stobasis.hpp"

namespace Nukexc {

struct XC_result {
  double energy = 0.0;
  DeviceView2DLeft potential;  // (N_bf, N_bf); symmetric
  double n_el = 0.0;           // sum_g w rho (batched build only)
};

struct XC_result_polarized {
  double energy = 0.0;
  DeviceView2DLeft potential_alpha;  // (N_bf, N_bf); symmetric
  DeviceView2DLeft potential_beta;
  double n_el[2] = {0.0, 0.0};       // per spin (batched build only)
};

// libxc-numbered functional list (LDA_X = 1, GGA_X_PBE = 101, GGA_C_PBE = 130,
// LDA_C_PW = 12, LDA_C_PW_MOD = 13, LDA_C_VWN = 7, GGA_X_B88 = 106, GGA_C_LYP = 131); family
// as func.info->family.
struct Functional {
  std::vector<int> ids;
  int family = NKX_FAMILY_LDA;
  explicit Functional(std::vector<int> list) : ids(std::move(list)) {
    for (int id : ids) {
      int fam = 0;
      check(nkx_functional_info(id, &fam, nullptr));
      if (fam == NKX_FAMILY_GGA) family = NKX_FAMILY_GGA;
    }
  }
};

inline XC_result compute_lda(const DeviceView2DLeft &collocation_values,
                             const DeviceView1D &weights, const DeviceView2DLeft &mo_orbitals,
                             const DeviceView1D &mo_coeff, const Functional &func) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  XC_result r;
  r.potential.allocate(nbf, nbf);
  check(nkx_xc_from_collocation(context().get(), (int)nbf, (int64_t)npts,
                                collocation_values.data(), nullptr, nullptr, nullptr,
                                weights.data(), (int)mo_orbitals.extent(1), mo_orbitals.data(),
                                mo_coeff.data(), (int)func.ids.size(), func.ids.data(),
                                NKX_FAMILY_LDA, &r.energy, r.potential.data()),
        context().get());
  return r;
}

inline XC_result compute_gga(const DeviceView2DLeft &collocation_values,
                             const DeviceView2DLeft &collocation_gx,
                             const DeviceView2DLeft &collocation_gy,
                             const DeviceView2DLeft &collocation_gz,
                             const DeviceView1D &weights, const DeviceView2DLeft &mo_orbitals,
                             const DeviceView1D &mo_coeff, const Functional &func) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  XC_result r;
  r.potential.allocate(nbf, nbf);
  check(nkx_xc_from_collocation(context().get(), (int)nbf, (int64_t)npts,
                                collocation_values.data(), collocation_gx.data(),
                                collocation_gy.data(), collocation_gz.data(), weights.data(),
                                (int)mo_orbitals.extent(1), mo_orbitals.data(), mo_coeff.data(),
                                (int)func.ids.size(), func.ids.data(), NKX_FAMILY_GGA,
                                &r.energy, r.potential.data()),
        context().get());
  return r;
}

inline XC_result_polarized compute_lsda(const DeviceView2DLeft &collocation_values,
                                        const DeviceView1D &weights,
                                        const DeviceView2DLeft &mo_alpha,
                                        const DeviceView1D &occ_alpha,
                                        const DeviceView2DLeft &mo_beta,
                                        const DeviceView1D &occ_beta, const Functional &func) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  XC_result_polarized r;
  r.potential_alpha.allocate(nbf, nbf);
  r.potential_beta.allocate(nbf, nbf);
  check(nkx_xc_from_collocation_polarized(
            context().get(), (int)nbf, (int64_t)npts, collocation_values.data(), nullptr, nullptr,
            nullptr, weights.data(), (int)mo_alpha.extent(1), mo_alpha.data(), occ_alpha.data(),
            (int)mo_beta.extent(1), mo_beta.data(), occ_beta.data(), (int)func.ids.size(),
            func.ids.data(), NKX_FAMILY_LDA, &r.energy, r.potential_alpha.data(),
            r.potential_beta.data()),
        context().get());
  return r;
}

inline XC_result_polarized compute_gga_lsda(
    const DeviceView2DLeft &collocation_values, const DeviceView2DLeft &collocation_gx,
    const DeviceView2DLeft &collocation_gy, const DeviceView2DLeft &collocation_gz,
    const DeviceView1D &weights, const DeviceView2DLeft &mo_alpha, const DeviceView1D &occ_alpha,
    const DeviceView2DLeft &mo_beta, const DeviceView1D &occ_beta, const Functional &func) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  XC_result_polarized r;
  r.potential_alpha.allocate(nbf, nbf);
  r.potential_beta.allocate(nbf, nbf);
  check(nkx_xc_from_collocation_polarized(
            context().get(), (int)nbf, (int64_t)npts, collocation_values.data(),
            collocation_gx.data(), collocation_gy.data(), collocation_gz.data(), weights.data(),
            (int)mo_alpha.extent(1), mo_alpha.data(), occ_alpha.data(), (int)mo_beta.extent(1),
            mo_beta.data(), occ_beta.data(), (int)func.ids.size(), func.ids.data(),
            NKX_FAMILY_GGA, &r.energy, r.potential_alpha.data(), r.potential_beta.data()),
        context().get());
  return r;
}

// Batched build from a density matrix (row-major = LayoutLeft for symmetric
// P): the per-SCF-iteration XC call, without materialising N_bf x N_pts
// collocation matrices.
enum class Precision { FP64 = NKX_PREC_FP64, Mixed = NKX_PREC_MIXED, MixedBF16 = NKX_PREC_MIXED_BF16 };

namespace detail {
// The grid is handed over once; repeated builds on it reuse the launch plan.
inline void use_grid(nkx_ctx *ctx, const FlatGrid &grid) {
  thread_local const void *last_grid = nullptr;
  thread_local size_t last_npts = 0;
  thread_local unsigned last_epoch = 0;
  if (last_grid != grid.weights.data() || last_npts != grid.weights.size() ||
      last_epoch != grid_epoch()) {
    check(nkx_set_grid(ctx, (int)grid.atom_centers.size(),
                       reinterpret_cast<const double *>(grid.atom_centers.data()),
                       (int64_t)grid.weights.size(),
                       reinterpret_cast<const double *>(grid.quad_points.data()),
                       grid.weights.data(), grid.point_owner.data()),
          ctx);
    last_grid = grid.weights.data();
    last_npts = grid.weights.size();
    last_epoch = grid_epoch();
  }
}
}  // namespace detail

inline XC_result compute_vxc(const FlatGrid &grid, const STOBasisSet &basis,
                             const DeviceView2DLeft &P, const Functional &func,
                             Precision prec = Precision::FP64) {
  nkx_ctx *ctx = context().get();
  detail::use_grid(ctx, grid);
  set_basis(basis);
  XC_result r;
  r.potential.allocate(basis.nbf(), basis.nbf());
  check(nkx_vxc(ctx, P.data(), (int)func.ids.size(), func.ids.data(), (int)prec, &r.energy,
                &r.n_el, r.potential.data()),
        ctx);
  return r;
}

// Open-shell batched build from the spin density matrices (FP64 or the int8
// mixed form).
inline XC_result_polarized compute_vxc_polarized(const FlatGrid &grid, const STOBasisSet &basis,
                                                 const DeviceView2DLeft &P_alpha,
                                                 const DeviceView2DLeft &P_beta,
                                                 const Functional &func,
                                                 Precision prec = Precision::FP64) {
  nkx_ctx *ctx = context().get();
  detail::use_grid(ctx, grid);
  set_basis(basis);
  XC_result_polarized r;
  r.potential_alpha.allocate(basis.nbf(), basis.nbf());
  r.potential_beta.allocate(basis.nbf(), basis.nbf());
  check(nkx_vxc_polarized(ctx, P_alpha.data(), P_beta.data(), (int)func.ids.size(),
                          func.ids.data(), (int)prec, &r.energy, r.n_el,
                          r.potential_alpha.data(), r.potential_beta.data()),
        ctx);
  return r;
}

// XC nuclear gradient at fixed P (no counterpart in the reference): E_xc and
// dE_xc/dR_A for every atom of the grid (atom-major x, y, z), with every basis
// function centred on a grid atom.  partition says how the grid weights were
// made (capped Becke for make_flat_grid).  FP64 or the int8 mixed form.
enum class Partition { Fixed = 0, BeckeCapped = 1, Becke = 2 };

struct XC_gradient {
  double energy = 0.0;
  std::vector<double> gradient;  // 3 natoms
};

inline XC_gradient compute_xc_gradient(const FlatGrid &grid, const STOBasisSet &basis,
                                       const DeviceView2DLeft &P, const Functional &func,
                                       Partition partition = Partition::BeckeCapped,
                                       Precision prec = Precision::FP64) {
  if (prec == Precision::MixedBF16)
    throw std::runtime_error("the nuclear gradient is FP64 or the int8 mixed form");
  nkx_ctx *ctx = context().get();
  detail::use_grid(ctx, grid);
  set_basis(basis);
  XC_gradient r;
  r.gradient.assign(3 * grid.atom_centers.size(), 0.0);
  const auto fn = prec == Precision::FP64 ? nkx_xc_grad : nkx_xc_grad_mixed;
  check(fn(ctx, P.data(), (int)func.ids.size(), func.ids.data(), (int)partition, &r.energy,
           r.gradient.data()),
        ctx);
  return r;
}

// ---- screened ("sparse") forms ---------------------------------------------
// The batched build screens on its own 64-point batches; the neighbor list
// only has to belong to the grid (same number of points).
namespace detail {
inline void sparse_setup(const FlatGrid &grid, const STOBasisSet &basis, const NeighborList &nl) {
  if ((size_t)nl.total_points != grid.weights.size())
    throw std::runtime_error("the neighbor list was built for another grid");
  use_grid(context().get(), grid);
  set_basis(basis);
}

// P = C diag(f) C^T on the device
inline DeviceView2DLeft density_matrix(const DeviceView2DLeft &C, const DeviceView1D &f) {
  const size_t nbf = C.extent(0);
  DeviceView2DLeft P(nbf, nbf);
  check(nkx_density_matrix(context().get(), (int)nbf, (int)C.extent(1), C.data(), f.data(),
                           P.data()),
        context().get());
  return P;
}
}  // namespace detail

// density.hpp:97-148
inline DeviceView1D compute_density_sparse(const STOBasisSet &basis, const FlatGrid &grid,
                                           const NeighborList &nl,
                                           const DeviceView2DLeft &mo_orbitals,
                                           const DeviceView1D &mo_coeff) {
  detail::sparse_setup(grid, basis, nl);
  DeviceView2DLeft P = detail::density_matrix(mo_orbitals, mo_coeff);
  DeviceView1D rho(grid.weights.size());
  check(nkx_density_points(context().get(), P.data(), rho.data(), nullptr, nullptr, nullptr,
                           nullptr),
        context().get());
  context().fence();
  return rho;
}

// density.hpp:261-328
inline void compute_density_and_sigma_sparse(const STOBasisSet &basis, const FlatGrid &grid,
                                             const NeighborList &nl,
                                             const DeviceView2DLeft &mo_orbitals,
                                             const DeviceView1D &mo_coeff, DeviceView1D &rho,
                                             DeviceView1D &gx_rho, DeviceView1D &gy_rho,
                                             DeviceView1D &gz_rho, DeviceView1D &sigma) {
  detail::sparse_setup(grid, basis, nl);
  DeviceView2DLeft P = detail::density_matrix(mo_orbitals, mo_coeff);
  check(nkx_density_points(context().get(), P.data(), rho.data(), gx_rho.data(), gy_rho.data(),
                           gz_rho.data(), sigma.data()),
        context().get());
  context().fence();
}

namespace detail {
inline XC_result_polarized lsda_sparse(const STOBasisSet &basis, const FlatGrid &grid,
                                       const NeighborList &nl, const DeviceView2DLeft &mo_alpha,
                                       const DeviceView1D &occ_alpha,
                                       const DeviceView2DLeft &mo_beta,
                                       const DeviceView1D &occ_beta, const Functional &func) {
  sparse_setup(grid, basis, nl);
  DeviceView2DLeft Pa = density_matrix(mo_alpha, occ_alpha);
  DeviceView2DLeft Pb = density_matrix(mo_beta, occ_beta);
  XC_result_polarized r = compute_vxc_polarized(grid, basis, Pa, Pb, func);
  context().fence();
  return r;
}
}  // namespace detail

// xc_integrals.hpp:263-422
inline XC_result_polarized compute_lsda_sparse(const STOBasisSet &basis, const FlatGrid &grid,
                                               const NeighborList &nl,
                                               const DeviceView2DLeft &mo_alpha,
                                               const DeviceView1D &occ_alpha,
                                               const DeviceView2DLeft &mo_beta,
                                               const DeviceView1D &occ_beta,
                                               const Functional &func) {
  if (func.family != NKX_FAMILY_LDA)
    throw std::runtime_error("Provided functional is not part of the LDA family");
  return detail::lsda_sparse(basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, func);
}

// xc_integrals.hpp:728-977
inline XC_result_polarized compute_gga_lsda_sparse(const STOBasisSet &basis,
                                                   const FlatGrid &grid, const NeighborList &nl,
                                                   const DeviceView2DLeft &mo_alpha,
                                                   const DeviceView1D &occ_alpha,
                                                   const DeviceView2DLeft &mo_beta,
                                                   const DeviceView1D &occ_beta,
                                                   const Functional &func) {
  if (func.family != NKX_FAMILY_GGA)
    throw std::runtime_error("Provided funtional is not a part of the GGA family");
  for (int id : func.ids) {
    int fam = 0;
    check(nkx_functional_info(id, &fam, nullptr));
    if (fam != NKX_FAMILY_GGA)
      throw std::runtime_error("Provided funtional is not a part of the GGA family");
  }
  return detail::lsda_sparse(basis, grid, nl, mo_alpha, occ_alpha, mo_beta, occ_beta, func);
}

}  // namespace Nukexc
