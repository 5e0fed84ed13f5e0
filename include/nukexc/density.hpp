// nukexc/density.hpp -- density and its gradient from collocation values and
// orbitals (density.hpp:150-255 of the reference): rho = sum_i f_i psi_i^2,
// grad rho = 2 sum_i f_i psi_i grad psi_i, sigma = |grad rho|^2,
// psi = C^T Phi.  C is LayoutLeft (N_bf, N_occ).
#pragma once

#include "grid.hpp"
#include "nukexc_config.hpp"
#include "stobasis.hpp"

namespace Nukexc {

inline DeviceView1D compute_density(const DeviceView2DLeft &collocation_values,
                                    const DeviceView2DLeft &mo_orbitals,
                                    const DeviceView1D &mo_coeff) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  DeviceView1D rho(npts);
  check(nkx_density(context().get(), (int)nbf, (int64_t)npts, collocation_values.data(), nullptr,
                    nullptr, nullptr, (int)mo_orbitals.extent(1), mo_orbitals.data(),
                    mo_coeff.data(), rho.data(), nullptr, nullptr, nullptr, nullptr),
        context().get());
  context().fence();
  return rho;
}

// compute_density(basis, grid, mo_orbitals, mo_coeff) (density.hpp:55-95):
// collocation on the grid's points, then rho from (C, f) as above.  The
// (N_bf, N_pts) collocation matrix is a temporary, as in the reference.
inline DeviceView1D compute_density(const STOBasisSet &basis, const FlatGrid &grid,
                                    const DeviceView2DLeft &mo_orbitals,
                                    const DeviceView1D &mo_coeff) {
  ExecSpace space;
  DeviceView2DLeft collocation_values(basis.nbf(), grid.quad_points.size());
  fill_collocation(space, basis, grid.quad_points, collocation_values);
  return compute_density(collocation_values, mo_orbitals, mo_coeff);
}

inline void compute_density_and_sigma(const DeviceView2DLeft &collocation_values,
                                      const DeviceView2DLeft &collocation_gx,
                                      const DeviceView2DLeft &collocation_gy,
                                      const DeviceView2DLeft &collocation_gz,
                                      const DeviceView2DLeft &mo_orbitals,
                                      const DeviceView1D &mo_coeff, DeviceView1D &rho,
                                      DeviceView1D &gx_rho, DeviceView1D &gy_rho,
                                      DeviceView1D &gz_rho, DeviceView1D &sigma) {
  const size_t nbf = collocation_values.extent(0), npts = collocation_values.extent(1);
  check(nkx_density(context().get(), (int)nbf, (int64_t)npts, collocation_values.data(),
                    collocation_gx.data(), collocation_gy.data(), collocation_gz.data(),
                    (int)mo_orbitals.extent(1), mo_orbitals.data(), mo_coeff.data(), rho.data(),
                    gx_rho.data(), gy_rho.data(), gz_rho.data(), sigma.data()),
        context().get());
  context().fence();
}

}  // namespace Nukexc
