// nukexc/partitioning.hpp -- Becke partition (partitioning.hpp:63-213 of the
// reference) on device arrays: weights(g) *= P_owner(g) / sum_i P_i.
#pragma once

#include "nukexc_config.hpp"

namespace Nukexc {

// Plain Becke partition (partitioning.hpp:63-120).
inline void partition_becke(const DeviceArray<Point> &atom_centers,
                            const DeviceArray<Point> &quadrature_points,
                            const DeviceArray<int> &point_owner, DeviceArray<double> &weights) {
  check(nkx_becke(context().get(), (int)atom_centers.size(),
                  reinterpret_cast<const double *>(atom_centers.data()),
                  (int64_t)quadrature_points.size(),
                  reinterpret_cast<const double *>(quadrature_points.data()),
                  point_owner.data(), weights.data(), 0),
        context().get());
  context().fence();
}

// Becke partition with R_ij capped at 5 bohr (partitioning.hpp:126-213).
inline void partition_becke_team(const DeviceArray<Point> &atom_centers,
                                 const DeviceArray<Point> &quadrature_points,
                                 const DeviceArray<int> &point_owner,
                                 DeviceArray<double> &weights) {
  check(nkx_becke(context().get(), (int)atom_centers.size(),
                  reinterpret_cast<const double *>(atom_centers.data()),
                  (int64_t)quadrature_points.size(),
                  reinterpret_cast<const double *>(quadrature_points.data()),
                  point_owner.data(), weights.data(), 1),
        context().get());
  context().fence();
}

}  // namespace Nukexc
