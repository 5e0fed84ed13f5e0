// nukexc/grid.hpp -- FlatGrid and make_flat_grid (grid.hpp:150-348 of the
// reference): per-atom Treutler-Ahlrichs (or Becke) radial x Lebedev-Laikov
// grids, capped Becke partition, removal of weights <= weight_threshold.
// Unlike the reference's atomic-counter compaction, the surviving points keep
// their generation order.
#pragma once

#include <type_traits>

#include "molecule.hpp"
#include "partitioning.hpp"

namespace Nukexc {

// Radial / angular scheme tags (IntegratorXX types in the reference).
struct TreutlerAhlrichs {};
struct Becke {};
struct LebedevLaikov {};

enum class PruningScheme { Unpruned, Treutler, Robust };
enum class RadialSizing { Uniform, PySCF };

constexpr double TA_M4 = 0.6;
constexpr double TA_M3 = 0.0;

struct FlatGrid {
  DeviceArray<Point> quad_points;
  DeviceArray<double> weights;
  DeviceArray<Point> atom_centers;
  DeviceArray<unsigned> Z;
  DeviceArray<int> point_owner;
};

namespace detail {
// Bumped whenever a grid's points are reordered in place (create_bounding_boxes),
// so that a ctx holding that grid is handed the new order.
inline unsigned &grid_epoch() {
  thread_local unsigned epoch = 0;
  return epoch;
}
}  // namespace detail

template <typename radial_type, typename angular_type>
FlatGrid make_flat_grid(const Molecule &mol, const size_t nrad = 50,
                        const size_t nang_order = 30, const double weight_threshold = 1e-30,
                        const double ta_alpha = TA_M4,
                        const PruningScheme pruning = PruningScheme::Unpruned,
                        const RadialSizing radial_sizing = RadialSizing::Uniform) {
  static_assert(std::is_same_v<angular_type, LebedevLaikov>, "Lebedev-Laikov angular grids only");
  constexpr int scheme = std::is_same_v<radial_type, TreutlerAhlrichs>
                             ? NKX_RADIAL_TREUTLER_AHLRICHS
                             : NKX_RADIAL_BECKE;
  nkx_ctx *ctx = context().get();
  std::vector<double> xyz;
  std::vector<int> Z;
  for (unsigned i = 0; i < mol.natoms; ++i) {
    for (int d = 0; d < 3; ++d) xyz.push_back(mol.atom_centers[i][d]);
    Z.push_back((int)mol.Z[i]);
  }
  int64_t npts = 0;
  check(nkx_make_flat_grid_pruned(ctx, (int)mol.natoms, xyz.data(), Z.data(), (int)nrad,
                                  (int)nang_order, scheme, ta_alpha, weight_threshold,
                                  radial_sizing == RadialSizing::PySCF ? NKX_SIZING_PYSCF
                                                                       : NKX_SIZING_UNIFORM,
                                  (int)pruning, &npts),
        ctx);
  FlatGrid g;
  g.quad_points.allocate(npts);
  g.weights.allocate(npts);
  g.point_owner.allocate(npts);
  check(nkx_grid_export(ctx, reinterpret_cast<double *>(g.quad_points.data()), g.weights.data(),
                        g.point_owner.data()),
        ctx);
  g.atom_centers.allocate(mol.natoms);
  g.atom_centers.from_host(mol.atom_centers);
  g.Z.allocate(mol.natoms);
  g.Z.from_host(mol.Z);
  context().fence();
  return g;
}

}  // namespace Nukexc
