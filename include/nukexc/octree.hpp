// nukexc/octree.hpp -- screened-grid set-up (octree.hpp:60-389 of the
// reference) on the GPU without ArborX: Morton order of the grid points, the
// bounding box of each run of consecutive points, and the per-box list of
// basis functions whose cutoff sphere meets the box.
#pragma once

#include "grid.hpp"
#include "stobasis.hpp"

namespace Nukexc {

struct Box {  // ArborX::Box<3, double> in the reference
  Point lo, hi;
  Point &minCorner() { return lo; }
  Point &maxCorner() { return hi; }
  const Point &minCorner() const { return lo; }
  const Point &maxCorner() const { return hi; }
};
static_assert(sizeof(Box) == 6 * sizeof(double), "Box must be six packed doubles");

// octree.hpp:279-284
struct NeighborList {
  DeviceArray<int> neighbors;  // flat list of basis function indices
  DeviceArray<int> offsets;    // offsets[b] .. offsets[b + 1]: the neighbors of box b
  int max_points_per_box = 0;
  int total_points = 0;
};

// octree.hpp:71-138.  MUTATES the grid, as the reference does: quad_points,
// weights and point_owner are permuted in place into Morton order, so data
// derived from the old point order is invalid afterwards.
inline DeviceArray<Box> create_bounding_boxes(FlatGrid &grid, const int max_points_per_bb) {
  if (max_points_per_bb <= 0) throw std::runtime_error("max_points_per_bb must be positive");
  const size_t n = grid.quad_points.size();
  DeviceArray<Box> boxes((n + max_points_per_bb - 1) / max_points_per_bb);
  nkx_ctx *ctx = context().get();
  check(nkx_bounding_boxes(ctx, (int64_t)n, reinterpret_cast<double *>(grid.quad_points.data()),
                           grid.weights.data(), grid.point_owner.data(), max_points_per_bb,
                           reinterpret_cast<double *>(boxes.data())),
        ctx);
  ++detail::grid_epoch();  // the ctx must be handed the new point order
  return boxes;
}

// octree.hpp:317-389 (the histogram output of the reference is not kept).
inline void build_neighbor_list(const STOBasisSet &basis, const DeviceArray<Box> &bounding_boxes,
                                const int max_points_per_box, const int total_points,
                                NeighborList &neighbor_list) {
  nkx_ctx *ctx = context().get();
  std::vector<double> c;
  for (const auto &p : basis.O)
    for (int d = 0; d < 3; ++d) c.push_back(p[d]);
  const int64_t nbox = (int64_t)bounding_boxes.size();
  const double *bx = reinterpret_cast<const double *>(bounding_boxes.data());
  neighbor_list.offsets.allocate(nbox + 1);
  int64_t total = 0;
  check(nkx_neighbor_list(ctx, nbox, bx, (int)basis.nbf(), c.data(), basis.cutoff_radii.data(),
                          neighbor_list.offsets.data(), nullptr, &total),
        ctx);
  neighbor_list.neighbors.allocate((size_t)total);
  if (total > 0)
    check(nkx_neighbor_list(ctx, nbox, bx, (int)basis.nbf(), c.data(), basis.cutoff_radii.data(),
                            neighbor_list.offsets.data(), neighbor_list.neighbors.data(), &total),
          ctx);
  neighbor_list.max_points_per_box = max_points_per_box;
  neighbor_list.total_points = total_points;
}

}  // namespace Nukexc
