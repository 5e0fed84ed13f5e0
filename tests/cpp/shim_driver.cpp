// Exercises the C++ shim (include/nukexc/*.hpp) the way the reference's unit
// tests exercise its headers (test_hydrogen.cxx:815-912, test_helium.cxx:
// 67-121): grid, basis, collocation, compute_lda / compute_gga, the
// polarised compute_lsda / compute_gga_lsda, the batched compute_vxc(_polarized),
// the nuclear gradient (FP64 and mixed), the screened (sparse) forms and the
// error path.  Prints one line
// per check and
// exits nonzero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "nukexc/xc_integrals.hpp"

using namespace Nukexc;

static void expect_rel(const char *what, double got, double ref, double rel) {
  const double err = std::fabs(got - ref) / std::fabs(ref);
  std::printf("%-28s got % .15f ref % .15f rel %.2e\n", what, got, ref, err);
  if (!(err <= rel)) {
    std::printf("FAIL %s\n", what);
    std::exit(1);
  }
}

int main() {
  initialize(0);
  Molecule H(std::vector<std::vector<double>>{{0., 0., 0.}}, std::vector<unsigned>{1u});
  auto grid = make_flat_grid<Becke, LebedevLaikov>(H, 100, 20);
  STOBasisSet basis = make_manual_basis({{1, 0, 0, 1.0, 0., 0., 0.}});
  const size_t npts = grid.weights.size();

  // the reference's argument lists (standalone.cxx:413,424: an ExecSpace
  // instance first)
  ExecSpace space;
  DeviceView2DLeft phi(1, npts), gx(1, npts), gy(1, npts), gz(1, npts);
  fill_collocation(space, basis, grid.quad_points, phi);
  fill_grad_collocation(space, basis, grid.quad_points, gx, gy, gz);
  space.fence();
  DeviceView2DLeft C(1, 1);
  DeviceView1D f(1);
  C.from_host(std::vector<double>{1.0});
  f.from_host(std::vector<double>{1.0});

  // compute_density(basis, grid, C, f) (density.hpp:55-95): N_el = 1 and the
  // same values as from the collocation matrix
  {
    auto rho_g = compute_density(basis, grid, C, f).to_host();
    auto rho_c = compute_density(phi, C, f).to_host();
    const auto w = grid.weights.to_host();
    double n = 0.0, dmax = 0.0;
    for (size_t g = 0; g < npts; ++g) {
      n += w[g] * rho_g[g];
      dmax = std::fmax(dmax, std::fabs(rho_g[g] - rho_c[g]));
    }
    expect_rel("H compute_density(grid) N", n, 1.0, 1e-10);
    std::printf("%-28s max |diff| %.2e\n", "grid vs collocation rho", dmax);
    if (dmax != 0.0) {
      std::printf("FAIL grid-based compute_density differs\n");
      return 1;
    }
  }
  // a mis-sized output is refused before any device work
  {
    bool bad = false;
    DeviceView2DLeft wrong(2, npts);
    try {
      fill_collocation(space, basis, grid.quad_points, wrong);
    } catch (const std::runtime_error &) {
      bad = true;
    }
    if (!bad) {
      std::printf("FAIL fill_collocation accepted a (2, N_pts) output for one function\n");
      return 1;
    }
  }

  // test_hydrogen.cxx:815-857
  auto lda = compute_lda(phi, grid.weights, C, f, Functional({1}));
  expect_rel("H LDA_X energy", lda.energy, -0.2127415030860106, 1e-10);

  // test_hydrogen.cxx:859-912
  auto gga = compute_gga(phi, gx, gy, gz, grid.weights, C, f, Functional({101}));
  expect_rel("H GGA_X_PBE energy", gga.energy, -0.253995708307881, 1e-10);
  expect_rel("H GGA_X_PBE V00", gga.potential.to_host()[0], -0.320733669386709, 1e-10);

  // batched build from P = f C C^T gives the same numbers
  DeviceView2DLeft P(1, 1);
  P.from_host(std::vector<double>{1.0});
  auto vb = compute_vxc(grid, basis, P, Functional({101}));
  expect_rel("H batched GGA_X_PBE energy", vb.energy, -0.253995708307881, 1e-10);
  expect_rel("H batched N_el", vb.n_el, 1.0, 1e-10);
  auto vm = compute_vxc(grid, basis, P, Functional({101}), Precision::Mixed);
  expect_rel("H batched mixed energy", vm.energy, -0.253995708307881, 1e-6);

  // Spin-polarised H (compute_lsda, xc_integrals.hpp:169-262): one alpha
  // electron, the beta orbital unoccupied.  Spin scaling of exchange gives
  // E_x[rho, 0] = 2^(1/3) E_x^unpolarised[rho].
  DeviceView1D f0(1);
  f0.from_host(std::vector<double>{0.0});
  auto lsda = compute_lsda(phi, grid.weights, C, f, C, f0, Functional({1}));
  expect_rel("H LSDA_X energy", lsda.energy, std::cbrt(2.0) * -0.2127415030860106, 1e-10);
  DeviceView2DLeft P0(1, 1);
  P0.from_host(std::vector<double>{0.0});
  auto vp = compute_vxc_polarized(grid, basis, P, P0, Functional({1}));
  expect_rel("H batched LSDA_X energy", vp.energy, lsda.energy, 1e-12);
  expect_rel("H batched LSDA V_alpha", vp.potential_alpha.to_host()[0],
             lsda.potential_alpha.to_host()[0], 1e-12);
  auto glsda = compute_gga_lsda(phi, gx, gy, gz, grid.weights, C, f, C, f0, Functional({101, 130}));
  auto vpg = compute_vxc_polarized(grid, basis, P, P0, Functional({101, 130}));
  expect_rel("H batched PBE polarised", vpg.energy, glsda.energy, 1e-12);

  // xc_integrals.hpp:434-438: an LDA functional handed to compute_gga throws
  bool threw = false;
  try {
    (void)compute_gga(phi, gx, gy, gz, grid.weights, C, f, Functional({1}));
  } catch (const std::runtime_error &e) {
    threw = true;
    std::printf("compute_gga(LDA) threw: %s\n", e.what());
  }
  if (!threw) {
    std::printf("FAIL compute_gga accepted an LDA functional\n");
    return 1;
  }
  // nuclear gradient of H2 (1s on each atom, P = 0.6 everywhere, PBE):
  // translational invariance, the mirror symmetry of the bond, and the mixed
  // form against FP64
  Molecule H2(std::vector<std::vector<double>>{{0., 0., -0.7}, {0., 0., 0.7}},
              std::vector<unsigned>{1u, 1u});
  auto grid2 = make_flat_grid<Becke, LebedevLaikov>(H2, 60, 17);
  STOBasisSet basis2 =
      make_manual_basis({{1, 0, 0, 1.0, 0., 0., -0.7}, {1, 0, 0, 1.0, 0., 0., 0.7}});
  DeviceView2DLeft P2(2, 2);
  P2.from_host(std::vector<double>{0.6, 0.6, 0.6, 0.6});
  auto g64 = compute_xc_gradient(grid2, basis2, P2, Functional({101, 130}));
  auto gmx = compute_xc_gradient(grid2, basis2, P2, Functional({101, 130}),
                                 Partition::BeckeCapped, Precision::Mixed);
  const auto &g = g64.gradient;
  std::printf("H2 dE/dz: % .12f % .12f\n", g[2], g[5]);
  double worst = 0.0;
  for (int k = 0; k < 6; ++k) worst = std::fmax(worst, std::fabs(gmx.gradient[k] - g[k]));
  if (std::fabs(g[2] + g[5]) > 1e-10 || std::fabs(g[0]) > 1e-10 || std::fabs(g[1]) > 1e-10 ||
      std::fabs(g[2]) < 1e-4 || worst > 1e-7) {
    std::printf("FAIL H2 gradient (sum %.2e, mixed vs FP64 %.2e)\n", g[2] + g[5], worst);
    return 1;
  }
  // screened forms (octree.hpp:71-389, density.hpp:97-148,
  // xc_integrals.hpp:728-977) on the H2 grid, Morton-ordered in place: with
  // P_alpha = P_beta = P2 / 2 the polarised PBE energy is the closed-shell one,
  // and sum_g w rho = N_el of the batched build
  auto boxes = create_bounding_boxes(grid2, 64);
  NeighborList nl;
  build_neighbor_list(basis2, boxes, 64, (int)grid2.weights.size(), nl);
  const auto off = nl.offsets.to_host();
  const auto nbr = nl.neighbors.to_host();
  const auto bh = boxes.to_host();
  std::printf("H2: %zu boxes, %d neighbor entries\n", boxes.size(), off.back());
  // each box against both spheres on the host (the outermost radial shells
  // lie beyond the 1s cutoff radius)
  std::vector<int> expect;
  for (const Box &b : bh)
    for (int j = 0; j < 2; ++j) {
      double d2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double c = basis2.O[j][d];
        const double e = c < b.lo[d] ? b.lo[d] - c : (c > b.hi[d] ? c - b.hi[d] : 0.0);
        d2 += e * e;
      }
      if (std::sqrt(d2) <= basis2.cutoff_radii[j]) expect.push_back(j);
    }
  if (off.back() != (int)nbr.size() || nbr != expect) {
    std::printf("FAIL neighbor list against the host test\n");
    return 1;
  }
  DeviceView2DLeft Cs(2, 1);
  DeviceView1D f1(1);
  Cs.from_host(std::vector<double>{std::sqrt(0.3), std::sqrt(0.3)});
  f1.from_host(std::vector<double>{1.0});
  auto sp = compute_gga_lsda_sparse(basis2, grid2, nl, Cs, f1, Cs, f1, Functional({101, 130}));
  auto cs = compute_vxc(grid2, basis2, P2, Functional({101, 130}));
  expect_rel("H2 sparse polarised PBE", sp.energy, cs.energy, 1e-12);
  auto rho = compute_density_sparse(basis2, grid2, nl, Cs, f1).to_host();
  const auto w2 = grid2.weights.to_host();
  double ne = 0.0;
  for (size_t g = 0; g < w2.size(); ++g) ne += w2[g] * rho[g];
  expect_rel("H2 sparse density N_el", 2.0 * ne, cs.n_el, 1e-12);
  std::printf("shim OK\n");
  return 0;
}
