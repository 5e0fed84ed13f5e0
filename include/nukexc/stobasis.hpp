// nukexc/stobasis.hpp -- STO basis and collocation (stobasis.hpp:54-225,
// :494-584 of the reference).  The basis lives on the host (it is small) and
// is handed to the library when a collocation or XC call needs it.
#pragma once

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

#include "molecule.hpp"
#include "nukexc_config.hpp"

namespace Nukexc {

struct STOBasisSet {
  std::vector<int> n, l, m;
  std::vector<double> norm, zeta, cutoff_radii;
  std::vector<Point> O;
  size_t nbf() const { return O.size(); }
};

struct STOFunc {
  int n, l, m;
  double zeta;
  double ox, oy, oz;
};

// N = (2 zeta)^{n+1/2} / sqrt((2n)!), cutoff radius at cutoff_tol.
inline STOBasisSet make_manual_basis(const std::vector<STOFunc> &funcs,
                                     double cutoff_tol = 1e-10) {
  STOBasisSet b;
  for (const auto &f : funcs) {
    double nrm = 0, cut = 0;
    check(nkx_sto_norm_cutoff(f.n, f.zeta, cutoff_tol, &nrm, &cut));
    b.n.push_back(f.n);
    b.l.push_back(f.l);
    b.m.push_back(f.m);
    b.zeta.push_back(f.zeta);
    b.norm.push_back(nrm);
    b.cutoff_radii.push_back(cut);
    Point p;
    p[0] = f.ox;
    p[1] = f.oy;
    p[2] = f.oz;
    b.O.push_back(p);
  }
  return b;
}

// S, P, D, F, G, H, I -> 0..6 (any case); anything else -1 (stobasis.hpp:67-73).
inline int label_to_l(char label) {
  static const std::string order = "SPDFGHI";
  const size_t pos = order.find((char)std::toupper((unsigned char)label));
  return pos == std::string::npos ? -1 : (int)pos;
}

// ADF basis files (stobasis.hpp:227-289): per atom data_dir/<Symbol>; the rows
// after the first line holding BASIS (FIT when fit) up to the first line
// holding END; a row "nL zeta" adds the functions m = -l..l.  Atoms without a
// file are skipped.
inline STOBasisSet load_adf_basis(const Molecule &mol,
                                  const std::string &data_dir = "input/zorabasis/TZP",
                                  const double cutoff_tol = 1e-10, const bool fit = false) {
  std::vector<STOFunc> funcs;
  const std::string key = fit ? "FIT" : "BASIS";
  for (size_t a = 0; a < mol.natoms; ++a) {
    std::string sym = detail::symbols().at(mol.Z[a]);
    sym[0] = (char)std::toupper((unsigned char)sym[0]);
    std::ifstream file(data_dir + "/" + sym);
    if (!file.is_open()) continue;
    std::string line;
    while (std::getline(file, line))
      if (line.find(key) != std::string::npos) break;
    while (std::getline(file, line)) {
      if (line.find("END") != std::string::npos) break;
      if (line.empty()) continue;
      std::stringstream ss(line);
      std::string label;
      double zeta;
      if (!(ss >> label >> zeta)) continue;
      const int n = std::stoi(label.substr(0, 1)), l = label_to_l(label[1]);
      const Point &c = mol.atom_centers[a];
      for (int m = -l; m <= l; ++m) funcs.push_back({n, l, m, zeta, c[0], c[1], c[2]});
    }
  }
  return make_manual_basis(funcs, cutoff_tol);
}

// Thakkar Roothaan-Hartree-Fock tables (stobasis.hpp:291-371): per atom
// data_dir/<symbol> (lower case); a line whose first word is a shell letter
// sets l; lines holding BASIS or CUSP are headers; a row "nL zeta ..." whose
// label starts with a digit adds the functions m = -l..l.
inline STOBasisSet load_thakkar_basis(const Molecule &mol,
                                      const std::string &data_dir = "input/k99light/neutral",
                                      const double cutoff_tol = 1e-10) {
  std::vector<STOFunc> funcs;
  for (size_t a = 0; a < mol.natoms; ++a) {
    std::string sym = detail::symbols().at(mol.Z[a]);
    sym[0] = (char)std::tolower((unsigned char)sym[0]);
    std::ifstream file(data_dir + "/" + sym);
    if (!file.is_open()) continue;
    std::string line;
    int l = 0;
    while (std::getline(file, line)) {
      if (line.empty()) continue;
      std::stringstream first(line);
      std::string w;
      first >> w;
      if (w.size() == 1 && label_to_l(w[0]) != -1) l = label_to_l(w[0]);
      if (line.find("BASIS") != std::string::npos || line.find("CUSP") != std::string::npos)
        continue;
      std::stringstream ss(line);
      std::string label;
      double zeta;
      if (!(ss >> label >> zeta) || !std::isdigit((unsigned char)label[0])) continue;
      const int n = std::stoi(label.substr(0, 1));
      const Point &c = mol.atom_centers[a];
      for (int m = -l; m <= l; ++m) funcs.push_back({n, l, m, zeta, c[0], c[1], c[2]});
    }
  }
  return make_manual_basis(funcs, cutoff_tol);
}

inline bool same_basis(const STOBasisSet &a, const STOBasisSet &b) {
  if (a.nbf() != b.nbf() || a.n != b.n || a.l != b.l || a.m != b.m || a.zeta != b.zeta ||
      a.norm != b.norm)
    return false;
  for (size_t i = 0; i < a.nbf(); ++i)
    for (int d = 0; d < 3; ++d)
      if (a.O[i][d] != b.O[i][d]) return false;
  return true;
}

// Hands the basis to the thread's ctx; a basis equal to the last one is not
// re-sent, so the ctx keeps its launch plan.
inline void set_basis(const STOBasisSet &b) {
  thread_local STOBasisSet last;
  thread_local bool have = false;
  if (have && same_basis(last, b)) return;
  std::vector<double> c;
  for (const auto &p : b.O)
    for (int d = 0; d < 3; ++d) c.push_back(p[d]);
  check(nkx_set_basis(context().get(), (int)b.nbf(), b.n.data(), b.l.data(), b.m.data(),
                      b.zeta.data(), b.norm.data(), c.data(), b.cutoff_radii.data()),
        context().get());
  last = b;
  have = true;
}

namespace detail {
inline void need_extents(const STOBasisSet &basis, const DeviceArray<Point> &points,
                         const DeviceArray<double> &v, const char *what) {
  if (v.extent(0) != basis.nbf() || v.extent(1) != points.size())
    throw std::runtime_error(std::string(what) + ": output must be (N_bf, N_pts) = (" +
                             std::to_string(basis.nbf()) + ", " +
                             std::to_string(points.size()) + ")");
}
}  // namespace detail

// fill_collocation(space, basis, points, values) (stobasis.hpp:494-532):
// values(mu, g), LayoutLeft (N_bf, N_pts), allocated by the caller; enqueued
// on the space's stream (asynchronous until space.fence(), as the
// reference's parallel_for on the instance).
inline void fill_collocation(ExecSpace &space, const STOBasisSet &basis,
                             const DeviceArray<Point> &points, DeviceArray<double> &values) {
  detail::need_extents(basis, points, values, "fill_collocation");
  set_basis(basis);
  check(nkx_collocation(space.ctx(), (int64_t)points.size(),
                        reinterpret_cast<const double *>(points.data()), values.data(), nullptr,
                        nullptr, nullptr),
        space.ctx());
}

// fill_grad_collocation(space, basis, points, gx, gy, gz)
// (stobasis.hpp:534-584): d/dx, d/dy, d/dz, same layout and ordering.
inline void fill_grad_collocation(ExecSpace &space, const STOBasisSet &basis,
                                  const DeviceArray<Point> &points, DeviceArray<double> &gx,
                                  DeviceArray<double> &gy, DeviceArray<double> &gz) {
  detail::need_extents(basis, points, gx, "fill_grad_collocation");
  detail::need_extents(basis, points, gy, "fill_grad_collocation");
  detail::need_extents(basis, points, gz, "fill_grad_collocation");
  set_basis(basis);
  check(nkx_collocation(space.ctx(), (int64_t)points.size(),
                        reinterpret_cast<const double *>(points.data()), nullptr, gx.data(),
                        gy.data(), gz.data()),
        space.ctx());
}

// Forms without a space: the thread's default instance, fenced on return.
inline void fill_collocation(const STOBasisSet &basis, const DeviceArray<Point> &points,
                             DeviceArray<double> &values) {
  ExecSpace space;
  fill_collocation(space, basis, points, values);
  space.fence();
}

inline void fill_grad_collocation(const STOBasisSet &basis, const DeviceArray<Point> &points,
                                  DeviceArray<double> &gx, DeviceArray<double> &gy,
                                  DeviceArray<double> &gz) {
  ExecSpace space;
  fill_grad_collocation(space, basis, points, gx, gy, gz);
  space.fence();
}

}  // namespace Nukexc
