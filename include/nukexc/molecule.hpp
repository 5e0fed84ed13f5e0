// nukexc/molecule.hpp -- Molecule (molecule.hpp:50-91 of the reference):
// centres in bohr and nuclear charges, held on the host.
#pragma once

#include <vector>

#include "nukexc_config.hpp"

namespace Nukexc {

struct Molecule {
  std::vector<Point> atom_centers;
  std::vector<unsigned> Z;
  unsigned natoms = 0;

  Molecule() = default;
  Molecule(const std::vector<std::vector<double>> &xyz, const std::vector<unsigned> &charges)
      : Z(charges), natoms((unsigned)charges.size()) {
    if (xyz.size() != charges.size())
      throw std::runtime_error("Molecule: coordinate and charge counts differ");
    for (const auto &c : xyz) {
      if (c.size() != 3) throw std::runtime_error("Molecule: coordinates need 3 entries");
      Point p;
      p[0] = c[0];
      p[1] = c[1];
      p[2] = c[2];
      atom_centers.push_back(p);
    }
  }
};

}  // namespace Nukexc
