// nukexc/molecule.hpp -- Molecule and read_xyz (molecule.hpp:50-125 of the
// reference): centres in bohr and nuclear charges, held on the host.
#pragma once

#include <cctype>
#include <fstream>
#include <string>
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

namespace detail {
constexpr double ang_to_bohr = 1.8897259885681697;  // atomic_properties.hpp:49

inline const std::vector<std::string> &symbols() {
  static const std::vector<std::string> s = {
      "",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si",
      "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
      "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru",
      "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr",
      "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
      "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};
  return s;
}

// Symbol in any capitalisation -> Z (atomic_properties.hpp:65-96).
inline unsigned get_atomic_number(const std::string &symbol) {
  std::string key = symbol;
  for (char &c : key) c = (char)std::toupper((unsigned char)c);
  const auto &s = symbols();
  for (size_t z = 1; z < s.size(); ++z) {
    std::string u = s[z];
    for (char &c : u) c = (char)std::toupper((unsigned char)c);
    if (u == key) return (unsigned)z;
  }
  throw std::runtime_error("Unknown element symbol in geometry: '" + symbol + "'");
}
}  // namespace detail

// molecule.hpp:97-125: the atom count, a comment line, then "symbol x y z"
// rows in Angstrom, read until a row does not parse.
inline void read_xyz(const std::string &filename, Molecule &mol) {
  std::ifstream file(filename);
  if (!file.is_open()) throw std::runtime_error("Could not open file: " + filename);
  unsigned natoms = 0;
  if (!(file >> natoms)) return;
  std::string line;
  std::getline(file, line);  // rest of the count line
  std::getline(file, line);  // comment
  std::vector<std::vector<double>> xyz;
  std::vector<unsigned> Z;
  std::string sym;
  double x, y, z;
  while (file >> sym >> x >> y >> z) {
    Z.push_back(detail::get_atomic_number(sym));
    xyz.push_back({x * detail::ang_to_bohr, y * detail::ang_to_bohr, z * detail::ang_to_bohr});
  }
  mol = Molecule(xyz, Z);
}

}  // namespace Nukexc
