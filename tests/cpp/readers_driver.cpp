// Reads an XYZ file and both basis formats through the C++ shim and prints
// what it read, one value per line, for comparison with the Python readers.
// Host only: needs libnukexc_hip.so for the STO normalisation, no GPU.
#include <cstdio>

#include "nukexc/stobasis.hpp"

using namespace Nukexc;

static void dump(const char *tag, const STOBasisSet &b) {
  for (size_t i = 0; i < b.nbf(); ++i)
    std::printf("%s %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g\n", tag, b.n[i], b.l[i], b.m[i],
                b.zeta[i], b.norm[i], b.cutoff_radii[i], b.O[i][0], b.O[i][1], b.O[i][2]);
}

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: readers_driver file.xyz adf_dir thakkar_dir\n");
    return 2;
  }
  Molecule mol;
  read_xyz(argv[1], mol);
  for (unsigned a = 0; a < mol.natoms; ++a)
    std::printf("atom %u %.17g %.17g %.17g\n", mol.Z[a], mol.atom_centers[a][0],
                mol.atom_centers[a][1], mol.atom_centers[a][2]);
  dump("adf", load_adf_basis(mol, argv[2]));
  dump("fit", load_adf_basis(mol, argv[2], 1e-10, true));
  dump("thakkar", load_thakkar_basis(mol, argv[3]));
  return 0;
}
