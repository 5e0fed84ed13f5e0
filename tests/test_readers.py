"""Geometry and basis readers (molecule.hpp:97-125, stobasis.hpp:227-371) on
fixtures: the reference's water.xyz and its Thakkar tables for H and O
(input/k99light/neutral), and two small ADF-format files written for these
tests.  Python and C++ shim readers must agree exactly."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from nukexc_amd import readers as R

G = os.path.join(ROOT, "tests", "golden", "readers")
XYZ = os.path.join(G, "water.xyz")


def test_read_xyz_water():
    xyz, Z = R.read_xyz(XYZ)
    assert Z.tolist() == [8, 1, 1]
    ang = np.array([[0.0, 0.0, 0.11779], [0.0, 0.757, -0.4693], [0.0, -0.757, -0.4693]])
    assert np.array_equal(xyz, ang * 1.8897259885681697)


def test_atomic_number_any_case():
    assert R.atomic_number("cl") == R.atomic_number("Cl") == R.atomic_number("CL") == 17
    assert R.atomic_number("h") == 1
    with pytest.raises(ValueError):
        R.atomic_number("Xx")


def test_label_to_l():
    assert [R.label_to_l(c) for c in "spdfghiK"] == [0, 1, 2, 3, 4, 5, 6, -1]


def test_thakkar_water():
    xyz, Z = R.read_xyz(XYZ)
    b = R.load_thakkar_basis(xyz, Z, os.path.join(G, "k99light"))
    # O: 8 s rows and 7 p rows (x3 m), each H one 1s row with zeta 1
    assert b.nbf() == 8 + 21 + 1 + 1
    assert b.l.tolist() == [0] * 8 + [1] * 21 + [0, 0]
    assert b.n[:8].tolist() == [2, 1, 2, 1, 1, 2, 1, 1]
    assert b.zeta[0] == 24.975698 and b.zeta[8] == 20.797 and b.zeta[-1] == 1.0
    assert b.m[8:11].tolist() == [-1, 0, 1]
    assert np.array_equal(b.O[:29], np.repeat(xyz[:1], 29, 0))
    # the 1s function with zeta 1 is the hydrogen ground state: N = 2
    assert abs(b.norm[-1] - 2.0) < 1e-14


def test_adf_basis_and_fit_blocks():
    xyz, Z = R.read_xyz(XYZ)
    b = R.load_adf_basis(xyz, Z, os.path.join(G, "adf"))
    assert b.nbf() == (3 + 2 * 3 + 5) + 2 * (2 + 3)
    assert b.l.tolist()[:14] == [0, 0, 0] + [1] * 6 + [2] * 5
    assert b.m.tolist()[9:14] == [-2, -1, 0, 1, 2]
    f = R.load_adf_basis(xyz, Z, os.path.join(G, "adf"), fit=True)
    assert f.nbf() == 4 and f.zeta.tolist() == [15.32, 4.1, 4.1, 4.1]  # H has no FIT block


def test_missing_element_file_is_skipped(tmp_path):
    xyz, Z = R.read_xyz(XYZ)
    b = R.load_thakkar_basis(xyz, Z, str(tmp_path))
    assert b.nbf() == 0


def _parse(lines, tag):
    return [tuple(float(x) for x in ln.split()[1:]) for ln in lines if ln.startswith(tag + " ")]


def test_cpp_readers_match_python():
    exe = os.path.join(ROOT, "tests", "cpp", "readers_driver")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "readers_driver"],
                   check=True)
    out = subprocess.run([exe, XYZ, os.path.join(G, "adf"), os.path.join(G, "k99light")],
                         capture_output=True, text=True, check=True).stdout.splitlines()
    xyz, Z = R.read_xyz(XYZ)
    assert _parse(out, "atom") == [(float(z), *map(float, p)) for z, p in zip(Z, xyz)]
    for tag, b in (("adf", R.load_adf_basis(xyz, Z, os.path.join(G, "adf"))),
                   ("fit", R.load_adf_basis(xyz, Z, os.path.join(G, "adf"), fit=True)),
                   ("thakkar", R.load_thakkar_basis(xyz, Z, os.path.join(G, "k99light")))):
        py = [(float(b.n[i]), float(b.l[i]), float(b.m[i]), float(b.zeta[i]), float(b.norm[i]),
               float(b.cutoff_radii[i]), *map(float, b.O[i])) for i in range(b.nbf())]
        assert _parse(out, tag) == py, tag
