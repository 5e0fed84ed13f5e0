"""Geometry and STO basis readers (molecule.hpp:97-125, stobasis.hpp:227-371
of the reference): XYZ files in Angstrom, ADF basis files (BASIS / FIT
blocks of "nL zeta" rows) and Thakkar's Roothaan-Hartree-Fock tables (rows
"nL zeta coefficients..." under single-letter shell headers).  Host code;
the resulting basis goes to the GPU through Context.set_basis."""
import os

import numpy as np

from .api import make_manual_basis

ANG_TO_BOHR = 1.8897259885681697  # atomic_properties.hpp:49

SYMBOLS = [
    "", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S",
    "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge",
    "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
    "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd",
    "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg",
    "Tl", "Pb", "Bi", "Po", "At", "Rn"]
_Z = {s.upper(): z for z, s in enumerate(SYMBOLS) if s}


def atomic_number(symbol):
    """Element symbol in any capitalisation -> Z (atomic_properties.hpp:65-96);
    unknown symbols raise."""
    z = _Z.get(symbol.upper())
    if z is None:
        raise ValueError(f"Unknown element symbol in geometry: '{symbol}'")
    return z


def label_to_l(c):
    """S, P, D, F, G, H, I -> 0..6 (any case); anything else -1
    (stobasis.hpp:67-73)."""
    pos = "SPDFGHI".find(c.upper())
    return pos


def read_xyz(path):
    """XYZ file -> (xyz in bohr (natoms, 3), Z): the atom count, a comment
    line, then "symbol x y z" rows in Angstrom, read until the first row that
    does not parse (molecule.hpp:97-125)."""
    with open(path) as f:
        text = f.read()
    # the count is the first token (leading blank lines skipped); the rest of
    # its line and the next line (the comment) are dropped
    head = text.lstrip()
    first, _, rest = head.partition("\n")
    try:
        int(first.split()[0])
    except (IndexError, ValueError):
        return np.zeros((0, 3)), np.zeros(0, dtype=np.int32)
    toks = rest.partition("\n")[2].split()
    xyz, Z = [], []
    for i in range(0, len(toks) - 3, 4):
        sym = toks[i]
        try:
            x, y, z = (float(t) for t in toks[i + 1:i + 4])
        except ValueError:
            break
        Z.append(atomic_number(sym))
        xyz.append([x * ANG_TO_BOHR, y * ANG_TO_BOHR, z * ANG_TO_BOHR])
    return np.array(xyz, dtype=np.float64).reshape(-1, 3), np.array(Z, dtype=np.int32)


def _rows(path):
    try:
        with open(path) as f:
            return f.read().split("\n")
    except OSError:
        return None  # elements without a file are skipped, as in the reference


def _label_zeta(line):
    t = line.split()
    if len(t) < 2:
        return None
    try:
        return t[0], float(t[1])
    except ValueError:
        return None


def load_adf_basis(xyz, Z, data_dir, cutoff_tol=1e-10, fit=False):
    """ADF basis (stobasis.hpp:227-289): per atom the file data_dir/<Symbol>;
    rows after the first line holding BASIS (FIT when fit) up to the first
    line holding END; row "nL zeta" adds the 2l + 1 functions m = -l..l."""
    funcs = []
    key = "FIT" if fit else "BASIS"
    for a, z in enumerate(Z):
        rows = _rows(os.path.join(data_dir, SYMBOLS[z][:1].upper() + SYMBOLS[z][1:]))
        if rows is None:
            continue
        it = iter(rows)
        for line in it:
            if key in line:
                break
        for line in it:
            if "END" in line:
                break
            if not line:
                continue
            lz = _label_zeta(line)
            if lz is None:
                continue
            lab, zeta = lz
            n, l = int(lab[0]), label_to_l(lab[1])
            funcs += [(n, l, m, zeta, *xyz[a]) for m in range(-l, l + 1)]
    return make_manual_basis(funcs, cutoff_tol)


def load_thakkar_basis(xyz, Z, data_dir, cutoff_tol=1e-10):
    """Thakkar RHF tables (stobasis.hpp:291-371): per atom the file
    data_dir/<symbol> (lower case); a line whose first word is a shell letter
    sets l; rows "nL zeta ..." whose label starts with a digit add the 2l + 1
    functions; lines holding BASIS or CUSP are headers."""
    funcs = []
    for a, z in enumerate(Z):
        rows = _rows(os.path.join(data_dir, SYMBOLS[z][:1].lower() + SYMBOLS[z][1:]))
        if rows is None:
            continue
        l = 0
        for line in rows:
            if not line:
                continue
            w = line.split()
            if w and len(w[0]) == 1 and label_to_l(w[0]) != -1:
                l = label_to_l(w[0])
            if "BASIS" in line or "CUSP" in line:
                continue
            lz = _label_zeta(line)
            if lz is None or not lz[0][0].isdigit():
                continue
            funcs += [(int(lz[0][0]), l, m, lz[1], *xyz[a]) for m in range(-l, l + 1)]
    return make_manual_basis(funcs, cutoff_tol)
