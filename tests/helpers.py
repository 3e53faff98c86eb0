"""Shared test helpers: molecules for both the product and the oracle."""
import numpy as np

import oracle_lib as O
from dft_gansu_amd import molecule as gm
from dft_gansu_amd.molecule import Molecule, basis_path


def ethane():
    """Staggered ethane from the alkane generator (n_carbon = 2)."""
    return gm.alkane(2, jitter=0.02, seed=7)


def methane_water():
    """A small two-molecule system with s, p and d shells."""
    Zw, w = gm.water(jitter=0.0)
    Z = np.concatenate([Zw, [6, 1, 1, 1, 1]])
    c = np.array([3.2, 0.4, -0.3]) * gm.ANGSTROM_TO_BOHR
    d = 1.09 * gm.ANGSTROM_TO_BOHR / np.sqrt(3.0)
    h = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) * d
    return Z, np.vstack([w, c, c + h])


def both(Z, xyz, basis, normalization="reference"):
    """(product Molecule, oracle (prims, types, norm)) for the same input."""
    path = basis_path(basis)
    mol = Molecule.from_atoms(Z, xyz, path, normalization=normalization)
    ora = O.build_molecule(path, Z, xyz, quirk=(normalization == "reference"))
    return mol, ora


def random_density(M, seed=3):
    """A symmetric positive semidefinite density-like matrix."""
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(M, max(1, M // 3))) / np.sqrt(M)
    return 2.0 * C @ C.T


def write_even_tempered_aux(path):
    """A small even-tempered fitting basis for H, C, O with s, p, d and f
    (C, O) shells, ratio 4 between exponents, so that (P|Q) stays well
    conditioned (cond ~ 1e3-1e4) while every auxiliary angular momentum the
    RI path supports appears.  Returns the path."""
    sets = {"H": [("S", [4.0, 1.0, 0.25]), ("P", [1.6, 0.4]), ("D", [1.0])],
            "C": [("S", [40.0, 10.0, 2.5, 0.625]), ("P", [8.0, 2.0, 0.5]), ("D", [3.0, 0.75]), ("F", [1.5])],
            "O": [("S", [64.0, 16.0, 4.0, 1.0]), ("P", [12.0, 3.0, 0.75]), ("D", [4.0, 1.0]), ("F", [2.0])]}
    lines = ["! even-tempered test fitting basis", ""]
    for el, shells in sets.items():
        lines.append(f"{el}     0")
        for L, exps in shells:
            for e in exps:
                lines += [f"{L}    1   1.00", f"      {e:.8f}             1.0000000"]
        lines.append("****")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return str(path)
