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
