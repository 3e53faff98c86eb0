"""SAD guess densities (read_density_from_sad, utils.cpp:21-58, block placement
of InitialGuess_RHF_SAD::guess, rhf.hpp:551-584) through the C ABI, host only.

Pins: the library's blocks equal an independent parse of the same .sad file
(oracle_lib.read_sad) placed on the atom blocks; and, in unit normalisation
(the normalisation the STO-3G .sad densities were generated in), each atom's
block holds its electrons, tr(D_s S) = (4, 4) for O, (3, 3) for C (the
spin-averaged atoms of the file), (1, 0) for H and (2, 1) for Li.
"""
import os
import shutil

import numpy as np
import pytest

import oracle_lib as O
from helpers import methane_water
from dft_gansu_amd import molecule as gm
from dft_gansu_amd._lib import GansuError

SYMBOL = {1: "H", 3: "Li", 6: "C", 8: "O"}


def _assembled(path, Z, sizes):
    M = sum(sizes)
    A, B = np.zeros((M, M)), np.zeros((M, M))
    s0 = 0
    for z, n in zip(Z, sizes):
        a, b = O.read_sad(path[:-4] + ".sad", SYMBOL[int(z)])
        assert a.shape == (n, n)
        A[s0:s0 + n, s0:s0 + n], B[s0:s0 + n, s0:s0 + n] = a, b
        s0 += n
    return A, B


def _functions_per_atom(mol):
    """Cartesian functions of each atom: ncart(L) per contracted shell (one
    basis_index per contracted shell)."""
    out = []
    for i in range(mol.n_atoms):
        sel = mol.prims[mol.prims["atom_index"] == i]
        shells = {int(b): int(l) for b, l in zip(sel["basis_index"], sel["shell_type"])}
        out.append(sum((l + 1) * (l + 2) // 2 for l in shells.values()))
    return out


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g_st"])
def test_sad_blocks_match_independent_parse(basis):
    Z, xyz = methane_water()
    path = gm.basis_path(basis)
    mol = gm.Molecule.from_atoms(Z, xyz, path)
    Da, Db = mol.sad_density()
    A, B = _assembled(path, Z, _functions_per_atom(mol))
    assert np.array_equal(Da, A) and np.array_equal(Db, B)


def test_sad_blocks_hold_the_atom_electrons():
    path = gm.basis_path("sto-3g")
    Z = np.array([8, 1, 1, 6, 3])
    xyz = np.array([[0, 0, 0], [1.8, 0, 0], [0, 1.8, 0], [6, 0, 0], [0, 0, 7.0]], dtype=float)
    mol = gm.Molecule.from_atoms(Z, xyz, path, charge=1, normalization="unit")
    Da, Db = mol.sad_density()
    p, t, n = O.build_molecule(path, Z, xyz, quirk=False)
    S, _, _ = O.oneint(p, n, Z, xyz)
    expect = {8: (4, 4), 1: (1, 0), 6: (3, 3), 3: (2, 1)}
    s0 = 0
    for z in Z:
        a, _ = O.read_sad(path[:-4] + ".sad", SYMBOL[int(z)])
        k = a.shape[0]
        blk = slice(s0, s0 + k)
        got = (np.trace(Da[blk, blk] @ S[blk, blk]), np.trace(Db[blk, blk] @ S[blk, blk]))
        assert np.allclose(got, expect[int(z)], atol=1e-6), (z, got)
        s0 += k
    assert s0 == mol.n_basis
    # nothing outside the atom blocks
    mask = np.zeros_like(Da, dtype=bool)
    s0 = 0
    for z in Z:
        k = O.read_sad(path[:-4] + ".sad", SYMBOL[int(z)])[0].shape[0]
        mask[s0:s0 + k, s0:s0 + k] = True
        s0 += k
    assert not Da[~mask].any() and not Db[~mask].any()


def test_sad_errors(tmp_path):
    """A basis file with no .sad beside it is GANSU_E_IO; an element missing
    from the .sad file is GANSU_E_BASIS (utils.cpp:28-29, :56)."""
    gbs = tmp_path / "sto-3g.gbs"
    shutil.copy(gm.basis_path("sto-3g"), gbs)
    Z, xyz = gm.water(jitter=0.0)
    mol = gm.Molecule.from_atoms(Z, xyz, str(gbs))
    with pytest.raises(GansuError) as e:
        mol.sad_density()
    assert e.value.code == 2 and "SAD file" in str(e.value)
    a, b = O.read_sad(gm.basis_path("sto-3g")[:-4] + ".sad", "H")
    with open(tmp_path / "sto-3g.sad", "w") as fh:
        fh.write("H 1\n%.10e\n%.10e\n" % (a[0, 0], b[0, 0]))
    with pytest.raises(GansuError) as e:
        mol.sad_density()
    assert e.value.code == 3 and "O" in str(e.value)
    assert os.path.exists(tmp_path / "sto-3g.sad")


def test_sad_file_helpers():
    """Element discovery of the .gbs and .sad files (generate_sad_cache.py's
    extract_elements_from_basis_file / cache_has_element)."""
    from dft_gansu_amd import sad
    path = gm.basis_path("sto-3g")
    in_basis = sad.basis_elements(path)
    cached = sad.cached_elements(path[:-4] + ".sad")
    assert in_basis[:3] == ["H", "He", "Li"]
    assert len(cached) == 54 and set(cached) <= set(in_basis)
    for sym in ("H", "O", "Li"):
        assert O.read_sad(path[:-4] + ".sad", sym)[0].ndim == 2


def test_density_restart_needs_both_densities():
    """gansu_rhf_run_guess takes both spin densities or neither (host-side
    argument check, no device work)."""
    from dft_gansu_amd import Parameters
    from dft_gansu_amd._lib import lib, ptr
    Z, xyz = gm.water(jitter=0.0)
    mol = gm.Molecule.from_atoms(Z, xyz, gm.basis_path("sto-3g"))
    Da = np.zeros((mol.n_basis, mol.n_basis))
    en = np.zeros(2)
    rc = lib().gansu_rhf_run_guess(mol.handle, None, Parameters().handle, ptr(Da), None, 1, ptr(en), None, None)
    assert rc == 1 and b"both" in lib().gansu_last_error()
