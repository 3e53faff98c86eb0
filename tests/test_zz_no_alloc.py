"""gansu_oneint_noalloc: the one-electron integrals without device
allocation (the shell tables and Boys grid go into the library's file-scope
device array) against gansu_oneint, which allocates its tables: bit-equal
S, T, V on two calls in a row, the second with a smaller molecule (so the
array holds the first call's larger tables underneath), and a molecule
whose tables exceed the array is refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from dft_gansu_amd import GansuError, molecule as gm
from dft_gansu_amd._lib import lib, ptr
from dft_gansu_amd.engine import oneint

pytestmark = pytest.mark.gpu


def _noalloc(mol):
    M = mol.n_basis
    S, T, V = (torch.full((M, M), np.nan, dtype=torch.float64, device="cuda") for _ in range(3))
    rc = lib().gansu_oneint_noalloc(mol.handle, ptr(S), ptr(T), ptr(V), None)
    if rc:
        raise GansuError(rc, lib().gansu_last_error().decode())
    return [x.cpu().numpy() for x in (S, T, V)]


def test_oneint_noalloc_matches_oneint(cuda):
    big = gm.config_molecule(1)[0]  # C10H22/cc-pVDZ, 330 primitives
    Z, xyz = gm.water()
    small = gm.Molecule.from_atoms(Z, xyz, gm.basis_path("sto-3g"))
    for mol in (big, small):
        ref = [x.cpu().numpy() for x in oneint(mol)]
        got = _noalloc(mol)
        for a, b, name in zip(got, ref, "STV"):
            assert np.array_equal(a, b), (mol.n_basis, name, np.abs(a - b).max())
    # the refusal comes before any device work: output pointers are not touched
    n = 4000  # waters with STO-3G: 60,000 primitives, tables beyond 4 MiB
    rng = np.random.default_rng(1)
    Zs, xs = [], []
    for k in range(n):
        zw, xw = gm.water(jitter=0.0)
        Zs.extend(zw)
        xs.append(xw + np.array([(k % 12) * 6.0, ((k // 12) % 12) * 6.0, (k // 144) * 6.0]))
    huge = gm.Molecule.from_atoms(np.array(Zs), np.vstack(xs), gm.basis_path("sto-3g"))
    dummy = ctypes.c_void_p(16)
    rc = lib().gansu_oneint_noalloc(huge.handle, dummy, dummy, dummy, None)
    assert rc == 1 and b"beyond" in lib().gansu_last_error()
    del rng
