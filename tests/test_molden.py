"""Molden output (rhf.cu:363-425 layout) and the RHF orbitals it is written from."""
import numpy as np
import pytest

import oracle_lib as O
from dft_gansu_amd import molecule as gm
from dft_gansu_amd import write_molden
from helpers import both


def _parse(path):
    lines = open(path).read().split("\n")
    sec = {}
    cur = None
    for ln in lines:
        if ln.startswith("["):
            cur = ln.split("]")[0] + "]"
            sec[cur] = []
        elif cur:
            sec[cur].append(ln)
    return lines, sec


def test_molden_layout_and_round_trip(tmp_path):
    """Water/STO-3G with synthetic orbitals: header, atoms in bohr, one [GTO]
    entry per contracted shell with the parser's exponents and raw contraction
    coefficients, and every MO coefficient back to 1e-15 (17 digits)."""
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, (prims, types, norm) = both(Z, xyz, "sto-3g")
    M = mol.n_basis
    rng = np.random.default_rng(3)
    C = rng.normal(size=(M, M))
    eps = np.sort(rng.normal(size=M))
    occ = np.array([2.0] * 5 + [0.0] * (M - 5))
    path = tmp_path / "w.molden"
    write_molden(mol, C, eps, occ, path)
    lines, sec = _parse(path)
    assert lines[0] == "[Molden Format]" and lines[1] == "generated by GANSU" and lines[2] == "[Atoms] (AU)"
    atoms = [ln.split() for ln in sec["[Atoms]"] if ln.strip()]
    assert [a[0] for a in atoms] == ["O", "H", "H"] and [int(a[2]) for a in atoms] == [8, 1, 1]
    np.testing.assert_allclose(np.array([[float(v) for v in a[3:6]] for a in atoms]), xyz, rtol=1e-5, atol=1e-5)
    # [GTO]: atom headers "i 0", shell lines " s n 1.00", primitive lines "\texp coef"
    shells, prims_out = [], []
    for ln in sec["[GTO]"]:
        if ln.startswith("\t"):
            prims_out.append(tuple(float(v) for v in ln.split()))
        elif ln.startswith(" ") and ln.split():
            shells.append((ln.split()[0], int(ln.split()[1])))
    assert shells == [("s", 3), ("s", 3), ("p", 3), ("s", 3), ("s", 3)]
    ref = sorted((float(p["exponent"]), float(p["coefficient"])) for p in prims)
    got = sorted(prims_out)
    assert len(got) == len(ref)
    np.testing.assert_allclose(np.array(got), np.array(ref), rtol=1e-5)
    # [MO]
    mo = sec["[MO]"]
    blocks = [i for i, ln in enumerate(mo) if ln == "Sym= A"]
    assert len(blocks) == M
    for k, b in enumerate(blocks):
        assert mo[b + 2] == "Spin= Alpha"
        assert abs(float(mo[b + 1].split()[1]) - eps[k]) <= 1e-5 * max(1.0, abs(eps[k]))
        assert float(mo[b + 3].split()[1]) == occ[k]
        vals = np.array([float(mo[b + 4 + j].split()[1]) for j in range(M)])
        np.testing.assert_allclose(vals, C[:, k], rtol=1e-15, atol=0)


@pytest.mark.gpu
def test_rhf_orbitals_and_molden(tmp_path):
    """The RHF driver's final orbitals: S-orthonormal, ascending energies, and
    2 C_occ C_occ^T equal to the converged density; written to Molden."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dft_gansu_amd import Parameters, oneint, rhf, rhf_orbitals
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, _ = both(Z, xyz, "cc-pvdz")
    prm = Parameters(convergence_energy_threshold=1e-10)
    r = rhf_orbitals(mol, prm)
    assert r["converged"]
    C, eps = r["C"], r["eps"]
    S = oneint(mol)[0].cpu().numpy()
    assert np.abs(C.T @ S @ C - np.eye(mol.n_basis)).max() < 1e-10
    assert np.all(np.diff(eps) >= 0)
    D = rhf(mol, prm)["D"]
    assert np.abs(2.0 * C[:, :5] @ C[:, :5].T - D).max() < 1e-6
    occ = np.array([2.0] * 5 + [0.0] * (mol.n_basis - 5))
    path = tmp_path / "h2o.molden"
    write_molden(mol, C, eps, occ, path)
    assert open(path).read().count("Sym= A") == mol.n_basis
