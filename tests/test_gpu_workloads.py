"""The full-size workloads on the GPU: C24H12/cc-pVDZ + cc-pVDZ-RIFIT,
64 waters/cc-pVDZ + cc-pVDZ-RIFIT and the C10H22/6-31G* gradient.

At these sizes the FP64 host oracle is too slow for a full SCF (C10H22/6-31G*
aside, whose oracle gradient is a committed fixture), so parity is checked
through properties that do not depend on size: the stored and direct
four-centre builds agree, mixed precision moves SCF energies by at most
1 micro-Eh, gradients are translationally invariant and match central
differences of GPU SCF energies (h = 1e-4 bohr, 1e-6 Eh/bohr).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import random_density
from dft_gansu_amd import FockEngine, Parameters, rhf, rhf_gradient
from dft_gansu_amd import molecule as gm
from dft_gansu_amd.geometry import gradient_params

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _torch(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


# ---------------------------------------------------- C24H12 (cfg 2) -------
def test_c24h12_stored_equals_direct(cuda):
    """C24H12/cc-pVDZ (M = 420 Cartesian, every class up to (dd|dd)): stored and
    direct J/K agree to 1e-11 of max|J|, max|K| for a random density."""
    mol, _ = gm.config_molecule(2)
    assert mol.n_basis == 420
    D = _torch(random_density(mol.n_basis, seed=7), cuda)
    res = {}
    for m in ("direct", "stored"):
        eng = FockEngine(mol, eri_method=m)
        eng.precomputation()
        res[m] = [x.cpu().numpy() for x in eng.compute_jk(D)]
        del eng
    for a, b in zip(res["direct"], res["stored"]):
        err = np.abs(a - b).max() / np.abs(a).max()
        print("C24H12 stored vs direct rel err %.2e" % err)
        assert err <= 1e-11
        np.testing.assert_allclose(a, a.T, atol=1e-12 * np.abs(a).max())


def test_c24h12_scf_mixed_vs_fp64_and_fitting_error(cuda):
    """C24H12: the stored build with FP32 storage of quartets bounded below
    1e-6, and the RI build with precision=mixed, each within 1 micro-Eh of
    their FP64 SCF energies.  The RI energy differs from the four-centre
    energy by the fitting error of cc-pVDZ-RIFIT (an MP2 fitting set used for
    HF), reported here and bounded by 30 mEh; the 1 micro-Eh claims are
    mixed against FP64 within one method, not RI against four-centre."""
    mol, aux = gm.config_molecule(2)
    e = {}
    for key, kw, a in (("stored_fp64", dict(eri_method="stored"), None),
                       ("stored_mixed", dict(eri_method="stored", precision="mixed"), None),
                       ("ri_fp64", dict(eri_method="ri"), aux),
                       ("ri_mixed", dict(eri_method="ri", precision="mixed"), aux)):
        r = rhf(mol, Parameters(convergence_energy_threshold=1e-10, maxiter=100, **kw), aux=a)
        assert r["converged"], key
        e[key] = r["e_total"]
    d_stored = e["stored_mixed"] - e["stored_fp64"]
    d_ri = e["ri_mixed"] - e["ri_fp64"]
    fit = e["ri_fp64"] - e["stored_fp64"]
    print("C24H12 E: %r; |dE| stored %.2e, RI %.2e; RI fitting error %.4e Eh" % (e, abs(d_stored), abs(d_ri), fit))
    assert abs(d_stored) <= 1e-6 and abs(d_ri) <= 1e-6
    assert 1e-4 < abs(fit) < 3e-2


# ------------------------------------------------- 64 waters (cfg 4) -------
def _water_box_aux(Z, xyz):
    return gm.Molecule.from_atoms(Z, xyz, gm.basis_path("cc-pvdz-rifit"))


def test_64water_mixed_ri_energy_gradient(cuda):
    """64 H2O/cc-pVDZ + cc-pVDZ-RIFIT (M = 1600, Maux = 6144): the mixed
    RI-JK SCF energy within 1 micro-Eh of the FP64 RI energy; its analytic
    gradient is translationally invariant and matches central differences of
    mixed RI SCF energies (h = 1e-4 bohr, each SCF started from the converged
    density) on two coordinates within 1e-6 Eh/bohr."""
    mol, aux = gm.config_molecule(4)
    assert (mol.n_basis, aux.n_basis) == (1600, 6144)
    p64 = Parameters(eri_method="ri", convergence_energy_threshold=1e-10, maxiter=100)
    r64 = rhf(mol, p64, aux=aux)
    pmx = gradient_params(eri_method="ri", precision="mixed", convergence_energy_threshold=1e-11)
    rmx = rhf_gradient(mol, pmx, aux=aux)
    assert r64["converged"] and rmx["converged"]
    dE = rmx["e_total"] - r64["e_total"]
    print("64 waters: E fp64 %.10f mixed %.10f |dE| %.2e" % (r64["e_total"], rmx["e_total"], abs(dE)))
    assert abs(dE) <= 1e-6
    g = rmx["grad"]
    assert np.max(np.abs(g.sum(axis=0))) < 1e-7
    D = rmx["D"]
    h = 1e-4
    for a, d in [(0, 2), (100, 0)]:
        e2 = []
        for s in (1.0, -1.0):
            x = mol.xyz.copy()
            x[a, d] += s * h
            r = rhf(mol.with_coordinates(x), pmx, aux=_water_box_aux(mol.Z, x), density=(0.5 * D, 0.5 * D),
                    force_density=True)
            assert r["converged"]
            e2.append(r["e_total"])
        fd = (e2[0] - e2[1]) / (2 * h)
        print("64 waters: atom %d dir %d analytic %.9f fd %.9f" % (a, d, g[a, d], fd))
        assert abs(fd - g[a, d]) < 1e-6, (a, d, fd, g[a, d])


# --------------------------------------------- C10H22 / 6-31G* (cfg 3) -----
@pytest.fixture(scope="module")
def c10h22_631gs(cuda):
    mol, _ = gm.config_molecule(3)
    prm = gradient_params(convergence_energy_threshold=1e-12)
    r = rhf_gradient(mol, prm)
    assert r["converged"]
    return mol, prm, r


def test_c10h22_gradient_matches_host_oracle(c10h22_631gs):
    """The GPU analytic gradient against the FP64 host oracle's (committed
    fixture written by tests/golden/make_c10h22_631gs_gradient.py from the
    oracle's SCF, one-electron, two-electron and nuclear gradient routines):
    energy within 1e-9 Eh, every component within 1e-6 Eh/bohr."""
    mol, _, r = c10h22_631gs
    with open(os.path.join(GOLDEN, "c10h22_631gs_gradient.json")) as fh:
        ref = json.load(fh)
    digest = hashlib.sha256(np.ascontiguousarray(mol.Z, dtype=np.int64).tobytes() +
                            np.ascontiguousarray(mol.xyz, dtype=np.float64).tobytes()).hexdigest()
    assert digest == ref["geometry_sha256"]
    g_ref = np.array(ref["gradient"])
    err = np.abs(r["grad"] - g_ref).max()
    print("C10H22/6-31G*: |E - E_oracle| %.2e, max|g - g_oracle| %.2e" % (abs(r["e_total"] - ref["e_total"]), err))
    assert abs(r["e_total"] - ref["e_total"]) < 1e-9
    assert err < 1e-6


@pytest.mark.parametrize("chunk", range(4))
def test_c10h22_gradient_matches_finite_differences(c10h22_631gs, chunk):
    """Central differences (h = 1e-4 bohr) of GPU SCF energies on all 96
    coordinates (24 per case), each SCF started from the converged density,
    within 1e-6 Eh/bohr of the analytic gradient."""
    mol, prm, r = c10h22_631gs
    g, D = r["grad"], r["D"]
    h = 1e-4
    worst = 0.0
    for k in range(24 * chunk, 24 * (chunk + 1)):
        a, d = divmod(k, 3)
        e2 = []
        for s in (1.0, -1.0):
            x = mol.xyz.copy()
            x[a, d] += s * h
            rr = rhf(mol.with_coordinates(x), prm, density=(0.5 * D, 0.5 * D), force_density=True)
            assert rr["converged"]
            e2.append(rr["e_total"])
        fd = (e2[0] - e2[1]) / (2 * h)
        worst = max(worst, abs(fd - g[a, d]))
        assert abs(fd - g[a, d]) < 1e-6, (a, d, fd, g[a, d])
    print("C10H22/6-31G* coordinates %d..%d: max |fd - analytic| %.2e" % (24 * chunk, 24 * chunk + 23, worst))
