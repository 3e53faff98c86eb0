"""The device-resident RHF iteration (scf_solver=purify): every iteration after
the guess is one captured HIP graph -- X^T F X and the other products on the
FP64 matrix cores, SP2 purification in place of the eigensolver, the
four-centre Fock build, the energy and convergence test and DIIS as device
kernels -- launched once per iteration.

Parity: the SCF energy equals the oracle's diagonalisation-loop energy within
1e-9 Eh (the purified density differs from the eigenvector one only by
rounding at convergence); the density is the S-metric projector
(D S D = 2 D, tr(D S) = N) to 1e-8; the iteration counts of the two loops
agree within one; the final orbitals are S-orthonormal and reproduce D.
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import both, ethane, methane_water
from dft_gansu_amd import Parameters, oneint, rhf, rhf_orbitals
from dft_gansu_amd import molecule as gm

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "h2o_sto3g": lambda: (gm.water(), "sto-3g"),
    "h2o_ccpvdz": lambda: (gm.water(), "cc-pvdz"),
    "ch4h2o_ccpvdz": lambda: (methane_water(), "cc-pvdz"),
    "ethane_631gs": lambda: (ethane(), "6-31g_st"),
}


def _projector_checks(mol, D):
    S = oneint(mol)[0].cpu().numpy()
    n_el = mol.n_electrons
    assert abs(np.trace(D @ S) - n_el) < 1e-8
    assert np.abs(D @ S @ D - 2.0 * D).max() < 1e-8
    assert np.abs(D - D.T).max() < 1e-12


@pytest.mark.parametrize("name,method,diis_x", [("h2o_sto3g", "stored", 0), ("h2o_sto3g", "direct", 1),
                                                ("h2o_ccpvdz", "stored", 0), ("ch4h2o_ccpvdz", "stored", 1),
                                                ("ethane_631gs", "direct", 0)])
def test_purify_scf_matches_oracle(name, method, diis_x, cuda):
    (Z, xyz), basis = SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, basis)
    prm = Parameters(eri_method=method, scf_solver="purify", convergence_energy_threshold=1e-10, maxiter=100,
                     diis_include_transform=diis_x)
    r = rhf(mol, prm)
    _, e_tot, _, _, it_ref = O.rhf(p, t, n, Z, xyz, mol.n_electrons // 2, thr=1e-12, conv=1e-10,
                                   diis_include_transform=bool(diis_x))
    assert r["converged"]
    assert abs(r["e_total"] - e_tot) < 1e-9, (r["e_total"], e_tot)
    # one graph launch per iteration, the converging one included
    assert r["graph_launches"] == r["iterations"] + 1
    assert abs(r["iterations"] - it_ref) <= 1, (r["iterations"], it_ref)
    _projector_checks(mol, r["D"])


def test_purify_matches_diagonalisation_config1(cuda):
    """Full-size C10H22/cc-pVDZ (stored ERIs): the graph loop reaches the
    diagonalisation loop's energy and density."""
    mol, _ = gm.config_molecule(1)
    base = dict(eri_method="stored", convergence_energy_threshold=1e-10)
    rd = rhf(mol, Parameters(**base))
    rp = rhf(mol, Parameters(scf_solver="purify", **base))
    assert rd["converged"] and rp["converged"]
    assert abs(rp["e_total"] - rd["e_total"]) < 1e-9
    assert abs(rp["iterations"] - rd["iterations"]) <= 1
    assert rp["graph_launches"] == rp["iterations"] + 1 and rd["graph_launches"] == 0
    assert np.abs(rp["D"] - rd["D"]).max() < 1e-6
    _projector_checks(mol, rp["D"])


def test_purify_orbitals(cuda):
    """The final orbitals of the graph loop (one eigensolve of the last
    iteration's X^T F X): S-orthonormal, ascending, 2 C_occ C_occ^T = D."""
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, _ = both(Z, xyz, "cc-pvdz")
    prm = Parameters(convergence_energy_threshold=1e-10, scf_solver="purify")
    r = rhf_orbitals(mol, prm)
    assert r["converged"]
    C, eps = r["C"], r["eps"]
    S = oneint(mol)[0].cpu().numpy()
    assert np.abs(C.T @ S @ C - np.eye(mol.n_basis)).max() < 1e-10
    assert np.all(np.diff(eps) >= 0)
    D = rhf(mol, prm)["D"]
    assert np.abs(2.0 * C[:, :5] @ C[:, :5].T - D).max() < 1e-6


def test_purify_gradient_matches_diagonalisation(cuda):
    """The analytic gradient after the graph loop equals the one after the
    diagonalisation loop (same converged density and Fock matrix)."""
    from dft_gansu_amd import rhf_gradient
    Z, xyz = ethane()
    mol, _ = both(Z, xyz, "6-31g_st")
    base = dict(eri_method="stored", convergence_energy_threshold=1e-11)
    gd = rhf_gradient(mol, Parameters(**base))
    gp = rhf_gradient(mol, Parameters(scf_solver="purify", **base))
    assert abs(gd["e_total"] - gp["e_total"]) < 1e-9
    assert np.abs(gd["grad"] - gp["grad"]).max() < 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ri_purify_matches_diagonalisation(precision, cuda):
    """RI engines with scf_solver=purify: after the first iteration the
    occupied orbitals come from SP2 purification of X^T F X and a Cholesky QR
    of the previous orbitals projected by it, with no eigensolver in the loop
    (the RI exchange needs an occupied factor, not eigenvectors).  The energy
    equals the diagonalisation loop's within 1e-9 Eh (FP64; 1e-8 Eh for the
    fp16-increment builds, whose rounding follows the orbital sequence), the
    iteration counts agree within one, the density is the S-metric projector
    and the final orbitals (one eigensolve of the last X^T F X) reproduce D."""
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    base = dict(eri_method="ri", ri_k="occupied", precision=precision, convergence_energy_threshold=1e-10)
    rd = rhf(mol, Parameters(**base), aux=aux)
    rp = rhf(mol, Parameters(scf_solver="purify", **base), aux=aux)
    assert rd["converged"] and rp["converged"]
    print("%s: E_diag %.12f, E_purify %.12f, iterations %d / %d" % (precision, rd["e_total"], rp["e_total"],
                                                                   rd["iterations"], rp["iterations"]))
    assert abs(rp["e_total"] - rd["e_total"]) < (1e-9 if precision == "fp64" else 1e-8)
    assert abs(rp["iterations"] - rd["iterations"]) <= 1
    assert rp["graph_launches"] == 0
    _projector_checks(mol, rp["D"])
    ro = rhf_orbitals(mol, Parameters(scf_solver="purify", **base), aux=aux)
    C, nocc = ro["C"], mol.n_electrons // 2
    S = oneint(mol)[0].cpu().numpy()
    assert np.abs(C.T @ S @ C - np.eye(mol.n_basis)).max() < 1e-10
    assert np.abs(2.0 * C[:, :nocc] @ C[:, :nocc].T - rp["D"]).max() < 1e-6
