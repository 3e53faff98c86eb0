"""Analytic RHF nuclear gradient on the GPU (backward kernels through the C ABI)
against the FP64 host oracle and against finite differences of GPU energies."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from dft_gansu_amd import FockEngine, Parameters, molecule as gm, nuclear_repulsion_gradient, oneint_gradient
from dft_gansu_amd import rhf_gradient
from dft_gansu_amd.geometry import gradient_params, optimize_geometry, rhf_energy
from helpers import both, ethane, methane_water

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda")


def _sym(M, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(M, M)) * 0.2
    return A + A.T


def _dev(a):
    return torch.tensor(a, dtype=torch.float64, device="cuda").contiguous()


@pytest.mark.parametrize("method", ["direct", "stored"])
@pytest.mark.parametrize("system,basis", [("water", "6-31g_st"), ("methane_water", "cc-pvdz"), ("ethane", "6-31g_st")])
def test_two_electron_gradient_matches_oracle(cuda, method, system, basis):
    Z, xyz = {"water": lambda: gm.water(jitter=0.01, seed=11), "methane_water": methane_water,
              "ethane": ethane}[system]()
    mol, (prims, types, norm) = both(Z, xyz, basis)
    D = _sym(mol.n_basis, 3)
    ref = O.grad_2e(prims, types, norm, O.schwarz(prims, types), D, len(Z), 1e-12)
    eng = FockEngine(mol, eri_method=method)
    eng.precomputation()
    g = eng.gradient_2e(_dev(D)).cpu().numpy()
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(g - ref)) < 1e-10 * scale, np.max(np.abs(g - ref))
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9 * scale  # translational invariance


@pytest.mark.parametrize("system,basis", [("water", "sto-3g"), ("methane_water", "cc-pvdz")])
def test_one_electron_gradient_matches_oracle(cuda, system, basis):
    Z, xyz = {"water": lambda: gm.water(jitter=0.01, seed=11), "methane_water": methane_water}[system]()
    mol, (prims, types, norm) = both(Z, xyz, basis)
    D, W = _sym(mol.n_basis, 4), _sym(mol.n_basis, 5)
    ref = O.grad_1e(prims, norm, Z, xyz, D, W)
    g = oneint_gradient(mol, _dev(D), _dev(W)).cpu().numpy()
    assert np.max(np.abs(g - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
    assert np.allclose(nuclear_repulsion_gradient(mol), O.grad_nuc(Z, xyz), rtol=0, atol=1e-12)


def test_rhf_gradient_matches_oracle(cuda):
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, _ = both(Z, xyz, "sto-3g")
    E_ref, g_ref = O.rhf_gradient(gm.basis_path("sto-3g"), Z, xyz, 5)
    r = rhf_gradient(mol, gradient_params())
    assert r["converged"]
    assert abs(r["e_total"] - E_ref) < 1e-9
    assert np.max(np.abs(r["grad"] - g_ref)) < 1e-7


def test_rhf_gradient_matches_finite_differences_c10h22_631gs(cuda):
    """The C10H22/6-31G* workload: central differences (h = 1e-4 bohr) of GPU
    SCF energies on a sample of coordinates, tolerance 1e-6 Eh/bohr."""
    mol, _ = gm.config_molecule(3)
    prm = gradient_params(convergence_energy_threshold=1e-12)
    r = rhf_gradient(mol, prm)
    assert r["converged"]
    g = r["grad"]
    assert np.max(np.abs(g.sum(axis=0))) < 1e-7
    h = 1e-4
    for a, d in [(0, 0), (1, 2), (4, 1), (12, 0), (31, 2)]:
        xp, xm = mol.xyz.copy(), mol.xyz.copy()
        xp[a, d] += h
        xm[a, d] -= h
        ep = rhf_gradient(mol.with_coordinates(xp), prm)["e_total"]
        em = rhf_gradient(mol.with_coordinates(xm), prm)["e_total"]
        fd = (ep - em) / (2 * h)
        assert abs(fd - g[a, d]) < 1e-6, (a, d, fd, g[a, d])


def test_autograd_binding_returns_the_analytic_gradient(cuda):
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, _ = both(Z, xyz, "sto-3g")
    prm = gradient_params()
    x = torch.tensor(mol.xyz, dtype=torch.float64, requires_grad=True)
    e = rhf_energy(x, mol, prm)
    e.backward()
    r = rhf_gradient(mol, prm)
    assert abs(float(e.detach()) - r["e_total"]) < 1e-12
    assert np.max(np.abs(x.grad.numpy() - r["grad"])) < 1e-12


def test_geometry_optimiser_reaches_the_sto3g_water_minimum(cuda):
    Z, xyz = gm.water(jitter=0.05, seed=5)
    mol, _ = both(Z, xyz, "sto-3g")
    out = optimize_geometry(mol, gtol=1e-5)
    assert out["converged"]
    e = [h[0] for h in out["history"]]
    assert all(b <= a + 1e-10 for a, b in zip(e, e[1:]))
    assert abs(out["e_total"] - (-74.965901)) < 2e-6  # RHF/STO-3G water at its minimum
    x = out["mol"].xyz
    roh = [np.linalg.norm(x[k] - x[0]) for k in (1, 2)]
    assert abs(roh[0] - roh[1]) < 1e-3
    assert abs(roh[0] / gm.ANGSTROM_TO_BOHR - 0.989) < 2e-3
