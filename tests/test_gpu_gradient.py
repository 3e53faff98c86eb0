"""Analytic RHF nuclear gradient on the GPU (backward kernels through the C ABI)
against the FP64 host oracle and against finite differences of GPU energies."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from dft_gansu_amd import FockEngine, Parameters, molecule as gm, nuclear_repulsion_gradient, oneint_gradient
from dft_gansu_amd import rhf_gradient
from dft_gansu_amd.geometry import gradient_params, optimize_geometry, rhf_energy
from helpers import both, ethane, methane_water, write_even_tempered_aux

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
    # each SCF started from the previous converged density (the default) reaches
    # the same minimum as cold starts, in fewer SCF iterations
    cold = optimize_geometry(mol, gtol=1e-5, warm_start=False)
    assert cold["converged"] and abs(cold["e_total"] - out["e_total"]) < 1e-8
    assert out["scf_iterations"] < cold["scf_iterations"], (out["scf_iterations"], cold["scf_iterations"])


# ------------------------------------------------------------ RI gradient --
def _orbitals(M, nocc, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(M, M)) / np.sqrt(M)
    return C, 2.0 * C[:, :nocc] @ C[:, :nocc].T


@pytest.mark.parametrize("system,aux_basis", [
    ("water", "6-31g_st"),
    ("water", "even-tempered"),
    ("methane_water", "even-tempered"),
    ("water", "cc-pvdz-rifit"),
    ("methane_water", "cc-pvdz-rifit"),
])
def test_ri_two_electron_gradient_matches_oracle(cuda, tmp_path, system, aux_basis):
    """GPU RI-JK gradient for a fixed density against the oracle's.  The
    2-centre weight V^-1 M V^-1 carries cond(V) in its rounding: with
    cc-pVDZ-RIFIT (cond 4e8 on water, 1e10 on methane+water) the oracle itself
    moves by 7..16 cond(V) eps relative between np.linalg.inv(V) and a Cholesky
    solve, so the bound is 50 cond(V) eps relative (floor 1e-10); the 6-31G*
    and even-tempered fitting sets (cond 1e3..4e5, f shells in the latter)
    hold it to ~1e-10."""
    Z, xyz = {"water": lambda: gm.water(jitter=0.01, seed=11), "methane_water": methane_water}[system]()
    mol, (prims, types, norm) = both(Z, xyz, "cc-pvdz")
    if aux_basis == "even-tempered":
        path = write_even_tempered_aux(tmp_path / "aux.gbs")
        aux = gm.Molecule.from_atoms(Z, xyz, path)
        pa, ta, na = O.build_molecule(path, Z, xyz)
    else:
        aux, (pa, ta, na) = both(Z, xyz, aux_basis)
    M, nocc = mol.n_basis, mol.n_electrons // 2
    C, D = _orbitals(M, nocc, 17)
    ref = O.ri_grad_2e(prims, types, norm, pa, na, D, len(Z))
    cond = np.linalg.cond(O.int2c2e(pa, na, O.schwarz_aux(pa), 1e-12))
    tol = max(1e-10, 50.0 * cond * np.finfo(np.float64).eps)
    eng = FockEngine(mol, aux=aux, eri_method="ri")
    eng.precomputation()
    g = eng.gradient_2e(_dev(D), C=_dev(C), nocc=nocc).cpu().numpy()
    scale = max(1.0, np.max(np.abs(ref)))
    print(system, aux_basis, "cond(V) %.2e" % cond, "max |g - ref| / scale = %.2e" % (np.max(np.abs(g - ref)) / scale))
    assert np.max(np.abs(g - ref)) < tol * scale, np.max(np.abs(g - ref))
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9 * scale  # translational invariance
    with pytest.raises(RuntimeError):  # the RI gradient needs the occupied orbitals
        eng.gradient_2e(_dev(D))


def test_ri_rhf_gradient_matches_finite_differences(cuda):
    """RI-JK SCF (cc-pVDZ with the well-conditioned 6-31G* as the fitting basis,
    so that energy differences resolve 1e-6 Eh/bohr): analytic gradient of the
    converged RI energy against central differences of GPU SCF energies."""
    Z, xyz = gm.water(jitter=0.02, seed=3)
    mol, _ = both(Z, xyz, "cc-pvdz")

    def aux_at(x):
        return both(Z, x, "6-31g_st")[0]

    prm = gradient_params(eri_method="ri", convergence_energy_threshold=1e-12)
    r = rhf_gradient(mol, prm, aux=aux_at(xyz))
    assert r["converged"]
    g = r["grad"]
    assert np.max(np.abs(g.sum(axis=0))) < 1e-7
    h = 1e-4
    for a in range(3):
        for d in range(3):
            xp, xm = mol.xyz.copy(), mol.xyz.copy()
            xp[a, d] += h
            xm[a, d] -= h
            ep = rhf_gradient(mol.with_coordinates(xp), prm, aux=aux_at(xp))["e_total"]
            em = rhf_gradient(mol.with_coordinates(xm), prm, aux=aux_at(xm))["e_total"]
            fd = (ep - em) / (2 * h)
            assert abs(fd - g[a, d]) < 1e-6, (a, d, fd, g[a, d])
