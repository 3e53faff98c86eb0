"""GPU parity of the Fock-build path against the FP64 host oracle.

Tolerances: integer/index work (pair order, screening masks) bit-exact;
FP64 integrals and J/K within 1e-10 of max|J|, max|K| (summation order of
FP64 atomics differs from the oracle's deterministic sums); SCF energies
within 1e-9 Eh.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from helpers import both, ethane, methane_water, random_density
from dft_gansu_amd import FockEngine, Parameters, oneint, rhf
from dft_gansu_amd import molecule as gm

pytestmark = pytest.mark.gpu

THR = 1e-12


def _torch(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=cuda)


SYSTEMS = {
    "h2o_sto3g": lambda: (gm.water(), "sto-3g"),
    "h2o_ccpvdz": lambda: (gm.water(), "cc-pvdz"),
    "ch4h2o_ccpvdz": lambda: (methane_water(), "cc-pvdz"),
    "ethane_631gs": lambda: (ethane(), "6-31g_st"),
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_schwarz_factors_and_masks(name, cuda):
    (Z, xyz), basis = SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, basis)
    eng = FockEngine(mol, eri_method="direct")
    eng.precomputation()
    q_gpu = eng.schwarz_factors()
    q_ref = O.schwarz(p, t)
    assert q_gpu.shape == q_ref.shape
    # Q = sqrt(K^2 I): where Q^2 is below the smallest normal double, the
    # oracle's products (K^2, pref K^2, ...) pass through the subnormal range
    # and keep only a few significant digits (measured: 1e-4 relative at
    # Q = 1e-161); the kernel takes K out of the root.  Those pairs lie more
    # than 140 orders of magnitude below the screening threshold, so there
    # both must merely vanish at that scale
    normal = q_ref ** 2 >= np.finfo(np.float64).tiny
    np.testing.assert_allclose(q_gpu[normal], q_ref[normal], rtol=1e-13, atol=1e-300)
    assert np.all(np.abs(q_gpu[~normal] - q_ref[~normal]) <= 1e-150)
    # screening mask Q_x Q_y >= thr: bit-exact outside a 1e-12 relative band
    prod_g = np.outer(q_gpu, q_gpu)
    prod_r = np.outer(q_ref, q_ref)
    band = np.abs(prod_r / THR - 1.0) < 1e-12
    assert np.array_equal((prod_g >= THR)[~band], (prod_r >= THR)[~band])
    assert band.sum() == 0


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("method", ["direct", "stored"])
def test_jk_matches_oracle(name, method, cuda):
    (Z, xyz), basis = SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, basis)
    M = mol.n_basis
    D = random_density(M)
    eng = FockEngine(mol, eri_method=method)
    eng.precomputation()
    J, K = eng.compute_jk(_torch(D, cuda))
    J, K = J.cpu().numpy(), K.cpu().numpy()
    Q = O.schwarz(p, t)
    Jr, Kr = O.jk(p, t, n, Q, D, THR)
    assert np.abs(J - Jr).max() <= 1e-10 * max(1.0, np.abs(Jr).max())
    assert np.abs(K - Kr).max() <= 1e-10 * max(1.0, np.abs(Kr).max())
    np.testing.assert_allclose(J, J.T, atol=1e-13 * np.abs(J).max())


def test_fock_rhf_matches_dense_oracle(cuda):
    (Z, xyz), basis = SYSTEMS["h2o_ccpvdz"]()
    mol, (p, t, n) = both(Z, xyz, basis)
    M = mol.n_basis
    D = random_density(M, seed=11)
    H = random_density(M, seed=12) - 1.0
    eng = FockEngine(mol, eri_method="stored")
    eng.precomputation()
    F = eng.compute_fock_matrix(_torch(D, cuda), _torch(H, cuda)).cpu().numpy()
    Q = O.schwarz(p, t)
    eri = O.eri_dense(p, t, n, Q, THR)
    F_ref = np.zeros((M, M))
    O.lib().orc_fock_from_dense(O._ptr(eri), O._ptr(np.ascontiguousarray(D)), O._ptr(np.ascontiguousarray(H)), M,
                                O._ptr(F_ref))
    assert np.abs(F - F_ref).max() <= 1e-10 * np.abs(F_ref).max()
    assert eng.get_algorithm_name() == "Stored"


@pytest.mark.parametrize("name", ["h2o_sto3g", "ch4h2o_ccpvdz"])
def test_oneint_matches_oracle(name, cuda):
    (Z, xyz), basis = SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, basis)
    S, T, V = (x.cpu().numpy() for x in oneint(mol))
    Sr, Tr, Vr = O.oneint(p, n, Z, xyz)
    for a, b in ((S, Sr), (T, Tr), (V, Vr)):
        assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("name,method", [("h2o_sto3g", "stored"), ("h2o_sto3g", "direct"), ("h2o_ccpvdz", "stored"),
                                         ("ethane_631gs", "direct")])
def test_scf_energy_matches_oracle(name, method, cuda):
    (Z, xyz), basis = SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, basis)
    prm = Parameters(eri_method=method, convergence_energy_threshold=1e-10, maxiter=100)
    r = rhf(mol, prm)
    e_elec, e_tot, _, _, _ = O.rhf(p, t, n, Z, xyz, mol.n_electrons // 2, thr=THR, conv=1e-10)
    assert r["converged"]
    assert abs(r["e_total"] - e_tot) < 1e-9, (r["e_total"], e_tot)


def test_config0_energy(cuda):
    mol, aux = gm.config_molecule(0)
    p, t, n = O.build_molecule(gm.basis_path("sto-3g"), mol.Z, mol.xyz)
    r = rhf(mol, Parameters(convergence_energy_threshold=1e-10))
    _, e_tot, _, _, _ = O.rhf(p, t, n, mol.Z, mol.xyz, 5, conv=1e-10)
    assert abs(r["e_total"] - e_tot) < 1e-9
    assert -75.05 < r["e_total"] < -74.85


def test_stored_equals_direct_config1(cuda):
    """Full-size C10H22/cc-pVDZ: the stored and direct builds agree and J, K are
    symmetric (size-independent properties; the oracle is too slow here)."""
    mol, _ = gm.config_molecule(1)
    M = mol.n_basis
    D = _torch(random_density(M, seed=5), cuda)
    res = {}
    for m in ("direct", "stored"):
        eng = FockEngine(mol, eri_method=m)
        eng.precomputation()
        res[m] = [x.cpu().numpy() for x in eng.compute_jk(D)]
        del eng
    for a, b in zip(res["direct"], res["stored"]):
        assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()
        np.testing.assert_allclose(a, a.T, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("fp32_thr", [1e-6, 1e-2])
def test_stored_mixed_precision_jk(fp32_thr, cuda):
    """precision=mixed on the stored path: quartets bounded below fp32_thr are
    kept in FP32.  Each such integral has absolute error <= fp32_thr * 2^-24,
    so J and K move by a correspondingly small amount; the FP32 share grows
    with the threshold."""
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    D = _torch(random_density(mol.n_basis), cuda)
    e64 = FockEngine(mol, eri_method="stored")
    e64.precomputation()
    J64, K64 = [x.cpu().numpy() for x in e64.compute_jk(D)]
    emx = FockEngine(mol, eri_method="stored", precision="mixed", fp32_bound_threshold=fp32_thr)
    emx.precomputation()
    J, K = [x.cpu().numpy() for x in emx.compute_jk(D)]
    st, st64 = emx.stats(), e64.stats()
    assert st["contracted_quartets"] == st64["contracted_quartets"]
    assert 0 < st["fp32_quartets"] < st["contracted_quartets"]
    tol = 1e-4 * fp32_thr
    assert np.abs(J - J64).max() <= tol and np.abs(K - K64).max() <= tol
    assert np.abs(J - J64).max() > 0.0  # the FP32 part is really used


def test_stored_mixed_precision_scf_config1(cuda):
    """C10H22/cc-pVDZ, FP32 for quartets bounded < 1e-6: |E_mixed - E_fp64| <= 1 uEh."""
    mol, _ = gm.config_molecule(1)
    r64 = rhf(mol, Parameters(eri_method="stored", convergence_energy_threshold=1e-9))
    rmx = rhf(mol, Parameters(eri_method="stored", precision="mixed", convergence_energy_threshold=1e-9))
    assert r64["converged"] and rmx["converged"]
    assert abs(rmx["e_total"] - r64["e_total"]) < 1e-6


# ------------------------------------------------------------------- RI ----
RI_SYSTEMS = {"h2o": lambda: gm.water(), "ch4h2o": lambda: methane_water()}


@pytest.mark.parametrize("name", list(RI_SYSTEMS))
def test_ri_jk_matches_oracle(name, cuda):
    Z, xyz = RI_SYSTEMS[name]()
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    aux, (pa, ta, na) = both(Z, xyz, "cc-pvdz-rifit")
    B = O.ri_tensor(p, t, n, pa, na, THR)
    D = random_density(mol.n_basis, seed=21)
    Jr, Kr = O.ri_jk(B, D)
    eng = FockEngine(mol, aux=aux, eri_method="ri")
    eng.precomputation()
    J, K = (x.cpu().numpy() for x in eng.compute_jk(_torch(D, cuda)))
    assert np.abs(J - Jr).max() <= 1e-10 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 1e-10 * np.abs(Kr).max()
    assert eng.get_algorithm_name() == "RI"


def test_ri_scf_energy_matches_oracle(cuda):
    Z, xyz = gm.water()
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    aux, (pa, ta, na) = both(Z, xyz, "cc-pvdz-rifit")
    r = rhf(mol, Parameters(eri_method="ri", convergence_energy_threshold=1e-10), aux=aux)
    B = O.ri_tensor(p, t, n, pa, na, THR)
    e_tot = O.rhf(p, t, n, Z, xyz, 5, conv=1e-10, riB=B)[1]
    assert abs(r["e_total"] - e_tot) < 1e-9
    # RI error against the exact 4-centre energy stays small
    e_exact = O.rhf(p, t, n, Z, xyz, 5, conv=1e-10)[1]
    assert abs(e_tot - e_exact) < 2e-3  # cc-pVDZ-RIFIT is an MP2 fitting set: ~1 mEh in HF


def test_ri_occupied_form_equals_density_form(cuda):
    import torch
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    M, nocc = mol.n_basis, mol.n_electrons // 2
    rng = np.random.default_rng(8)
    C = rng.normal(size=(M, M)) / np.sqrt(M)
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    H = random_density(M, seed=2)
    Dt, Ct, Ht = _torch(D, cuda), _torch(C, cuda), _torch(H, cuda)
    e_den = FockEngine(mol, aux=aux, eri_method="ri")
    e_den.precomputation()
    F_den = e_den.compute_fock_matrix(Dt, Ht).cpu().numpy()
    e_occ = FockEngine(mol, aux=aux, eri_method="ri", ri_k="occupied")
    e_occ.precomputation()
    F_occ = e_occ.compute_fock_matrix_occ(Dt, Ct, nocc, Ht).cpu().numpy()
    assert np.abs(F_occ - F_den).max() <= 1e-12 * np.abs(F_den).max()
    # without orbitals the occupied-form engine falls back to the density form
    F_fallback = e_occ.compute_fock_matrix(Dt, Ht).cpu().numpy()
    assert np.abs(F_fallback - F_den).max() <= 1e-12 * np.abs(F_den).max()
    del torch


def test_ri_scf_occupied_form_energy(cuda):
    Z, xyz = gm.water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    r1 = rhf(mol, Parameters(eri_method="ri", convergence_energy_threshold=1e-10), aux=aux)
    r2 = rhf(mol, Parameters(eri_method="ri", ri_k="occupied", convergence_energy_threshold=1e-10), aux=aux)
    assert abs(r1["e_total"] - r2["e_total"]) < 1e-10


def test_ri_mixed_precision_energy(cuda):
    """precision=mixed (K and J as increments on the fp16 matrix cores from
    the orbital change, over FP64 builds) stays within 1 micro-Eh of the
    FP64 density-form RI energy."""
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    r64 = rhf(mol, Parameters(eri_method="ri", convergence_energy_threshold=1e-10), aux=aux)
    rmx = rhf(mol, Parameters(eri_method="ri", precision="mixed", convergence_energy_threshold=1e-10), aux=aux)
    assert rmx["converged"]
    assert abs(rmx["e_total"] - r64["e_total"]) < 1e-6


@pytest.mark.parametrize("method", ["stored", "direct"])
def test_graph_replay_matches_eager(method, cuda):
    """The four-centre build is captured as a HIP graph once per buffer set and
    replayed.  Every replay, with eager work and synchronisations in between,
    must reproduce the eager build; new buffers trigger a new capture."""
    import torch
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    M = mol.n_basis
    H = _torch(random_density(M, seed=9), cuda)
    eager = FockEngine(mol, eri_method=method, graph=False)
    eager.precomputation()
    graph = FockEngine(mol, eri_method=method, graph=True)
    graph.precomputation()
    F = torch.empty((M, M), dtype=torch.float64, device=cuda)
    D = torch.empty_like(F)
    for it in range(4):
        D.copy_(_torch(random_density(M, seed=20 + it), cuda))  # new density, same buffer
        F.fill_(float("nan"))
        graph.compute_fock_matrix(D, H, F)
        ref = eager.compute_fock_matrix(D, H)
        x = torch.ones(1000, device=cuda).sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 1000.0
        assert torch.isfinite(F).all()
        assert float((F - ref).abs().max()) <= 1e-12 * float(ref.abs().max())
    assert graph.stats()["graph_captures"] == 1
    F2 = graph.compute_fock_matrix(D, H)  # other output buffer: a second capture
    assert float((F2 - ref).abs().max()) <= 1e-12 * float(ref.abs().max())
    assert graph.stats()["graph_captures"] == 2


def test_class_stats_counts_primitive_quartets(cuda):
    """gansu_fock_class_stats: per quartet class, the contracted quartets the
    builds visit and the primitive quartets among them that pass the
    primitive Schwarz screen.  Water/STO-3G has three primitives in every
    shell, so with no screening (threshold 1e-30) each contracted quartet
    holds 81 primitive quartets; at 1e-12 at most that; the class totals
    match the engine statistics."""
    mol = gm.Molecule.from_atoms(*gm.water(), gm.basis_path("sto-3g"))
    for thr in (1e-30, 1e-12):
        eng = FockEngine(mol, eri_method="direct", schwarz_screening_threshold=thr)
        eng.precomputation()
        cs = eng.class_stats()
        st = eng.stats()
        assert len(cs) == 21 and {c["L"] for c in cs} >= {(0, 0, 0, 0), (1, 1, 1, 1), (1, 0, 0, 0)}
        assert sum(c["contracted"] for c in cs) == st["contracted_quartets"]
        assert sum(c["primitive"] for c in cs) == st["primitive_quartets"] > 0
        for c in cs:
            if thr < 1e-20:
                assert c["primitive"] == 81 * c["contracted"], c
            else:
                assert c["primitive"] <= 81 * c["contracted"], c


def test_direct_incremental_fock_sequence(cuda):
    """incremental_fock=N (direct path): each build digests D - D_prev with
    density-weighted screening and adds to the previous J, K; every N-th is
    a full build.  Over a sequence of densities that move less and less (an
    SCF tail), J and K stay within 1e-9 of max|J|, max|K| of full builds, the
    full builds (k = 0, 4, 8: density-weighted screening only) within 1e-11,
    and a UHF build (two densities) on the same engine is a full one."""
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    M = mol.n_basis
    full = FockEngine(mol, eri_method="direct")
    full.precomputation()
    incr = FockEngine(mol, eri_method="direct", incremental_fock=4)
    incr.precomputation()
    D0 = random_density(M, seed=41)
    rng = np.random.default_rng(42)
    D = D0.copy()
    for k in range(9):
        if k:
            E = rng.normal(size=(M, M)) * 10.0 ** (-2 - k)
            D = D + (E + E.T) / 2
        Dt = _torch(D, cuda)
        Jf, Kf = (x.cpu().numpy() for x in full.compute_jk(Dt))
        Ji, Ki = (x.cpu().numpy() for x in incr.compute_jk(Dt))
        ej = np.abs(Ji - Jf).max() / np.abs(Jf).max()
        ek = np.abs(Ki - Kf).max() / np.abs(Kf).max()
        print("build %d: |J_incr - J| %.2e  |K_incr - K| %.2e" % (k, ej, ek))
        tol = 1e-11 if k % 4 == 0 else 1e-9
        assert ej < tol and ek < tol, (k, ej, ek)
    Da, Db, H = random_density(M, seed=43) / 2, random_density(M, seed=44) / 2, random_density(M, seed=45)
    Fa, Fb = (x.cpu().numpy() for x in full.compute_fock_matrix_uhf(_torch(Da, cuda), _torch(Db, cuda), _torch(H, cuda)))
    Ga, Gb = (x.cpu().numpy() for x in incr.compute_fock_matrix_uhf(_torch(Da, cuda), _torch(Db, cuda), _torch(H, cuda)))
    assert np.abs(Ga - Fa).max() <= 1e-13 * np.abs(Fa).max() and np.abs(Gb - Fb).max() <= 1e-13 * np.abs(Fb).max()


@pytest.mark.parametrize("solver", ["diagonalize", "purify"])
def test_direct_incremental_fock_scf_config1(solver, cuda):
    """C10H22/cc-pVDZ direct RHF with delta-density builds (a full build every
    8) converges to the energy of full builds within 1e-9 Eh, with the
    eigensolver loop and with the graph-captured purification loop (whose
    captured build replays the device-side increment state)."""
    mol, _ = gm.config_molecule(1)
    kw = dict(eri_method="direct", convergence_energy_threshold=1e-10, scf_solver=solver)
    r0 = rhf(mol, Parameters(**kw))
    r1 = rhf(mol, Parameters(incremental_fock=8, **kw))
    assert r0["converged"] and r1["converged"]
    print("%s: E_full %.12f, E_incremental %.12f" % (solver, r0["e_total"], r1["e_total"]))
    assert abs(r1["e_total"] - r0["e_total"]) < 1e-9


@pytest.mark.parametrize("method", ["stored", "direct", "ri"])
def test_uhf_fock_matches_oracle(method, cuda):
    """UHF build (gpu_kernels.cu:289-340): F_a = H + J[Da] + J[Db] - K[Da]."""
    Z, xyz = methane_water()
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    M = mol.n_basis
    Da, Db, H = random_density(M, seed=31) / 2, random_density(M, seed=32) / 2, random_density(M, seed=33)
    if method == "ri":
        aux, (pa, ta, na) = both(Z, xyz, "cc-pvdz-rifit")
        B = O.ri_tensor(p, t, n, pa, na, THR)
        JKa, JKb = O.ri_jk(B, Da), O.ri_jk(B, Db)
        eng = FockEngine(mol, aux=aux, eri_method="ri")
    else:
        Q = O.schwarz(p, t)
        JKa, JKb = O.jk(p, t, n, Q, Da, THR), O.jk(p, t, n, Q, Db, THR)
        eng = FockEngine(mol, eri_method=method)
    eng.precomputation()
    Fa, Fb = eng.compute_fock_matrix_uhf(_torch(Da, cuda), _torch(Db, cuda), _torch(H, cuda))
    Fa_ref = H + JKa[0] + JKb[0] - JKa[1]
    Fb_ref = H + JKa[0] + JKb[0] - JKb[1]
    scale = max(1.0, np.abs(Fa_ref).max())
    assert np.abs(Fa.cpu().numpy() - Fa_ref).max() <= 1e-10 * scale
    assert np.abs(Fb.cpu().numpy() - Fb_ref).max() <= 1e-10 * scale


@pytest.mark.parametrize("method", ["stored", "direct"])
def test_update_geometry_equals_fresh_engine(method, cuda):
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    xyz2 = xyz + np.random.default_rng(4).normal(scale=0.05, size=xyz.shape)
    mol2, _ = both(Z, xyz2, "cc-pvdz")
    D = _torch(random_density(mol.n_basis, seed=8), cuda)
    H = _torch(np.zeros((mol.n_basis, mol.n_basis)), cuda)
    eng = FockEngine(mol, eri_method=method)
    eng.precomputation()
    F1 = eng.compute_fock_matrix(D, H).clone()
    eng.update_geometry(xyz2)
    F2 = eng.compute_fock_matrix(D, H)
    fresh = FockEngine(mol2, eri_method=method)
    fresh.precomputation()
    Fr = fresh.compute_fock_matrix(D, H)
    assert float((F2 - Fr).abs().max()) <= 1e-12 * float(Fr.abs().max())
    assert float((F2 - F1).abs().max()) > 1e-6  # the geometry really moved


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ri_update_geometry_equals_fresh_engine(precision, cuda):
    """RI engines keep B (and its fp16 planes, the increment state) across a
    new geometry: the builds after update_geometry equal a fresh engine's,
    for the mixed path over a sequence that starts with the FP64 build and
    continues with increments on the reused buffers."""
    Z, xyz = methane_water()
    xyz2 = xyz + np.random.default_rng(5).normal(scale=0.05, size=xyz.shape)
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    mol2, _ = both(Z, xyz2, "cc-pvdz")
    aux2, _ = both(Z, xyz2, "cc-pvdz-rifit")
    M, nocc = mol.n_basis, mol.n_electrons // 2
    H = _torch(np.zeros((M, M)), cuda)
    kw = dict(eri_method="ri", ri_k="occupied") if precision == "fp64" else dict(eri_method="ri", precision="mixed")
    seq = [_torch(C, cuda) for C in _orbital_sequence(M, nocc, 3, seed=6)]
    dens = [(2.0 * C[:, :nocc] @ C[:, :nocc].T).contiguous() for C in seq]
    eng = FockEngine(mol, aux=aux, **kw)
    eng.precomputation()
    for C, D in zip(seq, dens):
        F1 = eng.compute_fock_matrix_occ(D, C, nocc, H).clone()
    eng.update_geometry(xyz2)
    fresh = FockEngine(mol2, aux=aux2, **kw)
    fresh.precomputation()
    for C, D in zip(seq, dens):
        F2 = eng.compute_fock_matrix_occ(D, C, nocc, H)
        Fr = fresh.compute_fock_matrix_occ(D, C, nocc, H)
        assert float((F2 - Fr).abs().max()) <= 1e-12 * float(Fr.abs().max())
    assert float((F2 - F1).abs().max()) > 1e-6  # the geometry really moved
    if precision == "mixed":
        # the counts restart with the new geometry's state
        assert eng.stats()["ri_incremental_builds"] == fresh.stats()["ri_incremental_builds"] == 2


@pytest.mark.parametrize("method", ["stored", "direct", "ri"])
def test_rohf_fock_matches_oracle(method, cuda):
    """ROHF build (gpu_manager.cu:733-817, RI :1800-1961): F_c, F_o and the
    unified F = S C R C^T S against the oracle for two coupling sets, and the
    closed-shell limit (no open shell, S-orthonormal C) equal to the engine's
    own RHF Fock matrix."""
    Z, xyz = methane_water()
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    M = mol.n_basis
    S = O.oneint(p, n, Z, xyz)[0]
    w, V = np.linalg.eigh(S)
    C = V @ np.diag(w ** -0.5) @ V.T @ np.linalg.qr(np.random.default_rng(41).normal(size=(M, M)))[0]
    Dc, Do, H = random_density(M, seed=34), 0.3 * random_density(M, seed=35), random_density(M, seed=36)
    if method == "ri":
        aux, (pa, ta, na) = both(Z, xyz, "cc-pvdz-rifit")
        B = O.ri_tensor(p, t, n, pa, na, THR)
        JKc, JKo = O.ri_jk(B, Dc), O.ri_jk(B, Do)
        eng = FockEngine(mol, aux=aux, eri_method="ri")
    else:
        Q = O.schwarz(p, t)
        JKc, JKo = O.jk(p, t, n, Q, Dc, THR), O.jk(p, t, n, Q, Do, THR)
        eng = FockEngine(mol, eri_method=method)
    eng.precomputation()
    args = [_torch(x, cuda) for x in (Dc, Do, H, C, S)]
    nc, no = mol.n_electrons // 2 - 1, 2
    for coupling in ("roothaan", "guest-saunders"):
        Fc, Fo, F = (x.cpu().numpy() for x in eng.compute_fock_matrix_rohf(*args, nc, no, coupling))
        Fc_ref, Fo_ref, F_ref = O.rohf_fock(H, JKc, JKo, C, S, nc, no, coupling)
        for got, ref in ((Fc, Fc_ref), (Fo, Fo_ref), (F, F_ref)):
            assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), coupling
    # closed-shell limit: F = F_RHF of the same engine
    zero = _torch(np.zeros((M, M)), cuda)
    F_rhf = eng.compute_fock_matrix(args[0], args[2]).cpu().numpy()
    F = eng.compute_fock_matrix_rohf(args[0], zero, args[2], args[3], args[4], mol.n_electrons // 2, 0)[2]
    assert np.abs(F.cpu().numpy() - F_rhf).max() <= 1e-9 * np.abs(F_rhf).max()
    with pytest.raises(RuntimeError):
        eng.compute_fock_matrix_rohf(*args, nc, no, "no-such-set")


def test_uhf_scf_matches_oracle(cuda):
    """UHF SCF (uhf.cu, per-spin DIIS): the Li atom and the water cation against
    the oracle's UHF; closed-shell water UHF equals RHF on the stored and RI
    engines.

    DIIS per spin (two independent extrapolations, as the reference's
    Convergence_UHF_DIIS does) converges H2O+ slowly: the oracle stops 5e-8 Eh
    above the limit at |dE| < 1e-10, and where it stops depends on round-off.
    The cation is therefore converged to |dE| < 1e-12 on both sides, where the
    oracle's own stop points spread by 1e-9 Eh and 1.5e-5 in the densities
    (measured against |dE| < 1e-14), and compared within 1e-8 Eh and 1e-4."""
    from dft_gansu_amd import uhf
    prm = Parameters(convergence_energy_threshold=1e-10)
    origin = np.zeros((1, 3))
    Z = np.array([3])
    mol, (p, t, n) = both(Z, origin, "sto-3g")
    r = uhf(mol, prm)
    assert r["converged"]
    assert abs(r["e_total"] - O.uhf(p, t, n, Z, origin, 2, 1)[1]) < 1e-9
    assert abs(r["e_total"] - (-7.315526)) < 2e-6
    Z, xyz = gm.water(jitter=0.01, seed=11)
    path = gm.basis_path("cc-pvdz")
    cation = gm.Molecule.from_atoms(Z, xyz, path, charge=1)
    p, t, n = O.build_molecule(path, Z, xyz)
    r = uhf(cation, Parameters(convergence_energy_threshold=1e-12, maxiter=200))
    ref = O.uhf(p, t, n, Z, xyz, 5, 4, conv=1e-12)
    assert r["converged"]
    assert abs(r["e_total"] - ref[1]) < 1e-8
    assert np.abs(r["Da"] - ref[2]).max() < 1e-4 and np.abs(r["Db"] - ref[3]).max() < 1e-4
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    for kw in (dict(eri_method="stored"), dict(eri_method="ri")):
        p2 = Parameters(convergence_energy_threshold=1e-10, **kw)
        a = aux if kw["eri_method"] == "ri" else None
        assert abs(uhf(mol, p2, aux=a)["e_total"] - rhf(mol, p2, aux=a)["e_total"]) < 1e-9, kw


def test_rohf_scf_matches_oracle(cuda):
    """ROHF SCF (rohf.cu): the O triplet (STO-3G) and the water cation
    (cc-pVDZ) against the oracle's ROHF; closed-shell water ROHF equals RHF."""
    from dft_gansu_amd import rohf
    prm = Parameters(convergence_energy_threshold=1e-10)
    origin = np.zeros((1, 3))
    Z = np.array([8])
    path = gm.basis_path("sto-3g")
    mol = gm.Molecule.from_atoms(Z, origin, path, beta_to_alpha=1)
    p, t, n = O.build_molecule(path, Z, origin)
    r = rohf(mol, prm)
    assert r["converged"]
    assert abs(r["e_total"] - O.rohf(p, t, n, Z, origin, 3, 2)[1]) < 1e-9
    Z, xyz = gm.water(jitter=0.01, seed=11)
    path = gm.basis_path("cc-pvdz")
    cation = gm.Molecule.from_atoms(Z, xyz, path, charge=1)
    p, t, n = O.build_molecule(path, Z, xyz)
    for coupling in ("roothaan", "guest-saunders"):
        r = rohf(cation, Parameters(convergence_energy_threshold=1e-10, rohf_parameter_name=coupling))
        ref = O.rohf(p, t, n, Z, xyz, 4, 1, coupling)
        assert r["converged"]
        assert abs(r["e_total"] - ref[1]) < 1e-9, coupling
    mol, _ = both(Z, xyz, "cc-pvdz")
    assert abs(rohf(mol, prm)["e_total"] - rhf(mol, prm)["e_total"]) < 1e-9


def test_gwh_guess_reaches_the_core_guess_solution(cuda):
    """initial_guess=gwh (F_pq = 1.75 S_pq (H_pp + H_qq)/2, gpu_kernels.cu:535-545)
    converges RHF and UHF to the same energies as the core guess (RHF within
    1e-9 Eh; the open-shell H2O+ runs stop on |dE| < 1e-10 on slightly
    different paths and agree within 1e-6 Eh, measured 5e-8).  ROHF of H2O+
    from the GWH guess settles in another state (0.08 Eh higher), as aufbau
    occupation of the unified matrix allows, so it is not compared."""
    from dft_gansu_amd import uhf
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, _ = both(Z, xyz, "cc-pvdz")
    cation = gm.Molecule.from_atoms(Z, xyz, gm.basis_path("cc-pvdz"), charge=1)
    for run, m in ((rhf, mol), (uhf, cation)):
        e = [run(m, Parameters(convergence_energy_threshold=1e-10, initial_guess=g))["e_total"] for g in ("core", "gwh")]
        assert abs(e[0] - e[1]) < (1e-9 if run is rhf else 1e-6), (run.__name__, e)


def test_sad_guess_scf_matches_oracle(cuda):
    """initial_guess=sad (rhf.hpp:551-584, uhf.hpp:668-705): RHF of water
    (cc-pVDZ) matches the oracle's RHF started from the same atomic densities,
    energy within 1e-9 Eh and the iteration count within one, and reaches the
    core-guess energy; the RI engine, whose first build takes a density that is
    not 2 C_occ C_occ^T, does too.  UHF of H2O+ from the SAD guess (C_beta's
    last row and column zeroed) matches the oracle's SAD UHF, both converged to
    |dE| < 1e-12 as the per-spin DIIS needs (within 1e-8 Eh).  ROHF refuses
    the guess (the reference's ROHF SAD reads C before anything writes it)."""
    from dft_gansu_amd import rohf, uhf
    from dft_gansu_amd._lib import GansuError
    Z, xyz = gm.water(jitter=0.01, seed=11)
    path = gm.basis_path("cc-pvdz")
    mol = gm.Molecule.from_atoms(Z, xyz, path)
    Da, Db = mol.sad_density()
    p, t, n = O.build_molecule(path, Z, xyz)
    ref = O.rhf(p, t, n, Z, xyz, 5, conv=1e-10, D0=Da + Db)
    r = rhf(mol, Parameters(convergence_energy_threshold=1e-10, initial_guess="sad"))
    assert r["converged"]
    assert abs(r["e_total"] - ref[1]) < 1e-9
    assert abs(r["iterations"] - ref[4]) <= 1, (r["iterations"], ref[4])
    core = rhf(mol, Parameters(convergence_energy_threshold=1e-10))
    assert abs(r["e_total"] - core["e_total"]) < 1e-9
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    e_ri = [rhf(mol, Parameters(convergence_energy_threshold=1e-10, eri_method="ri", initial_guess=g), aux=aux)["e_total"]
            for g in ("core", "sad")]
    assert abs(e_ri[0] - e_ri[1]) < 1e-9, e_ri
    cation = gm.Molecule.from_atoms(Z, xyz, path, charge=1)
    ca, cb = cation.sad_density()
    refu = O.uhf(p, t, n, Z, xyz, 5, 4, conv=1e-12, sad=(ca, cb))
    ru = uhf(cation, Parameters(convergence_energy_threshold=1e-12, maxiter=200, initial_guess="sad"))
    assert ru["converged"]
    assert abs(ru["e_total"] - refu[1]) < 1e-8, (ru["e_total"], refu[1])
    with pytest.raises(GansuError) as e:
        rohf(cation, Parameters(initial_guess="sad"))
    assert e.value.code == 9


def test_export_sad_cache_reproduces_the_shipped_blocks(cuda, tmp_path):
    """The SAD cache export (HF::generate_sad_cache after RHF for even Z, ROHF
    for odd Z) regenerates the STO-3G blocks of the bundled .sad file (made in
    unit normalisation): H and Li (ROHF) within 1e-6, O (RHF; its occupied p
    pair is one of the degenerate choices) in its s block within 1e-6 (the
    shipped Li beta block differs by 2.1e-7 from a run converged to 1e-12: the
    file comes from a looser stop), with
    tr(D_s S) = 4 per spin and idempotent spin densities (D_s S D_s = D_s).
    The written file is what the SAD guess then reads."""
    import shutil
    from dft_gansu_amd import sad
    gbs = tmp_path / "sto-3g.gbs"
    shutil.copy(gm.basis_path("sto-3g"), gbs)
    shipped = gm.basis_path("sto-3g")[:-4] + ".sad"
    added = sad.generate_sad_file(str(gbs), elements=["H", "Li", "O"], normalization="unit",
                                  convergence_energy_threshold=1e-12)
    assert added == ["H", "Li", "O"]
    assert sad.cached_elements(str(tmp_path / "sto-3g.sad")) == ["H", "Li", "O"]
    for sym in ("H", "Li"):
        a, b = O.read_sad(str(tmp_path / "sto-3g.sad"), sym)
        ra, rb = O.read_sad(shipped, sym)
        assert np.abs(a - ra).max() < 1e-6 and np.abs(b - rb).max() < 1e-6, sym
    a, b = O.read_sad(str(tmp_path / "sto-3g.sad"), "O")
    ra, rb = O.read_sad(shipped, "O")
    assert np.abs(a[:2, :2] - ra[:2, :2]).max() < 1e-6 and np.abs(b[:2, :2] - rb[:2, :2]).max() < 1e-6
    p, t, n = O.build_molecule(str(gbs), np.array([8]), np.zeros((1, 3)), quirk=False)
    S = O.oneint(p, n, np.array([8]), np.zeros((1, 3)))[0]
    for d in (a, b):
        assert abs(np.trace(d @ S) - 4.0) < 1e-8
        assert np.abs(d @ S @ d - d).max() < 1e-8
    Z, xyz = gm.water(jitter=0.0)
    mol = gm.Molecule.from_atoms(Z, xyz, str(gbs), normalization="unit")
    Da, _ = mol.sad_density()
    assert np.array_equal(Da[:5, :5], a)


def test_rhf_restart_from_densities(cuda):
    """HF::solve with density matrices (rhf.cu:127-135, InitialGuess_RHF_Density,
    rhf.hpp:441-454): the loop starts from F(Da + Db).  From the converged
    density (D/2, D/2) the SCF stops at once on the same energy; from the SAD
    densities it reaches the oracle's energy; initial_guess=density without
    densities is the core guess.  UHF restarts from (Da, Db) the same way
    (F(Da, Db), then C and D per spin, uhf.hpp:557-568); ROHF refuses the key
    (the reference's ROHF density guess forms S C R C^T S from an unwritten C)."""
    from dft_gansu_amd import rohf, uhf
    from dft_gansu_amd._lib import GansuError
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    prm = Parameters(convergence_energy_threshold=1e-10)
    r0 = rhf(mol, prm)
    r1 = rhf(mol, prm, density=(0.5 * r0["D"], 0.5 * r0["D"]), force_density=True)
    assert r1["converged"] and r1["iterations"] <= 2, r1["iterations"]
    assert abs(r1["e_total"] - r0["e_total"]) < 1e-9
    # without force_density the parameters' guess (core) runs, as HF::solve
    # with densities does unless initial_guess=density (rhf.cu:122-134)
    r1c = rhf(mol, prm, density=(0.5 * r0["D"], 0.5 * r0["D"]))
    assert r1c["iterations"] == r0["iterations"] and abs(r1c["e_total"] - r0["e_total"]) < 1e-12
    Da, Db = mol.sad_density()
    r2 = rhf(mol, prm, density=(Da, Db), force_density=True)
    assert abs(r2["e_total"] - O.rhf(p, t, n, Z, xyz, 5, conv=1e-10, D0=Da + Db)[1]) < 1e-9
    r3 = rhf(mol, Parameters(convergence_energy_threshold=1e-10, initial_guess="density"))
    assert abs(r3["iterations"] - r0["iterations"]) <= 1 and abs(r3["e_total"] - r0["e_total"]) < 1e-9
    cation = gm.Molecule.from_atoms(Z, xyz, gm.basis_path("cc-pvdz"), charge=1)
    pu = Parameters(convergence_energy_threshold=1e-12, maxiter=200)
    u0 = uhf(cation, pu)
    u1 = uhf(cation, pu, density=(u0["Da"], u0["Db"]), force_density=True)
    # the density guess ends with the symmetry break (the last row and column
    # of C_beta zeroed, uhf.hpp:557-568), so the restart perturbs the
    # converged state and per-spin DIIS restarts with an empty history
    assert u1["converged"] and u1["iterations"] < u0["iterations"], (u1["iterations"], u0["iterations"])
    assert abs(u1["e_total"] - u0["e_total"]) < 1e-9
    with pytest.raises(GansuError) as e:
        rohf(cation, Parameters(initial_guess="density"))
    assert e.value.code == 9


# ----------------------------------------- guesses, damping, recipes ------
def test_uhf_symmetry_break_stretched_h2(cuda):
    """Every UHF guess zeroes the last row and column of C_beta
    (uhf.hpp:429-436): H2/STO-3G at 4 bohr reaches the broken-symmetry UHF
    state far below RHF, as the oracle does, from the core and GWH guesses."""
    from dft_gansu_amd import uhf
    Z = np.array([1, 1])
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 4.0]])
    mol, (p, t, n) = both(Z, xyz, "sto-3g")
    ref = O.uhf(p, t, n, Z, xyz, 1, 1, conv=1e-12, maxiter=300)
    r = uhf(mol, Parameters(convergence_energy_threshold=1e-12, maxiter=300))
    assert r["converged"]
    assert abs(r["e_total"] - ref[1]) < 1e-9
    e_rhf = rhf(mol, Parameters(convergence_energy_threshold=1e-12))["e_total"]
    assert r["e_total"] < e_rhf - 0.1
    rg = uhf(mol, Parameters(convergence_energy_threshold=1e-12, maxiter=300, initial_guess="gwh"))
    assert rg["converged"] and rg["e_total"] < e_rhf - 0.1


def test_damping_drivers_match_oracle(cuda):
    """Constant and optimal damping (rhf.hpp:182-262) on RHF water/STO-3G,
    constant damping per spin on UHF Li (uhf.hpp:196-258), constant and
    optimal damping on the ROHF O triplet (rohf.hpp:216-306): energies within
    1e-9 Eh of the oracle's damped SCF and the same number of iterations
    (within one), so the path is the same, not only the end point."""
    from dft_gansu_amd import rohf, uhf
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, (p, t, n) = both(Z, xyz, "sto-3g")
    for optimal in (False, True):
        ref = O.rhf_damped(p, t, n, Z, xyz, 5, optimal=optimal, conv=1e-10)
        r = rhf(mol, Parameters(convergence_energy_threshold=1e-10, maxiter=300,
                                convergence_method="optimaldamping" if optimal else "damping"))
        assert r["converged"]
        assert abs(r["e_total"] - ref[1]) < 1e-9, optimal
        assert abs(r["iterations"] - ref[3]) <= 1, (optimal, r["iterations"], ref[3])
    origin = np.zeros((1, 3))
    Zl = np.array([3])
    mol, (p, t, n) = both(Zl, origin, "sto-3g")
    ref = O.uhf(p, t, n, Zl, origin, 2, 1, damping=0.9, conv=1e-10, maxiter=500)
    r = uhf(mol, Parameters(convergence_energy_threshold=1e-10, maxiter=500, convergence_method="damping"))
    assert r["converged"] and abs(r["e_total"] - ref[1]) < 1e-9 and abs(r["iterations"] - ref[4]) <= 1
    Zo = np.array([8])
    mol = gm.Molecule.from_atoms(Zo, origin, gm.basis_path("sto-3g"), beta_to_alpha=1)
    p, t, n = O.build_molecule(gm.basis_path("sto-3g"), Zo, origin)
    for kw, cm in ((dict(damping=0.9), "damping"), (dict(optimal=True), "optimaldamping")):
        ref = O.rohf(p, t, n, Zo, origin, 3, 2, conv=1e-10, maxiter=500, **kw)
        r = rohf(mol, Parameters(convergence_energy_threshold=1e-10, maxiter=500, convergence_method=cm))
        assert r["converged"] and abs(r["e_total"] - ref[1]) < 1e-9, cm
        assert abs(r["iterations"] - ref[4]) <= 1, (cm, r["iterations"], ref[4])
    with pytest.raises(Exception):  # uhf.cu:57-69 has no optimal damping
        uhf(mol, Parameters(convergence_method="optimaldamping"))


RECIPES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "recipes")


@pytest.mark.parametrize("recipe", sorted(f for f in os.listdir(RECIPES) if f.endswith(".txt")))
def test_reference_recipes_run(cuda, recipe):
    """Each of the reference's nine parameter recipes (parameter_recipe/*.txt)
    loads and runs its method and convergence scheme on H2O/STO-3G (their
    molecule): closed-shell water gives the oracle's RHF energy whatever the
    method (RHF, UHF, ROHF) and scheme (DIIS, damping, optimal damping)."""
    from dft_gansu_amd import rohf, uhf
    prm = Parameters(os.path.join(RECIPES, recipe))
    prm["convergence_energy_threshold"] = "1e-11"
    prm["maxiter"] = "1000"
    Z, xyz = gm.water(jitter=0.01, seed=11)
    mol, (p, t, n) = both(Z, xyz, "sto-3g")
    e_ref = O.rhf(p, t, n, Z, xyz, 5, conv=1e-12)[1]
    run = {"rhf": rhf, "uhf": uhf, "rohf": rohf}[prm["method"]]
    r = run(mol, prm)
    assert r["converged"], recipe
    assert abs(r["e_total"] - e_ref) < 1e-8, (recipe, r["e_total"], e_ref)


def test_ri_nan_in_fock_is_an_error(cuda):
    """The RI build fails with GANSU_E_NUMERIC when F holds a NaN, as
    ERI_RI_RHF::compute_fock_matrix throws (rhf.hpp:665-674); nan_check=0
    skips the test (and its one flag read-back)."""
    from dft_gansu_amd._lib import GansuError
    Z, xyz = gm.water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    M = mol.n_basis
    D = _torch(random_density(M, seed=3), cuda)
    H = random_density(M, seed=4)
    H[3, 5] = np.nan
    eng = FockEngine(mol, aux=aux, eri_method="ri")
    eng.precomputation()
    with pytest.raises(GansuError) as e:
        eng.compute_fock_matrix(D, _torch(H, cuda))
    assert e.value.code == 8 and "NaN" in str(e.value)
    H[3, 5] = 0.0
    F = eng.compute_fock_matrix(D, _torch(H, cuda)).cpu().numpy()
    assert np.all(np.isfinite(F))
    off = FockEngine(mol, aux=aux, eri_method="ri", nan_check=0)
    off.precomputation()
    H[3, 5] = np.nan
    F = off.compute_fock_matrix(D, _torch(H, cuda)).cpu().numpy()
    assert np.isnan(F[3, 5])


def _orbital_sequence(M, nocc, n, seed, amp=1e-2, decay=0.5):
    """Occupied-orbital matrices C_k (M x M, first nocc columns occupied)
    whose occupied space moves by amp decay^k per step, as in an SCF tail,
    with an arbitrary rotation of the occupied block at every step (which
    the build's alignment has to undo)."""
    rng = np.random.default_rng(seed)
    C = np.linalg.qr(rng.normal(size=(M, M)))[0]
    out = [C]
    for k in range(1, n):
        C = C + amp * decay ** (k - 1) * rng.normal(size=(M, M)) / np.sqrt(M)
        C = np.linalg.qr(C)[0]
        Q = np.linalg.qr(rng.normal(size=(nocc, nocc)))[0]
        C[:, :nocc] = C[:, :nocc] @ Q
        out.append(C)
    return out


@pytest.mark.parametrize("system", ["ch4h2o", "c24h12"])
def test_ri_incremental_exchange_tracks_fp64(system, cuda):
    """precision=mixed RI builds given the occupied orbitals form K and J as
    increments on the fp16 matrix cores (three-product hi/lo split) from the
    aligned orbital change; over a sequence of moving orbitals their Fock
    matrices stay within 1e-8 (relative to max|F|) of the FP64 occupied-form
    builds, and every build after the first is incremental (no refresh due
    within ri_refresh, and ri_tol = 1 so that the summed change never asks
    for one).  C24H12 (nocc 78) runs three 32-column blocks, the
    methane-water pair (nocc 10) one."""
    if system == "ch4h2o":
        Z, xyz = methane_water()
        mol, _ = both(Z, xyz, "cc-pvdz")
        aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    else:
        mol, aux = gm.config_molecule(2)
    M, nocc = mol.n_basis, mol.n_electrons // 2
    H = _torch(random_density(M, seed=2), cuda)
    e64 = FockEngine(mol, aux=aux, eri_method="ri", ri_k="occupied")
    e64.precomputation()
    emx = FockEngine(mol, aux=aux, eri_method="ri", precision="mixed", ri_refresh=20, ri_tol=1.0)
    emx.precomputation()
    worst = 0.0
    seq = _orbital_sequence(M, nocc, 7, seed=4)
    for k, C in enumerate(seq):
        Ct = _torch(C, cuda)
        D = _torch(2.0 * C[:, :nocc] @ C[:, :nocc].T, cuda)
        F64 = e64.compute_fock_matrix_occ(D, Ct, nocc, H).cpu().numpy()
        Fmx = emx.compute_fock_matrix_occ(D, Ct, nocc, H).cpu().numpy()
        err = np.abs(Fmx - F64).max() / np.abs(F64).max()
        worst = max(worst, err)
        print("%s build %d: max|F_mixed - F_fp64| / max|F| = %.2e" % (system, k, err))
        assert err < 1e-8, (k, err)
    assert emx.stats()["ri_incremental_builds"] == len(seq) - 1
    print("%s worst %.2e" % (system, worst))


def test_ri_mixed_fresh_build_and_refresh_policy(cuda):
    """A mixed RI build whose orbitals are unrelated to the previous ones
    (the alignment cannot follow them) starts afresh on the fp16 matrix cores
    (K = T T^T from T = B X at the accuracy of the split products: within
    1e-7 of max|F| of the FP64 build); small changes after it are increments,
    and once the summed change is large and the next change small, the build
    is an FP64 one (exact to rounding)."""
    Z, xyz = methane_water()
    mol, _ = both(Z, xyz, "cc-pvdz")
    aux, _ = both(Z, xyz, "cc-pvdz-rifit")
    M, nocc = mol.n_basis, mol.n_electrons // 2
    H = _torch(random_density(M, seed=2), cuda)
    e64 = FockEngine(mol, aux=aux, eri_method="ri", ri_k="occupied")
    e64.precomputation()
    emx = FockEngine(mol, aux=aux, eri_method="ri", precision="mixed")
    emx.precomputation()
    rng = np.random.default_rng(9)
    seq = [np.linalg.qr(rng.normal(size=(M, M)))[0]]  # FP64 (no state yet)
    seq.append(np.linalg.qr(rng.normal(size=(M, M)))[0])  # unrelated: fresh
    seq += _orbital_sequence(M, nocc, 4, seed=5, amp=1e-5)[1:]  # small moves: increments
    seq[2:] = [np.linalg.qr(seq[1] + 1e-5 * (k + 1) * rng.normal(size=(M, M)) / np.sqrt(M))[0] for k in range(3)]
    kinds = []
    for k, C in enumerate(seq):
        Ct = _torch(C, cuda)
        D = _torch(2.0 * C[:, :nocc] @ C[:, :nocc].T, cuda)
        before = emx.stats()["ri_incremental_builds"]
        F64 = e64.compute_fock_matrix_occ(D, Ct, nocc, H).cpu().numpy()
        Fmx = emx.compute_fock_matrix_occ(D, Ct, nocc, H).cpu().numpy()
        err = np.abs(Fmx - F64).max() / np.abs(F64).max()
        kinds.append(emx.stats()["ri_incremental_builds"] - before)
        print("build %d: incremental %d, max|F_mixed - F_fp64| / max|F| = %.2e" % (k, kinds[-1], err))
        assert err < (1e-7 if k == 1 else 1e-12 if k == 0 else 1e-7), (k, err)
    # build 0: FP64; build 1: fresh (not an increment); then increments until the
    # summed change (1 from the fresh build) asks for the FP64 wipe on a small change
    assert kinds[0] == 0 and kinds[1] == 0
    assert kinds[2] == 0  # small change after a fresh build: the FP64 wipe
    assert kinds[3] == 1  # then increments
