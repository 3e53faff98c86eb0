"""Pins the FP64 host oracle (CPU only).

Every oracle function is checked against something that does not share its
code: the reference's own Boys grid (golden fixture), closed forms, an
independent Obara-Saika recursion, textbook energies, symmetry/invariance
properties and the reference's unit-test known answers.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import special

import oracle_lib as O
from helpers import both, random_density
from dft_gansu_amd import molecule as gm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ Boys ----
def dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def boys_closed(n, x):
    """F_n(x) = 1F1(n+1/2; n+3/2; -x)/(2n+1) = Gamma(n+1/2) P(n+1/2, x) / (2 x^(n+1/2))."""
    if x < 30.0:  # hyp1f1 is the more accurate scipy routine below ~30
        return special.hyp1f1(n + 0.5, n + 1.5, -x) / (2 * n + 1)
    return special.gamma(n + 0.5) * special.gammainc(n + 0.5, x) / (2.0 * x ** (n + 0.5))


@pytest.mark.parametrize("x", [0.0, 1e-8, 0.03, 0.5, 1.7, 5.0, 12.3, 25.0, 31.9, 40.0, 80.0, 150.0])
def test_boys_exact_matches_closed_form(x):
    F = O.boys(24, x)
    for n in range(25):
        ref = boys_closed(n, x)
        assert abs(F[n] - ref) <= 2e-15 * max(ref, 1e-300) + 1e-300, (n, x, F[n], ref)


def test_boys_zero_and_asymptote():
    F = O.boys(20, 0.0)
    np.testing.assert_allclose(F, [1.0 / (2 * n + 1) for n in range(21)], rtol=0, atol=1e-16)
    x = 200.0
    F = O.boys(10, x)
    for n in range(11):
        assert F[n] == pytest.approx(dfact(2 * n - 1) / 2 ** (n + 1) * math.sqrt(math.pi / x ** (2 * n + 1)),
                                     rel=1e-14)


def test_boys_table_matches_reference_grid():
    """The oracle's F_n(i/32) grid, rounded to 15 decimals, against the
    reference's boys_30720.h (fixture: digest + 640 sampled entries)."""
    fx = json.load(open(os.path.join(GOLDEN, "boys_ref_samples.json")))
    s = np.array(fx["samples"])
    idx = (s[:, 0] * 1024 + s[:, 1]).astype(int)
    t15 = O.boys_table(round15=True)
    exact = O.boys_table(round15=False)
    assert np.abs(t15[idx] - s[:, 2]).max() <= 1.01e-15
    assert np.mean(t15[idx] == s[:, 2]) > 0.99
    assert np.abs(exact[idx] - s[:, 2]).max() <= 1e-15


@pytest.mark.parametrize("x", [1e-6, 0.004, 0.2, 0.9, 3.3, 9.99, 18.0, 27.0, 29.5, 35.0])
def test_boys_reference_style_close_to_exact(x):
    tab = O.boys_table(round15=True)
    Fr = O.boys(16, x, reference_style=True, table=tab)
    Fe = O.boys(16, x)
    np.testing.assert_allclose(Fr, Fe, rtol=5e-12, atol=1e-16)


# --------------------------------------------------------------- ERIs ------
def test_ssss_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(20):
        ctr = rng.uniform(-2, 2, (4, 3))
        ex = rng.uniform(0.1, 5.0, 4)
        p, q = ex[0] + ex[1], ex[2] + ex[3]
        P = (ex[0] * ctr[0] + ex[1] * ctr[1]) / p
        Q = (ex[2] * ctr[2] + ex[3] * ctr[3]) / q
        Kab = math.exp(-ex[0] * ex[1] / p * np.sum((ctr[0] - ctr[1]) ** 2))
        Kcd = math.exp(-ex[2] * ex[3] / q * np.sum((ctr[2] - ctr[3]) ** 2))
        T = p * q / (p + q) * np.sum((P - Q) ** 2)
        F0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T)) if T > 0 else 1.0
        ref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0
        assert O.eri_prim_md(ctr, ex, [0] * 12) == pytest.approx(ref, rel=1e-13)


def _components(L):
    return [(i, j, L - i - j) for i in range(L, -1, -1) for j in range(L - i, -1, -1)]


@pytest.mark.parametrize("ls", [(1, 0, 0, 0), (1, 1, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 0, 0, 0), (2, 1, 1, 0),
                                (2, 2, 0, 0), (2, 0, 2, 0), (2, 1, 2, 1), (2, 2, 2, 2), (3, 0, 2, 0), (3, 2, 0, 0)])
def test_md_matches_obara_saika(ls):
    """Two unrelated recursions (McMurchie-Davidson vs Obara-Saika/HGP)."""
    rng = np.random.default_rng(sum(ls) * 7 + ls[0])
    ctr = rng.uniform(-1.5, 1.5, (4, 3))
    ex = rng.uniform(0.2, 3.0, 4)
    comps = [_components(l) for l in ls]
    picks = rng.integers(0, 10 ** 6, size=(24, 4))
    for pk in picks:
        ang = []
        for k in range(4):
            ang.extend(comps[k][pk[k] % len(comps[k])])
        md = O.eri_prim_md(ctr, ex, ang)
        os_ = O.eri_prim_os(ctr, ex, ang)
        assert md == pytest.approx(os_, rel=1e-11, abs=1e-14), ang


@pytest.fixture(scope="module")
def water_ccpvdz():
    (Z, xyz) = gm.water()
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz", normalization="unit")
    Q = O.schwarz(p, t)
    eri = O.eri_dense(p, t, n, Q, 0.0)
    return Z, xyz, p, t, n, eri


def test_eri_eightfold_symmetry(water_ccpvdz):
    eri = water_ccpvdz[-1]
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        np.testing.assert_allclose(eri, eri.transpose(perm), rtol=0, atol=1e-13)


def test_eri_schwarz_inequality_and_psd(water_ccpvdz):
    eri = water_ccpvdz[-1]
    M = eri.shape[0]
    G = eri.reshape(M * M, M * M)
    d = np.sqrt(np.clip(np.diag(G), 0, None))
    assert np.all(np.abs(G) <= np.outer(d, d) * (1 + 1e-12) + 1e-14)
    w = np.linalg.eigvalsh(G)
    assert w.min() > -1e-10 * w.max()


def test_overlap_unit_normalisation(water_ccpvdz):
    Z, xyz, p, t, n, _ = water_ccpvdz
    S, T, V = O.oneint(p, n, Z, xyz)
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-12)
    assert np.all(np.linalg.eigvalsh(S) > 0)


def test_overlap_ss_closed_form():
    """Two normalised s Gaussians: S = (2 sqrt(ab)/(a+b))^(3/2) exp(-ab/(a+b) R^2)."""
    import tempfile
    a, b, R = 0.8, 1.9, 1.3
    gbs = "!\n\nH     0\nS    1   1.00\n      %.10f       1.0\n****\nHe     0\nS    1   1.00\n      %.10f       1.0\n****\n" % (a, b)
    with tempfile.NamedTemporaryFile("w", suffix=".gbs", delete=False) as f:
        f.write(gbs)
        path = f.name
    try:
        p, t, n = O.build_molecule(path, [1, 2], [[0, 0, 0], [0, 0, R]], quirk=False)
        S, _, _ = O.oneint(p, n, [1, 2], [[0, 0, 0], [0, 0, R]])
        ref = (2 * math.sqrt(a * b) / (a + b)) ** 1.5 * math.exp(-a * b / (a + b) * R * R)
        assert S[0, 1] == pytest.approx(ref, rel=1e-13)
    finally:
        os.unlink(path)


# ---------------------------------------------------------------- SCF ------
def test_h2_sto3g_textbook_energy():
    """Szabo & Ostlund, H2/STO-3G at R = 1.4 bohr: E = -1.1167 Eh."""
    p, t, n = O.build_molecule(gm.basis_path("sto-3g"), [1, 1], [[0, 0, 0], [0, 0, 1.4]])
    e_elec, e_tot, D, F, it = O.rhf(p, t, n, [1, 1], [[0, 0, 0], [0, 0, 1.4]], 1)
    assert e_tot == pytest.approx(-1.1167, abs=1e-4)
    assert e_elec == pytest.approx(-1.8310, abs=1e-4)


def test_h2o_sto3g_textbook_energy():
    """Szabo & Ostlund table 3.13, H2O/STO-3G at R = 1.809 bohr, 104.52 deg: -74.963 Eh."""
    th = math.radians(104.52)
    R = 1.809
    xyz = [[0, 0, 0], [R, 0, 0], [R * math.cos(th), R * math.sin(th), 0]]
    p, t, n = O.build_molecule(gm.basis_path("sto-3g"), [8, 1, 1], xyz)
    e = O.rhf(p, t, n, [8, 1, 1], xyz, 5, conv=1e-10)[1]
    assert e == pytest.approx(-74.963, abs=1e-3)


def test_energy_invariant_to_normalisation_and_rotation():
    Z, xyz = gm.water(jitter=0.0)
    e = []
    for q in (True, False):
        p, t, n = O.build_molecule(gm.basis_path("cc-pvdz"), Z, xyz, quirk=q)
        e.append(O.rhf(p, t, n, Z, xyz, 5, conv=1e-11)[1])
    R = gm._random_rotation(np.random.default_rng(3))
    xr = (R @ xyz.T).T + np.array([0.3, -1.1, 2.0])
    p, t, n = O.build_molecule(gm.basis_path("cc-pvdz"), Z, xr, quirk=True)
    e.append(O.rhf(p, t, n, Z, xr, 5, conv=1e-11)[1])
    assert abs(e[0] - e[1]) < 1e-9
    assert abs(e[0] - e[2]) < 1e-8
    assert -76.1 < e[0] < -75.9


def test_jk_direct_equals_dense(water_ccpvdz):
    Z, xyz, p, t, n, eri = water_ccpvdz
    D = random_density(len(n))
    J, K = O.jk(p, t, n, O.schwarz(p, t), D, 0.0)
    np.testing.assert_allclose(J, np.einsum("ijkl,kl->ij", eri, D), atol=1e-11)
    np.testing.assert_allclose(K, np.einsum("ikjl,kl->ij", eri, D), atol=1e-11)


# ----------------------------------------------- reference known answers ----
def test_reference_cholesky_known_answer():
    """test/test_gpu_manager.cu:332-343."""
    A = np.array([[25, 15, -5], [15, 18, 0], [-5, 0, 11]], dtype=float)
    np.testing.assert_allclose(O.cholesky(A), [[5, 0, 0], [3, 3, 0], [-1, 1, 3]], atol=1e-12)


def test_reference_eigen_known_answer():
    """test/test_gpu_manager.cu:32-36."""
    w, V = O.eigh(np.array([[4.0, 1.0], [1.0, 3.0]]))
    np.testing.assert_allclose(w, [2.381966, 4.618034], atol=1e-6)
    A = random_density(30, seed=9) - 0.3
    w, V = O.eigh(A)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
    np.testing.assert_allclose(V @ np.diag(w) @ V.T, A, atol=1e-12)


def test_reference_optimal_damping_known_answer():
    """test/test_gpu_manager.cu:468-492: alpha = 0.613207547."""
    F = np.array([[2.0, 2.0, 1.0], [2.0, 3.0, 1.0], [1.0, 1.0, 2.0]])
    Fp = np.array([[1.0, 2.0, 3.0], [4.0, 6.0, 6.0], [7.0, 8.0, 10.0]])
    D = np.full((3, 3), 3.0)
    Dp = np.array([[2.0, 4.0, 6.0], [4.0, 1.0, 1.0], [6.0, 1.0, 10.0]])
    assert O.optimal_damping(F, Fp, D, Dp) == pytest.approx(0.613207547, abs=1e-8)


# ----------------------------------------------------------------- RI ------
def test_ri_B_reproduces_fitted_integrals():
    """sum_P B^P_mn B^P_ls == (mn|P) (P|Q)^-1 (Q|ls), and RI-J/K approach exact J/K."""
    Z, xyz = gm.water(jitter=0.0)
    mol, (p, t, n) = both(Z, xyz, "cc-pvdz")
    aux, (pa, ta, na) = both(Z, xyz, "cc-pvdz-rifit")
    Q = O.schwarz(p, t)
    Qa = O.schwarz_aux(pa)
    V = O.int2c2e(pa, na, Qa, 0.0)
    T = O.int3c2e(p, t, n, Q, pa, na, Qa, 0.0)
    np.testing.assert_allclose(V, V.T, atol=1e-12)
    np.testing.assert_allclose(T, T.transpose(0, 2, 1), atol=1e-12)
    L = O.cholesky(V)
    B = O.ri_B(L, T)
    M, Ma = len(n), len(na)
    lhs = B.reshape(Ma, -1).T @ B.reshape(Ma, -1)
    Tm = T.reshape(Ma, -1)
    rhs = Tm.T @ np.linalg.solve(V, Tm)
    assert np.abs(lhs - rhs).max() < 1e-9 * np.abs(rhs).max()
    D = random_density(M, seed=4)
    Jri, Kri = O.ri_jk(B, D)
    J, K = O.jk(p, t, n, Q, D, 0.0)
    assert np.abs(Jri - J).max() < 2e-3 * np.abs(J).max()
    assert np.abs(Kri - K).max() < 2e-2 * np.abs(K).max()


# ------------------------------------------------------------------ ROHF ----
def _s_orthonormal(S, seed):
    """Columns orthonormal in the S metric: C^T S C = 1, hence C C^T = S^-1."""
    w, V = np.linalg.eigh(S)
    X = V @ np.diag(w ** -0.5) @ V.T
    Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=S.shape))
    return X @ Q


def test_rohf_closed_shell_limit_is_rhf(water_ccpvdz):
    """No open shell: F_o = F_c / 2, so the unified matrix is (Acc + Bcc) A in
    the closed block, A between closed and virtual and (Avv + Bvv) A in the
    virtual block, A = C^T F_RHF C; for S-orthonormal C (C C^T = S^-1) the
    sets whose sums are 1 (all but davidson, whose table entry has
    Acc + Bcc = 3/2) return F = F_RHF -- a transposed C breaks it."""
    Z, xyz, p, t, n, _ = water_ccpvdz
    S, T, V = O.oneint(p, n, Z, xyz)
    H = T + V
    D = random_density(len(n))
    JK = O.jk(p, t, n, O.schwarz(p, t), D, 0.0)
    C = _s_orthonormal(S, 3)
    assert np.abs(C.T @ S @ C - np.eye(len(n))).max() < 1e-10
    F_rhf = H + JK[0] - 0.5 * JK[1]
    zero = (np.zeros_like(D), np.zeros_like(D))
    A = C.T @ F_rhf @ C
    nc = 5
    c, v = slice(0, nc), slice(nc, len(n))
    for name, (Acc, Bcc, Aoo, Boo, Avv, Bvv) in O.ROHF_COUPLING.items():
        Fc, Fo, F = O.rohf_fock(H, JK, zero, C, S, nc, 0, name)
        np.testing.assert_allclose(Fc, F_rhf, atol=1e-12)
        np.testing.assert_allclose(Fo, 0.5 * F_rhf, atol=1e-12)
        R = C.T @ F @ C
        tol = 1e-8 * np.abs(A).max()
        assert np.abs(R[c, c] - (Acc + Bcc) * A[c, c]).max() < tol, name
        assert np.abs(R[c, v] - A[c, v]).max() < tol, name
        assert np.abs(R[v, v] - (Avv + Bvv) * A[v, v]).max() < tol, name
        if abs(Acc + Bcc - 1) < 1e-12 and abs(Avv + Bvv - 1) < 1e-12:
            assert np.abs(F - F_rhf).max() < 1e-8 * np.abs(F_rhf).max(), name


def test_rohf_couplings_only_change_the_diagonal_blocks(water_ccpvdz):
    """The coupling set enters the closed-closed, open-open and virtual-virtual
    blocks only: R = C^T F C (S-orthonormal C) agrees between sets in the
    closed-open, closed-virtual and open-virtual blocks and differs in the
    diagonal ones; R is symmetric."""
    Z, xyz, p, t, n, _ = water_ccpvdz
    S, T, V = O.oneint(p, n, Z, xyz)
    H = T + V
    Q = O.schwarz(p, t)
    Dc, Do = random_density(len(n), seed=5), 0.3 * random_density(len(n), seed=6)
    JKc, JKo = O.jk(p, t, n, Q, Dc, 0.0), O.jk(p, t, n, Q, Do, 0.0)
    C = _s_orthonormal(S, 4)
    nc, no = 4, 2
    Rs = {}
    for name in O.ROHF_COUPLING:
        F = O.rohf_fock(H, JKc, JKo, C, S, nc, no, name)[2]
        R = C.T @ F @ C
        assert np.abs(R - R.T).max() < 1e-9 * np.abs(R).max()
        Rs[name] = R
    c, o, v = slice(0, nc), slice(nc, nc + no), slice(nc + no, len(n))
    ref = Rs["roothaan"]
    for name, R in Rs.items():
        for a, b in ((c, o), (c, v), (o, v)):
            assert np.abs(R[a, b] - ref[a, b]).max() < 1e-9 * np.abs(ref).max(), name
    assert np.abs(Rs["guest-saunders"][c, c] - ref[c, c]).max() > 1e-3
    assert np.abs(Rs["guest-saunders"][v, v] - ref[v, v]).max() > 1e-3


# ------------------------------------------------------------------- UHF ----
def test_uhf_oracle_known_answers():
    """UHF/STO-3G of the H atom (-0.466582 Eh, the STO-3G 1s energy) and the Li
    atom (-7.315526 Eh), and the closed-shell limit equal to RHF."""
    p = gm.basis_path("sto-3g")
    origin = np.zeros((1, 3))
    for Zv, na, nb, ref in ((1, 1, 0, -0.466582), (3, 2, 1, -7.315526)):
        Z = np.array([Zv])
        pr, t, n = O.build_molecule(p, Z, origin)
        assert abs(O.uhf(pr, t, n, Z, origin, na, nb)[1] - ref) < 2e-6, Zv
    Z, xyz = gm.water(jitter=0.01, seed=11)
    pr, t, n = O.build_molecule(p, Z, xyz)
    # from the symmetry-broken guess the closed shell returns to RHF along a
    # soft spin mode: the per-iteration energy change falls below the stop
    # threshold while the energy still lies 1e-11..5e-10 above the RHF one
    # (the stop point moves with the BLAS summation order), so the UHF energy
    # is bounded from both sides and the spin densities must have merged
    _, e_u, Da, Db, _ = O.uhf(pr, t, n, Z, xyz, 5, 5, conv=1e-12)
    e_r = O.rhf(pr, t, n, Z, xyz, 5, conv=1e-12)[1]
    assert -1e-11 < e_u - e_r < 2e-9
    assert np.abs(Da - Db).max() < 1e-4


# ------------------------------------------------------------------ ROHF ----
def test_rohf_oracle_pins():
    """ROHF: the H atom at the STO-3G 1s energy; closed-shell water equal to
    RHF; the converged energy of the Li doublet and the O triplet independent
    of the coupling set (the sets differ only in the diagonal blocks of the
    unified matrix, which do not enter the stationarity conditions)."""
    p = gm.basis_path("sto-3g")
    origin = np.zeros((1, 3))
    Z = np.array([1])
    pr, t, n = O.build_molecule(p, Z, origin)
    assert abs(O.rohf(pr, t, n, Z, origin, 0, 1)[1] - (-0.466582)) < 2e-6
    for Zv, nc, no in ((3, 1, 1), (8, 3, 2)):
        Z = np.array([Zv])
        pr, t, n = O.build_molecule(p, Z, origin)
        e = [O.rohf(pr, t, n, Z, origin, nc, no, c)[1] for c in ("roothaan", "guest-saunders", "mcweeny-diercksen")]
        assert max(e) - min(e) < 1e-9, (Zv, e)
    Z, xyz = gm.water(jitter=0.01, seed=11)
    pr, t, n = O.build_molecule(p, Z, xyz)
    assert abs(O.rohf(pr, t, n, Z, xyz, 5, 0)[1] - O.rhf(pr, t, n, Z, xyz, 5)[1]) < 1e-10


# --------------------------------------------------------------- Schwarz ----
# The primitive-pair factor of schwarz.txt:239-478 is
#   Q_ab = sqrt(max over Cartesian components of c_a^2 c_b^2 N_a^2 N_b^2 [ab|ab])
# with the contraction coefficients and primitive norms included and the
# contracted norms left out.  These tests recompute it through the
# Obara-Saika recursion (orc_eri_prim_os, which shares no code with the
# McMurchie-Davidson path of orc_schwarz) and the textbook primitive norm,
# and check it bounds every primitive quartet.

def _components(L):
    return [(l, m, L - l - m) for l in range(L, -1, -1) for m in range(L - l, -1, -1)]


def _prim_norm(a, c):
    """Textbook normalisation of x^l y^m z^n exp(-a r^2)."""
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (sum(c) / 2) / math.sqrt(
        dfact(2 * c[0] - 1) * dfact(2 * c[1] - 1) * dfact(2 * c[2] - 1))


def _reference_pairs(types):
    """(a, b) primitive indices in the reference's pair order
    (gpu_manager.cu:413-432): pair types s0 <= s1 row-major over types; within
    a type the upper triangle r1 <= r2 (r2 outer) for s0 = s1, row-major
    (r1 outer) otherwise."""
    out = []
    for s0 in range(len(types)):
        for s1 in range(s0, len(types)):
            c0, c1 = int(types[s0]["count"]), int(types[s1]["count"])
            b0, b1 = int(types[s0]["start_index"]), int(types[s1]["start_index"])
            if s0 == s1:
                out += [(b0 + r1, b0 + r2) for r2 in range(c0) for r1 in range(r2 + 1)]
            else:
                out += [(b0 + r1, b1 + r2) for r1 in range(c0) for r2 in range(c1)]
    return out


def _os_block(prs, idx, exps_override=None):
    """c N [ab|cd] for all components, primitives idx = (a, b, c, d); an index
    of None is an s function of exponent 0 at its partner's centre (the
    missing index of a 2-centre integral), with coefficient and norm 1."""
    cen, ex, ls, cf = [], [], [], []
    for k, i in enumerate(idx):
        if i is None:
            j = idx[k - 1]
            cen.append([prs[j]["x"], prs[j]["y"], prs[j]["z"]])
            ex.append(0.0)
            ls.append(None)
            cf.append(1.0)
        else:
            cen.append([prs[i]["x"], prs[i]["y"], prs[i]["z"]])
            ex.append(float(prs[i]["exponent"]))
            ls.append(int(prs[i]["shell_type"]))
            cf.append(float(prs[i]["coefficient"]))
    comps = [[(0, 0, 0)] if L is None else _components(L) for L in ls]
    out = {}
    for ca in comps[0]:
        for cb in comps[1]:
            for cc in comps[2]:
                for cd in comps[3]:
                    v = O.eri_prim_os(np.array(cen), np.array(ex), np.array(ca + cb + cc + cd))
                    nrm = 1.0
                    for k, c in enumerate((ca, cb, cc, cd)):
                        if ls[k] is not None:
                            nrm *= _prim_norm(ex[k], c) * cf[k]
                    out[(ca, cb, cc, cd)] = nrm * v
    return out


def test_schwarz_factor_is_the_tight_primitive_bound():
    """Q_ab equals sqrt(max_comp c^2 N^2 [ab|ab]) recomputed by Obara-Saika,
    for pairs of every pair type (s..d on H2O/cc-pVDZ, quirk normalisation:
    Q does not involve the contracted norm)."""
    Z, xyz = gm.water(jitter=0.01, seed=3)
    p, t, n = O.build_molecule(gm.basis_path("cc-pvdz"), Z, xyz)
    Q = O.schwarz(p, t)
    pairs = _reference_pairs(t)
    assert len(pairs) == len(Q) == O.n_pairs(t)
    rng = np.random.default_rng(5)
    sample = sorted(set(rng.choice(len(pairs), 160, replace=False).tolist()) | {0, len(pairs) - 1})
    checked_types = set()
    for k in sample:
        a, b = pairs[k]
        blk = _os_block(p, (a, b, a, b))
        best = max(v for (ca, cb, cc, cd), v in blk.items() if ca == cc and cb == cd)
        assert best > 0.0
        assert abs(Q[k] - math.sqrt(best)) <= 1e-12 * math.sqrt(best), (k, Q[k], math.sqrt(best))
        checked_types.add((int(p[a]["shell_type"]), int(p[b]["shell_type"])))
    assert {(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)} <= checked_types


def test_schwarz_factor_bounds_every_primitive_quartet():
    """|c N [ab|cd]| <= Q_ab Q_cd for every component of sampled primitive
    quartets (the screening test of int2e.cu:1681 drops a quartet only when
    this product is below the threshold), with equality reached on the
    diagonal (ab|ab) at the maximising component."""
    Z, xyz = gm.water(jitter=0.01, seed=3)
    p, t, n = O.build_molecule(gm.basis_path("cc-pvdz"), Z, xyz)
    Q = O.schwarz(p, t)
    pairs = _reference_pairs(t)
    rng = np.random.default_rng(9)
    for _ in range(60):
        x, y = rng.integers(len(pairs), size=2)
        (a, b), (c, d) = pairs[x], pairs[y]
        blk = _os_block(p, (a, b, c, d))
        assert max(abs(v) for v in blk.values()) <= Q[x] * Q[y] * (1 + 1e-12)
        diag = _os_block(p, (a, b, a, b))
        assert abs(max(diag.values()) - Q[x] ** 2) <= 1e-12 * Q[x] ** 2


def test_schwarz_aux_factor_closed_form_and_os():
    """Auxiliary factor sqrt(max_comp c^2 N^2 (a|a)) (schwarz.txt:479+): for an
    s primitive the closed form (a|a) = 2 pi^(5/2) / (a^2 sqrt(2a)) times
    c^2 N^2; for every shell type Obara-Saika with exponent-0 partners."""
    Z, xyz = gm.water(jitter=0.01, seed=3)
    ap, at, an = O.build_molecule(gm.basis_path("cc-pvdz-rifit"), Z, xyz, quirk=False)
    Qa = O.schwarz_aux(ap)
    for i in range(len(ap)):
        a, L, c = float(ap[i]["exponent"]), int(ap[i]["shell_type"]), float(ap[i]["coefficient"])
        if L == 0:
            ref = c * c * _prim_norm(a, (0, 0, 0)) ** 2 * 2 * math.pi ** 2.5 / (a * a * math.sqrt(2 * a))
            assert abs(Qa[i] - math.sqrt(ref)) <= 1e-12 * math.sqrt(ref)
    rng = np.random.default_rng(2)
    for L in range(4):
        idx = [i for i in range(len(ap)) if int(ap[i]["shell_type"]) == L]
        for i in rng.choice(idx, min(5, len(idx)), replace=False):
            blk = _os_block(ap, (int(i), None, int(i), None))
            best = max(v for (ca, cb, cc, cd), v in blk.items() if ca == cc)
            assert abs(Qa[i] - math.sqrt(best)) <= 1e-12 * math.sqrt(best), (L, i)


def test_screened_jk_within_the_dropped_bound():
    """J and K with the threshold raised to 1e-6 stay within the bound the
    dropped primitive quartets allow: each holds |value| <= Q_ab Q_cd < thr
    per component, so an element moves by at most
    8 thr max(N)^4 max|D| x (number of dropped quartets).  A factor that is too
    small (a missing coefficient, a wrong norm power) drops larger
    integrals and breaks it."""
    Z, xyz = gm.water(jitter=0.01, seed=3)
    p, t, n = O.build_molecule(gm.basis_path("cc-pvdz"), Z, xyz)
    Q = O.schwarz(p, t)
    D = random_density(len(n), seed=4)
    J0, K0 = O.jk(p, t, n, Q, D, thr=0.0)
    thr = 1e-6
    J1, K1 = O.jk(p, t, n, Q, D, thr=thr)
    # S = sum of Q_x Q_y over unordered pairs of primitive pairs with
    # Q_x Q_y < thr; each such quartet moves one element of J or K by at
    # most 8 (permutations) x 36 (components of the other pair) x Q_x Q_y x
    # max(N)^4 x max|D|
    qs = np.sort(Q)
    pre = np.concatenate(([0.0], np.cumsum(qs)))
    below = np.searchsorted(qs, thr / np.maximum(qs, 1e-300), side="left")
    k = np.minimum(below, np.arange(1, len(qs) + 1))
    S = float(np.sum(qs * pre[k]))
    assert S > 0.0
    bound = 8.0 * 36.0 * np.abs(n).max() ** 4 * np.abs(D).max() * S
    dj, dk = np.abs(J1 - J0).max(), np.abs(K1 - K0).max()
    assert dj <= bound and dk <= bound, (dj, dk, bound)
    # the screening does drop something visible at this threshold
    assert dj > 0.0 and dk > 0.0


def test_uhf_symmetry_break_reaches_the_broken_symmetry_state():
    """Every UHF guess zeroes the last row and column of C_beta
    (break_symmetry, uhf.hpp:429-436).  For H2/STO-3G stretched to 4 bohr
    that takes the core guess off the restricted solution: the UHF energy
    lies far below RHF and near two STO-3G hydrogen atoms (2 x -0.466582 Eh),
    with unequal spin densities; at 1.4 bohr UHF returns to RHF
    (Szabo-Ostlund -1.1167 Eh)."""
    p = gm.basis_path("sto-3g")
    Z = np.array([1, 1])
    for R, below in ((4.0, True), (1.4, False)):
        xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, R]])
        pr, t, n = O.build_molecule(p, Z, xyz)
        er = O.rhf(pr, t, n, Z, xyz, 1, conv=1e-12)[1]
        eu, Da, Db = (lambda r: (r[1], r[2], r[3]))(O.uhf(pr, t, n, Z, xyz, 1, 1, conv=1e-12, maxiter=300))
        if below:
            assert eu < er - 0.1, (eu, er)
            assert abs(eu - 2 * -0.466582) < 5e-3, eu
            assert np.abs(Da - Db).max() > 0.1
        else:
            assert abs(eu - er) < 1e-8 and abs(er - (-1.1167)) < 1e-4, (eu, er)


def test_damped_scf_reaches_the_diis_energy():
    """Constant and optimal damping (rhf.hpp:182-262, rohf.hpp:216-306,
    uhf.hpp:196-258) change the path, not the fixed point: each damped SCF
    ends on the DIIS energy (RHF water, UHF Li, ROHF O triplet; STO-3G)."""
    p = gm.basis_path("sto-3g")
    Z, xyz = gm.water(jitter=0.01, seed=11)
    pr, t, n = O.build_molecule(p, Z, xyz)
    e_diis = O.rhf(pr, t, n, Z, xyz, 5, conv=1e-12)[1]
    for optimal in (False, True):
        e, _, _, it = (lambda r: (r[1], r[2], r[2], r[3]))(O.rhf_damped(pr, t, n, Z, xyz, 5, optimal=optimal,
                                                                          conv=1e-12))
        assert abs(e - e_diis) < 1e-9, (optimal, e, e_diis, it)
    origin = np.zeros((1, 3))
    Zli = np.array([3])
    pr, t, n = O.build_molecule(p, Zli, origin)
    assert abs(O.uhf(pr, t, n, Zli, origin, 2, 1, damping=0.9, conv=1e-12, maxiter=500)[1]
               - O.uhf(pr, t, n, Zli, origin, 2, 1, conv=1e-12)[1]) < 1e-9
    Zo = np.array([8])
    pr, t, n = O.build_molecule(p, Zo, origin)
    e_ref = O.rohf(pr, t, n, Zo, origin, 3, 2, conv=1e-12)[1]
    for kw in (dict(damping=0.9), dict(optimal=True)):
        assert abs(O.rohf(pr, t, n, Zo, origin, 3, 2, conv=1e-12, maxiter=500, **kw)[1] - e_ref) < 1e-9, kw


def test_md_flop_counter_hand_counts():
    """The op counter of the MD routine (orc_md_flops), pinned by hand counts
    of the routine's operations.  (ss|ss) at x = 1 (Boys closed form with
    exp and erf): alpha 3, P - Q 3, Boys argument 6, Boys 5, R^0 seed 3,
    prefactor 6, the contraction term 8 and its scaling 1, accumulation 2:
    37 operations; special calls sqrt, exp, sqrt, erf in Boys and sqrt in the
    prefactor: 5.  At x = 30 (beyond the exp/erf range) one multiply and
    exp/erf/one sqrt drop out: 36 and 2.  (ps|ss) at x = 1: Boys 8, R^0..1
    seeds 6, three first-order R entries 3, per p component two Hermite
    terms of 8 plus the scaling: 92."""
    tab = O.boys_table()
    assert O.md_flops(0, 0, 0, 0, 1.0, tab) == (37, 5)
    assert O.md_flops(0, 0, 0, 0, 30.0, tab) == (36, 2)
    assert O.md_flops(1, 0, 0, 0, 1.0, tab) == (92, 5)
    # derivative routine, (ss|ss): the (ps|ss)-order R table 3 + 3 + 23,
    # prefactor 6, coefficient product 3, norms 8, Gamma 7, factor 4, and 12
    # derivatives: a shifted bra index gives 2 Hermite terms of 8 ops, a
    # shifted ket index 2 + 2 x 6, each plus 3 to scale and accumulate:
    # 6 x 19 + 6 x 17 = 216; 273 in all
    assert O.grad_md_flops(0, 0, 0, 0, 1.0, tab) == (273, 5)


def test_md_flops_fixture_is_current():
    """tests/golden/md_flops.json (bench.py's per-class algorithmic flops)
    holds what the oracle counts now, for all 21 classes up to (dd|dd), and
    the count grows with the class: adding angular momentum on any centre
    never lowers it."""
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "md_flops.json")))
    tab = O.boys_table()
    assert len(fx["classes"]) == 21 and len(fx["gradient_classes"]) == 21
    for key, (arith, special) in fx["classes"].items():
        la, lb, lc, ld = map(int, key.split())
        assert O.md_flops(la, lb, lc, ld, fx["x_boys"], tab) == (arith, special), key
        assert O.grad_md_flops(la, lb, lc, ld, fx["x_boys"], tab) == tuple(fx["gradient_classes"][key]), key
        assert fx["gradient_classes"][key][0] > 6 * arith  # 12 derivatives of every component
    c = {tuple(map(int, k.split())): v[0] for k, v in fx["classes"].items()}
    assert c[(1, 1, 1, 1)] > c[(1, 1, 1, 0)] > c[(1, 0, 1, 0)] > c[(1, 0, 0, 0)] > c[(0, 0, 0, 0)]
    assert c[(2, 2, 2, 2)] > c[(2, 2, 2, 1)] > c[(2, 2, 2, 0)]
