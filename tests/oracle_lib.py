"""ctypes binding of the FP64 host oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP library and is
loaded from tests, the smoke check and the CPU-baseline leg of the benchmark.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB = None

PRIM_DTYPE = np.dtype([("exponent", "f8"), ("coefficient", "f8"), ("x", "f8"), ("y", "f8"), ("z", "f8"),
                       ("shell_type", "i4"), ("_pad0", "i4"), ("basis_index", "u8"), ("atom_index", "i4"),
                       ("_pad1", "i4")])
TYPE_DTYPE = np.dtype([("count", "i4"), ("_pad", "i4"), ("start_index", "u8")])
assert PRIM_DTYPE.itemsize == 64 and TYPE_DTYPE.itemsize == 16

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(so):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        _LIB = ctypes.CDLL(so)
        L = _LIB
        L.orc_cgto_norm.restype = ctypes.c_double
        L.orc_schwarz.restype = ctypes.c_size_t
        L.orc_optimal_damping.restype = ctypes.c_double
        L.orc_eri_prim_os.restype = ctypes.c_double
        L.orc_eri_prim_md.restype = ctypes.c_double
        L.orc_last_error.restype = ctypes.c_char_p
    return _LIB


def boys(n, x, reference_style=False, table=None):
    out = np.zeros(n + 1)
    tab = table if table is not None else np.zeros(1)
    lib().orc_boys(ctypes.c_int(n), ctypes.c_double(x), ctypes.c_int(int(reference_style)), _ptr(tab), _ptr(out))
    return out


def boys_table(round15=True):
    out = np.zeros(30 * 1024)
    lib().orc_boys_table(ctypes.c_int(int(round15)), _ptr(out))
    return out


def build_molecule(gbs_path, Z, xyz_bohr, quirk=True):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
    npr, nb, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_build_molecule(gbs_path.encode(), ctypes.c_int(len(Z)), _ptr(Z), _ptr(xyz), ctypes.c_int(int(quirk)),
                                  ctypes.byref(npr), ctypes.byref(nb), ctypes.byref(nt))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode())
    prims = np.zeros(npr.value, PRIM_DTYPE)
    types = np.zeros(nt.value, TYPE_DTYPE)
    norm = np.zeros(nb.value)
    lib().orc_get_molecule(_ptr(prims), _ptr(types), _ptr(norm))
    return prims, types, norm


def n_pairs(types):
    c = [int(t["count"]) for t in types]
    n = 0
    for i in range(len(c)):
        for j in range(i, len(c)):
            n += c[i] * (c[i] + 1) // 2 if i == j else c[i] * c[j]
    return n


def schwarz(prims, types, aux_order=False):
    Q = np.zeros(n_pairs(types))
    lib().orc_schwarz(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), ctypes.c_int(int(aux_order)), _ptr(Q))
    return Q


def schwarz_aux(prims):
    Q = np.zeros(len(prims))
    lib().orc_schwarz_aux(_ptr(prims), ctypes.c_int(len(prims)), _ptr(Q))
    return Q


def eri_dense(prims, types, norm, Q, thr=1e-12):
    M = len(norm)
    out = np.zeros((M, M, M, M))
    lib().orc_eri_dense(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), _ptr(norm), ctypes.c_int(M), _ptr(Q),
                        ctypes.c_double(thr), _ptr(out))
    return out


def jk(prims, types, norm, Q, D, thr=1e-12, sample_stride=1):
    M = len(norm)
    D = np.ascontiguousarray(D, dtype=np.float64)
    J = np.zeros((M, M))
    K = np.zeros((M, M))
    lib().orc_jk(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), _ptr(norm), ctypes.c_int(M), _ptr(Q),
                 ctypes.c_double(thr), _ptr(D), _ptr(J), _ptr(K), ctypes.c_int(sample_stride))
    return J, K


def oneint(prims, norm, Z, xyz_bohr):
    M = len(norm)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
    S, T, V = np.zeros((M, M)), np.zeros((M, M)), np.zeros((M, M))
    lib().orc_oneint(_ptr(prims), ctypes.c_int(len(prims)), _ptr(norm), ctypes.c_int(M), ctypes.c_int(len(Z)), _ptr(Z),
                     _ptr(xyz), _ptr(S), _ptr(T), _ptr(V))
    return S, T, V


def eigh(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_eigh(ctypes.c_int(n), _ptr(A), _ptr(w), _ptr(V))
    return w, V


def cholesky(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((n, n))
    if lib().orc_cholesky(ctypes.c_int(n), _ptr(A), _ptr(L)):
        raise np.linalg.LinAlgError("not SPD")
    return L


def optimal_damping(F, Fp, D, Dp):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (F, Fp, D, Dp)]
    return lib().orc_optimal_damping(ctypes.c_int(arrs[0].shape[0]), *[_ptr(a) for a in arrs])


def read_sad(path, symbol):
    """(Da, Db) of one element of a .sad file: a header "<symbol> <n>", then the
    n x n alpha and n x n beta densities as whitespace-separated rows
    (the format read_density_from_sad parses, utils.cpp:21-58)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if len(tok) == 2 and not tok[0][0].isdigit() and tok[0][0] != "-":
            n = int(tok[1])
            if tok[0] == symbol:
                vals = np.array(" ".join(lines[i + 1:i + 1 + 2 * n]).split(), dtype=np.float64)
                return vals[:n * n].reshape(n, n), vals[n * n:2 * n * n].reshape(n, n)
            i += 1 + 2 * n
            continue
        i += 1
    raise KeyError(symbol)


def nuclear_repulsion(Z, xyz_bohr):
    xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for a in range(len(Z)):
        for b in range(a):
            e += Z[a] * Z[b] / np.linalg.norm(xyz[a] - xyz[b])
    return e


def rhf(prims, types, norm, Z, xyz_bohr, nocc, thr=1e-12, conv=1e-10, maxiter=100, diis_size=8, dense=True,
        diis_include_transform=False, riB=None, D0=None):
    """Full FP64 RHF on the host.  Returns (E_elec, E_total, D, F, iterations).
    riB (Maux x M x M) selects RI-J/K Fock builds; D0 (M x M) replaces the
    core guess by F(D0) (the SAD guess with D0 = Da + Db, rhf.hpp:551-584)."""
    M = len(norm)
    S, T, V = oneint(prims, norm, Z, xyz_bohr)
    H = np.ascontiguousarray(T + V)
    Q = schwarz(prims, types)
    eri = eri_dense(prims, types, norm, Q, thr) if (dense and riB is None) else None
    if riB is not None:
        riB = np.ascontiguousarray(riB)
    enuc = nuclear_repulsion(Z, xyz_bohr)
    en = np.zeros(2)
    D = np.zeros((M, M))
    F = np.zeros((M, M))
    if D0 is not None:
        D0 = np.ascontiguousarray(D0, dtype=np.float64)
    it = lib().orc_rhf_guess(ctypes.c_int(M), ctypes.c_int(nocc), _ptr(S), _ptr(H),
                       _ptr(eri) if eri is not None else None, _ptr(prims), _ptr(types), ctypes.c_int(len(types)),
                       _ptr(norm), _ptr(Q), ctypes.c_double(thr), ctypes.c_double(enuc), ctypes.c_int(maxiter),
                       ctypes.c_double(conv), ctypes.c_int(diis_size), ctypes.c_int(int(diis_include_transform)),
                       _ptr(en), _ptr(D), _ptr(F), _ptr(riB) if riB is not None else None,
                       ctypes.c_int(riB.shape[0] if riB is not None else 0),
                       _ptr(D0) if D0 is not None else None)
    return en[0], en[1], D, F, it


def eri_prim_os(ctr, ex, ang):
    ctr = np.ascontiguousarray(ctr, dtype=np.float64).reshape(-1)
    ex = np.ascontiguousarray(ex, dtype=np.float64)
    ang = np.ascontiguousarray(ang, dtype=np.int32)
    return lib().orc_eri_prim_os(_ptr(ctr), _ptr(ex), _ptr(ang))


def md_flops(la, lb, lc, ld, x_boys=1.0, table=None):
    """(arithmetic ops, special-function calls) of one primitive quartet of
    class (la lb | lc ld) in the oracle's MD routine (orc_md_flops)."""
    tab = boys_table() if table is None else table
    out = np.zeros(2)
    rc = lib().orc_md_flops(ctypes.c_int(la), ctypes.c_int(lb), ctypes.c_int(lc), ctypes.c_int(ld),
                            ctypes.c_double(x_boys), _ptr(tab), _ptr(out))
    if rc:
        raise ValueError("orc_md_flops: bad class")
    return int(out[0]), int(out[1])


def grad_md_flops(la, lb, lc, ld, x_boys=1.0, table=None):
    """(arithmetic ops, special-function calls) of one primitive quartet of
    class (la lb | lc ld) in the oracle's derivative routine (orc_grad_md_flops)."""
    tab = boys_table() if table is None else table
    out = np.zeros(2)
    rc = lib().orc_grad_md_flops(ctypes.c_int(la), ctypes.c_int(lb), ctypes.c_int(lc), ctypes.c_int(ld),
                                 ctypes.c_double(x_boys), _ptr(tab), _ptr(out))
    if rc:
        raise ValueError("orc_grad_md_flops: bad class")
    return int(out[0]), int(out[1])


def eri_prim_md(ctr, ex, ang):
    ctr = np.ascontiguousarray(ctr, dtype=np.float64).reshape(-1)
    ex = np.ascontiguousarray(ex, dtype=np.float64)
    ang = np.ascontiguousarray(ang, dtype=np.int32)
    return lib().orc_eri_prim_md(_ptr(ctr), _ptr(ex), _ptr(ang))


def int2c2e(aux_prims, aux_norm, Qaux, thr=1e-12):
    Ma = len(aux_norm)
    out = np.zeros((Ma, Ma))
    lib().orc_int2c2e(_ptr(aux_prims), ctypes.c_int(len(aux_prims)), _ptr(aux_norm), ctypes.c_int(Ma), _ptr(Qaux),
                      ctypes.c_double(thr), _ptr(out))
    return out


def int3c2e(prims, types, norm, Q, aux_prims, aux_norm, Qaux, thr=1e-12):
    M, Ma = len(norm), len(aux_norm)
    out = np.zeros((Ma, M, M))
    lib().orc_int3c2e(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), _ptr(norm), ctypes.c_int(M), _ptr(Q),
                      _ptr(aux_prims), ctypes.c_int(len(aux_prims)), _ptr(aux_norm), ctypes.c_int(Ma), _ptr(Qaux),
                      ctypes.c_double(thr), _ptr(out))
    return out


def ri_B(L, T):
    T = np.ascontiguousarray(T, dtype=np.float64).copy()
    Ma = L.shape[0]
    ncols = T.size // Ma
    lib().orc_ri_B(ctypes.c_int(Ma), _ptr(np.ascontiguousarray(L)), ctypes.c_size_t(ncols), _ptr(T))
    return T


# ROHF coupling sets (Acc, Bcc, Aoo, Boo, Avv, Bvv) by rohf_parameter_name
# (rohf.cu:30-39).
ROHF_COUPLING = {
    "roothaan": (-0.5, 1.5, 0.5, 0.5, 1.5, -0.5),
    "mcweeny-diercksen": (1 / 3, 2 / 3, 1 / 3, 1 / 3, 2 / 3, 1 / 3),
    "davidson": (0.5, 1.0, 0.0, 1.0, 1.0, 0.0),
    "guest-saunders": (0.5, 0.5, 0.5, 0.5, 0.5, 0.5),
    "binkley-pople-dobosh": (0.5, 0.5, 1.0, 0.0, 0.0, 1.0),
    "faegri-manne": (0.5, 0.5, 1.0, 0.0, 0.5, 0.5),
    "goddard": (0.5, 0.5, 0.5, 0.0, 1.0, 0.0),
    "plakhutin-gorelik-breslavskaya": (0.0, 1.0, 1.0, 0.0, 1.0, 0.0),
}


def rohf_fock(H, JKc, JKo, C, S, nc, no, coupling="roothaan"):
    """ROHF Fock matrices from the J/K of the closed and open densities:
    F_c, F_o (gpu_kernels.cu:351-418), the unified MO-basis matrix built per
    orbital block from C^T F_c C and C^T F_o C (gpu_kernels.cu:421-497, the
    upper element serving both halves) and F = S C R C^T S
    (gpu_manager.cu:733-817).  C holds the MOs in its columns."""
    (Jc, Kc), (Jo, Ko) = JKc, JKo
    Fc = H + Jc + Jo - 0.5 * (Kc + Ko)
    Fo = 0.5 * (H + Jc + Jo - 0.5 * Kc - Ko)
    A, B = C.T @ Fc @ C, C.T @ Fo @ C
    Acc, Bcc, Aoo, Boo, Avv, Bvv = ROHF_COUPLING[coupling]
    M = len(H)
    kind = [0 if i < nc else (1 if i < nc + no else 2) for i in range(M)]
    R = np.empty((M, M))
    for i in range(M):
        for j in range(i, M):
            fc, fo = A[i, j], B[i, j]
            blk = (kind[i], kind[j])
            if blk == (0, 0):
                d = 2.0 * (Acc * fo + Bcc * (fc - fo))
            elif blk == (0, 1):
                d = 2.0 * (fc - fo)
            elif blk == (0, 2):
                d = fc
            elif blk == (1, 1):
                d = 2.0 * (Aoo * fo + Boo * (fc - fo))
            elif blk == (1, 2):
                d = 2.0 * fo
            else:
                d = 2.0 * (Avv * fo + Bvv * (fc - fo))
            R[i, j] = R[j, i] = d
    return Fc, Fo, S @ C @ R @ C.T @ S


def _damp(state, F, D, factor, optimal):
    """One damping step (rhf.hpp:219-243, gpu::damping gpu_manager.cu:989-1007):
    the first call stores (F, D); later calls return the damped F (and D) and
    store them, the factor optimal (orc_optimal_damping) or constant."""
    if not state:
        state.extend([F.copy(), D.copy()])
        return F, D
    Fp, Dp = state
    a = optimal_damping(F, Fp, D, Dp) if optimal else factor
    F = (1.0 - a) * Fp + a * F
    if optimal:
        D = (1.0 - a) * Dp + a * D
        state[1] = D.copy()
    state[0] = F.copy()
    return F, D


def _diis(hist, e, F, size):
    """DIIS extrapolation of F from the stored (error, F) pairs (rhf.hpp:298-327)."""
    hist.append((e, F.copy()))
    del hist[:-size]
    n = len(hist)
    if n > 1:
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.0
        for a in range(n):
            for b in range(n):
                B[a, b] = np.sum(hist[a][0] * hist[b][0])
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        try:  # a singular system keeps F as it is
            c = np.linalg.solve(B, rhs)
            F = sum(c[k] * hist[k][1] for k in range(n))
        except np.linalg.LinAlgError:
            pass
    return F


def rhf_damped(prims, types, norm, Z, xyz_bohr, nocc, optimal=False, factor=0.9, thr=1e-12, conv=1e-10,
               maxiter=300):
    """FP64 RHF on the host with Convergence_RHF_Damping (rhf.hpp:182-262):
    core guess, then per iteration C, D = 2 C_occ C_occ^T, F = H + J - K/2,
    E = 1/2 D.(H + F), stop on |dE| < conv, else damp F (and D when
    optimal).  Returns (E_elec, E_total, D, iterations)."""
    S, T, V = oneint(prims, norm, Z, xyz_bohr)
    H = T + V
    Q = schwarz(prims, types)
    w, U = np.linalg.eigh(S)
    X = U[:, w >= 1e-6] / np.sqrt(w[w >= 1e-6])

    def dens(F):
        C = X @ np.linalg.eigh(X.T @ F @ X)[1]
        return 2.0 * C[:, :nocc] @ C[:, :nocc].T

    def fock(D):
        J, K = jk(prims, types, norm, Q, D, thr)
        return H + J - 0.5 * K

    F = fock(dens(H))
    state, prev, E, it = [], None, 0.0, 0
    while it < maxiter:
        D = dens(F)
        F = fock(D)
        E = 0.5 * np.sum(D * (H + F))
        if prev is not None and abs(E - prev) < conv:
            break
        it += 1
        prev = E
        F, D = _damp(state, F, D, factor, optimal)
    return E, E + nuclear_repulsion(Z, xyz_bohr), D, it


def uhf(prims, types, norm, Z, xyz_bohr, na, nb, thr=1e-12, conv=1e-10, maxiter=200, diis_size=8, sad=None,
        damping=None):
    """FP64 UHF on the host from the oracle's integrals and J/K (uhf.cu:118-300):
    core guess (C_alpha from H, C_beta a copy of it), the symmetry break of
    every UHF guess (break_symmetry, uhf.hpp:429-436: the last row and last
    column of C_beta zeroed), X = U s^-1/2 (eigenvalues of S below 1e-6
    dropped), per-spin D_s = C_s,occ C_s,occ^T, F_s = H + J[Da + Db] - K[D_s],
    E = 1/2 sum_s D_s.(H + F_s), convergence on |dE|, DIIS per spin with the
    error F_s D_s S - S D_s F_s, or, with damping = factor, constant damping
    per spin (uhf.hpp:236-258).  Returns (E_elec, E_total, Da, Db, iterations).
    sad = (Da0, Db0) replaces the core guess by InitialGuess_UHF_SAD
    (uhf.hpp:668-705) or InitialGuess_UHF_Density (uhf.hpp:557-568):
    F(Da0, Db0), C per spin, the symmetry break, then D and F."""
    S, T, V = oneint(prims, norm, Z, xyz_bohr)
    H = T + V
    Q = schwarz(prims, types)
    w, U = np.linalg.eigh(S)
    keep = w >= 1e-6
    X = U[:, keep] / np.sqrt(w[keep])

    def coeff(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return X @ Cp

    def dens(C, n):
        return C[:, :n] @ C[:, :n].T

    def fock(Da, Db):
        Ja, Ka = jk(prims, types, norm, Q, Da, thr)
        Jb, Kb = jk(prims, types, norm, Q, Db, thr)
        return H + Ja + Jb - Ka, H + Ja + Jb - Kb

    if sad is None:
        Ca = coeff(H)
        Cb = Ca.copy()
    else:
        Fa, Fb = fock(*sad)
        Ca, Cb = coeff(Fa), coeff(Fb).copy()
    Cb[-1, :] = 0.0
    Cb[:, -1] = 0.0
    Da, Db = dens(Ca, na), dens(Cb, nb)
    Fa, Fb = fock(Da, Db)
    hist = ([], [])
    damp_state = ([], [])
    prev, E, it = None, 0.0, 0
    while it < maxiter:
        Da, Db = dens(coeff(Fa), na), dens(coeff(Fb), nb)
        Fa, Fb = fock(Da, Db)
        E = 0.5 * (np.sum(Da * (H + Fa)) + np.sum(Db * (H + Fb)))
        if prev is not None and abs(E - prev) < conv:
            break
        it += 1
        prev = E
        if damping is not None:
            Fa = _damp(damp_state[0], Fa, Da, damping, False)[0]
            Fb = _damp(damp_state[1], Fb, Db, damping, False)[0]
        else:
            Fa = _diis(hist[0], Fa @ Da @ S - S @ Da @ Fa, Fa, diis_size)
            Fb = _diis(hist[1], Fb @ Db @ S - S @ Db @ Fb, Fb, diis_size)
    return E, E + nuclear_repulsion(Z, xyz_bohr), Da, Db, it


def rohf(prims, types, norm, Z, xyz_bohr, nc, no, coupling="roothaan", thr=1e-12, conv=1e-10, maxiter=300,
         diis_size=8, damping=None, optimal=False):
    """FP64 ROHF on the host (rohf.cu:59-300): core guess, C from the unified
    Fock matrix, D_c = 2 C_c C_c^T, D_o = C_o C_o^T (gpu_kernels.cu:108-134),
    rohf_fock for F_c, F_o, F, E = 1/2 [D_c.(H + F_c) + D_o.H + 2 D_o.F_o]
    (gpu_manager.cu:871-882), DIIS on F with the total density, or damping
    of F (constant factor `damping`, or optimal, which damps the total
    density as well: Convergence_ROHF_Damping, rohf.hpp:216-306).
    Returns (E_elec, E_total, Dc, Do, iterations)."""
    S, T, V = oneint(prims, norm, Z, xyz_bohr)
    H = T + V
    Q = schwarz(prims, types)
    w, U = np.linalg.eigh(S)
    keep = w >= 1e-6
    X = U[:, keep] / np.sqrt(w[keep])

    def coeff(F):
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        return X @ Cp

    def step(F):
        C = coeff(F)
        Dc = 2.0 * C[:, :nc] @ C[:, :nc].T
        Do = C[:, nc:nc + no] @ C[:, nc:nc + no].T
        Fc, Fo, Fu = rohf_fock(H, jk(prims, types, norm, Q, Dc, thr), jk(prims, types, norm, Q, Do, thr), C, S, nc, no,
                               coupling)
        E = 0.5 * (np.sum(Dc * (H + Fc)) + np.sum(Do * H) + 2.0 * np.sum(Do * Fo))
        return Dc, Do, Fu, E

    Dc, Do, F, E = step(H)
    hist, damp_state = [], []
    prev, it = None, 0
    while it < maxiter:
        Dc, Do, F, E = step(F)
        if prev is not None and abs(E - prev) < conv:
            break
        it += 1
        prev = E
        D = Dc + Do
        if damping is not None or optimal:
            F = _damp(damp_state, F, D, damping, optimal)[0]
        else:
            F = _diis(hist, F @ D @ S - S @ D @ F, F, diis_size)
    return E, E + nuclear_repulsion(Z, xyz_bohr), Dc, Do, it


def ri_jk(B, D):
    Ma, M = B.shape[0], B.shape[1]
    J, K = np.zeros((M, M)), np.zeros((M, M))
    lib().orc_ri_jk(ctypes.c_int(Ma), ctypes.c_int(M), _ptr(np.ascontiguousarray(B)),
                    _ptr(np.ascontiguousarray(D, dtype=np.float64)), _ptr(J), _ptr(K))
    return J, K


def ri_tensor(prims, types, norm, aux_prims, aux_norm, thr=1e-12):
    """B (Maux x M x M) = L^-1 (P|mn) with (P|Q) = L L^T."""
    Q = schwarz(prims, types)
    Qa = schwarz_aux(aux_prims)
    V = int2c2e(aux_prims, aux_norm, Qa, thr)
    T = int3c2e(prims, types, norm, Q, aux_prims, aux_norm, Qa, thr)
    return ri_B(cholesky(V), T).reshape(len(aux_norm), len(norm), len(norm))


def grad_2e(prims, types, norm, Q, D, natoms, thr=1e-12):
    """sum Gamma(D) d(ab|cd)/dX, Gamma = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)."""
    M = len(norm)
    D = np.ascontiguousarray(D, dtype=np.float64)
    g = np.zeros(3 * natoms)
    lib().orc_grad_2e(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), _ptr(norm), ctypes.c_int(M), _ptr(Q),
                      ctypes.c_double(thr), _ptr(D), ctypes.c_int(natoms), _ptr(g))
    return g.reshape(natoms, 3)


def grad_1e(prims, norm, Z, xyz_bohr, D, W):
    """sum D d(T+V)/dX - sum W dS/dX (basis-centre and Hellmann-Feynman terms)."""
    M = len(norm)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
    D = np.ascontiguousarray(D, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    g = np.zeros(3 * len(Z))
    lib().orc_grad_1e(_ptr(prims), ctypes.c_int(len(prims)), _ptr(norm), ctypes.c_int(M), ctypes.c_int(len(Z)), _ptr(Z),
                      _ptr(xyz), _ptr(D), _ptr(W), _ptr(g))
    return g.reshape(len(Z), 3)


def grad_nuc(Z, xyz_bohr):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
    g = np.zeros(3 * len(Z))
    lib().orc_grad_nuc(ctypes.c_int(len(Z)), _ptr(Z), _ptr(xyz), _ptr(g))
    return g.reshape(len(Z), 3)


def rhf_gradient(gbs_path, Z, xyz_bohr, nocc, quirk=True, conv=1e-12, thr=1e-12):
    """FP64 host RHF energy and analytic nuclear gradient (W = D F D / 2)."""
    prims, types, norm = build_molecule(gbs_path, Z, xyz_bohr, quirk=quirk)
    _, E, D, F, _ = rhf(prims, types, norm, Z, xyz_bohr, nocc, thr=thr, conv=conv, maxiter=200)
    W = 0.5 * D @ F @ D
    Q = schwarz(prims, types)
    g = grad_1e(prims, norm, Z, xyz_bohr, D, W) + grad_2e(prims, types, norm, Q, D, len(Z), thr) + grad_nuc(Z, xyz_bohr)
    return E, g


def grad_3c(prims, types, norm, aux_prims, aux_norm, G, natoms, thr=1e-12):
    """sum_P sum_mn G[P][m][n] d(mn|P)/dX."""
    M, Ma = len(norm), len(aux_norm)
    g = np.zeros(3 * natoms)
    lib().orc_grad_3c(_ptr(prims), _ptr(types), ctypes.c_int(len(types)), _ptr(norm), ctypes.c_int(M),
                      _ptr(schwarz(prims, types)), _ptr(aux_prims), ctypes.c_int(len(aux_prims)), _ptr(aux_norm),
                      ctypes.c_int(Ma), _ptr(schwarz_aux(aux_prims)), ctypes.c_double(thr),
                      _ptr(np.ascontiguousarray(G, dtype=np.float64)), ctypes.c_int(natoms), _ptr(g))
    return g.reshape(natoms, 3)


def grad_2c(aux_prims, aux_norm, Gam, natoms, thr=1e-12):
    """sum_PQ Gam[P][Q] d(P|Q)/dX."""
    Ma = len(aux_norm)
    g = np.zeros(3 * natoms)
    lib().orc_grad_2c(_ptr(aux_prims), ctypes.c_int(len(aux_prims)), _ptr(aux_norm), ctypes.c_int(Ma),
                      _ptr(schwarz_aux(aux_prims)), ctypes.c_double(thr),
                      _ptr(np.ascontiguousarray(Gam, dtype=np.float64)), ctypes.c_int(natoms), _ptr(g))
    return g.reshape(natoms, 3)


def ri_grad_2e(prims, types, norm, aux_prims, aux_norm, D, natoms, thr=1e-12):
    """Gradient of the RI two-electron energy 1/2 D.J_RI - 1/4 D.K_RI for a
    fixed density: G^P = c_P D - 1/2 D At^P D into the 3-centre derivatives,
    Gam = -1/2 c c^T + 1/4 V^-1 M V^-1 into the 2-centre ones, with
    gamma = A.D, c = V^-1 gamma, At^P = sum_Q V^-1_PQ A^Q, M_PQ = tr(A^P D A^Q D)."""
    Q = schwarz(prims, types)
    Qa = schwarz_aux(aux_prims)
    V = int2c2e(aux_prims, aux_norm, Qa, thr)
    A3 = int3c2e(prims, types, norm, Q, aux_prims, aux_norm, Qa, thr)
    Ma, M = A3.shape[0], A3.shape[1]
    Vinv = np.linalg.inv(V)
    gamma = np.einsum("pmn,mn->p", A3, D)
    c = Vinv @ gamma
    At = (Vinv @ A3.reshape(Ma, -1)).reshape(Ma, M, M)
    DAD = np.einsum("mk,pkl,ln->pmn", D, At, D)
    G = c[:, None, None] * D[None] - 0.5 * DAD
    AD = np.einsum("pmk,kn->pmn", A3, D)
    Mpq = np.einsum("pmn,qnm->pq", AD, AD)
    Gam = -0.5 * np.outer(c, c) + 0.25 * Vinv @ Mpq @ Vinv
    return grad_3c(prims, types, norm, aux_prims, aux_norm, G, natoms, thr) + grad_2c(aux_prims, aux_norm, Gam, natoms,
                                                                                     thr)
