"""Host-side mirror of the reference's ERI strategy interface and HF driver.

  FockEngine.precomputation()        == ERI::precomputation()      (eri.hpp:36-63)
  FockEngine.compute_fock_matrix()   == ERI::compute_fock_matrix() (RHF: F = H + J - K/2)
  FockEngine.get_algorithm_name()    == ERI::get_algorithm_name()
  Parameters                         == ParameterManager (parameter_manager.cpp)
  rhf(...)                           == HF::single_point_energy (hf.cu:245-314)

Matrices are torch tensors on the GPU (M x M, row-major FP64): torch supplies
device memory and streams; every computation runs in libgansu_fock.so.
"""
import ctypes

import numpy as np

from ._lib import PRIM_DTYPE, check, lib, ptr


class Parameters:
    """Same keys, defaults and recipe syntax as the reference's ParameterManager."""

    def __init__(self, recipe=None, **kv):
        self._h = ctypes.c_void_p()
        check(lib().gansu_params_create(ctypes.byref(self._h)))
        if recipe:
            check(lib().gansu_params_load_recipe(self._h, recipe.encode()))
        for k, v in kv.items():
            self[k] = v

    def __setitem__(self, key, value):
        if isinstance(value, bool):
            value = "1" if value else "0"
        check(lib().gansu_params_set(self._h, str(key).encode(), str(value).encode()))

    def __getitem__(self, key):
        buf = ctypes.create_string_buffer(4096)
        check(lib().gansu_params_get(self._h, str(key).encode(), buf, len(buf)))
        return buf.value.decode()

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().gansu_params_destroy(h)
            except Exception:
                pass
            self._h = None


class FockEngine:
    """Fock-build engine over a molecule (and an auxiliary basis for RI)."""

    def __init__(self, mol, aux=None, eri_method="stored", schwarz_screening_threshold=1e-12, precision="fp64",
                 fp32_bound_threshold=1e-6, ri_k="density", autotune=True, graph=True, nan_check=True,
                 ri_refresh=None, ri_tol=None, incremental_fock=None):
        self.mol = mol
        self.aux = aux
        self.M = mol.n_basis
        opts = (f"eri_method={eri_method};schwarz_screening_threshold={schwarz_screening_threshold!r};"
                f"precision={precision};fp32_bound_threshold={fp32_bound_threshold!r};ri_k={ri_k};"
                f"autotune={int(bool(autotune))};graph={int(bool(graph))};nan_check={int(bool(nan_check))}")
        if ri_refresh is not None:
            opts += f";ri_refresh={int(ri_refresh)}"
        if ri_tol is not None:
            opts += f";ri_tol={float(ri_tol)!r}"
        if incremental_fock is not None:
            opts += f";incremental_fock={int(incremental_fock)}"
        self._h = ctypes.c_void_p()
        a = aux
        check(lib().gansu_fock_create(
            ptr(mol.prims), len(mol.prims), ptr(mol.types), len(mol.types), ptr(mol.norm), mol.n_basis,
            ptr(a.prims) if a else None, len(a.prims) if a else 0, ptr(a.types) if a else None,
            len(a.types) if a else 0, ptr(a.norm) if a else None, a.n_basis if a else 0, opts.encode(),
            ctypes.byref(self._h)))
        self._precomputed = False

    def precomputation(self):
        check(lib().gansu_fock_precompute(self._h))
        self._precomputed = True

    @staticmethod
    def _stream(stream):
        if stream is None:
            return None
        return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))

    def _check_mat(self, t, name):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and tuple(t.shape) == (self.M, self.M)):
            raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of shape ({self.M}, {self.M})")

    def compute_fock_matrix(self, D, H, F=None, stream=None):
        """F = H + J[D] - K[D]/2 (RHF); returns F."""
        import torch
        self._check_mat(D, "D")
        self._check_mat(H, "H")
        if F is None:
            F = torch.empty_like(D)
        self._check_mat(F, "F")
        check(lib().gansu_fock_build_rhf(self._h, ptr(D), ptr(H), ptr(F), self._stream(stream)))
        return F

    def compute_fock_matrix_occ(self, D, C, nocc, H, F=None, stream=None):
        """F = H + J - K/2 with D = 2 C_occ C_occ^T given as well (C: M x M, first nocc columns)."""
        import torch
        for t, n in ((D, "D"), (C, "C"), (H, "H")):
            self._check_mat(t, n)
        if F is None:
            F = torch.empty_like(D)
        check(lib().gansu_fock_build_rhf_occ(self._h, ptr(D), ptr(C), int(nocc), ptr(H), ptr(F), self._stream(stream)))
        return F

    def compute_fock_matrix_uhf(self, Da, Db, H, Fa=None, Fb=None, stream=None):
        """UHF: F_a = H + J[Da + Db] - K[Da], F_b = H + J[Da + Db] - K[Db]; returns (Fa, Fb)."""
        import torch
        for t, n in ((Da, "Da"), (Db, "Db"), (H, "H")):
            self._check_mat(t, n)
        Fa = torch.empty_like(Da) if Fa is None else Fa
        Fb = torch.empty_like(Da) if Fb is None else Fb
        check(lib().gansu_fock_build_uhf(self._h, ptr(Da), ptr(Db), ptr(H), ptr(Fa), ptr(Fb), self._stream(stream)))
        return Fa, Fb

    def compute_fock_matrix_rohf(self, Dc, Do, H, C, S, num_closed, num_open, coupling="roothaan", stream=None):
        """ROHF: returns (Fc, Fo, F) with F_c = H + J[Dc + Do] - (K[Dc] + K[Do])/2,
        F_o = (H + J[Dc + Do] - K[Dc]/2 - K[Do])/2 and the unified AO Fock matrix
        F = S C R C^T S (C: MOs in columns, num_closed closed then num_open open)."""
        import torch
        for t, n in ((Dc, "Dc"), (Do, "Do"), (H, "H"), (C, "C"), (S, "S")):
            self._check_mat(t, n)
        Fc, Fo, F = (torch.empty_like(Dc) for _ in range(3))
        check(lib().gansu_fock_build_rohf(self._h, ptr(Dc), ptr(Do), ptr(H), ptr(C), ptr(S),
                                          coupling.encode() if coupling else None, int(num_closed), int(num_open),
                                          ptr(Fc), ptr(Fo), ptr(F), self._stream(stream)))
        return Fc, Fo, F

    def update_geometry(self, xyz_bohr):
        """Move the atoms (n_atoms x 3, bohr); a precomputed engine precomputes again."""
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        check(lib().gansu_fock_update_geometry(self._h, ptr(xyz), len(xyz) // 3))

    def compute_jk(self, D, stream=None):
        import torch
        self._check_mat(D, "D")
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        check(lib().gansu_fock_build_jk(self._h, ptr(D), ptr(J), ptr(K), self._stream(stream)))
        return J, K

    def get_algorithm_name(self):
        return lib().gansu_fock_algorithm_name(self._h).decode()

    def schwarz_factors(self):
        """Primitive-pair Schwarz factors in the reference's pair order."""
        n = lib().gansu_fock_num_primitive_pairs(self._h)
        q = np.zeros(n)
        check(lib().gansu_fock_schwarz(self._h, ptr(q), n))
        return q

    def stats(self):
        s = np.zeros(8)
        check(lib().gansu_fock_stats(self._h, ptr(s)))
        return dict(contracted_quartets=s[0], primitive_quartets=s[1], stored_bytes=s[2], fp32_quartets=s[3],
                    unique_quartets=s[4], ri_bytes=s[5], graph_captures=s[6], ri_incremental_builds=int(s[7]))

    def class_stats(self):
        """Per quartet class: dict(L=(LA, LB, LC, LD), contracted=..., primitive=...): the
        contracted quartets the builds visit and the primitive quartets among them that pass
        the primitive Schwarz screen (stored and direct engines, after precomputation)."""
        n = ctypes.c_int(0)
        check(lib().gansu_fock_class_stats(self._h, 0, ctypes.byref(n), None, None, None))
        lab = np.zeros(4 * n.value, dtype=np.int32)
        nc, npq = np.zeros(n.value), np.zeros(n.value)
        check(lib().gansu_fock_class_stats(self._h, n.value, ctypes.byref(n), ptr(lab), ptr(nc), ptr(npq)))
        return [dict(L=tuple(int(x) for x in lab[4 * c:4 * c + 4]), contracted=float(nc[c]), primitive=float(npq[c]))
                for c in range(n.value)]

    def set_profiling(self, enable=True):
        check(lib().gansu_fock_set_profiling(self._h, int(enable)))

    def last_profile(self):
        """Per-launch rows of the last build: (kernel id, ms, bytes, flops, units)."""
        n = ctypes.c_int()
        check(lib().gansu_fock_last_profile(self._h, None, 0, ctypes.byref(n)))
        rows = np.zeros((max(n.value, 1), 5))
        check(lib().gansu_fock_last_profile(self._h, ptr(rows), n.value, ctypes.byref(n)))
        return rows[:n.value]

    def gradient_2e(self, D, n_atoms=None, stream=None, C=None, nocc=None):
        """Two-electron term of the RHF nuclear gradient, sum Gamma(D) d(ab|cd)/dX,
        as an (n_atoms, 3) float64 CUDA tensor.  eri_method="ri" needs the
        orbitals C (M x M, first nocc columns occupied, D = 2 C_occ C_occ^T) and
        differentiates the RI energy 1/2 D.J_RI - 1/4 D.K_RI."""
        import torch
        self._check_mat(D, "D")
        n = int(n_atoms if n_atoms is not None else self.mol.n_atoms)
        g = torch.empty((n, 3), dtype=torch.float64, device=D.device)
        if C is None:
            check(lib().gansu_fock_gradient_rhf(self._h, ptr(D), n, ptr(g), self._stream(stream)))
        else:
            self._check_mat(C, "C")
            k = int(nocc if nocc is not None else self.mol.n_electrons // 2)
            check(lib().gansu_fock_gradient_rhf_occ(self._h, ptr(D), ptr(C), k, n, ptr(g), self._stream(stream)))
        return g

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().gansu_fock_destroy(h)
            except Exception:
                pass
            self._h = None


def oneint(mol):
    """S, T, V on the GPU (torch float64 tensors)."""
    import torch
    M = mol.n_basis
    S, T, V = (torch.empty((M, M), dtype=torch.float64, device="cuda") for _ in range(3))
    check(lib().gansu_oneint(mol.handle, ptr(S), ptr(T), ptr(V), None))
    torch.cuda.synchronize()
    return S, T, V


def oneint_gradient(mol, D, W, stream=None):
    """sum D d(T+V)/dX - sum W dS/dX as an (n_atoms, 3) float64 CUDA tensor."""
    import torch
    g = torch.empty((mol.n_atoms, 3), dtype=torch.float64, device=D.device)
    for t in (D, W):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and tuple(t.shape) == (mol.n_basis,) * 2):
            raise ValueError("D and W must be contiguous float64 CUDA tensors of shape (M, M)")
    check(lib().gansu_oneint_gradient(mol.handle, ptr(D), ptr(W), ptr(g), FockEngine._stream(stream)))
    return g


def nuclear_repulsion_gradient(mol):
    g = np.zeros((mol.n_atoms, 3))
    check(lib().gansu_molecule_nuclear_repulsion_gradient(mol.handle, ptr(g)))
    return g


def rhf_gradient(mol, params, aux=None, density=None, force_density=False):
    """SCF then the analytic nuclear gradient; returns dict(e_elec, e_total,
    iterations, fock_builds, converged, grad (n_atoms x 3, hartree/bohr), D).
    density = (Da, Db) with force_density=True (or initial_guess=density)
    starts the SCF from F(Da + Db) (a previous geometry's D/2, D/2, say)."""
    en = np.zeros(2)
    g = np.zeros((mol.n_atoms, 3))
    D = np.zeros((mol.n_basis, mol.n_basis))
    info = np.zeros(4, dtype=np.int32)
    Da = Db = None
    if density is not None:
        Da, Db = (np.ascontiguousarray(x, dtype=np.float64) for x in density)
        if Da.shape != D.shape or Db.shape != D.shape:
            raise ValueError(f"densities must be {D.shape}")
    check(lib().gansu_rhf_gradient_guess(mol.handle, aux.handle if aux else None, params.handle, ptr(Da), ptr(Db),
                                         int(bool(force_density)), ptr(en), ptr(g), ptr(D), ptr(info)))
    return dict(e_elec=float(en[0]), e_total=float(en[1]), iterations=int(info[0]), fock_builds=int(info[1]),
                converged=bool(info[2]), grad=g, D=D)


def uhf(mol, params, aux=None, density=None, force_density=False):
    """Run the UHF SCF; returns dict(e_elec, e_total, iterations, fock_builds, converged, Da, Db).
    density = (Da, Db) with force_density=True (or initial_guess=density)
    starts from F(Da, Db) (HF::solve(Da, Db, force_density), uhf.cu:118-130)."""
    en = np.zeros(2)
    Da = np.zeros((mol.n_basis, mol.n_basis))
    Db = np.zeros((mol.n_basis, mol.n_basis))
    info = np.zeros(4, dtype=np.int32)
    if density is not None:
        a0, b0 = (np.ascontiguousarray(x, dtype=np.float64) for x in density)
        if a0.shape != Da.shape or b0.shape != Da.shape:
            raise ValueError(f"densities must be {Da.shape}")
        check(lib().gansu_uhf_run_guess(mol.handle, aux.handle if aux else None, params.handle, ptr(a0), ptr(b0),
                                        int(bool(force_density)), ptr(en), ptr(Da), ptr(Db), ptr(info)))
    else:
        check(lib().gansu_uhf_run(mol.handle, aux.handle if aux else None, params.handle, ptr(en), ptr(Da),
                                  ptr(Db), ptr(info)))
    return dict(e_elec=float(en[0]), e_total=float(en[1]), iterations=int(info[0]), fock_builds=int(info[1]),
                converged=bool(info[2]), Da=Da, Db=Db)


def rohf(mol, params, aux=None):
    """Run the ROHF SCF; returns dict(e_elec, e_total, iterations, fock_builds, converged, Dc, Do)."""
    en = np.zeros(2)
    Dc = np.zeros((mol.n_basis, mol.n_basis))
    Do = np.zeros((mol.n_basis, mol.n_basis))
    info = np.zeros(4, dtype=np.int32)
    check(lib().gansu_rohf_run(mol.handle, aux.handle if aux else None, params.handle, ptr(en), ptr(Dc), ptr(Do),
                               ptr(info)))
    return dict(e_elec=float(en[0]), e_total=float(en[1]), iterations=int(info[0]), fock_builds=int(info[1]),
                converged=bool(info[2]), Dc=Dc, Do=Do)


def rhf_orbitals(mol, params, aux=None):
    """RHF SCF returning dict(e_elec, e_total, iterations, fock_builds, converged, C, eps): the
    final MO coefficients (M x M, MOs in columns) and orbital energies."""
    en = np.zeros(2)
    C = np.zeros((mol.n_basis, mol.n_basis))
    eps = np.zeros(mol.n_basis)
    info = np.zeros(4, dtype=np.int32)
    check(lib().gansu_rhf_run_mo(mol.handle, aux.handle if aux else None, params.handle, ptr(en), ptr(C), ptr(eps),
                                 ptr(info)))
    return dict(e_elec=float(en[0]), e_total=float(en[1]), iterations=int(info[0]), fock_builds=int(info[1]),
                converged=bool(info[2]), C=C, eps=eps)


def write_molden(mol, C, eps, occ, path):
    """Molden file (rhf.cu:363-425 layout) of orbitals C (M x n_mo, MOs in columns)."""
    C = np.ascontiguousarray(C, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    occ = np.ascontiguousarray(occ, dtype=np.float64)
    check(lib().gansu_write_molden(mol.handle, ptr(C), ptr(eps), ptr(occ), C.shape[1], str(path).encode()))


def rhf(mol, params, aux=None, density=None, force_density=False):
    """Run the RHF SCF; returns dict(e_elec, e_total, iterations, fock_builds, converged, graph_launches, D)
    (graph_launches: captured iterations launched with scf_solver=purify, else 0).
    density = (Da, Db) (M x M each, D/2 each for a closed shell) starts the
    loop from F(Da + Db) when force_density=True or initial_guess=density
    (HF::solve(Da, Db, force_density), rhf.cu:122-134); otherwise the
    parameters' guess runs, as in the reference."""
    en = np.zeros(2)
    D = np.zeros((mol.n_basis, mol.n_basis))
    info = np.zeros(4, dtype=np.int32)
    if density is not None:
        Da, Db = (np.ascontiguousarray(x, dtype=np.float64) for x in density)
        if Da.shape != D.shape or Db.shape != D.shape:
            raise ValueError(f"densities must be {D.shape}")
        check(lib().gansu_rhf_run_guess(mol.handle, aux.handle if aux else None, params.handle, ptr(Da), ptr(Db),
                                        int(bool(force_density)), ptr(en), ptr(D), ptr(info)))
    else:
        check(lib().gansu_rhf_run(mol.handle, aux.handle if aux else None, params.handle, ptr(en), ptr(D),
                                  ptr(info)))
    return dict(e_elec=float(en[0]), e_total=float(en[1]), iterations=int(info[0]), fock_builds=int(info[1]),
                converged=bool(info[2]), graph_launches=int(info[3]), D=D)


__all__ = ["FockEngine", "Parameters", "oneint", "rhf", "uhf", "rohf", "rhf_orbitals", "write_molden", "oneint_gradient", "nuclear_repulsion_gradient",
           "rhf_gradient", "PRIM_DTYPE"]
