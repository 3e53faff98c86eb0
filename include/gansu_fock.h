/*
 * gansu_fock.h -- C ABI of the MI355X (gfx950) Fock-build engine.
 *
 * Drop-in boundary for the reference's ERI strategy interface
 * (zhaokaifan02/DFT_GANSU include/eri.hpp:36-63):
 *
 *   ERI::precomputation()        -> gansu_fock_precompute      (called at hf.cu:265
 *                                                               via rhf.cu:110-114)
 *   ERI::compute_fock_matrix()   -> gansu_fock_build_rhf       (hf.cu:284, rhf.hpp:394,
 *                                                               :422, :450, :578, :583)
 *   ERI::get_algorithm_name()    -> gansu_fock_algorithm_name  (hf.cu:424-429, :460-472)
 *
 * plus the input path it is fed from (basis_set.cpp:32-124, parseXYZ.cpp:44-115,
 * molecular.hpp:117-190), the parameter recipes (parameter_manager.cpp:42-225)
 * and the RHF SCF driver around it (hf.cu:171-314).
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero GANSU_E_* code on
 *    failure; gansu_last_error() then holds a message (thread-local).  No C++
 *    exception crosses this boundary (the reference throws gansu::Exception,
 *    utils.hpp THROW_EXCEPTION; each throw site becomes an error code here).
 *  - Matrix pointers named d_* are caller-owned device pointers to M x M
 *    row-major FP64 matrices (the reference's DeviceHostMatrix layout).  The
 *    engine never frees caller memory.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream); build
 *    calls are ordered on it and do not synchronise the host.
 *  - One engine per host thread.
 */
#ifndef GANSU_FOCK_H
#define GANSU_FOCK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GANSU_OK = 0,
  GANSU_E_ARG = 1,        /* invalid argument / size                            */
  GANSU_E_IO = 2,         /* file cannot be opened or parsed                    */
  GANSU_E_BASIS = 3,      /* element missing from basis / bad shell label       */
  GANSU_E_ELECTRONS = 4,  /* electron count invalid (molecular.hpp:100-109)     */
  GANSU_E_PARAM = 5,      /* unknown key or bad value                           */
  GANSU_E_HIP = 6,        /* HIP runtime / library failure                      */
  GANSU_E_STATE = 7,      /* call out of order (e.g. build before precompute)   */
  GANSU_E_NUMERIC = 8,    /* NaN in F (rhf.hpp:665-674), SPD failure, DIIS      */
  GANSU_E_UNSUPPORTED = 9 /* angular momentum beyond the compiled classes       */
};

/* == gansu::PrimitiveShell (types.hpp:116-125), 64 bytes. */
typedef struct gansu_primitive_shell {
  double exponent;
  double coefficient;
  double x, y, z;       /* bohr */
  int shell_type;       /* 0=s 1=p 2=d 3=f */
  size_t basis_index;   /* first basis function of the contracted shell */
  int atom_index;
} gansu_primitive_shell;

/* == gansu::ShellTypeInfo (types.hpp:150-154). */
typedef struct gansu_shell_type_info {
  int count;
  size_t start_index;
} gansu_shell_type_info;

const char* gansu_last_error(void);
const char* gansu_version(void);

/* ---------------------------------------------------------------- input --
 * Molecule = atoms + basis -> primitive list sorted by shell type, contracted
 * normalisation factors, shell type infos (molecular.hpp:117-190).
 * normalization: 0 = reference (keeps the integer-division quirk of
 * basis_set.hpp:123 for odd L), 1 = unit-normalised contracted functions.   */
typedef struct gansu_molecule gansu_molecule;

/* From an .xyz file (Angstrom, parseXYZ.cpp). */
int gansu_molecule_from_xyz(const char* xyz_path, const char* gbs_path, int charge, int beta_to_alpha,
                            int normalization, gansu_molecule** out);
/* From atomic numbers and coordinates in bohr. */
int gansu_molecule_from_atoms(int n_atoms, const int* Z, const double* xyz_bohr, const char* gbs_path, int charge,
                              int beta_to_alpha, int normalization, gansu_molecule** out);
/* sizes[7] = {n_prims, n_basis, n_types, n_atoms, n_electrons, n_alpha, n_beta} */
int gansu_molecule_sizes(const gansu_molecule* mol, int* sizes);
/* Any output pointer may be NULL.  norm has n_basis entries, xyz 3*n_atoms. */
int gansu_molecule_get(const gansu_molecule* mol, gansu_primitive_shell* prims, gansu_shell_type_info* types,
                       double* norm, int* Z, double* xyz_bohr);
double gansu_molecule_nuclear_repulsion(const gansu_molecule* mol);
void gansu_molecule_destroy(gansu_molecule* mol);
/* Superposition-of-atomic-densities guess densities (replaces
 * read_density_from_sad, utils.cpp:21-58, and the block placement of
 * InitialGuess_RHF_SAD::guess, rhf.hpp:551-584): the per-spin atomic
 * densities of the .sad file beside the molecule's basis file (same name,
 * extension .sad) on the diagonal atom blocks of Da and Db (host, M*M
 * row-major, zero elsewhere).  Host only.  GANSU_E_IO when the file is
 * missing, GANSU_E_BASIS when an element is absent from it or its block size
 * differs from the atom's basis.                                           */
int gansu_sad_density(const gansu_molecule* mol, double* Da, double* Db);


/* ----------------------------------------------------------- parameters --
 * Same keys, defaults and recipe syntax as ParameterManager
 * (parameter_manager.cpp:42-88, :164-225); unknown keys are rejected.       */
typedef struct gansu_params gansu_params;
int gansu_params_create(gansu_params** out);
int gansu_params_load_recipe(gansu_params* p, const char* path);
int gansu_params_set(gansu_params* p, const char* key, const char* value);
/* Copies the value (NUL-terminated) into buf. */
int gansu_params_get(const gansu_params* p, const char* key, char* buf, int buflen);
void gansu_params_destroy(gansu_params* p);

/* ---------------------------------------------------------- Fock engine --
 * `options` is a ";"-separated "key=value" list using parameter keys
 * (eri_method=stored|direct|ri, schwarz_screening_threshold=..,
 * precision=fp64|mixed, fp32_bound_threshold=.., ri_k=density|occupied) and
 * engine keys autotune=1|0 (time digestion tiles at precompute, stored path)
 * and graph=1|0 (four-centre builds replay a HIP graph captured once per set
 * of D/J/K/H/F pointers; the graph is re-captured when the pointers change),
 * nan_check=1|0 (eri_method=ri: an RHF build that forms F reads one flag
 * word back and fails with GANSU_E_NUMERIC when F holds a NaN, as
 * ERI_RI_RHF::compute_fock_matrix does, rhf.hpp:665-674; this is the one
 * host synchronisation inside a build; 0 skips it), and for
 * eri_method=ri precision=mixed ri_refresh=N (default 16) and ri_tol=t
 * (default 0.01): an RHF build given the occupied orbitals
 * (gansu_fock_build_rhf_occ) after a previous one with the same nocc
 * (<= 320) forms K and J as increments on the fp16 matrix cores from the
 * aligned orbital change E (K_n = K_prev + T_A T_E^T + T_E T_A^T,
 * T = B X); the build is the FP64 one instead after N chained increments,
 * when the alignment fails or |E| / |X| > 0.5, and, once the changes
 * summed since the last FP64 build exceed t, for the first build whose
 * change is below t / 4 (so that the error of early, large increments
 * leaves K before the SCF converges); the
 * incremental build reads three numbers back (alignment residual, |E|,
 * |X|).  State for the increments (about 8 M Maux nocc bytes) is
 * allocated by the first such build.  For eri_method=direct,
 * incremental_fock=N (default 0: off) makes RHF builds (gansu_fock_build and
 * the RHF drivers; UHF and ROHF builds stay full) digest the change of the
 * density since the previous build and add it to that build's J and K,
 * skipping quartets whose bound times the largest density-change element of
 * their six shell blocks is below incremental_threshold (default 1e-14); every
 * N-th build is a full one.  The state lives on the device, so a captured
 * build replays it; any other eri_method with incremental_fock > 0 is
 * GANSU_E_PARAM.
 * The engine runs on the device that was current at gansu_fock_create; each
 * call makes it current and restores the caller's device on return.
 * aux* may be NULL/0 unless eri_method=ri.                                  */
typedef struct gansu_fock_engine gansu_fock_engine;

int gansu_fock_create(const gansu_primitive_shell* prims, size_t n_prims, const gansu_shell_type_info* types,
                      int n_types, const double* cgto_norm, int num_basis, const gansu_primitive_shell* aux,
                      size_t n_aux, const gansu_shell_type_info* aux_types, int n_aux_types, const double* aux_norm,
                      int num_aux, const char* options, gansu_fock_engine** out);
/* == ERI::precomputation (eri.cu:60-116 stored, :134-236 RI). */
int gansu_fock_precompute(gansu_fock_engine* e);
/* == ERI::compute_fock_matrix for RHF: F = H + J - K/2 (gpu_kernels.cu:230-287). */
int gansu_fock_build_rhf(gansu_fock_engine* e, const double* d_D, const double* d_H, double* d_F, void* stream);
/* J and K separately: J_ij = sum_kl (ij|kl) D_kl, K_ij = sum_kl (ik|jl) D_kl. */
int gansu_fock_build_jk(gansu_fock_engine* e, const double* d_D, double* d_J, double* d_K, void* stream);
/* RHF build given the occupied orbitals as well: D = 2 C_occ C_occ^T with C
 * the M x M row-major coefficient matrix (first nocc columns occupied), as
 * the reference's RHF holds it (rhf.hpp coefficient_matrix).  The RI path
 * then forms K from X = sqrt(2) C_occ, K = sum_P (B^P X)(B^P X)^T, in
 * 4 Maux M^2 nocc flops instead of the density form's 4 Maux M^3; the
 * four-centre paths use D only.  No counterpart in the reference, whose RI
 * K is the density form (gpu_manager.cu:1589-1618).                          */
int gansu_fock_build_rhf_occ(gansu_fock_engine* e, const double* d_D, const double* d_C, int nocc, const double* d_H,
                             double* d_F, void* stream);
/* UHF build == ERI::compute_fock_matrix for UHF (gpu_kernels.cu:289-340,
 * computeFockMatrix_RI_UHF gpu_manager.cu:1634-1800):
 *   F_a = H + J[D_a + D_b] - K[D_a],   F_b = H + J[D_a + D_b] - K[D_b]
 * with D_a, D_b the alpha and beta densities (C_occ C_occ^T each).          */
int gansu_fock_build_uhf(gansu_fock_engine* e, const double* d_Da, const double* d_Db, const double* d_H, double* d_Fa,
                         double* d_Fb, void* stream);
/* ROHF build == ERI::compute_fock_matrix for ROHF (rohf.hpp:660-760:
 * computeFockMatrix_ROHF gpu_manager.cu:733-817, computeFockMatrix_RI_ROHF
 * :1800-1961).  d_Dc, d_Do: closed- and open-shell densities; d_C: the MO
 * coefficients (M x M row-major, MOs in columns: num_closed closed, then
 * num_open open); d_S: overlap.  Outputs
 *   F_c = H + J[Dc + Do] - (K[Dc] + K[Do])/2,
 *   F_o = (H + J[Dc + Do] - K[Dc]/2 - K[Do])/2,
 *   F   = S C R C^T S,
 * R the unified MO-basis Fock matrix built blockwise from C^T F_c C and
 * C^T F_o C with the coupling set `coupling` (rohf_parameter_name: roothaan
 * when NULL, mcweeny-diercksen, davidson, guest-saunders,
 * binkley-pople-dobosh, faegri-manne, goddard,
 * plakhutin-gorelik-breslavskaya; GANSU_E_PARAM for another name).        */
int gansu_fock_build_rohf(gansu_fock_engine* e, const double* d_Dc, const double* d_Do, const double* d_H,
                          const double* d_C, const double* d_S, const char* coupling, int num_closed, int num_open,
                          double* d_Fc, double* d_Fo, double* d_F, void* stream);
/* Moves every primitive to the position of its atom (xyz_bohr[3*atom_index])
 * and, if the engine was precomputed, precomputes again (the reference's
 * HF::update_geometry, hf.cu:460-472, rebuilds its ERI object).             */
int gansu_fock_update_geometry(gansu_fock_engine* e, const double* xyz_bohr, int n_atoms);
/* == ERI::get_algorithm_name ("Stored", "Direct", "RI"). */
const char* gansu_fock_algorithm_name(const gansu_fock_engine* e);
/* Primitive-pair Schwarz factors in the reference's pair order
 * (gpu_manager.cu:413-432, schwarz.txt): host copy of n entries.            */
int gansu_fock_schwarz(gansu_fock_engine* e, double* host_Q, size_t n);
size_t gansu_fock_num_primitive_pairs(const gansu_fock_engine* e);
/* stats[8] = {surviving contracted quartets, surviving primitive quartets,
 *             stored bytes, fp32 quartets, unique contracted quartets,
 *             ri B bytes, graphs captured, incremental RI builds}           */
int gansu_fock_stats(const gansu_fock_engine* e, double* stats);
/* Per quartet class of a stored or direct engine (after precompute): lab[4c
 * .. 4c+3] = LA LB LC LD, n_contracted[c] = the contracted quartets the
 * builds visit, n_primitive[c] = the primitive quartets among them that pass
 * the primitive screen Q_ab Q_cd >= schwarz_screening_threshold (the unit of
 * work of the MD integrals, eri.cu:160-196 / int2e.cu:1646-1947).  At most
 * max_classes entries are written; *n_classes = the number of classes.  The
 * first call counts on the host (and fills stats[1]); RI engines:
 * GANSU_E_UNSUPPORTED.  No counterpart in the reference (measurement).      */
int gansu_fock_class_stats(gansu_fock_engine* e, int max_classes, int* n_classes, int* lab, double* n_contracted,
                           double* n_primitive);
/* Per-launch timing of the builds that follow, with HIP events recorded on the
 * build stream (no counterpart in the reference, whose profiler.hpp times host
 * wall clock only).  After a build, rows[5*i + k] for launch i hold
 * {kernel id, milliseconds, algorithmic bytes, algorithmic flops, work units}
 * where kernel id = 1000*LA + 100*LB + 10*LC + LD + 10000*mode for ERI class
 * kernels (mode 0 direct, 2 stored digestion) and 90000+ for helper kernels. */
int gansu_fock_set_profiling(gansu_fock_engine* e, int enable);
int gansu_fock_last_profile(gansu_fock_engine* e, double* rows, int max_rows, int* n_rows);
void gansu_fock_destroy(gansu_fock_engine* e);

/* ------------------------------------------------ one-electron / SCF -----
 * S, T, V on the device (M x M row-major, caller-owned device pointers).    */
int gansu_oneint(const gansu_molecule* mol, double* d_S, double* d_T, double* d_V, void* stream);
/* gansu_oneint without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The molecule's shell
 * tables and the Boys grid are copied into a 4 MiB device array of the
 * library (one per GPU, held until the process ends) and the same kernel
 * as gansu_oneint runs on them; the call synchronises `stream` before it
 * returns.  Molecules whose tables exceed the array (about 50,000
 * primitives) are GANSU_E_ARG.  Calls that overlap on one GPU share the
 * array and race.                                                          */
int gansu_oneint_noalloc(const gansu_molecule* mol, double* d_S, double* d_T, double* d_V, void* stream);

/* Solves the molecule with the params' method (rhf, uhf or rohf; for a SAD
 * cache, a single atom) and writes its per-spin densities to path in the
 * layout of HF::generate_sad_cache (hf.hpp:372-402, the --export_sad_cache
 * option of HF_main.cu:55-59): a line "M", then M rows of alpha and M rows of
 * beta values "%.10e ".  The spin densities are those of export_density_matrix:
 * RHF D/2 and D/2; UHF Da, Db; ROHF Dc/2 + Do and Dc/2.                     */
int gansu_export_sad_cache(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                           const char* path);
/* RHF SCF loop of hf.cu:245-314 (core, GWH or SAD guess, DIIS/damping as the recipe
 * says) with this engine as the ERI strategy.  energies[2] = {E_electronic, E_total};
 * D_out (host, M*M) may be NULL.  info[4] = {iterations, fock builds,
 * converged, graph launches}.
 * Engine key scf_solver (no reference counterpart): "diagonalize" (default,
 * the reference's eigensolver step) or "purify" (DIIS only).  With
 * eri_method stored or direct each iteration after the guess is one captured
 * HIP graph -- X^T F X on the FP64 matrix cores, SP2 density purification in
 * place of the eigensolver, D, the Fock build, the energy and convergence
 * test and the DIIS step as device kernels -- launched once per iteration
 * with one 56-byte read-back; info[3] counts those launches (0 otherwise).
 * With eri_method=ri the loop stays eager (the RI builds take an occupied
 * factor): each iteration after the first purifies X^T F X by SP2 and takes
 * the occupied orbitals as a Cholesky QR of the previous ones projected by
 * the result, falling back to the eigensolver for an iteration whose
 * purification or projection fails.  Either way the final orbitals come
 * from one eigensolve of the last iteration's X^T F X.                      */
int gansu_rhf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                  double* D_out, int* info);
/* HF::solve(Da0, Db0, force_density) for RHF (rhf.cu:122-134,
 * InitialGuess_RHF_Density, rhf.hpp:441-454): when force_density is
 * non-zero or initial_guess=density, the loop starts from F(Da0 + Db0)
 * (Da0, Db0 host M*M, for RHF each D/2); otherwise initial_guess (core, gwh,
 * sad) runs and the densities are not used.  With both NULL it is
 * gansu_rhf_run (initial_guess=density or force_density then fall back to
 * the core guess with a message, as the reference does).  One of them NULL
 * is GANSU_E_ARG.                                                           */
int gansu_rhf_run_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                        const double* Da0, const double* Db0, int force_density, double* energies, double* D_out,
                        int* info);
/* gansu_rhf_gradient started from caller densities as gansu_rhf_run_guess
 * (same force_density rule), also returning the converged density (D_out,
 * host M*M, may be NULL): a geometry optimiser passes each step's D/2, D/2
 * with force_density = 1 to the next step, as the reference's optimisers
 * pass densities to HF::solve (rhf.cu:122-134).                             */
int gansu_rhf_gradient_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                             const double* Da0, const double* Db0, int force_density, double* energies, double* grad,
                             double* D_out, int* info);
/* gansu_rhf_run returning the final orbitals: C_out (host, M*M row-major,
 * MOs in columns, those of the converged density) and eps_out (host, M,
 * ascending orbital energies); either may be NULL.                          */
int gansu_rhf_run_mo(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                     double* energies, double* C_out, double* eps_out, int* info);
/* Molden file of orbitals C (host, M x n_mo row-major, MOs in columns) with
 * energies eps[n_mo] and occupations occ[n_mo], in the layout of
 * RHF::export_molden_file (rhf.cu:363-425): [Atoms] (AU), [GTO] per atom and
 * contracted shell (raw contraction coefficients), [MO] blocks.             */
int gansu_write_molden(const gansu_molecule* mol, const double* C, const double* eps, const double* occ, int n_mo,
                       const char* path);
/* UHF SCF loop (uhf.cu:118-300 via hf.cu:245-314): core, GWH or SAD guess,
 * each followed by the symmetry break of uhf.hpp:429-436 (the last row and
 * column of C_beta zeroed), per-spin densities D_s = C_s,occ C_s,occ^T over
 * the molecule's n_alpha / n_beta electrons, E = 1/2 sum_s D_s.(H + F_s),
 * DIIS per spin (uhf.hpp:291-400) or damping per spin (uhf.hpp:196-258;
 * optimaldamping is GANSU_E_PARAM, as uhf.cu:57-69 rejects it).  Da_out,
 * Db_out (host, M*M) may be NULL; energies and info as gansu_rhf_run.      */
int gansu_uhf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                  double* Da_out, double* Db_out, int* info);
/* HF::solve(Da0, Db0, force_density) for UHF (uhf.cu:118-130): when
 * force_density is non-zero or initial_guess=density, F(Da0, Db0) (host,
 * M*M), then C per spin, the symmetry break, D and F
 * (InitialGuess_UHF_Density, uhf.hpp:544-568); otherwise initial_guess runs.
 * Both NULL is gansu_uhf_run, one NULL GANSU_E_ARG.                        */
int gansu_uhf_run_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                        const double* Da0, const double* Db0, int force_density, double* energies, double* Da,
                        double* Db, int* info);
/* ROHF SCF loop (rohf.cu:59-300 via hf.cu:245-314): n_beta closed and
 * n_alpha - n_beta open orbitals, core or GWH guess, D_c = 2 C_c C_c^T,
 * D_o = C_o C_o^T, the ROHF build with rohf_parameter_name,
 * E = 1/2 [D_c.(H + F_c) + D_o.H + 2 D_o.F_o], DIIS, damping or optimal
 * damping (rohf.hpp:216-306) on the unified Fock matrix and the total
 * density.  Dc_out, Do_out (host, M*M) may be NULL; energies and info as
 * gansu_rhf_run.                                                            */
int gansu_rohf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                   double* Dc_out, double* Do_out, int* info);

/* ------------------------------------------------------ nuclear gradient --
 * Analytic RHF gradient dE/dX (hartree/bohr, atom-major x,y,z).  No
 * counterpart in the reference, which has no gradient; the terms follow the
 * RHF energy of hf.cu:245-314:
 *   dE/dX = sum D d(T+V)/dX - sum W dS/dX
 *         + sum Gamma_abcd d(ab|cd)/dX + dV_nn/dX,
 *   Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc),
 *   W = 2 C_occ eps_occ C_occ^T (= D F D / 2 at convergence).
 * Device outputs are overwritten; calls are ordered on `stream`.            */
/* Two-electron term for an engine with eri_method=stored|direct, after
 * precompute; d_grad has 3*n_atoms entries.                                 */
int gansu_fock_gradient_rhf(gansu_fock_engine* e, const double* d_D, int n_atoms, double* d_grad, void* stream);
/* The same for any eri_method, given the occupied orbitals as in
 * gansu_fock_build_rhf_occ (d_C row-major M x M, the first nocc columns
 * occupied, d_D = 2 C_occ C_occ^T).  eri_method=ri differentiates the RI
 * energy 1/2 D.J_RI - 1/4 D.K_RI of the fitted integrals B = L^-1 (Q|mn)
 * (rhf.cu / gpu_manager.cu:1311-1349 build them): the 3-centre and 2-centre
 * derivative integrals contracted with
 *   G^P = c_P D - 1/2 D (V^-1 A)^P D,   Gam = -1/2 c c^T + 1/4 V^-1 M V^-1,
 *   c = V^-1 (A.D),  M_PQ = tr(A^P D A^Q D),  A^P_mn = (mn|P),  V = (P|Q);
 * stored|direct ignore d_C and nocc.                                         */
int gansu_fock_gradient_rhf_occ(gansu_fock_engine* e, const double* d_D, const double* d_C, int nocc, int n_atoms,
                                double* d_grad, void* stream);
/* One-electron and overlap terms (kinetic, nuclear attraction including the
 * Hellmann-Feynman term of every nucleus, -W dS).                          */
int gansu_oneint_gradient(const gansu_molecule* mol, const double* d_D, const double* d_W, double* d_grad,
                          void* stream);
/* dV_nn/dX into host grad[3*n_atoms]. */
int gansu_molecule_nuclear_repulsion_gradient(const gansu_molecule* mol, double* grad);
/* SCF (as gansu_rhf_run) then the full gradient into host grad[3*n_atoms];
 * energies[2] and info[4] as gansu_rhf_run.  Converge tightly
 * (convergence_energy_threshold <= 1e-10): the gradient is exact only at the
 * SCF solution.                                                             */
int gansu_rhf_gradient(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                       double* energies, double* grad, int* info);

#ifdef __cplusplus
}
#endif
#endif /* GANSU_FOCK_H */
