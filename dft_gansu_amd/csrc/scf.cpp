// RHF SCF driver around the Fock engine, on the device in FP64.
//
// Semantics of HF::single_point_energy (hf.cu:245-314):
//   core Hamiltonian, ERI precompute, X = U s^-1/2 (eigenvalues < 1e-6 dropped,
//   hf.cu:171-242), initial guess (core: rhf.hpp:384-396), then per iteration
//   C <- F (X^T F X, eigen, X C'), D = 2 C_occ C_occ^T, F <- D through the
//   engine, E = 1/2 sum D (H + F); stop when |E - E_prev| < threshold, else
//   update F by DIIS (rhf.hpp:298-327, gpu_manager.cu:1020-1194) or damping
//   (rhf.hpp:182-262, gpu_manager.cu:898-1007).
// Dense linear algebra: rocBLAS dgemm, rocSOLVER dsyevd; small DIIS systems
// on the host.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <limits>

#include "gradient_kernels.hpp"
#include "scf.hpp"
#include "scf_graph.hpp"

namespace gansu {

namespace {

void blas_check(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) throw Error(GANSU_E_HIP, std::string(what) + ": rocblas status " + std::to_string(st));
}

// The drivers' rocBLAS/rocSOLVER handle: one per host thread, created on
// first use and kept for the life of the process.  A fresh handle costs its
// first rocSOLVER call tens of milliseconds (workspace set-up), which a
// handle per SCF run paid on every run.
struct Blas {
  rocblas_handle h = nullptr;
  Blas() {
    thread_local rocblas_handle cached = nullptr;
    if (!cached) blas_check(rocblas_create_handle(&cached), "rocblas_create_handle");
    h = cached;
  }
  // row-major C = op(A) op(B), all n x n
  void gemm(bool ta, bool tb, int n, const double* A, const double* B, double* C, double alpha = 1.0,
            double beta = 0.0) {
    blas_check(rocblas_dgemm(h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                             ta ? rocblas_operation_transpose : rocblas_operation_none, n, n, n, &alpha, B, n, A, n,
                             &beta, C, n),
               "rocblas_dgemm");
  }
};

// Device workspace of the driver.
struct Work {
  int n;
  DevBuf<double> S, T, V, H, X, F, D, C, tmp, tmp2, evec, eval, part;
  DevBuf<int> info;
  explicit Work(int n_) : n(n_) {
    const size_t nn = (size_t)n * n;
    for (auto* b : {&S, &T, &V, &H, &X, &F, &D, &C, &tmp, &tmp2, &evec}) b->alloc(nn);
    eval.alloc(n);
    part.alloc(256);
    info.alloc(1);
  }
};

double dot(const double* x, const double* y, size_t n, DevBuf<double>& part) {
  GANSU_HIP(launch_dot_partial(x, y, n, part.ptr, 256, 0));
  double h[256];
  GANSU_HIP(hipMemcpy(h, part.ptr, sizeof h, hipMemcpyDeviceToHost));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

// symmetric eigendecomposition: evec (column-major eigenvectors), ascending eval
void syevd(rocblas_handle h, int n, const double* A, Work& w) {
  GANSU_HIP(hipMemcpy(w.evec.ptr, A, sizeof(double) * n * n, hipMemcpyDeviceToDevice));
  blas_check(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, w.evec.ptr, n, w.eval.ptr,
                              w.tmp2.ptr, w.info.ptr),
             "rocsolver_dsyevd");
  int info = 0;
  GANSU_HIP(hipMemcpy(&info, w.info.ptr, sizeof(int), hipMemcpyDeviceToHost));
  if (info != 0) throw Error(GANSU_E_NUMERIC, "dsyevd failed, info=" + std::to_string(info));
}

// Gaussian elimination with partial pivoting (the getrf/getrs of the reference's DIIS).
// Returns an empty vector when the system is singular, which happens when the
// stored error vectors vanish exactly (a converged atom in a tiny basis).
std::vector<double> solve_small(int n, std::vector<double> A, std::vector<double> b) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) return {};  // singular: the caller skips the extrapolation
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r * n + k] * b[k];
    b[r] = s / A[r * n + r];
  }
  return b;
}

// DIIS history (rhf.hpp:298-327): the last `size` error and Fock matrices in
// rings of slots, and the Gram matrix of the stored errors.  A step computes
// only the new error's row of it (one dgemv and one read-back, where forming
// every pair again cost one synchronising dot per pair), then extrapolates F
// from the stored Fock matrices once two are held.
struct Diis {
  int size = 0, count = 0;
  size_t nn = 0;
  DevBuf<double> errs, focks, row, coef;
  DevBuf<const double*> mats;
  std::vector<double> gram;  // size x size, by slot

  void init(int size_, size_t nn_) {
    size = size_;
    nn = nn_;
    count = 0;
    errs.alloc(nn * size);
    focks.alloc(nn * size);
    row.alloc(size);
    coef.alloc(size);
    gram.assign((size_t)size * size, 0.0);
    std::vector<const double*> m(size);
    for (int i = 0; i < size; ++i) m[i] = focks.ptr + i * nn;
    mats.upload(m.data(), m.size());
  }

  // stores (e, F) and overwrites F with the extrapolated matrix
  void step(rocblas_handle h, const double* e, double* F) {
    const int slot = count % size;
    GANSU_HIP(hipMemcpy(errs.ptr + slot * nn, e, nn * sizeof(double), hipMemcpyDeviceToDevice));
    GANSU_HIP(hipMemcpy(focks.ptr + slot * nn, F, nn * sizeof(double), hipMemcpyDeviceToDevice));
    ++count;
    const int np = std::min(count, size);
    // row[i] = <e_i, e_slot> over the held slots (errs as an nn x np column-major matrix)
    const double one = 1.0, zero = 0.0;
    blas_check(rocblas_dgemv(h, rocblas_operation_transpose, (rocblas_int)nn, np, &one, errs.ptr, (rocblas_int)nn,
                             errs.ptr + slot * nn, 1, &zero, row.ptr, 1),
               "rocblas_dgemv(DIIS)");
    std::vector<double> r(np);
    GANSU_HIP(hipMemcpy(r.data(), row.ptr, np * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i) gram[(size_t)slot * size + i] = gram[(size_t)i * size + slot] = r[i];
    if (np < 2) return;
    const int ns = np + 1;
    std::vector<double> B((size_t)ns * ns, 0.0), rhs(ns, 0.0);
    for (int i = 0; i < np; ++i) {
      for (int j = 0; j < np; ++j) B[i * ns + j] = gram[(size_t)i * size + j];
      B[i * ns + np] = B[np * ns + i] = -1.0;
    }
    rhs[np] = -1.0;
    const auto c = solve_small(ns, B, rhs);
    if (c.empty()) return;  // a singular system keeps F as it is
    GANSU_HIP(hipMemcpy(coef.ptr, c.data(), np * sizeof(double), hipMemcpyHostToDevice));
    GANSU_HIP(launch_lincomb(np, coef.ptr, mats.ptr, nn, F, 0));
  }
};

// initial_guess values of rhf.cu:127-145 that the drivers provide ("density",
// a restart from caller densities, is not)
void check_guess(const std::string& guess) {
  if (guess != "core" && guess != "gwh" && guess != "sad")
    throw Error(GANSU_E_UNSUPPORTED, "initial_guess=" + guess + " is not provided (core, gwh and sad only)");
}

// Whether the guess comes from caller densities (guess_initial_fock_matrix,
// rhf.cu:122-134, uhf.cu:118-130): when force_density is set or
// initial_guess=density; without both densities the core guess is used then.
bool density_guess(const std::string& guess, bool force_density, const double* Da0, const double* Db0) {
  if (!force_density && guess != "density") return false;
  if (Da0 && Db0) return true;
  std::fprintf(stderr, "initial_guess=density or force_density without density matrices: the core Hamiltonian "
                       "guess is used\n");
  return false;
}

// UHF symmetry break of every UHF guess (InitialGuess_UHF::break_symmetry,
// uhf.hpp:429-436): the last row and the last column of the row-major
// C_beta are zeroed before the densities are formed.
void break_symmetry(double* Cb, int n) {
  GANSU_HIP(hipMemset(Cb + (size_t)(n - 1) * n, 0, sizeof(double) * n));
  GANSU_HIP(hipMemset2D(Cb + (n - 1), sizeof(double) * n, 0, sizeof(double), n));
}

// Damping of a Fock matrix (and, for optimal damping, of the density) as
// Convergence_RHF_Damping::get_new_fock_matrix (rhf.hpp:219-243) and
// gpu::damping (gpu_manager.cu:989-1007): the first call stores F and D;
// each later call sets F <- (1 - a) F_prev + a F and F_prev <- F, with a the
// damping factor or, optimally (gpu_manager.cu:898-971),
// a = min(1, -s / (2 c)), s = tr(F_prev dD), c = tr(dF dD); the optimal form
// damps D the same way.
struct Damper {
  bool first = true, optimal = false;
  double factor = 0.0;
  size_t nn = 0;
  DevBuf<double> prevF, prevD;
  void init(size_t nn_, bool optimal_, double factor_) {
    nn = nn_;
    optimal = optimal_;
    factor = factor_;
    first = true;
    prevF.alloc(nn);
    prevD.alloc(nn);
  }
  void step(double* F, double* D, double* t1, double* t2, DevBuf<double>& part) {
    if (first) {
      first = false;
      GANSU_HIP(hipMemcpy(prevD.ptr, D, nn * sizeof(double), hipMemcpyDeviceToDevice));
      GANSU_HIP(hipMemcpy(prevF.ptr, F, nn * sizeof(double), hipMemcpyDeviceToDevice));
      return;
    }
    double alpha = factor;
    if (optimal) {
      GANSU_HIP(launch_axpby(1.0, F, -1.0, prevF.ptr, nn, t1, 0));
      GANSU_HIP(launch_axpby(1.0, D, -1.0, prevD.ptr, nn, t2, 0));
      // for symmetric matrices tr(AB) = sum A .* B
      const double sv = dot(prevF.ptr, t2, nn, part);
      const double cv = dot(t1, t2, nn, part);
      alpha = (cv <= -sv / 2.0) ? 1.0 : -0.5 * sv / cv;
      GANSU_HIP(launch_axpby(1.0 - alpha, prevD.ptr, alpha, D, nn, prevD.ptr, 0));
      GANSU_HIP(hipMemcpy(D, prevD.ptr, nn * sizeof(double), hipMemcpyDeviceToDevice));
    }
    GANSU_HIP(launch_axpby(1.0 - alpha, prevF.ptr, alpha, F, nn, prevF.ptr, 0));
    GANSU_HIP(hipMemcpy(F, prevF.ptr, nn * sizeof(double), hipMemcpyDeviceToDevice));
  }
};

// The RHF iteration of the loop below (hf.cu:278-307) as one captured HIP
// graph: X^T F X, SP2 purification instead of the eigensolver, D = 2 X P X^T,
// the engine's four-centre build, E = 1/2 D.(H + F) and the convergence
// test, then (unless converged) the DIIS error F D S - S D F and the DIIS
// step, all reading their control words from a device state.  The host
// launches the graph once per iteration and reads the control words back
// (56 bytes) to decide whether to go on.  The purification's step count is
// captured with a margin over what the first iteration needed; an
// iteration that runs out of steps doubles it (re-captured), and a
// convergence reached on such an iteration is not accepted.
struct GraphLoop {
  double E = 0.0, seconds = 0.0;
  int iterations = 0, converged = 0, launches = 0;
};

GraphLoop graph_loop(FockEngine& eng, Work& w, int nocc, int maxiter, double conv, int diis_size, bool diis_x,
                     bool timing, bool trace) {
  const int n = w.n;
  const size_t nn = (size_t)n * n;
  DevBuf<double> Fo, A, e, errs(nn * diis_size), focks(nn * diis_size), dpart((size_t)kDiisBlocks * diis_size);
  for (auto* b : {&Fo, &A, &e}) b->alloc(nn);
  DevBuf<double> P0(nn), P1(nn), part(sp2_part_doubles(n)), rows(2 * (size_t)n), epart(kEnergyBlocks);
  DevBuf<ScfDevState> st(1);
  ScfDevState h{};
  h.Eprev = std::numeric_limits<double>::max();
  GANSU_HIP(hipMemcpy(st.ptr, &h, sizeof h, hipMemcpyHostToDevice));
  const int* conv_flag = &st.ptr->converged;

  hipStream_t cs = nullptr;
  GANSU_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf<double> eps;
  DevBuf<int> done;
  auto release = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  };
  auto capture = [&](int steps) {
    release();
    eps.alloc(steps + 1);
    done.alloc(steps + 1);
    const Sp2Work sw{{P0.ptr, P1.ptr}, A.ptr, rows.ptr, part.ptr, eps.ptr, done.ptr, sp2_workgroups(n)};
    const double* P = A.ptr;
    GANSU_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      GANSU_HIP(launch_gemm64(true, false, n, 1.0, w.X.ptr, w.F.ptr, 0.0, w.tmp.ptr, nullptr, cs));
      GANSU_HIP(launch_gemm64(false, false, n, 1.0, w.tmp.ptr, w.X.ptr, 0.0, Fo.ptr, nullptr, cs));
      GANSU_HIP(launch_sp2(Fo.ptr, n, nocc, steps, sw, st.ptr, cs));
      GANSU_HIP(launch_gemm64(false, false, n, 1.0, w.X.ptr, P, 0.0, w.tmp.ptr, nullptr, cs));
      GANSU_HIP(launch_gemm64(false, true, n, 2.0, w.tmp.ptr, w.X.ptr, 0.0, w.D.ptr, nullptr, cs));
      eng.record_build(w.D.ptr, w.H.ptr, w.F.ptr, cs);
      GANSU_HIP(launch_scf_energy(w.D.ptr, w.H.ptr, w.F.ptr, nn, conv, epart.ptr, st.ptr, cs));
      // e = F D S - S D F = A - A^T with A = F D S (F, D, S symmetric)
      GANSU_HIP(launch_gemm64(false, false, n, 1.0, w.F.ptr, w.D.ptr, 0.0, w.tmp.ptr, conv_flag, cs));
      GANSU_HIP(launch_gemm64(false, false, n, 1.0, w.tmp.ptr, w.S.ptr, 0.0, A.ptr, conv_flag, cs));
      GANSU_HIP(launch_antisym(A.ptr, n, e.ptr, st.ptr, cs));
      if (diis_x) {
        GANSU_HIP(launch_gemm64(false, false, n, 1.0, w.X.ptr, e.ptr, 0.0, w.tmp.ptr, conv_flag, cs));
        GANSU_HIP(launch_gemm64(false, true, n, 1.0, w.tmp.ptr, w.X.ptr, 0.0, e.ptr, conv_flag, cs));
      }
      GANSU_HIP(launch_diis(e.ptr, w.F.ptr, errs.ptr, focks.ptr, dpart.ptr, nn, diis_size, st.ptr, cs));
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    GANSU_HIP(hipStreamEndCapture(cs, &graph));
    GANSU_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  };

  GraphLoop r;
  GANSU_HIP(hipDeviceSynchronize());  // the guess ran on the null stream
  int steps = 128;
  bool first = true;
  try {
    capture(steps);
    while (true) {
      const auto t0 = std::chrono::steady_clock::now();
      GANSU_HIP(hipGraphLaunch(exec, cs));
      // the control words only (not the DIIS Gram matrix behind them)
      GANSU_HIP(hipMemcpyAsync(&h, st.ptr, offsetof(ScfDevState, gram), hipMemcpyDeviceToHost, cs));
      GANSU_HIP(hipStreamSynchronize(cs));
      r.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      ++r.launches;
      if (trace)
        std::fprintf(stderr, "[scf] graph iteration %d: E = %.12f, purification steps %d (bounds %.4f %.4f), "
                             "DIIS steps %d\n", r.launches, h.E, h.sp2_used, h.bounds[0], h.bounds[1], h.diis_count);
      if (!std::isfinite(h.E)) throw Error(GANSU_E_NUMERIC, "SCF energy is not finite");
      if (h.sp2_used < 0) {
        // the purification ran out of steps: more steps, and an energy
        // reached with this density does not count as converged
        if (steps >= 1024)
          throw Error(GANSU_E_NUMERIC, "density purification did not reach idempotency (no gap at the Fermi level?)");
        steps *= 2;
        if (timing) std::fprintf(stderr, "[scf] purification out of steps; re-captured with %d\n", steps);
        if (h.converged) {
          const int zero = 0;
          GANSU_HIP(hipMemcpy(&st.ptr->converged, &zero, sizeof zero, hipMemcpyHostToDevice));
          h.converged = 0;
        }
        capture(steps);
      } else if (first) {
        // size the step count to the first density's need with a margin
        // (steps after the stop cost a launch each)
        const int want = std::max(48, (h.sp2_used * 3 / 2 + 16 + 7) / 8 * 8);
        if (want < steps) {
          steps = want;
          capture(steps);
        }
        first = false;
      }
      if (h.converged || h.iter >= maxiter) break;
    }
  } catch (...) {
    release();
    (void)hipStreamDestroy(cs);
    throw;
  }
  release();
  GANSU_HIP(hipStreamSynchronize(cs));
  (void)hipStreamDestroy(cs);
  // X^T F X of the last density's Fock matrix, for the final orbitals
  GANSU_HIP(hipMemcpy(w.tmp2.ptr, Fo.ptr, nn * sizeof(double), hipMemcpyDeviceToDevice));
  r.E = h.E;
  r.iterations = h.iter;
  r.converged = h.converged;
  return r;
}

}  // namespace

OneIntTables make_oneint_tables(const Molecule& mol) {
  OneIntTables t;
  t.M = mol.nbasis;
  t.natoms = (int)mol.atoms.size();
  const auto shells = contracted_shells(mol.prims.data(), mol.prims.size(), kMaxL);
  std::vector<int> o{0}, sp, l, bs, bf(mol.nbasis, -1);
  std::vector<double> wv(mol.nbasis);
  for (size_t i = 0; i < shells.size(); ++i) {
    for (int p : shells[i].prims) sp.push_back(p);
    o.push_back((int)sp.size());
    l.push_back(shells[i].L);
    bs.push_back(shells[i].basis);
    for (int c = 0; c < ncart(shells[i].L); ++c) bf[shells[i].basis + c] = (int)i;
  }
  for (int mu = 0; mu < mol.nbasis; ++mu) {
    if (bf[mu] < 0) throw Error(GANSU_E_ARG, "basis function without shell");
    const int L = l[bf[mu]], c = mu - bs[bf[mu]];
    // same factor as the engine: cgto norm / sqrt((2l-1)!!(2m-1)!!(2n-1)!!)
    int e[3];
    static const int d2[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for (int d = 0; d < 3; ++d) e[d] = L == 0 ? 0 : (L == 1 ? (c == d) : d2[c][d]);
    double df = 1.0;
    for (int d = 0; d < 3; ++d) df *= (e[d] == 2 ? 3.0 : 1.0);
    wv[mu] = mol.norm[mu] / std::sqrt(df);
  }
  for (const auto& a : mol.atoms) {
    t.xyz.insert(t.xyz.end(), a.xyz, a.xyz + 3);
    t.Z.push_back(a.Z);
    t.zsum += std::abs((double)a.Z);
  }
  t.prims = mol.prims;
  t.off = std::move(o);
  t.sprims = std::move(sp);
  t.sl = std::move(l);
  t.sb = std::move(bs);
  t.bfs = std::move(bf);
  t.w = std::move(wv);
  t.boys = boys_grid_xn();
  return t;
}

OneIntHost::OneIntHost(const Molecule& mol) : M(mol.nbasis), natoms((int)mol.atoms.size()) {
  const OneIntTables t = make_oneint_tables(mol);
  zsum = t.zsum;
  shell_L = t.sl;
  prims.upload(t.prims);
  off.upload(t.off); sprims.upload(t.sprims); sl.upload(t.sl); sb.upload(t.sb); bfs.upload(t.bfs); Z.upload(t.Z);
  w.upload(t.w); xyz.upload(t.xyz);
  boys.upload(t.boys);
}

OneIntArgs OneIntHost::args(double* S, double* T, double* V) const {
  return OneIntArgs{prims.ptr, off.ptr, sprims.ptr, sl.ptr, sb.ptr, bfs.ptr, w.ptr, M, natoms,
                    xyz.ptr, Z.ptr, boys.ptr, S, T, V, zsum, (int)shell_L.size()};
}

void compute_oneint_noalloc(const Molecule& mol, double* dS, double* dT, double* dV, hipStream_t s) {
  oneint_noalloc(make_oneint_tables(mol), dS, dT, dV, s);
}

// Build S, T, V of a molecule on the device.
void compute_oneint(const Molecule& mol, double* dS, double* dT, double* dV, hipStream_t s) {
  OneIntHost h(mol);
  GANSU_HIP(launch_oneint(h.args(dS, dT, dV), s));
  GANSU_HIP(hipStreamSynchronize(s));
}

void compute_oneint_gradient(const Molecule& mol, const double* dD, const double* dW, double* dgrad, hipStream_t s) {
  OneIntHost h(mol);
  if (h.natoms > kGradMaxAtoms)
    throw Error(GANSU_E_UNSUPPORTED, "gradient limited to " + std::to_string(kGradMaxAtoms) + " atoms");
  GANSU_HIP(launch_zero(dgrad, 3 * (size_t)h.natoms, s));
  const OneIntArgs args = h.args(nullptr, nullptr, nullptr);
  GANSU_HIP(launch_grad_1e(args, dD, dW, dgrad, s));
  // attraction terms by shell-pair class: every unordered pair once, the
  // shell of higher angular momentum first
  const int ns = (int)h.shell_L.size();
  std::vector<int2> pairs;
  int range[kMaxL + 1][kMaxL + 1][2] = {};
  for (int la = 0; la <= kMaxL; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      range[la][lb][0] = (int)pairs.size();
      for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 < ns; ++s2)
          if (h.shell_L[s1] == la && h.shell_L[s2] == lb && (la > lb || s1 >= s2)) pairs.push_back(make_int2(s1, s2));
      range[la][lb][1] = (int)pairs.size() - range[la][lb][0];
    }
  DevBuf<int2> dpairs;
  dpairs.upload(pairs);
  for (int la = 0; la <= kMaxL; ++la)
    for (int lb = 0; lb <= la; ++lb)
      GANSU_HIP(launch_grad_1e_nuc(la, lb, args, dpairs.ptr + range[la][lb][0], range[la][lb][1], dD, dgrad, s));
  GANSU_HIP(hipStreamSynchronize(s));
}

std::vector<double> nuclear_repulsion_gradient(const Molecule& mol) {
  const size_t n = mol.atoms.size();
  std::vector<double> g(3 * n, 0.0);
  for (size_t a = 0; a < n; ++a)
    for (size_t b = 0; b < n; ++b) {
      if (a == b) continue;
      double r[3], r2 = 0.0;
      for (int d = 0; d < 3; ++d) { r[d] = mol.atoms[a].xyz[d] - mol.atoms[b].xyz[d]; r2 += r[d] * r[d]; }
      const double f = double(mol.atoms[a].Z) * mol.atoms[b].Z / (r2 * std::sqrt(r2));
      for (int d = 0; d < 3; ++d) g[3 * a + d] -= f * r[d];
    }
  return g;
}

ScfResult rhf_run(const Molecule& mol, const Molecule* aux, const Params& prm, bool gradient, const double* Da0,
                  const double* Db0, bool force_density) {
  if (prm.get("method") != "rhf") throw Error(GANSU_E_PARAM, "only method=rhf is provided by this engine");
  if (mol.n_electrons % 2 != 0) throw Error(GANSU_E_ELECTRONS, "RHF requires an even number of electrons");
  const int n = mol.nbasis;
  const int nocc = mol.n_electrons / 2;
  if (nocc > n) throw Error(GANSU_E_ELECTRONS, "more occupied orbitals than basis functions");
  const std::string guess = prm.get("initial_guess");
  if (guess != "density") check_guess(guess);
  // caller densities start the loop when force_density is set or
  // initial_guess=density (rhf.cu:122-134); otherwise the named guess runs
  const bool from_density = density_guess(guess, force_density, Da0, Db0);
  const std::string conv_method = prm.get("convergence_method");
  if (conv_method != "diis" && conv_method != "damping" && conv_method != "optimaldamping")
    throw Error(GANSU_E_PARAM, "Unknown convergence method: " + conv_method);
  const int maxiter = prm.geti("maxiter");
  const double conv = prm.getd("convergence_energy_threshold");
  const int diis_size = prm.geti("diis_size");
  const bool diis_x = prm.geti("diis_include_transform") != 0;
  const double damp = prm.getd("damping_factor");
  const std::string solver = prm.get("scf_solver");
  if (solver != "diagonalize" && solver != "purify") throw Error(GANSU_E_PARAM, "Unknown scf_solver: " + solver);
  const bool purify = solver == "purify";
  if (purify && conv_method != "diis") throw Error(GANSU_E_UNSUPPORTED, "scf_solver=purify runs DIIS only");
  if (purify && (diis_size < 1 || diis_size > kGraphDiisMax))
    throw Error(GANSU_E_PARAM, "scf_solver=purify needs 1 <= diis_size <= " + std::to_string(kGraphDiisMax));

  // GANSU_SCF_TIMING=1 prints the wall time of each phase (with a device
  // synchronisation at every phase boundary) to stderr.
  const char* tenv = std::getenv("GANSU_SCF_TIMING");
  const bool timing = tenv && tenv[0] == '1';
  // GANSU_SCF_TRACE=1 prints the energy of every iteration to stderr.
  const char* trenv = std::getenv("GANSU_SCF_TRACE");
  const bool trace = trenv && trenv[0] == '1';
  auto wall = [&]() {
    if (timing) GANSU_HIP(hipDeviceSynchronize());
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  const double ts0 = wall();
  Work w(n);
  Blas blas;
  const size_t nn = (size_t)n * n;
  compute_oneint(mol, w.S.ptr, w.T.ptr, w.V.ptr, 0);
  GANSU_HIP(launch_axpby(1.0, w.T.ptr, 1.0, w.V.ptr, nn, w.H.ptr, 0));

  std::map<std::string, std::string> opts{{"eri_method", prm.get("eri_method")},
                                          {"schwarz_screening_threshold", prm.get("schwarz_screening_threshold")},
                                          {"precision", prm.get("precision")},
                                          {"fp32_bound_threshold", prm.get("fp32_bound_threshold")},
                                          {"ri_k", prm.get("ri_k")},
                                          {"incremental_fock", prm.get("incremental_fock")},
                                          {"incremental_threshold", prm.get("incremental_threshold")}};
  const bool ri = prm.get("eri_method") == "ri";
  if (ri && !aux) throw Error(GANSU_E_ARG, "eri_method=ri needs auxiliary_gbsfilename");
  FockEngine eng(mol.prims.data(), mol.prims.size(), mol.types.data(), (int)mol.types.size(), mol.norm.data(), n,
                 ri ? aux->prims.data() : nullptr, ri ? aux->prims.size() : 0, ri ? aux->types.data() : nullptr,
                 ri ? (int)aux->types.size() : 0, ri ? aux->norm.data() : nullptr, ri ? aux->nbasis : 0, opts);
  const double ts1 = wall();
  eng.precompute();
  const double ts2 = wall();

  // X = U s^-1/2
  syevd(blas.h, n, w.S.ptr, w);
  GANSU_HIP(launch_transform(w.evec.ptr, w.eval.ptr, n, 1e-6, w.X.ptr, 0));

  ScfResult res;
  double t_diag = 0, t_dens = 0, t_fock = 0, t_energy = 0, t_conv = 0;
  auto now = [&]() {
    if (timing) GANSU_HIP(hipDeviceSynchronize());
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  auto coefficients = [&](const double* Fm) {
    blas.gemm(true, false, n, w.X.ptr, Fm, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.X.ptr, w.tmp2.ptr);
    syevd(blas.h, n, w.tmp2.ptr, w);
    // evec column-major: row view is V^T, so C = X V = X (row view)^T
    blas.gemm(false, true, n, w.X.ptr, w.evec.ptr, w.C.ptr);
  };
  // RI engines with scf_solver=purify: the occupied orbitals of an iteration
  // come from SP2 purification of Fo = X^T F X instead of the eigensolver
  // (their builds need an occupied factor, so the captured loop of the
  // four-centre engines does not apply).  The projector P takes the previous
  // iteration's occupied orbitals Z (orthogonal basis, column-major n x nocc)
  // into its range, Q0 = P Z, and a Cholesky QR (Q0^T Q0 = R^T R, Q = Q0 R^-1)
  // makes them orthonormal: C_occ = X Q.  Any orthonormal basis of the
  // occupied space gives the same D and RI exchange (the incremental RI-K
  // aligns successive factors itself).  The first iteration, one whose
  // purification does not reach idempotency and one whose Q0 is
  // ill-conditioned (sigma_min / sigma_max < 1e-4) diagonalise instead.
  const bool ri_purify = purify && ri;
  DevBuf<double> sFo, sP, sP2, srows, spart, strfr, sZ, sG;
  int sp2_iters = 0, sp2_fallbacks = 0, sp2_steps = 0;
  bool have_z = false, last_sp2 = false;
  if (ri_purify) {
    for (auto* b : {&sFo, &sP, &sP2}) b->alloc(nn);
    srows.alloc(2 * (size_t)n);
    spart.alloc(sp2e_part_doubles());
    strfr.alloc(2);
    sZ.alloc((size_t)n * nocc);
    sG.alloc((size_t)nocc * nocc);
  }
  // SP2 (Niklasson, Phys. Rev. B 66, 155115 (2002)) of sFo into sP, the
  // products on rocBLAS, the branch and stop decisions on the host from
  // (tr P, tr P^2) as in the captured loop's kernels: P <- P^2 while tr P >
  // nocc, else 2P - P^2, until the idempotency error tr P - tr P^2 is below
  // 1e-4 and no longer falls between every second step.  Returns the steps
  // taken, -1 if 256 did not reach that floor.
  auto sp2_eager = [&]() -> int {
    double h[2];
    GANSU_HIP(launch_sp2e_init(sFo.ptr, n, srows.ptr, sP.ptr, spart.ptr, strfr.ptr, 0));
    GANSU_HIP(hipMemcpy(h, strfr.ptr, sizeof h, hipMemcpyDeviceToHost));
    std::vector<double> eps;
    const double one = 1.0, zero = 0.0;
    for (int k = 1; k <= 256; ++k) {
      const double e = h[0] - h[1];  // idempotency error of P_(k-1)
      const bool stop = !(e > 0.0) || (k >= 3 && e < 1e-4 && e >= eps[k - 3]);
      eps.push_back(e);
      if (stop) return k - 1;
      blas_check(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, sP.ptr, n,
                               sP.ptr, n, &zero, sP2.ptr, n),
                 "rocblas_dgemm(P P)");
      GANSU_HIP(launch_sp2e_apply(sP.ptr, sP2.ptr, n, h[0] > nocc, spart.ptr, strfr.ptr, 0));
      GANSU_HIP(hipMemcpy(h, strfr.ptr, sizeof h, hipMemcpyDeviceToHost));
    }
    return -1;
  };
  auto keep_z = [&]() {  // the occupied eigenvectors of the last diagonalisation
    GANSU_HIP(hipMemcpy(sZ.ptr, w.evec.ptr, sizeof(double) * n * nocc, hipMemcpyDeviceToDevice));
    have_z = true;
  };
  auto coefficients_sp2 = [&](const double* Fm) -> bool {
    blas.gemm(true, false, n, w.X.ptr, Fm, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.X.ptr, sFo.ptr);
    const int used = sp2_eager();
    if (used < 0) return false;
    sp2_steps = std::max(sp2_steps, used);
    const double one = 1.0, zero = 0.0;
    double* Q = w.tmp.ptr;  // n x nocc, column-major
    blas_check(rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, n, nocc, n, &one, sP.ptr, n,
                             sZ.ptr, n, &zero, Q, n),
               "rocblas_dgemm(P Z)");
    blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, nocc, nocc, n, &one, Q, n,
                             Q, n, &zero, sG.ptr, nocc),
               "rocblas_dgemm(Q0^T Q0)");
    blas_check(rocsolver_dpotrf(blas.h, rocblas_fill_upper, nocc, sG.ptr, nocc, w.info.ptr), "rocsolver_dpotrf");
    int info = 0;
    GANSU_HIP(hipMemcpy(&info, w.info.ptr, sizeof info, hipMemcpyDeviceToHost));
    if (info != 0) return false;
    std::vector<double> rd(nocc);
    GANSU_HIP(hipMemcpy2D(rd.data(), sizeof(double), sG.ptr, (size_t)(nocc + 1) * sizeof(double), sizeof(double),
                          nocc, hipMemcpyDeviceToHost));
    double rmin = std::numeric_limits<double>::max(), rmax = 0.0;
    for (double v : rd) {
      rmin = std::min(rmin, std::fabs(v));
      rmax = std::max(rmax, std::fabs(v));
    }
    if (!(rmin > 1e-4 * rmax)) return false;
    blas_check(rocblas_dtrsm(blas.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                             rocblas_diagonal_non_unit, n, nocc, &one, sG.ptr, nocc, Q, n),
               "rocblas_dtrsm(Q0 R^-1)");
    GANSU_HIP(hipMemcpy(sZ.ptr, Q, sizeof(double) * n * nocc, hipMemcpyDeviceToDevice));
    // rows i < nocc of the column-major view of row-major C: C^T[i][mu] = (Q^T X^T)[i][mu]
    blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, nocc, n, n, &one, sZ.ptr,
                             n, w.X.ptr, n, &zero, w.C.ptr, n),
               "rocblas_dgemm(C_occ)");
    return true;
  };
  auto density = [&]() {
    // D = 2 C_occ C_occ^T (gpu_kernels.cu:54-69).  The column-major view of
    // row-major C is C^T, whose first nocc rows are the occupied orbitals.
    const double two = 2.0, zero = 0.0;
    blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, n, n, nocc, &two, w.C.ptr,
                             n, w.C.ptr, n, &zero, w.D.ptr, n),
               "rocblas_dgemm(D)");
  };
  // occ: D = 2 C_occ C_occ^T holds (every build but the SAD guess's first),
  // so the RI exchange may take its occupied form
  // (the RI build checks F for NaN itself, as ERI_RI_RHF does, rhf.hpp:665-674)
  auto fock = [&](bool occ = true) {
    eng.build_jk(w.D.ptr, nullptr, nullptr, w.H.ptr, w.F.ptr, 0, occ ? w.C.ptr : nullptr, occ ? nocc : 0);
    ++res.fock_builds;
  };
  auto energy = [&]() { return 0.5 * (dot(w.D.ptr, w.H.ptr, nn, w.part) + dot(w.D.ptr, w.F.ptr, nn, w.part)); };

  const double ts3 = wall();
  DevBuf<double> guess_F;
  if (from_density) {
    // D = Da + Db and F(D) only (InitialGuess_RHF_Density::guess, rhf.hpp:441-454)
    std::vector<double> d(nn);
    for (size_t i = 0; i < nn; ++i) d[i] = Da0[i] + Db0[i];
    GANSU_HIP(hipMemcpy(w.D.ptr, d.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    fock(false);
  } else {
    if (guess == "sad") {
      // D = Da + Db of the atoms, F(D), then C, D and F as after any guess
      // (InitialGuess_RHF_SAD::guess, rhf.hpp:551-584)
      std::vector<double> da, db;
      sad_density(mol, da, db);
      for (size_t i = 0; i < nn; ++i) da[i] += db[i];
      GANSU_HIP(hipMemcpy(w.D.ptr, da.data(), nn * sizeof(double), hipMemcpyHostToDevice));
      fock(false);
      coefficients(w.F.ptr);
    } else {
      if (guess == "gwh") {
        guess_F.alloc(nn);
        GANSU_HIP(launch_gwh_guess(w.H.ptr, w.S.ptr, n, guess_F.ptr, 0));
      }
      coefficients(guess == "gwh" ? guess_F.ptr : w.H.ptr);
    }
    if (ri_purify) keep_z();
    density();
    fock();
  }
  if (timing)
    std::fprintf(stderr, "[scf] setup: one-electron integrals + engine %.2f ms, precompute %.2f ms, X %.2f ms, "
                         "core guess + first build %.2f ms\n", 1e3 * (ts1 - ts0), 1e3 * (ts2 - ts1),
                 1e3 * (ts3 - ts2), 1e3 * (wall() - ts3));

  // convergence state
  Diis diis;
  Damper damper;
  const bool graph = purify && !ri;  // the captured loop of the four-centre engines
  if (conv_method == "diis" && !graph) diis.init(diis_size, nn);
  else if (!graph) damper.init(nn, conv_method == "optimaldamping", damp);
  double prev = std::numeric_limits<double>::max(), E = 0.0;
  int it = 0;
  if (graph) {
    GraphLoop gl = graph_loop(eng, w, nocc, maxiter, conv, diis_size, diis_x, timing, trace);
    E = gl.E;
    it = gl.iterations;
    res.converged = gl.converged;
    res.fock_builds += gl.launches;
    res.graph_launches = gl.launches;
    t_fock = gl.seconds;
    // orbitals of the last density: eigenvectors of its X^T F X (left in w.tmp2)
    syevd(blas.h, n, w.tmp2.ptr, w);
    blas.gemm(false, true, n, w.X.ptr, w.evec.ptr, w.C.ptr);
  }
  while (!graph && it < maxiter) {
    double t0 = now();
    if (ri_purify && have_z && coefficients_sp2(w.F.ptr)) {
      ++sp2_iters;
      last_sp2 = true;
    } else {
      coefficients(w.F.ptr);
      if (ri_purify) {
        keep_z();
        ++sp2_fallbacks;
        last_sp2 = false;
      }
    }
    double t1 = now();
    density();
    double t2 = now();
    fock();
    double t3 = now();
    E = energy();
    double t4 = now();
    t_diag += t1 - t0; t_dens += t2 - t1; t_fock += t3 - t2; t_energy += t4 - t3;
    if (trace) std::fprintf(stderr, "[scf] iteration %d: E = %.12f\n", it + 1, E);
    if (std::fabs(E - prev) < conv) { res.converged = 1; break; }
    ++it;
    prev = E;
    const double tc0 = now();
    if (conv_method == "diis") {
      // e = F D S - S D F (optionally X e X^T)
      double* e = w.tmp2.ptr;
      blas.gemm(false, false, n, w.F.ptr, w.D.ptr, w.tmp.ptr);
      blas.gemm(false, false, n, w.tmp.ptr, w.S.ptr, e);
      blas.gemm(false, false, n, w.S.ptr, w.D.ptr, w.tmp.ptr);
      blas.gemm(false, false, n, w.tmp.ptr, w.F.ptr, e, -1.0, 1.0);
      if (diis_x) {
        blas.gemm(false, false, n, w.X.ptr, e, w.tmp.ptr);
        blas.gemm(false, true, n, w.tmp.ptr, w.X.ptr, e);
      }
      diis.step(blas.h, e, w.F.ptr);
    } else {
      damper.step(w.F.ptr, w.D.ptr, w.tmp.ptr, w.tmp2.ptr, w.part);
    }
    t_conv += now() - tc0;
  }
  if (ri_purify && last_sp2) {
    // canonical orbitals and energies of the Fock matrix that gave the last D
    syevd(blas.h, n, sFo.ptr, w);
    blas.gemm(false, true, n, w.X.ptr, w.evec.ptr, w.C.ptr);
  }
  if (timing && ri_purify)
    std::fprintf(stderr, "[scf] purification for %d iterations, eigensolver for %d (at most %d SP2 steps)\n",
                 sp2_iters, sp2_fallbacks, sp2_steps);
  if (timing && graph)
    std::fprintf(stderr, "[scf] %d iterations in %d graph launches: %.3f ms per iteration\n", it + 1,
                 res.graph_launches, 1e3 * t_fock / std::max(1, res.graph_launches));
  if (timing && !graph)
    std::fprintf(stderr, "[scf] %d iterations: diag %.2f ms, density %.2f ms, fock %.2f ms, energy %.2f ms, "
                         "convergence %.2f ms (per iteration)\n", it + 1, 1e3 * t_diag / (it + 1),
                 1e3 * t_dens / (it + 1), 1e3 * t_fock / (it + 1), 1e3 * t_energy / (it + 1), 1e3 * t_conv / (it + 1));
  GANSU_HIP(hipDeviceSynchronize());
  res.e_elec = E;
  res.e_total = E + mol.nuclear_repulsion();
  res.iterations = it;
  res.D = w.D.download();
  res.C = w.C.download();     // the orbitals of the last diagonalisation (those of D)
  res.eps = w.eval.download();
  if (gradient) {
    // W = 2 C_occ eps C_occ^T = D F D / 2 for the converged pair (D, F(D))
    blas.gemm(false, false, n, w.F.ptr, w.D.ptr, w.tmp.ptr);
    blas.gemm(false, false, n, w.D.ptr, w.tmp.ptr, w.tmp2.ptr, 0.5);
    const int na = (int)mol.atoms.size();
    DevBuf<double> g1(3 * (size_t)na), g2(3 * (size_t)na);
    compute_oneint_gradient(mol, w.D.ptr, w.tmp2.ptr, g1.ptr, 0);
    eng.gradient_2e(w.D.ptr, na, g2.ptr, 0, w.C.ptr, nocc);  // D = 2 C_occ C_occ^T at the converged step
    GANSU_HIP(hipDeviceSynchronize());
    const auto h1 = g1.download(), h2 = g2.download();
    res.grad = nuclear_repulsion_gradient(mol);
    for (size_t i = 0; i < res.grad.size(); ++i) res.grad[i] += h1[i] + h2[i];
  }
  return res;
}

// UHF (uhf.cu:118-300, uhf.hpp:291-400): core-Hamiltonian guess; per spin
// C from X^T F X, D_s = C_s,occ C_s,occ^T (n_alpha / n_beta orbitals), the
// Fock pair from the engine's UHF build, E = 1/2 sum_s D_s.(H + F_s),
// convergence on |dE|, and DIIS run separately for each spin with the error
// F_s D_s S - S D_s F_s.
UhfResult uhf_run(const Molecule& mol, const Molecule* aux, const Params& prm, const double* Da0, const double* Db0,
                  bool force_density) {
  const int n = mol.nbasis, na = mol.n_alpha, nb = mol.n_beta;
  if (na > n || nb > n)
    throw Error(GANSU_E_ELECTRONS, "more alpha or beta electrons than basis functions");
  const std::string guess = prm.get("initial_guess");
  if (guess != "density") check_guess(guess);
  const bool from_density = density_guess(guess, force_density, Da0, Db0);
  // uhf.cu:57-69: damping (constant factor) or DIIS
  const std::string conv_method = prm.get("convergence_method");
  if (conv_method != "diis" && conv_method != "damping")
    throw Error(GANSU_E_PARAM, "Invalid convergence_method: " + conv_method);
  const int maxiter = prm.geti("maxiter");
  const double conv = prm.getd("convergence_energy_threshold");
  const int diis_size = prm.geti("diis_size");
  const bool diis_x = prm.geti("diis_include_transform") != 0;
  const double damp = prm.getd("damping_factor");

  Work w(n);
  Blas blas;
  const size_t nn = (size_t)n * n;
  compute_oneint(mol, w.S.ptr, w.T.ptr, w.V.ptr, 0);
  GANSU_HIP(launch_axpby(1.0, w.T.ptr, 1.0, w.V.ptr, nn, w.H.ptr, 0));
  std::map<std::string, std::string> opts{{"eri_method", prm.get("eri_method")},
                                          {"schwarz_screening_threshold", prm.get("schwarz_screening_threshold")},
                                          {"precision", prm.get("precision")},
                                          {"fp32_bound_threshold", prm.get("fp32_bound_threshold")},
                                          {"ri_k", prm.get("ri_k")}};
  const bool ri = prm.get("eri_method") == "ri";
  if (ri && !aux) throw Error(GANSU_E_ARG, "eri_method=ri needs auxiliary_gbsfilename");
  FockEngine eng(mol.prims.data(), mol.prims.size(), mol.types.data(), (int)mol.types.size(), mol.norm.data(), n,
                 ri ? aux->prims.data() : nullptr, ri ? aux->prims.size() : 0, ri ? aux->types.data() : nullptr,
                 ri ? (int)aux->types.size() : 0, ri ? aux->norm.data() : nullptr, ri ? aux->nbasis : 0, opts);
  eng.precompute();
  syevd(blas.h, n, w.S.ptr, w);
  GANSU_HIP(launch_transform(w.evec.ptr, w.eval.ptr, n, 1e-6, w.X.ptr, 0));

  // spin 0 = alpha (the driver's F, D, C), spin 1 = beta
  DevBuf<double> Fb(nn), Db(nn), Cb(nn);
  double* F[2] = {w.F.ptr, Fb.ptr};
  double* D[2] = {w.D.ptr, Db.ptr};
  double* C[2] = {w.C.ptr, Cb.ptr};
  const int nocc[2] = {na, nb};
  auto coefficients = [&](const double* Fm, double* Cm) {
    blas.gemm(true, false, n, w.X.ptr, Fm, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.X.ptr, w.tmp2.ptr);
    syevd(blas.h, n, w.tmp2.ptr, w);
    blas.gemm(false, true, n, w.X.ptr, w.evec.ptr, Cm);
  };
  auto density = [&](int sp) {
    // D_s = C_s,occ C_s,occ^T (gpu_kernels.cu computeDensityMatrix_UHF_kernel)
    if (nocc[sp] == 0) {
      GANSU_HIP(launch_zero(D[sp], nn, 0));
      return;
    }
    const double one = 1.0, zero = 0.0;
    blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, n, n, nocc[sp], &one,
                             C[sp], n, C[sp], n, &zero, D[sp], n),
               "rocblas_dgemm(D_s)");
  };
  UhfResult res;
  auto fock = [&]() {
    eng.build_uhf(D[0], D[1], w.H.ptr, F[0], F[1], 0);
    ++res.fock_builds;
  };
  auto energy = [&]() {
    double e = 0.0;
    for (int sp = 0; sp < 2; ++sp) e += dot(D[sp], w.H.ptr, nn, w.part) + dot(D[sp], F[sp], nn, w.part);
    return 0.5 * e;
  };
  // Every guess ends with the symmetry break (the last row and column of
  // C_beta zeroed), then D per spin and F (uhf.hpp:494-568, :668-705).
  DevBuf<double> guess_F;
  if (from_density || guess == "sad") {
    // caller (InitialGuess_UHF_Density, uhf.hpp:557-568) or atomic
    // (InitialGuess_UHF_SAD) D_a, D_b, F(D), C per spin
    std::vector<double> dh[2];
    if (!from_density) sad_density(mol, dh[0], dh[1]);
    for (int sp = 0; sp < 2; ++sp)
      GANSU_HIP(hipMemcpy(D[sp], from_density ? (sp ? Db0 : Da0) : dh[sp].data(), nn * sizeof(double),
                          hipMemcpyHostToDevice));
    fock();
    for (int sp = 0; sp < 2; ++sp) coefficients(F[sp], C[sp]);
  } else {
    // core or GWH: C_alpha from H (or the GWH matrix), C_beta a copy of it
    if (guess == "gwh") {
      guess_F.alloc(nn);
      GANSU_HIP(launch_gwh_guess(w.H.ptr, w.S.ptr, n, guess_F.ptr, 0));
    }
    coefficients(guess == "gwh" ? guess_F.ptr : w.H.ptr, C[0]);
    GANSU_HIP(hipMemcpy(C[1], C[0], nn * sizeof(double), hipMemcpyDeviceToDevice));
  }
  break_symmetry(C[1], n);
  for (int sp = 0; sp < 2; ++sp) density(sp);
  fock();
  Diis diis[2];
  Damper damper[2];
  for (int sp = 0; sp < 2; ++sp) {
    if (conv_method == "diis") diis[sp].init(diis_size, nn);
    else damper[sp].init(nn, false, damp);
  }
  double prev = std::numeric_limits<double>::max(), E = 0.0;
  int it = 0;
  while (it < maxiter) {
    for (int sp = 0; sp < 2; ++sp) {
      coefficients(F[sp], C[sp]);
      density(sp);
    }
    fock();
    E = energy();
    if (std::fabs(E - prev) < conv) { res.converged = 1; break; }
    ++it;
    prev = E;
    for (int sp = 0; sp < 2; ++sp) {
      if (conv_method == "damping") {
        // Convergence_UHF_Damping (uhf.hpp:236-258), each spin on its own
        damper[sp].step(F[sp], D[sp], w.tmp.ptr, w.tmp2.ptr, w.part);
        continue;
      }
      double* e = w.tmp2.ptr;
      blas.gemm(false, false, n, F[sp], D[sp], w.tmp.ptr);
      blas.gemm(false, false, n, w.tmp.ptr, w.S.ptr, e);
      blas.gemm(false, false, n, w.S.ptr, D[sp], w.tmp.ptr);
      blas.gemm(false, false, n, w.tmp.ptr, F[sp], e, -1.0, 1.0);
      if (diis_x) {
        blas.gemm(false, false, n, w.X.ptr, e, w.tmp.ptr);
        blas.gemm(false, true, n, w.tmp.ptr, w.X.ptr, e);
      }
      diis[sp].step(blas.h, e, F[sp]);
    }
  }
  GANSU_HIP(hipDeviceSynchronize());
  res.e_elec = E;
  res.e_total = E + mol.nuclear_repulsion();
  res.iterations = it;
  res.Da = w.D.download();
  res.Db = Db.download();
  return res;
}

// ROHF (rohf.cu:59-300, rohf.hpp:330-360): n_beta closed and
// n_alpha - n_beta open orbitals; core guess; C from the unified Fock matrix;
// D_c = 2 C_c C_c^T, D_o = C_o C_o^T (gpu_kernels.cu:108-134); the engine's
// ROHF build with the coupling set rohf_parameter_name;
// E = 1/2 [D_c.(H + F_c) + D_o.H + 2 D_o.F_o] (gpu_manager.cu:871-882);
// DIIS on the unified matrix with the total density.
UhfResult rohf_run(const Molecule& mol, const Molecule* aux, const Params& prm) {
  const int n = mol.nbasis, nc = mol.n_beta, no = mol.n_alpha - mol.n_beta;
  if (no < 0 || nc + no > n) throw Error(GANSU_E_ELECTRONS, "invalid closed/open orbital counts");
  const std::string guess = prm.get("initial_guess");
  check_guess(guess);
  // InitialGuess_ROHF_SAD (rohf.hpp:600-640) forms the unified Fock matrix
  // S C R C^T S before anything has written C, so it has no defined result
  if (guess == "sad")
    throw Error(GANSU_E_UNSUPPORTED, "initial_guess=sad is not provided for ROHF (core and gwh only)");
  // rohf.cu:73-88: damping, optimal damping or DIIS
  const std::string conv_method = prm.get("convergence_method");
  if (conv_method != "diis" && conv_method != "damping" && conv_method != "optimaldamping")
    throw Error(GANSU_E_PARAM, "Invalid convergence_method: " + conv_method);
  RohfCoupling cp{};
  if (!rohf_coupling(prm.get("rohf_parameter_name"), &cp))
    throw Error(GANSU_E_PARAM, "unknown rohf_parameter_name: " + prm.get("rohf_parameter_name"));
  const int maxiter = prm.geti("maxiter");
  const double conv = prm.getd("convergence_energy_threshold");
  const int diis_size = prm.geti("diis_size");
  const bool diis_x = prm.geti("diis_include_transform") != 0;
  const double damp = prm.getd("damping_factor");

  Work w(n);
  Blas blas;
  const size_t nn = (size_t)n * n;
  compute_oneint(mol, w.S.ptr, w.T.ptr, w.V.ptr, 0);
  GANSU_HIP(launch_axpby(1.0, w.T.ptr, 1.0, w.V.ptr, nn, w.H.ptr, 0));
  std::map<std::string, std::string> opts{{"eri_method", prm.get("eri_method")},
                                          {"schwarz_screening_threshold", prm.get("schwarz_screening_threshold")},
                                          {"precision", prm.get("precision")},
                                          {"fp32_bound_threshold", prm.get("fp32_bound_threshold")},
                                          {"ri_k", prm.get("ri_k")}};
  const bool ri = prm.get("eri_method") == "ri";
  if (ri && !aux) throw Error(GANSU_E_ARG, "eri_method=ri needs auxiliary_gbsfilename");
  FockEngine eng(mol.prims.data(), mol.prims.size(), mol.types.data(), (int)mol.types.size(), mol.norm.data(), n,
                 ri ? aux->prims.data() : nullptr, ri ? aux->prims.size() : 0, ri ? aux->types.data() : nullptr,
                 ri ? (int)aux->types.size() : 0, ri ? aux->norm.data() : nullptr, ri ? aux->nbasis : 0, opts);
  eng.precompute();
  syevd(blas.h, n, w.S.ptr, w);
  GANSU_HIP(launch_transform(w.evec.ptr, w.eval.ptr, n, 1e-6, w.X.ptr, 0));

  DevBuf<double> Dc(nn), Do(nn), Fc(nn), Fo(nn);
  auto coefficients = [&](const double* Fm) {
    blas.gemm(true, false, n, w.X.ptr, Fm, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.X.ptr, w.tmp2.ptr);
    syevd(blas.h, n, w.tmp2.ptr, w);
    blas.gemm(false, true, n, w.X.ptr, w.evec.ptr, w.C.ptr);
  };
  auto density = [&]() {
    // the column-major view of row-major C is C^T: rows 0..nc-1 closed, nc.. open
    const double two = 2.0, one = 1.0, zero = 0.0;
    if (nc > 0)
      blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, n, n, nc, &two, w.C.ptr,
                               n, w.C.ptr, n, &zero, Dc.ptr, n),
                 "rocblas_dgemm(D_c)");
    else
      GANSU_HIP(launch_zero(Dc.ptr, nn, 0));
    if (no > 0)
      blas_check(rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none, n, n, no, &one,
                               w.C.ptr + nc, n, w.C.ptr + nc, n, &zero, Do.ptr, n),
                 "rocblas_dgemm(D_o)");
    else
      GANSU_HIP(launch_zero(Do.ptr, nn, 0));
    GANSU_HIP(launch_axpby(1.0, Dc.ptr, 1.0, Do.ptr, nn, w.D.ptr, 0));
  };
  UhfResult res;
  auto fock = [&]() {
    eng.build_rohf(Dc.ptr, Do.ptr, w.H.ptr, w.C.ptr, w.S.ptr, cp, nc, no, Fc.ptr, Fo.ptr, w.F.ptr, 0);
    ++res.fock_builds;
  };
  auto energy = [&]() {
    return 0.5 * (dot(Dc.ptr, w.H.ptr, nn, w.part) + dot(Dc.ptr, Fc.ptr, nn, w.part) + dot(Do.ptr, w.H.ptr, nn, w.part) +
                  2.0 * dot(Do.ptr, Fo.ptr, nn, w.part));
  };
  DevBuf<double> guess_F;
  if (guess == "gwh") {
    guess_F.alloc(nn);
    GANSU_HIP(launch_gwh_guess(w.H.ptr, w.S.ptr, n, guess_F.ptr, 0));
  }
  coefficients(guess == "gwh" ? guess_F.ptr : w.H.ptr);
  density();
  fock();
  Diis diis;
  Damper damper;
  if (conv_method == "diis") diis.init(diis_size, nn);
  else damper.init(nn, conv_method == "optimaldamping", damp);
  double prev = std::numeric_limits<double>::max(), E = 0.0;
  int it = 0;
  while (it < maxiter) {
    coefficients(w.F.ptr);
    density();
    fock();
    E = energy();
    if (std::fabs(E - prev) < conv) { res.converged = 1; break; }
    ++it;
    prev = E;
    if (conv_method != "diis") {
      // Convergence_ROHF_Damping (rohf.hpp:254-276) on the unified Fock
      // matrix and the total density
      damper.step(w.F.ptr, w.D.ptr, w.tmp.ptr, w.tmp2.ptr, w.part);
      continue;
    }
    double* e = w.tmp2.ptr;
    blas.gemm(false, false, n, w.F.ptr, w.D.ptr, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.S.ptr, e);
    blas.gemm(false, false, n, w.S.ptr, w.D.ptr, w.tmp.ptr);
    blas.gemm(false, false, n, w.tmp.ptr, w.F.ptr, e, -1.0, 1.0);
    if (diis_x) {
      blas.gemm(false, false, n, w.X.ptr, e, w.tmp.ptr);
      blas.gemm(false, true, n, w.tmp.ptr, w.X.ptr, e);
    }
    diis.step(blas.h, e, w.F.ptr);
  }
  GANSU_HIP(hipDeviceSynchronize());
  res.e_elec = E;
  res.e_total = E + mol.nuclear_repulsion();
  res.iterations = it;
  res.Da = Dc.download();  // closed-shell density
  res.Db = Do.download();  // open-shell density
  return res;
}

}  // namespace gansu
