// RI (density fitting) path of the Fock engine.
//
// Precompute (ERI_RI::precomputation, eri.cu:134-236, and
// compute_RI_IntermediateMatrixB, gpu_manager.cu:1448-1526):
//   aux Schwarz factors -> (P|Q) -> Cholesky (P|Q) = L L^T (rocSOLVER dpotrf,
//   gpu_manager.cu:1311-1349) -> (mn|P) -> B = L^-1 (P|mn) (rocBLAS dtrsm).
// B is kept as B[m][P][n]: for each m an Maux x M slab.  On that layout
//   J = sum_P (B^P . D) B^P            two HBM-bound passes (ri_kernels.hip)
//   K = sum_P B^P D B^P^T              T[m][P][l] = B[m][P][:] D  (batched DGEMM)
//                                      K = T_cat B_cat^T          (one DGEMM, k = Maux M)
// which is the reference's D-form K (gpu_manager.cu:1589-1618: 2 Maux small
// DGEMMs per build) as two large GEMMs, chunked over P to bound the workspace.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "ri_kernels.hpp"
#include "scf_kernels.hpp"

namespace gansu {

namespace {
void blas_ok(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) throw Error(GANSU_E_HIP, std::string(what) + ": rocblas status " + std::to_string(st));
}

double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// cgto norm x 1/sqrt((2l-1)!!(2m-1)!!(2n-1)!!) in the RI component order
double aux_component_factor(int L, int c) {
  static const int d2[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  static const int f3r[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1},
                                 {0, 2, 1}, {1, 2, 0}, {0, 1, 2}, {1, 0, 2}, {1, 1, 1}};
  int e[3];
  for (int d = 0; d < 3; ++d) {
    if (L == 0) e[d] = 0;
    else if (L == 1) e[d] = (c == d);
    else if (L == 2) e[d] = d2[c][d];
    else e[d] = f3r[c][d];
  }
  return 1.0 / std::sqrt(dfact(2 * e[0] - 1) * dfact(2 * e[1] - 1) * dfact(2 * e[2] - 1));
}
}  // namespace

struct FockEngine::RIState {
  rocblas_handle h = nullptr;
  std::vector<AuxShell> shells_h;
  DevBuf<AuxShell> shells;
  DevBuf<AuxPrim> prims;
  DevBuf<double> waux, L, B, W, J, K, T;
  DevBuf<double> Kpart;
  // precision=mixed: incremental builds (ri_mfma.hip).  B as fp16 hi/lo
  // planes (B sB = hi + lo, the size of an FP32 copy), the previous build's
  // factor X_prev, K, J, W, and T_A, T_E (M Maux nocc each) with
  // B X_prev = T_A + T_E / 2.
  DevBuf<uint16_t> Bp;
  float sB = 1.0f;
  int rik_nocc = 0;   // nocc of the held state (0: none)
  int rik_since = 0;  // incremental builds since the last FP64 build
  double rik_drift = 0.0;  // sum of |E| / |X| over them
  int rik_incremental = 0, rik_fresh = 0, rik_full = 0;  // counts, for the caller's statistics
  DevBuf<float> TA, TE, Zpart;
  DevBuf<double> Xprev, Xn, Ework, Y, Y2, G, Kprev, Jprev, Wprev, dW, dWrow, nrm, Zsum;
  DevBuf<uint16_t> Ehi, Elo;
  DevBuf<RikState> rst;
  DevBuf<unsigned> wgmax;  // per-workgroup maxima of the T update
  double* nrm_host = nullptr;
  std::vector<int> aux_atom_h;  // atom of each auxiliary shell
  DevBuf<int> aux_atom;
  DevBuf<int> nanflag;       // NaN check of F (device flag, pinned host copy)
  int* nan_host = nullptr;
  int pchunk = 0;
  ~RIState() {
    if (h) rocblas_destroy_handle(h);
    if (nan_host) (void)hipHostFree(nan_host);
    if (nrm_host) (void)hipHostFree(nrm_host);
  }
};

void ri_destroy(FockEngine::RIState* p) { delete p; }

// Output blocking of K = A^T A (A = kdim x M column-major): nb x nb blocks of
// side bs, of which the upper ones (bi <= bj) are computed; each block's k is
// split so that one product has about 1024 64x64 output tiles in flight.
static void ri_k_blocking(int M, int* nb, int* bs) {
  *nb = M >= 1024 ? (M + 799) / 800 : 1;
  *bs = (M + *nb - 1) / *nb;
}

// Split-k depth of one K block: about 4608 64x64 output tiles in flight
// (the three 768 x 768 x 2e6 upper blocks of 64 waters: 161 ms with 7 parts,
// 150 ms with 32).
static int ri_k_max_split(int nr, int nc) {
  const int tiles = ((nr + 63) / 64) * ((nc + 63) / 64);
  return std::min(64, std::max(1, 4608 / tiles));
}

void FockEngine::precompute_ri() {
  // GANSU_RI_TRACE=1: wall time of each precompute phase (device-synchronised)
  auto t_last = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (!rik_trace_) return;
    GANSU_HIP(hipDeviceSynchronize());
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[ri] precompute %s: %.1f ms\n", what,
                 1e3 * std::chrono::duration<double>(now - t_last).count());
    t_last = now;
  };
  // A repeated precompute (a new geometry) keeps the large tensors of the
  // previous state when their sizes match: reallocating a freed buffer of
  // this size costs about 2 s (the driver clears the memory again), the
  // recomputation fills every element anyway.
  RIState* old = ri_;
  ri_ = new RIState();
  RIState& r = *ri_;
  if (old) {
    r.B = std::move(old->B);
    r.Bp = std::move(old->Bp);
    r.TA = std::move(old->TA);
    r.TE = std::move(old->TE);
    r.T = std::move(old->T);
    r.Zpart = std::move(old->Zpart);
    ri_destroy(old);
  }
  blas_ok(rocblas_create_handle(&r.h), "rocblas_create_handle");
  mark("handle");
  const int M = M_, Ma = Maux_;
  // auxiliary shells and primitive Schwarz factors (schwarz.txt:479+)
  const auto ash = contracted_shells(aux_prims_.data(), aux_prims_.size(), kMaxLAux);
  DevBuf<gansu_primitive_shell> dprims;
  dprims.upload(aux_prims_);
  std::vector<int> pa(aux_prims_.size()), pb(aux_prims_.size(), -1);
  for (size_t i = 0; i < pa.size(); ++i) pa[i] = (int)i;
  std::vector<double> qaux;
  {
    DevBuf<int> dpa, dpb;
    DevBuf<double> dq(aux_prims_.size());
    dpa.upload(pa);
    dpb.upload(pb);
    GANSU_HIP(launch_schwarz_ref(dprims.ptr, dpa.ptr, dpb.ptr, pa.size(), dq.ptr, 0));
    GANSU_HIP(hipDeviceSynchronize());
    qaux = dq.download();
  }
  std::vector<AuxPrim> aprims;
  std::vector<double> waux(Ma, 0.0);
  for (const auto& s : ash) {
    AuxShell as{};
    as.L = s.L;
    as.basis = s.basis;
    as.prim_begin = (int)aprims.size();
    as.nprim = (int)s.prims.size();
    double qmax = 0.0;
    for (int ip : s.prims) {
      const auto& p = aux_prims_[ip];
      AuxPrim ap{};
      ap.alpha = p.exponent;
      ap.C[0] = p.x; ap.C[1] = p.y; ap.C[2] = p.z;
      ap.Kc = p.coefficient * std::pow(2.0 / M_PI, 0.75) * std::pow(2.0, s.L) *
              std::pow(p.exponent, (2.0 * s.L + 3.0) / 4.0);
      ap.Q = qaux[ip];
      qmax = std::max(qmax, ap.Q);
      aprims.push_back(ap);
    }
    as.qmax = qmax;
    r.shells_h.push_back(as);
    r.aux_atom_h.push_back(s.atom);
    if (s.basis + ncart(s.L) > Ma) throw Error(GANSU_E_ARG, "aux basis_index out of range");
    for (int c = 0; c < ncart(s.L); ++c) waux[s.basis + c] = aux_norm_[s.basis + c] * aux_component_factor(s.L, c);
  }
  r.shells.upload(r.shells_h);
  r.aux_atom.upload(r.aux_atom_h);
  r.prims.upload(aprims);
  r.waux.upload(waux);

  // (P|Q) and its Cholesky factor.  Row-major L (lower) is the column-major
  // upper factor of dpotrf (the reference's convention, gpu_manager.cu:1311-1349).
  mark("aux shells");
  r.L.alloc((size_t)Ma * Ma);
  GANSU_HIP(launch_zero(r.L.ptr, r.L.n, 0));
  const size_t nsh = r.shells_h.size();
  RI2cArgs a2{r.shells.ptr, r.prims.ptr, nsh * (nsh + 1) / 2, r.waux.ptr, Ma, thr_, boys_.ptr, r.L.ptr};
  GANSU_HIP(launch_int2c2e(a2, 0));
  DevBuf<int> info(1);
  blas_ok(rocsolver_dpotrf(r.h, rocblas_fill_upper, Ma, r.L.ptr, Ma, info.ptr), "rocsolver_dpotrf");
  int hinfo = 0;
  GANSU_HIP(hipMemcpy(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost));
  if (hinfo != 0) throw Error(GANSU_E_NUMERIC, "(P|Q) is not positive definite (potrf info " + std::to_string(hinfo) + ")");
  mark("metric + Cholesky");

  // (mn|P) into B[m][P][n], then B_m = L^-1 T_m for every m
  if (r.B.n != (size_t)M * Ma * M) r.B.alloc((size_t)M * Ma * M);
  mark("B allocation");
  GANSU_HIP(launch_zero(r.B.ptr, r.B.n, 0));
  mark("B fill");
  DevBuf<double> dw;
  dw.upload(w_h_);
  RI3cArgs a3{d_sp_.ptr, d_pr_.ptr, (int)sp_h_.size(), r.shells.ptr, r.prims.ptr, (int)nsh, dw.ptr, r.waux.ptr,
              M, Ma, thr_, boys_.ptr, r.B.ptr};
  GANSU_HIP(launch_int3c2e(a3, 0));
  mark("3-centre integrals");
  const double one = 1.0;
  // column-major view of slab m is T_m^T (M x Maux, ld M): T_m^T <- T_m^T U^-1,
  // U = L^T, row n by row n.  (mn|P) is symmetric in m and n, so slab m is
  // solved only for its rows n >= m0, the first m of its block of kSymBlock
  // slabs, and the rows below are copied from the slabs solved for them:
  // 0.54 of the solve of all M^2 rows at M = 1600.
  constexpr int kSymBlock = 128;
  for (int m0 = 0; m0 < M; m0 += kSymBlock) {
    const int cnt = std::min(kSymBlock, M - m0);
    blas_ok(rocblas_dtrsm_strided_batched(r.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                          rocblas_diagonal_non_unit, M - m0, Ma, &one, r.L.ptr, Ma, 0,
                                          r.B.ptr + (size_t)m0 * Ma * M + m0, M, (rocblas_stride)Ma * M, cnt),
            "rocblas_dtrsm_strided_batched");
  }
  mark("B = L^-1 (mn|P), rows n >= m block");
  GANSU_HIP(launch_ri_mirror(r.B.ptr, M, Ma, kSymBlock, 0));
  mark("B mirror (n < m block)");
  if (mixed_) {
    const int kpad = rik_kpad(M);
    const size_t np = (size_t)M * Ma * kpad * 2;
    if (r.Bp.n != np) r.Bp.alloc(np);
    mark("B planes allocation");
    DevBuf<unsigned> mx(1);
    GANSU_HIP(launch_rik_absmax(r.B.ptr, r.B.n, mx.ptr, 0));
    unsigned bits = 0;
    GANSU_HIP(hipMemcpy(&bits, mx.ptr, sizeof bits, hipMemcpyDeviceToHost));
    float bmax;
    std::memcpy(&bmax, &bits, sizeof bmax);
    int e = 0;
    if (bmax > 0.0f) std::frexp(bmax, &e);
    r.sB = bmax > 0.0f ? std::ldexp(1.0f, 14 - e) : 1.0f;
    GANSU_HIP(launch_rik_planes(r.B.ptr, (long)M * Ma, M, kpad, r.sB, r.Bp.ptr, 0));
    mark("B planes (fp16 hi/lo)");
    r.Wprev.alloc(Ma);
    r.dW.alloc(Ma);
    r.Jprev.alloc((size_t)M * M);
    r.Kprev.alloc((size_t)M * M);
    r.rst.alloc(1);
    r.nrm.alloc(4);
    GANSU_HIP(hipHostMalloc((void**)&r.nrm_host, 4 * sizeof(double), hipHostMallocDefault));
  }
  r.W.alloc(Ma);
  r.J.alloc((size_t)M * M);
  r.K.alloc((size_t)M * M);
  r.nanflag.alloc(1);
  GANSU_HIP(hipHostMalloc((void**)&r.nan_host, sizeof(int), hipHostMallocDefault));
  // K workspace: a P-chunk of T[m][P][l], at most ~4 GB
  const size_t budget = size_t(4) << 30;
  r.pchunk = (int)std::max<size_t>(1, std::min<size_t>(Ma, budget / ((size_t)M * M * sizeof(double))));
  if (r.T.n != (size_t)M * r.pchunk * M) r.T.alloc((size_t)M * r.pchunk * M);
  {
    int nb, bs;
    ri_k_blocking(M, &nb, &bs);
    r.Kpart.alloc((size_t)(ri_k_max_split(bs, bs) + 1) * bs * bs);
  }
  GANSU_HIP(hipDeviceSynchronize());
  mark("work buffers");
}

void FockEngine::build_ri(const double* d_D, double* d_J, double* d_K, const double* d_H, double* d_F,
                          hipStream_t s, const double* d_C, int nocc) {
  RIState& r = *ri_;
  const int M = M_, Ma = Maux_;
  blas_ok(rocblas_set_stream(r.h, s), "rocblas_set_stream");
  prof_clear();
  double* J = d_J ? d_J : r.J.ptr;
  double* K = d_K ? d_K : r.K.ptr;
  const double bytesB = double(M) * Ma * M * sizeof(double);
  const bool occ = d_C && (ri_k_occ_ || mixed_);
  if (mixed_ && occ && rik_blocks(nocc) <= kRikMaxBlocks && build_ri_incremental(d_C, nocc, J, K, s)) {
    // J, K and W formed incrementally from the previous build
  } else {
    if (mixed_) ri_->rik_nocc = 0;  // a build the increments cannot follow
    if (occ) {
      // K and W from the occupied-orbital intermediate T = B X
      const bool seed = mixed_ && rik_blocks(nocc) <= kRikMaxBlocks;
      if (seed) rik_prepare(nocc);
      build_ri_k_occ(d_C, nocc, K, s, seed);
    } else {
      prof_begin(s, 80001, bytesB, 2.0 * M * Ma * M, (double)Ma);
      GANSU_HIP(launch_ri_w(r.B.ptr, d_D, M, Ma, r.W.ptr, s));
      prof_end(s);
    }
    // J reads the rows of B up to the diagonal (B is symmetric in m, n)
    prof_begin(s, 80002, 0.5 * double(M) * (M + 1) * Ma * sizeof(double), 2.0 * M * Ma * M, (double)M);
    GANSU_HIP(launch_ri_j(r.B.ptr, r.W.ptr, M, Ma, J, s));
    prof_end(s);
    if (mixed_ && occ && r.rik_nocc == nocc) {
      // the state the next increments start from
      GANSU_HIP(hipMemcpyAsync(r.Kprev.ptr, K, sizeof(double) * M * M, hipMemcpyDeviceToDevice, s));
      GANSU_HIP(hipMemcpyAsync(r.Jprev.ptr, J, sizeof(double) * M * M, hipMemcpyDeviceToDevice, s));
      GANSU_HIP(hipMemcpyAsync(r.Wprev.ptr, r.W.ptr, sizeof(double) * Ma, hipMemcpyDeviceToDevice, s));
      GANSU_HIP(launch_rik_xocc(d_C, M, nocc, r.Xprev.ptr, s));
      r.rik_since = 0;
      r.rik_drift = 0.0;
      ++r.rik_full;
    }
  }
  const double one = 1.0, zero = 0.0;
  if (!occ) {
  for (int p0 = 0; p0 < Ma; p0 += r.pchunk) {
    const int pc = std::min(r.pchunk, Ma - p0);
    // T_m (pc x M) = B_m[p0:p0+pc] D  for every m; column-major: T_m^T = D B_m^T
    prof_begin(s, 80003, 0.0, 2.0 * M * (double)pc * M * M, (double)pc);
    blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_none, rocblas_operation_none, M, pc, M, &one, d_D, M,
                                          0, r.B.ptr + (size_t)p0 * M, M, (rocblas_stride)Ma * M, &zero, r.T.ptr, M,
                                          (rocblas_stride)pc * M, M),
            "rocblas_dgemm_strided_batched(T)");
    prof_end(s);
    // K += T_cat B_cat^T (row-major); column-major: K^T = B_cat T_cat^T
    prof_begin(s, 80004, 0.0, 2.0 * M * (double)M * pc * M, (double)pc);
    const double beta = p0 == 0 ? 0.0 : 1.0;
    blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, M, M, pc * M, &one,
                          r.B.ptr + (size_t)p0 * M, Ma * M, r.T.ptr, pc * M, &beta, K, M),
            "rocblas_dgemm(K)");
    prof_end(s);
  }
  }
  if (d_F) {
    GANSU_HIP(launch_fock_ri(d_H, J, K, (size_t)M * M, d_F, s));
    if (nan_check_) {
      // ERI_RI_RHF::compute_fock_matrix reads F back and fails on a NaN
      // (rhf.hpp:665-674); here one flag word comes back instead of F
      GANSU_HIP(launch_set_int(r.nanflag.ptr, 0, s));
      GANSU_HIP(launch_has_nan(d_F, (size_t)M * M, r.nanflag.ptr, s));
      GANSU_HIP(hipMemcpyAsync(r.nan_host, r.nanflag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
      GANSU_HIP(hipStreamSynchronize(s));
      if (*r.nan_host) throw Error(GANSU_E_NUMERIC, "NaN detected in the Fock matrix");
    }
  }
}

// ---- precision=mixed: incremental builds --------------------------------
// State for nocc occupied orbitals: T_A, T_E (M Maux nocc FP32 each), X_prev,
// the fp16 split of E, the split-k partials of dK.  Allocated on the first
// build with this nocc (the one allocation of the mixed path; later builds
// reuse it).
void FockEngine::rik_prepare(int nocc) {
  RIState& r = *ri_;
  const int M = M_, Ma = Maux_;
  const size_t tn = (size_t)M * Ma * nocc;
  if (r.dWrow.n != (size_t)M * Ma || r.TA.n != tn) {
    if (r.TA.n != tn) {
      r.TA.alloc(tn);
      r.TE.alloc(tn);
    }
    r.dWrow.alloc((size_t)M * Ma);
    r.wgmax.alloc(2 * rik_tstep_workgroups((long)M * Ma));
    r.Xprev.alloc((size_t)M * nocc);
    r.Xn.alloc((size_t)M * nocc);
    r.Ework.alloc((size_t)M * nocc);
    r.Y.alloc((size_t)nocc * nocc);
    r.Y2.alloc((size_t)nocc * nocc);
    r.G.alloc((size_t)nocc * nocc);
    const int kpad = rik_kpad(M);
    r.Ehi.alloc((size_t)rik_cols(nocc) * kpad);
    r.Elo.alloc((size_t)rik_cols(nocc) * kpad);
    const size_t zn = (size_t)rik_chunks(M, (long)Ma * nocc) * M * M;
    if (r.Zpart.n != zn) r.Zpart.alloc(zn);
    if (r.Zsum.n != (size_t)M * M) r.Zsum.alloc((size_t)M * M);
  }
  r.rik_nocc = nocc;
}

// split-k depth of the dK product: enough workgroups to fill the chip
// several times over, and at most 4096 k per chunk, so that each FP32
// partial sums no more than 256 MFMA steps before the FP64 reduction
// (a two-level FP32 sum inside the kernel was as accurate and, at one wave
// per SIMD, half as fast)
int FockEngine::rik_chunks(int M, long Kd) {
  const long nt2 = rik_dk_tiles(M);
  const long fill = (4096 + nt2 - 1) / nt2;
  const long acc = (Kd + 4095) / 4096;
  return (int)std::max<long>(1, std::min<long>({std::max(fill, acc), 1024, std::max<long>(1, Kd / 32)}));
}

// One incremental or fresh build (ri_mfma.hip): align the new occupied factor to the
// previous one (orthogonal Procrustes by Newton-Schulz), E = X_n R - X_prev,
// then T_E = B E with the T_A / T_E update and dW, J = J_prev + B dW,
// K = K_prev + Z + Z^T with Z = T_A T_E^T.  Returns false (nothing changed)
// when a full FP64 build is due: no state for this nocc, `rik_refresh`
// increments since the last one, an alignment that did not converge or a
// change |E| / |X| above `rik_accept`, or a small change after increments
// whose summed |E| / |X| exceeds `rik_tol`.  The rounding error of an
// increment is proportional to its |E|, so the sum bounds the error K
// carries between FP64 builds.
bool FockEngine::build_ri_incremental(const double* d_C, int nocc, double* J, double* K, hipStream_t s) {
  RIState& r = *ri_;
  const int M = M_, Ma = Maux_;
  if (r.rik_nocc != nocc || r.rik_since >= rik_refresh_) return false;
  const double one = 1.0, zero = 0.0, mone = -1.0;
  const size_t mn = (size_t)M * nocc;
  prof_begin(s, 80020, 0.0, 0.0, 1.0);
  GANSU_HIP(launch_rik_xocc(d_C, M, nocc, r.Xn.ptr, s));
  // column-major views: X (M x nocc row-major) is X^T (nocc x M, ld nocc)
  // A = X_n^T X_prev, scaled by 1/|A|_F into Y
  blas_ok(rocblas_dgemm(r.h, rocblas_operation_none, rocblas_operation_transpose, nocc, nocc, M, &one, r.Xn.ptr,
                        nocc, r.Xprev.ptr, nocc, &zero, r.Y.ptr, nocc),
          "rocblas_dgemm(A)");
  double fro = 0.0;
  blas_ok(rocblas_set_pointer_mode(r.h, rocblas_pointer_mode_host), "rocblas_set_pointer_mode");
  blas_ok(rocblas_dnrm2(r.h, nocc * nocc, r.Y.ptr, 1, &fro), "rocblas_dnrm2");
  if (!(fro > 0.0)) {
    prof_end(s);
    return false;
  }
  const double inv = 1.0 / fro;
  blas_ok(rocblas_dscal(r.h, nocc * nocc, &inv, r.Y.ptr, 1), "rocblas_dscal");
  // Newton-Schulz: Y <- 1.5 Y - 0.5 Y (Y^T Y) converges to the orthogonal
  // polar factor R of A (singular values grow by 1.5 per step while small)
  const double c15 = 1.5, cm05 = -0.5;
  for (int it = 0; it < rik_ns_iters_; ++it) {
    blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, nocc, nocc, nocc, &one, r.Y.ptr,
                          nocc, r.Y.ptr, nocc, &zero, r.G.ptr, nocc),
            "rocblas_dgemm(NS G)");
    GANSU_HIP(hipMemcpyAsync(r.Y2.ptr, r.Y.ptr, sizeof(double) * nocc * nocc, hipMemcpyDeviceToDevice, s));
    blas_ok(rocblas_dgemm(r.h, rocblas_operation_none, rocblas_operation_none, nocc, nocc, nocc, &cm05, r.Y.ptr, nocc,
                          r.G.ptr, nocc, &c15, r.Y2.ptr, nocc),
            "rocblas_dgemm(NS Y)");
    std::swap(r.Y.ptr, r.Y2.ptr);
  }
  // orthogonality residual of R, then E = X_n R - X_prev (column-major
  // E^T = R^T X_n^T - X_prev^T)
  blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, nocc, nocc, nocc, &one, r.Y.ptr,
                        nocc, r.Y.ptr, nocc, &zero, r.G.ptr, nocc),
          "rocblas_dgemm(R^T R)");
  GANSU_HIP(hipMemcpyAsync(r.Ework.ptr, r.Xprev.ptr, sizeof(double) * mn, hipMemcpyDeviceToDevice, s));
  blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, nocc, M, nocc, &one, r.Y.ptr, nocc,
                        r.Xn.ptr, nocc, &mone, r.Ework.ptr, nocc),
          "rocblas_dgemm(E)");
  GANSU_HIP(launch_rik_norms(r.Ework.ptr, r.Xprev.ptr, mn, r.G.ptr, nocc, r.nrm.ptr, s));
  GANSU_HIP(hipMemcpyAsync(r.nrm_host, r.nrm.ptr, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  GANSU_HIP(hipStreamSynchronize(s));
  prof_end(s);
  const double rel = std::sqrt(r.nrm_host[0] / std::max(r.nrm_host[1], 1e-300));
  last_rik_rel_ = rel;
  // an increment is taken when the alignment converged and the change is at
  // most rik_accept_; otherwise the build starts afresh on the same kernels
  // (X_prev = 0, E = X_n: K = T T^T at the accuracy of the split products),
  // counted as a change of 1.  Once the changes summed since the last FP64
  // build exceed rik_tol_, the first build whose own change is below
  // rik_tol_ / 4 is an FP64 one, so that the error of the early, large
  // changes does not stay in K while the SCF converges.
  const bool aligned = std::sqrt(r.nrm_host[2]) < 1e-10 && rel <= rik_accept_;
  const double step = aligned ? rel : 1.0;
  const bool wipe = r.rik_drift + step > rik_tol_ && step < 0.25 * rik_tol_;
  if (rik_trace_)
    std::fprintf(stderr, "[ri] |E|/|X| = %.3e, drift %.3e, orthogonality %.1e -> %s\n", rel, r.rik_drift + step,
                 std::sqrt(r.nrm_host[2]), wipe ? "FP64" : (aligned ? "incremental" : "fresh (fp16 split)"));
  if (wipe) return false;
  r.rik_drift += step;
  const bool fresh = !aligned;
  if (fresh) {
    // E = X_n, and the previous J, W, K are zero
    GANSU_HIP(hipMemcpyAsync(r.Ework.ptr, r.Xn.ptr, sizeof(double) * mn, hipMemcpyDeviceToDevice, s));
    GANSU_HIP(launch_zero(r.Xprev.ptr, mn, s));
    GANSU_HIP(launch_zero(r.Jprev.ptr, (size_t)M * M, s));
    GANSU_HIP(launch_zero(r.Wprev.ptr, (size_t)Ma, s));
  }

  const int kpad = rik_kpad(M);
  prof_begin(s, 80021, 0.0, 0.0, 1.0);
  GANSU_HIP(launch_rik_split_e(r.Ework.ptr, M, nocc, kpad, r.rst.ptr, r.Ehi.ptr, r.Elo.ptr, s));
  prof_end(s);
  // T_E = B E (+ T_A update, dW rows): 2 R M nocc flops, three fp16 products each
  const long R = (long)M * Ma;
  RikTArgs ta{r.Bp.ptr, R, M, Ma, nocc, kpad, r.sB, r.Ehi.ptr, r.Elo.ptr, r.Ework.ptr, r.rst.ptr,
              r.TA.ptr, r.TE.ptr, r.dWrow.ptr, fresh ? 1 : 0, r.wgmax.ptr, 0, 0};
  prof_begin(s, 80022, double(R) * kpad * 4 + double(R) * nocc * (fresh ? 8 : 16), 2.0 * R * M * nocc * 3.0,
             (double)R);
  GANSU_HIP(launch_rik_tstep(ta, s));
  prof_end(s);
  prof_begin(s, 80023, double(M) * Ma * 8, 0.0, (double)Ma);
  GANSU_HIP(launch_rik_dw(r.dWrow.ptr, M, Ma, r.dW.ptr, r.W.ptr, r.Wprev.ptr, s));
  prof_end(s);
  double jgroups = 0.0;  // 8-column groups the J increment reads per P (rows to their diagonal group)
  for (int m = 0; m < M; ++m) jgroups += std::min(kpad / 8, (m >> 3) + 1);
  prof_begin(s, 80024, jgroups * Ma * 32.0, 2.0 * M * Ma * M, (double)M);
  GANSU_HIP(launch_rik_j(r.Bp.ptr, r.sB, kpad, r.dW.ptr, M, Ma, r.Jprev.ptr, J, s));
  prof_end(s);
  // Z = T_A T_E^T over Kd = Maux nocc: 2 M^2 Kd flops, three fp16 products each
  const long Kd = (long)Ma * nocc;
  const int nch = rik_chunks(M, Kd);
  const long kchunk = ((Kd + nch - 1) / nch + 31) / 32 * 32;
  RikDkArgs da{r.TA.ptr, r.TE.ptr, M, Kd, kchunk, nch, r.rst.ptr, r.Zpart.ptr};
  prof_begin(s, 80025, 2.0 * double(M) * Kd * 4, 2.0 * M * (double)M * Kd * 3.0, (double)nch);
  GANSU_HIP(launch_rik_dk(da, nch, s));
  prof_end(s);
  prof_begin(s, 80026, 0.0, 0.0, 1.0);
  GANSU_HIP(launch_rik_reduce(r.Zpart.ptr, nch, M, r.Zsum.ptr, fresh ? nullptr : r.Kprev.ptr, K, s));
  GANSU_HIP(hipMemcpyAsync(r.Kprev.ptr, K, sizeof(double) * M * M, hipMemcpyDeviceToDevice, s));
  // X_prev <- X_prev + E = X_n R, the factor that B X_prev = T_A + T_E / 2 now holds
  blas_ok(rocblas_daxpy(r.h, (int)mn, &one, r.Ework.ptr, 1, r.Xprev.ptr, 1), "rocblas_daxpy(X)");
  prof_end(s);
  ++r.rik_since;
  ++(fresh ? r.rik_fresh : r.rik_incremental);
  return true;
}

int FockEngine::ri_incremental_builds() const { return ri_ ? ri_->rik_incremental : 0; }

// K = sum_P (B^P X)(B^P X)^T with X = sqrt(2) C_occ (M x nocc, row-major,
// leading dimension M).  Per P-chunk: T_m = B_m[p0:p0+pc] X for every m
// (batched DGEMM), then K += T_cat T_cat^T with T_cat = M x (pc nocc).
void FockEngine::build_ri_k_occ(const double* d_C, int nocc, double* K, hipStream_t s, bool seed) {
  RIState& r = *ri_;
  const int M = M_, Ma = Maux_;
  const int pc_max = (int)std::max<size_t>(1, std::min<size_t>(Ma, r.T.n / ((size_t)M * nocc)));
  const double two = std::sqrt(2.0), zero = 0.0;
  for (int p0 = 0; p0 < Ma; p0 += pc_max) {
    const int pc = std::min(pc_max, Ma - p0);
    // column-major: T_m^T (nocc x pc) = X^T (nocc x M) B_m^T (M x pc); X^T is the
    // column-major view of row-major C restricted to its first nocc columns
    prof_begin(s, 80005, 0.0, 2.0 * M * (double)pc * M * nocc, (double)pc);
    blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_none, rocblas_operation_none, nocc, pc, M, &two, d_C,
                                          M, 0, r.B.ptr + (size_t)p0 * M, M, (rocblas_stride)Ma * M, &zero, r.T.ptr,
                                          nocc, (rocblas_stride)pc * nocc, M),
            "rocblas_dgemm_strided_batched(T occ)");
    prof_end(s);
    prof_begin(s, 80007, 0.0, 2.0 * M * (double)pc * nocc, (double)pc);
    GANSU_HIP(launch_ri_w_from_t(r.T.ptr, d_C, M, pc, nocc, p0, r.W.ptr, s));
    prof_end(s);
    if (seed) GANSU_HIP(launch_rik_seed(r.T.ptr, M, pc, nocc, p0, Ma, r.TA.ptr, r.TE.ptr, s));
    ri_k_product(r.T.ptr, pc * nocc, p0 > 0, K, s);
  }
  GANSU_HIP(launch_mirror_upper(K, M, s));
}

// K (+)= A^T A into the column-major upper block triangle of K (the diagonal
// blocks in full); the caller mirrors.  A = T_cat, kdim x M, leading dim kdim.
void FockEngine::ri_k_product(const double* A, int kdim, bool accumulate, double* K, hipStream_t s) {
  RIState& r = *ri_;
  const int M = M_;
  int nb, bs;
  ri_k_blocking(M, &nb, &bs);
  const double one = 1.0, zero = 0.0;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = bi; bj < nb; ++bj) {
      const int r0 = bi * bs, c0 = bj * bs;
      const int nr = std::min(bs, M - r0), nc = std::min(bs, M - c0);
      if (nr <= 0 || nc <= 0) continue;
      double* Kb = K + r0 + (size_t)c0 * M;
      const int nsplit = std::min(ri_k_max_split(nr, nc), std::max(1, kdim / 2048));
      prof_begin(s, 80008, 0.0, 2.0 * nr * (double)nc * kdim, 1.0);
      if (nsplit <= 1) {
        const double beta = accumulate ? 1.0 : 0.0;
        blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, nr, nc, kdim, &one,
                              A + (size_t)r0 * kdim, kdim, A + (size_t)c0 * kdim, kdim, &beta, Kb, M),
                "rocblas_dgemm(K block)");
      } else {
        const int kb = kdim / nsplit;  // the remainder goes into one more part
        const size_t blk = (size_t)nr * nc;
        blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_transpose, rocblas_operation_none, nr, nc, kb,
                                              &one, A + (size_t)r0 * kdim, kdim, kb, A + (size_t)c0 * kdim, kdim, kb,
                                              &zero, r.Kpart.ptr, nr, (rocblas_stride)blk, nsplit),
                "rocblas_dgemm_strided_batched(K split)");
        int nparts = nsplit;
        const int rem = kdim - kb * nsplit;
        if (rem > 0) {
          const size_t ko = (size_t)kb * nsplit;
          blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, nr, nc, rem, &one,
                                A + (size_t)r0 * kdim + ko, kdim, A + (size_t)c0 * kdim + ko, kdim, &zero,
                                r.Kpart.ptr + (size_t)nsplit * blk, nr),
                  "rocblas_dgemm(K split rest)");
          nparts += 1;
        }
        GANSU_HIP(launch_sum_parts(r.Kpart.ptr, nparts, nr, nc, accumulate ? 1 : 0, Kb, M, s));
      }
      prof_end(s);
    }
}

// Gradient of the RI-JK two-electron energy 1/2 D.J_RI - 1/4 D.K_RI
// (B = L^-1 A, A^P_mn = (mn|P), V = (P|Q) = L L^T, column-major U = L^T):
//   W = L^-1 gamma, c = V^-1 gamma = U^-1 W,
//   S^P = X^T B^P X (X = sqrt2 C_occ, D = X X^T), N_RS = tr(S^R S^S),
//   St = S U^-T along P (St^P = sum_Q (L^-T)_PQ S^Q),
//   Gam = -1/2 c c^T + 1/4 U^-1 N U^-T         -> d(P|Q) contraction,
//   G^P = c_P D - 1/2 X St^P X^T  (P-chunks)   -> d(mn|P) contraction,
// which equal the oracle's G^P = c_P D - 1/2 D V^-1 A^P D and
// Gam = -1/2 c c^T + 1/4 V^-1 M V^-1 with M_PQ = tr(A^P D A^Q D).
void FockEngine::gradient_ri(const double* d_D, const double* d_C, int nocc, int natoms, double* d_grad,
                             hipStream_t s) {
  RIState& r = *ri_;
  const int M = M_, Ma = Maux_;
  if (!d_C || nocc <= 0 || nocc > M) throw Error(GANSU_E_ARG, "the RI gradient needs the occupied orbitals (C, nocc)");
  blas_ok(rocblas_set_stream(r.h, s), "rocblas_set_stream");
  if (!d_alpha_.ptr) d_alpha_.upload(pair_alpha_h_);
  if (!d_shell_atom_.ptr) {
    std::vector<int> sa(shells_.size());
    for (size_t i = 0; i < shells_.size(); ++i) sa[i] = shells_[i].atom;
    d_shell_atom_.upload(sa);
  }
  for (int a : r.aux_atom_h)
    if (a < 0 || a >= natoms) throw Error(GANSU_E_ARG, "auxiliary basis on an atom beyond n_atoms");
  GANSU_HIP(launch_zero(d_grad, 3 * (size_t)natoms, s));
  const double one = 1.0, zero = 0.0, two = std::sqrt(2.0);
  // W and c
  GANSU_HIP(launch_ri_w(r.B.ptr, d_D, M, Ma, r.W.ptr, s));
  DevBuf<double> c(Ma);
  GANSU_HIP(hipMemcpyAsync(c.ptr, r.W.ptr, sizeof(double) * Ma, hipMemcpyDeviceToDevice, s));
  blas_ok(rocblas_dtrsm(r.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                        Ma, 1, &one, r.L.ptr, Ma, c.ptr, Ma),
          "rocblas_dtrsm(c)");
  // S^P for all P, through T = B X in P-chunks
  const size_t no2 = (size_t)nocc * nocc;
  DevBuf<double> S(no2 * Ma);
  const int pc_max = (int)std::max<size_t>(1, std::min<size_t>(Ma, r.T.n / ((size_t)M * nocc)));
  for (int p0 = 0; p0 < Ma; p0 += pc_max) {
    const int pc = std::min(pc_max, Ma - p0);
    blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_none, rocblas_operation_none, nocc, pc, M, &two, d_C,
                                          M, 0, r.B.ptr + (size_t)p0 * M, M, (rocblas_stride)Ma * M, &zero, r.T.ptr,
                                          nocc, (rocblas_stride)pc * nocc, M),
            "rocblas_dgemm_strided_batched(T grad)");
    blas_ok(rocblas_dgemm(r.h, rocblas_operation_none, rocblas_operation_transpose, nocc, pc * nocc, M, &two, d_C, M,
                          r.T.ptr, pc * nocc, &zero, S.ptr + (size_t)p0 * no2, nocc),
            "rocblas_dgemm(S)");
  }
  // N = S_flat^T S_flat, Gam = 1/4 U^-1 N U^-T - 1/2 c c^T
  DevBuf<double> N((size_t)Ma * Ma);
  blas_ok(rocblas_dgemm(r.h, rocblas_operation_transpose, rocblas_operation_none, Ma, Ma, (int)no2, &one, S.ptr,
                        (int)no2, S.ptr, (int)no2, &zero, N.ptr, Ma),
          "rocblas_dgemm(N)");
  blas_ok(rocblas_dtrsm(r.h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                        Ma, Ma, &one, r.L.ptr, Ma, N.ptr, Ma),
          "rocblas_dtrsm(N left)");
  blas_ok(rocblas_dtrsm(r.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, Ma, Ma, &one, r.L.ptr, Ma, N.ptr, Ma),
          "rocblas_dtrsm(N right)");
  const double quarter = 0.25, mhalf = -0.5;
  blas_ok(rocblas_dscal(r.h, Ma * Ma, &quarter, N.ptr, 1), "rocblas_dscal(Gam)");
  blas_ok(rocblas_dger(r.h, Ma, Ma, &mhalf, c.ptr, 1, c.ptr, 1, N.ptr, Ma), "rocblas_dger(Gam)");
  // St = S U^-T
  blas_ok(rocblas_dtrsm(r.h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, (int)no2, Ma, &one, r.L.ptr, Ma, S.ptr, (int)no2),
          "rocblas_dtrsm(St)");
  // auxiliary shells in basis-function order, so that P-chunks are contiguous
  const int nsh = (int)r.shells_h.size();
  std::vector<int> order(nsh);
  for (int i = 0; i < nsh; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return r.shells_h[a].basis < r.shells_h[b].basis; });
  RIGradArgs g{d_sp_.ptr, d_pr_.ptr, d_alpha_.ptr, (int)sp_h_.size(), r.shells.ptr, r.prims.ptr, nullptr,
               nsh, d_w_.ptr, r.waux.ptr, M, Ma, 0, d_shell_atom_.ptr, r.aux_atom.ptr, natoms, thr_,
               boys_.ptr, N.ptr, d_grad};
  GANSU_HIP(launch_ri_grad_2c(g, s));
  // G^P chunks of whole auxiliary shells (basis order), then the 3-centre
  // derivatives class by class: orbital pair class (la, lb) x auxiliary L
  const size_t mm = (size_t)M * M;
  int maxc = 1;
  for (const auto& as : r.shells_h) maxc = std::max(maxc, ncart(as.L));
  // the K workspace holds a chunk of G; a separate buffer of up to 12 GB
  // (and a quarter of the free memory) when that gives larger chunks
  int pg_max = (int)std::min<size_t>(Ma, r.T.n / mm);
  DevBuf<double> Gown;
  double* G = r.T.ptr;
  {
    size_t fr = 0, tot = 0;
    GANSU_HIP(hipMemGetInfo(&fr, &tot));
    const size_t want = std::min<size_t>(Ma, std::min<size_t>(size_t(12) << 30, fr / 4) / (mm * sizeof(double)));
    if ((int)want > pg_max || pg_max < maxc) {
      pg_max = std::max<int>((int)want, maxc);
      Gown.alloc((size_t)pg_max * mm);
      G = Gown.ptr;
    }
  }
  DevBuf<double> Yt((size_t)pg_max * nocc * M);
  struct Part {
    int P0, pc, off[kMaxLAux + 1], cnt[kMaxLAux + 1];
  };
  std::vector<Part> parts;
  std::vector<int> grouped(nsh);
  for (int sh0 = 0; sh0 < nsh;) {
    const int P0 = r.shells_h[order[sh0]].basis;
    int sh1 = sh0, pc = 0;
    while (sh1 < nsh && r.shells_h[order[sh1]].basis == P0 + pc && pc + ncart(r.shells_h[order[sh1]].L) <= pg_max) {
      pc += ncart(r.shells_h[order[sh1]].L);
      ++sh1;
    }
    if (sh1 == sh0) throw Error(GANSU_E_STATE, "auxiliary basis functions are not covered contiguously by shells");
    Part pt{P0, pc, {}, {}};
    int pos = sh0;
    for (int l = 0; l <= kMaxLAux; ++l) {
      pt.off[l] = pos;
      for (int i = sh0; i < sh1; ++i)
        if (r.shells_h[order[i]].L == l) grouped[pos++] = order[i];
      pt.cnt[l] = pos - pt.off[l];
    }
    parts.push_back(pt);
    sh0 = sh1;
  }
  DevBuf<int> d_grouped;
  d_grouped.upload(grouped);
  g.aux_list = d_grouped.ptr;
  // orbital shell pairs by class (la >= lb)
  std::vector<int> pairs_h;
  int pcls[kMaxL + 1][kMaxL + 1][2] = {};
  for (int la = 0; la <= kMaxL; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      pcls[la][lb][0] = (int)pairs_h.size();
      for (int id = 0; id < (int)sp_h_.size(); ++id)
        if (sp_h_[id].la == la && sp_h_[id].lb == lb) pairs_h.push_back(id);
      pcls[la][lb][1] = (int)pairs_h.size() - pcls[la][lb][0];
    }
  if (pairs_h.size() != sp_h_.size()) throw Error(GANSU_E_STATE, "shell pair with L_i < L_j");
  DevBuf<int> d_pairs;
  d_pairs.upload(pairs_h);
  for (const Part& pt : parts) {
    GANSU_HIP(launch_fill_cD(c.ptr, d_D, pt.P0, pt.pc, M, G, s));
    blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_none, rocblas_operation_none, nocc, M, nocc, &two,
                                          S.ptr + (size_t)pt.P0 * no2, nocc, (rocblas_stride)no2, d_C, M, 0, &zero,
                                          Yt.ptr, nocc, (rocblas_stride)nocc * M, pt.pc),
            "rocblas_dgemm_strided_batched(Y)");
    const double alpha = -0.5 * two;
    blas_ok(rocblas_dgemm_strided_batched(r.h, rocblas_operation_transpose, rocblas_operation_none, M, M, nocc, &alpha,
                                          Yt.ptr, nocc, (rocblas_stride)nocc * M, d_C, M, 0, &one, G, M,
                                          (rocblas_stride)mm, pt.pc),
            "rocblas_dgemm_strided_batched(G)");
    g.P0 = pt.P0;
    g.G = G;
    for (int la = 0; la <= kMaxL; ++la)
      for (int lb = 0; lb <= la; ++lb)
        for (int l = 0; l <= kMaxLAux; ++l)
          GANSU_HIP(launch_ri_grad_3c_class(la, lb, l, g, d_pairs.ptr + pcls[la][lb][0], pcls[la][lb][1], pt.off[l],
                                            pt.cnt[l], s));
  }
  GANSU_HIP(hipStreamSynchronize(s));  // the temporaries above are released on return
}

}  // namespace gansu
