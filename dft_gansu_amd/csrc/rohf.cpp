// ROHF Fock build through the same J/K engine (rohf.hpp:660-760 calls
// gpu_manager.cu:733-817, and :1800-1961 on the RI path):
//   F_c = H + J[D_c + D_o] - (K[D_c] + K[D_o])/2,
//   F_o = (H + J[D_c + D_o] - K[D_c]/2 - K[D_o])/2,
//   R   = the unified MO-basis Fock matrix built blockwise (closed / open /
//         virtual orbitals) from C^T F_c C and C^T F_o C with a named set of
//         coupling coefficients,
//   F   = S C R C^T S.
#include <rocblas/rocblas.h>

#include <map>
#include <string>

#include "engine.hpp"
#include "fock_kernels.hpp"

namespace gansu {

namespace {

void blas_check(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw Error(GANSU_E_HIP, std::string(what) + " failed: " + rocblas_status_to_string(st));
}

}  // namespace

// The coupling sets of rohf.cu:30-39 (Acc, Bcc, Aoo, Boo, Avv, Bvv).
bool rohf_coupling(const std::string& name, RohfCoupling* out) {
  static const std::map<std::string, RohfCoupling> sets = {
      {"roothaan", {-0.5, 1.5, 0.5, 0.5, 1.5, -0.5}},
      {"mcweeny-diercksen", {1.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 1.0 / 3.0}},
      {"davidson", {0.5, 1.0, 0.0, 1.0, 1.0, 0.0}},
      {"guest-saunders", {0.5, 0.5, 0.5, 0.5, 0.5, 0.5}},
      {"binkley-pople-dobosh", {0.5, 0.5, 1.0, 0.0, 0.0, 1.0}},
      {"faegri-manne", {0.5, 0.5, 1.0, 0.0, 0.5, 0.5}},
      {"goddard", {0.5, 0.5, 0.5, 0.0, 1.0, 0.0}},
      {"plakhutin-gorelik-breslavskaya", {0.0, 1.0, 1.0, 0.0, 1.0, 0.0}},
  };
  const auto it = sets.find(name);
  if (it == sets.end()) return false;
  *out = it->second;
  return true;
}

void FockEngine::rohf_release() {
  if (rohf_blas_) (void)rocblas_destroy_handle((rocblas_handle)rohf_blas_);
  rohf_blas_ = nullptr;
}

void FockEngine::build_rohf(const double* d_Dc, const double* d_Do, const double* d_H, const double* d_C,
                            const double* d_S, const RohfCoupling& cp, int nc, int no, double* d_Fc, double* d_Fo,
                            double* d_F, hipStream_t s) {
  if (!precomputed_) throw Error(GANSU_E_STATE, "build before precompute");
  if (!d_Dc || !d_Do || !d_H || !d_C || !d_S || !d_Fc || !d_Fo || !d_F) throw Error(GANSU_E_ARG, "null device pointer");
  if (nc < 0 || no < 0 || nc + no > M_) throw Error(GANSU_E_ARG, "closed/open orbital counts out of range");
  const int M = M_;
  const size_t nn = (size_t)M * M;
  if (uhf_.n < 4 * nn) uhf_.alloc(4 * nn);
  if (rohf_.n < 3 * nn) rohf_.alloc(3 * nn);
  double *Jc = uhf_.ptr, *Kc = Jc + nn, *Jo = Kc + nn, *Ko = Jo + nn;
  multi_density_ = true;  // two densities per build: full builds
  try {
    build_jk(d_Dc, Jc, Kc, nullptr, nullptr, s);
    build_jk(d_Do, Jo, Ko, nullptr, nullptr, s);
  } catch (...) {
    multi_density_ = false;
    throw;
  }
  multi_density_ = false;
  GANSU_HIP(launch_rohf_combine(d_H, Jc, Kc, Jo, Ko, nn, d_Fc, d_Fo, s));
  if (!rohf_blas_) {
    rocblas_handle h = nullptr;
    blas_check(rocblas_create_handle(&h), "rocblas_create_handle");
    rohf_blas_ = h;
  }
  const rocblas_handle h = (rocblas_handle)rohf_blas_;
  blas_check(rocblas_set_stream(h, s), "rocblas_set_stream");
  const double one = 1.0, zero = 0.0;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  auto gemm = [&](rocblas_operation ta, rocblas_operation tb, const double* A, const double* B, double* Cm) {
    blas_check(rocblas_dgemm(h, ta, tb, M, M, M, &one, A, M, B, M, &zero, Cm, M), "rocblas_dgemm(rohf)");
  };
  // Column-major views: a row-major X is the column-major X^T, so with
  // Cb = C^T the MO matrix C^T F C (row-major) is Cb F^T Cb^T column-major.
  double *W = rohf_.ptr, *Fc_mo = W + nn, *Fo_mo = Fc_mo + nn;
  gemm(N, T, d_Fc, d_C, W);
  gemm(N, N, d_C, W, Fc_mo);
  gemm(N, T, d_Fo, d_C, W);
  gemm(N, N, d_C, W, Fo_mo);
  GANSU_HIP(launch_rohf_unified(Fc_mo, Fo_mo, cp, nc, no, M, W, s));
  // F^T = S C R^T C^T S = S Cb^T R^T Cb S in the column-major views
  gemm(T, N, d_C, W, Fc_mo);
  gemm(N, N, Fc_mo, d_C, Fo_mo);
  gemm(N, N, d_S, Fo_mo, Fc_mo);
  gemm(N, N, Fc_mo, d_S, d_F);
}

}  // namespace gansu
