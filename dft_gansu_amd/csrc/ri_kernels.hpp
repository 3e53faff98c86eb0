// RI kernel launchers and device layouts.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

#include "fock_kernels.hpp"

namespace gansu {

// One auxiliary primitive: exponent, centre, Kc = c N'(alpha, L), Schwarz factor.
struct AuxPrim {
  double alpha, C[3], Kc, Q;
};

struct AuxShell {
  int L, nprim, prim_begin, basis;
  double qmax, pad;
};

struct RI2cArgs {
  const AuxShell* shells;
  const AuxPrim* prims;
  size_t npairs;  // nshell (nshell + 1) / 2
  const double* w;  // aux scaling per function
  int Maux;
  double thr;
  const double* boys;
  double* V;  // Maux x Maux
};

struct RI3cArgs {
  const SPair* sp;
  const PairRec* pr;
  int nsp;
  const AuxShell* shells;
  const AuxPrim* prims;
  int naux_shells;
  const double* w;     // orbital scaling per function
  const double* waux;  // aux scaling per function
  int M, Maux;
  double thr;
  const double* boys;
  double* B;  // [M][Maux][M]
};

hipError_t launch_int2c2e(const RI2cArgs& a, hipStream_t s);
hipError_t launch_int3c2e(const RI3cArgs& a, hipStream_t s);
hipError_t launch_ri_w(const double* B, const double* D, int M, int Maux, double* W, hipStream_t s);
hipError_t launch_ri_j(const double* B, const double* W, int M, int Maux, double* J, hipStream_t s);
hipError_t launch_ri_j_f32(const float* B, const double* W, int M, int Maux, double* J, hipStream_t s);
hipError_t launch_ri_w_from_t(const double* T, const double* C, int M, int pc, int nocc, int p0, double* W,
                              hipStream_t s);
hipError_t launch_mirror_upper(double* K, int M, hipStream_t s);
hipError_t launch_to_f32(const double* x, size_t n, float* y, hipStream_t s);
hipError_t launch_sum_parts(const double* part, int nparts, int nr, int nc, int accumulate, double* K, int ldk,
                            hipStream_t s);
// precision=mixed J: FP64 J when the fitted density moved by more than rel
// (max-norm of W) or after `refresh` increments, else Jprev + FP32-B increment.
hipError_t launch_ri_j_incremental(const double* B, const float* B32, const double* W, double* Wprev, double* dW,
                                   int M, int Maux, double rel, int refresh, int* st, double* Jprev, double* J,
                                   hipStream_t s);
hipError_t launch_fock_ri(const double* H, const double* J, const double* K, size_t n, double* F, hipStream_t s);

}  // namespace gansu
