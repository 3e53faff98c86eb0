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

// RI-JK energy gradient: sum_P sum_mn G^P_mn d(mn|P)/dX over the chunk
// [P0, P0 + pc) of G (G[p - P0][m][n], real basis, symmetric in m, n), and
// sum_PQ Gam_PQ d(P|Q)/dX over all nsh auxiliary shells; per-atom sums are
// added into grad.
struct RIGradArgs {
  const SPair* sp;
  const PairRec* pr;
  const double* alpha;  // exponent of the first primitive of each pair record
  int nsp;
  const AuxShell* shells;
  const AuxPrim* prims;
  const int* aux_list;  // 3-centre term: auxiliary shell ids, grouped per chunk and L
  int nsh;              // 2-centre term: number of auxiliary shells
  const double* w;      // orbital raw-basis factors
  const double* waux;   // auxiliary raw-basis factors
  int M, Maux, P0;
  const int* shell_atom;
  const int* aux_atom;
  int natoms;
  double thr;
  const double* boys;
  const double* G;    // 3-centre gradient density chunk, or Gam (Maux x Maux) for the 2-centre term
  double* grad;       // 3 natoms, accumulated
};
// Per-class instance of the 3-centre term: orbital shell pairs plist[0..npairs)
// of class (la, lb), la >= lb, and auxiliary shells aux_list[aoff .. aoff+naux)
// of angular momentum lc, all inside the chunk of G.
hipError_t launch_ri_grad_3c_class(int la, int lb, int lc, const RIGradArgs& a, const int* plist, int npairs,
                                   int aoff, int naux, hipStream_t s);
hipError_t launch_ri_grad_2c(const RIGradArgs& a, hipStream_t s);
// G[p][m][n] = c[P0 + p] D[m][n] for p < pc.
hipError_t launch_fill_cD(const double* c, const double* D, int P0, int pc, int M, double* G, hipStream_t s);

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
