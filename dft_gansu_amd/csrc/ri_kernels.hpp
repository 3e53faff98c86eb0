// RI kernel launchers and device layouts.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

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
// B[m][P][n] = B[n][P][m] for n in an earlier block of bs (bs % 32 == 0) than m
hipError_t launch_ri_mirror(double* B, int M, int Maux, int bs, hipStream_t s);
hipError_t launch_ri_w(const double* B, const double* D, int M, int Maux, double* W, hipStream_t s);
hipError_t launch_ri_j(const double* B, const double* W, int M, int Maux, double* J, hipStream_t s);
hipError_t launch_ri_w_from_t(const double* T, const double* C, int M, int pc, int nocc, int p0, double* W,
                              hipStream_t s);
hipError_t launch_mirror_upper(double* K, int M, hipStream_t s);
hipError_t launch_sum_parts(const double* part, int nparts, int nr, int nc, int accumulate, double* K, int ldk,
                            hipStream_t s);
hipError_t launch_fock_ri(const double* H, const double* J, const double* K, size_t n, double* F, hipStream_t s);

// ---- incremental RI exchange on the fp16 matrix cores (ri_mfma.hip) ----
// Device-side scalars of one incremental build.
struct RikState {
  float sE, emax;            // scale and max|E| of the orbital change
  unsigned ta_max, te_max;   // max|T_A|, max|T_E| (float bits, atomicMax)
};
struct RikTArgs {
  const uint16_t* Bp; // fp16 hi/lo planes of B, [R][kpad / 8][hi 8 | lo 8] (rows m Maux + P)
  long R;
  int M, Maux, nocc, kpad;
  float sB;           // power-of-two scale of the planes (B sB = hi + lo)
  const uint16_t* Ehi;  // fp16 bits, E^T [rik_cols(nocc)][kpad]
  const uint16_t* Elo;
  const double* E;    // M x nocc, row-major
  RikState* st;
  float* TA;          // R x nocc
  float* TE;
  double* dWrow;      // R
  int fresh;          // 1: the previous T_A, T_E count as zero (not read)
  unsigned* wgmax;    // 2 per workgroup: max|T_A|, max|T_E| (float bits)
  long row0, wg0;     // first row and workgroup of one launch (set by launch_rik_tstep)
};
size_t rik_tstep_workgroups(long R);
int rik_kpad(int M);  // k extent of the planes and of E^T: M rounded up to 64
// Bp from the FP64 B (R x M): the hi/lo split of (float)B sB
hipError_t launch_rik_planes(const double* B, long R, int M, int kpad, float sB, uint16_t* Bp, hipStream_t s);
struct RikDkArgs {
  const float* TA;    // M x Kd (Kd = Maux nocc)
  const float* TE;
  int M;
  long Kd, kchunk;
  int nchunk;         // k chunks (split-k partials)
  const RikState* st;
  float* Zpart;       // nchunk x M x M
};
int rik_blocks(int nocc);  // 32-column blocks of the orbital change (at most 10)
int rik_cols(int nocc);    // columns of the fp16 E^T split (whole 64-column pairs of blocks)
hipError_t launch_rik_split_e(const double* E, int M, int nocc, int kpad, RikState* st, uint16_t* Ehi, uint16_t* Elo,
                              hipStream_t s);
hipError_t launch_rik_tstep(const RikTArgs& a, hipStream_t s);
hipError_t launch_rik_dk(const RikDkArgs& a, int nchunk, hipStream_t s);
int rik_dk_tiles(int M);  // output tiles of one k chunk of the dK product
// S = sum of the Zpart chunks (FP64, M x M scratch), then K = Kprev + S + S^T
hipError_t launch_rik_reduce(const float* Zpart, int nchunk, int M, double* S, const double* Kprev, double* K,
                             hipStream_t s);
hipError_t launch_rik_dw(const double* dWrow, int M, int Maux, double* dW, double* W, double* Wprev, hipStream_t s);
hipError_t launch_rik_j(const uint16_t* Bp, float sB, int kpad, const double* dW, int M, int Maux, double* Jprev,
                        double* J, hipStream_t s);
hipError_t launch_rik_seed(const double* T, int M, int pc, int nocc, int p0, int Maux, float* TA, float* TE,
                           hipStream_t s);
hipError_t launch_rik_xocc(const double* C, int M, int nocc, double* X, hipStream_t s);
// out[0] = |x|^2, out[1] = |y|^2 (n entries), out[2] = |G - I|^2 (G no x no)
hipError_t launch_rik_norms(const double* x, const double* y, size_t n, const double* G, int no, double* out,
                            hipStream_t s);
hipError_t launch_rik_absmax(const double* x, size_t n, unsigned* w, hipStream_t s);  // max |(float)x|

}  // namespace gansu
