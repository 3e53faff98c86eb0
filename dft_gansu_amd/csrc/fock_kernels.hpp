// Device data layouts and launch entry points of the Fock-build kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace gansu {

// One primitive pair (a, b) of a contracted shell pair, L_a >= L_b.
// Kc = exp(-ab/p |AB|^2) c_a c_b N'_a N'_b with N'(x, L) = (2/pi)^3/4 2^L x^((2L+3)/4),
// the component-independent part of the primitive norm (int2e.hpp:127-130).
struct PairRec {
  double p, P[3], PA[3], PB[3], Kc, Q;
};

// One contracted shell pair (i, j), L_i >= L_j (i >= j when L_i == L_j).
struct SPair {
  int prim_begin, nprim;  // range in the PairRec array
  int bi, bj;             // first basis function of shells i, j
  int si, sj;             // shell ids
  int la, lb;
  double qmax;            // max over primitive pairs of Q
  double pad;
};

// A wavefront's work: one bra pair (position in the bra class list) and up to
// 64 consecutive kets of the ket class list starting at ket_start.  slot
// locates its stored integrals: >= 0 in the FP64 store, -slot-1 in the FP32
// store (precision=mixed).
struct BlockRec {
  int bra_pos, ket_start, count, slot;
};

enum EriMode { kModeDirect = 0, kModeStore = 1, kModeStoredDigest = 2 };

// One bra pair with its kets at positions 0..count-1 of the ket class list.
// Kets below `split` are stored in FP64 from block first_block on, the rest
// (precision=mixed: bound below fp32_bound_threshold) in FP32 from block
// first_block32 on, 64 kets per block counted from the start of each range.
struct BraRec {
  int bra_pos, first_block, count, split;
  int first_block32, pad0, pad1, pad2;
};

// A slab workgroup's share: nb consecutive bras starting at bras[bra0] and
// the ket positions [k_begin, k_end) of each (clipped to the bra's count).
struct SlabWork {
  int bra0, nb, k_begin, k_end;
};

struct SlabLaunch {
  int LA, LB, LC, LD;
  const BraRec* bras;
  int nbras;
  const SlabWork* work;
  int nwork;
  const int* bra_list;
  const int* ket_list;
  const double* store;
  int max_count;  // largest ket count of any bra (ket-major J pass)
  const float* store32;  // FP32 blocks of the kets at or above each bra's split
  const int4* ket_meta;  // per ket position: {first basis fn i, first basis fn j, i != j, shell-pair id}
  int bpw;               // bras per workgroup of the SlabWork records (1 .. slab_max_bpw)
};

// Largest bras-per-workgroup the slab kernel of a class is compiled for.
int slab_max_bpw(int LA, int LB, int LC, int LD, int mode);

// Geometry of the slab kernel of a class: bras per workgroup and kets a
// workgroup visits per pass (used by the host to cut SlabWork records).
void slab_geometry(int LA, int LB, int LC, int LD, int mode, int* bras_per_wg, int* kets_per_pass);
// The class has a slab variant in this mode whose LDS footprint at basis
// size M fits one workgroup of the current device.
bool slab_fits(int LA, int LB, int LC, int LD, int mode, int M);

struct ClassLaunch {
  int LA, LB, LC, LD;
  const BlockRec* blocks;
  int nblocks;
  const int* bra_list;
  const int* ket_list;
  double* store;  // stored blocks of this class (kModeStore / kModeStoredDigest)
  float* store32 = nullptr;  // FP32 stored blocks (BlockRec::slot < 0)
};

struct EriArgs {
  const SPair* sp;
  const PairRec* pr;
  const double* boys;  // [1024][32]
  double thr;
  const double* Dp;  // density scaled to the raw basis (M x M)
  int M;
  double* Jr;  // raw-basis J accumulator (M x M)
  double* Kr;  // raw-basis K accumulator (M x M)
  // density-weighted screening of the direct path (null: off): max |Dp| over
  // each shell block, [nsh][nsh], then the largest of them at [nsh * nsh]; a
  // quartet (ij|kl) is skipped when
  // qmax_ij qmax_kl max(d_ij, d_kl, d_ik, d_il, d_jk, d_jl) < dthr
  const float* dmax = nullptr;
  int nsh = 0;
  double dthr = 0.0;
};

// Host-callable launchers (fock_kernels.hip).
// Shell types up to f (3); a pair_b entry < 0 marks an auxiliary primitive (a|a).
hipError_t launch_schwarz_ref(const void* prims, const int* pair_a, const int* pair_b, size_t n, double* Q,
                              hipStream_t s);
hipError_t launch_pair_data(const void* prims, const int* pa, const int* pb, const double* Qg, size_t n, PairRec* out,
                            hipStream_t s);
hipError_t launch_eri_class(const ClassLaunch& c, const EriArgs& a, int mode, hipStream_t s);
// Slab digestion: one workgroup per bra pair, K rows of the bra shells
// accumulated in LDS.  Returns hipErrorInvalidValue when the class/mode has no
// slab variant or its LDS footprint exceeds the device limit (callers then
// use launch_eri_class).
hipError_t launch_fock_slab(const SlabLaunch& c, const EriArgs& a, int mode, hipStream_t s);
bool slab_available(int LA, int LB, int LC, int LD, int mode);
// J of the ket blocks from stored integrals, ket-major (no atomics per quartet).
hipError_t launch_jket(const SlabLaunch& c, const EriArgs& a, hipStream_t s);
hipError_t launch_scale_density(const double* D, const double* w, int M, double* Dp, hipStream_t s);
hipError_t launch_zero(double* p, size_t n, hipStream_t s);
hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s);
// J = w w sym(Jr), K = w w sym(Kr); if F: F = H + J - K/2 (J, K may be null then)
// ROHF coupling coefficients of the unified Fock matrix (types.hpp:165-173).
struct RohfCoupling {
  double Acc, Bcc, Aoo, Boo, Avv, Bvv;
};
hipError_t launch_rohf_combine(const double* H, const double* Jc, const double* Kc, const double* Jo, const double* Ko,
                               size_t n, double* Fc, double* Fo, hipStream_t s);
hipError_t launch_rohf_unified(const double* Fc_mo, const double* Fo_mo, const RohfCoupling& cp, int nc, int no, int M,
                               double* R, hipStream_t s);
hipError_t launch_uhf_combine(const double* H, const double* Ja, const double* Ka, const double* Jb, const double* Kb,
                              size_t n, double* Fa, double* Fb, hipStream_t s);
hipError_t launch_finalize(const double* Jr, const double* Kr, const double* w, int M, const double* H, double* J,
                           double* K, double* F, hipStream_t s);

// Incremental (delta-density) direct builds.  st[0] counts builds since the
// last full one; launch_incr_step sets st[1] = 1 (full build: the previous
// density counts as zero and Jr, Kr are cleared) when st[0] % period == 0,
// else 0, and advances st[0].  launch_delta_density: Dp = scaled D,
// dDp = Dp - Dprev (Dp when st[1]), Dprev = Dp.  launch_zero_if_full clears
// Jr and Kr when st[1].  launch_shell_dmax: dmax[i][j] = max |dDp| over the
// basis functions of shells i, j (first function bf[i], nf[i] functions) and
// dmax[nsh * nsh] = the largest of them.  All device-side, so a captured
// build graph replays them.
hipError_t launch_incr_step(int* st, int period, hipStream_t s);
hipError_t launch_delta_density(const double* D, const double* w, int M, const int* st, double* Dp, double* Dprev,
                                double* dDp, hipStream_t s);
hipError_t launch_zero_if_full(double* Jr, double* Kr, size_t n, const int* st, hipStream_t s);
hipError_t launch_shell_dmax(const double* dDp, int M, const int* bf, const int* nf, int nsh, float* dmax,
                             hipStream_t s);

}  // namespace gansu
