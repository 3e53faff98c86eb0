// Fock-build kernels for gfx950: Schwarz factors, primitive-pair data and the
// shell-quartet ERI kernels with fused J/K digestion (or stored blocks).
//
// Work decomposition.  Contracted shell pairs of one angular-momentum class are
// sorted by their largest primitive Schwarz factor (descending).  For a bra
// pair at sorted position s, the kets whose bound survives screening form a
// prefix of the ket class list, further cut to positions <= s when bra and
// ket class coincide (8-fold symmetry).  Each 64-lane wavefront takes one bra
// pair and 64 consecutive kets: bra data is wave-uniform, the bra block of J
// is reduced across the wave, and every lane digests its own quartet.
//
// Screening keeps the reference's semantics: a primitive quartet is skipped
// when Q_ab * Q_cd < threshold (int2e.cu:1681), with Q the primitive-pair
// factor of schwarz.txt:239-478.
//
// Digestion.  A lane holding the contracted block v = (ab|cd) of a unique
// shell quartet (i j | k l) adds, in the raw (un-normalised) basis,
//   Jr[a][b] += f_i f_k v D[c][d]        Jr[c][d] += f_i f_k v D[a][b]  (ij != kl)
//   Kr[a][c] += f_o v D[b][d]   Kr[b][c] += f_o v D[a][d] (i != j)
//   Kr[a][d] += f_o v D[b][c] (k != l)   Kr[b][d] += f_o v D[a][c] (i != j, k != l)
// with f_i = 2 if i != j, f_k = 2 if k != l, f_o = 2 if (ij) != (kl); the
// final J, K are the symmetrised accumulators (gpu_kernels.cu:230-287 defines
// F = H + J - K/2 on the dense tensor).
#include <hip/hip_runtime.h>
#include <algorithm>

#include "fock_kernels.hpp"
#include "md_device.hpp"

namespace gansu {

using namespace dev;

namespace {

__device__ __forceinline__ double dfact2(int n) {  // (n)!! for n >= -1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// primitive norm of component (l,m,n), int2e.hpp:127-130:
// prim_norm_shape(l, m, n) * a^((2L+3)/4)
__device__ __forceinline__ double prim_norm_shape(int l, int m, int n) {
  return double(1 << (l + m + n)) / sqrt(dfact2(2 * l - 1) * dfact2(2 * m - 1) * dfact2(2 * n - 1)) *
         0.71270547035499016035;
}

struct DPrim {  // == gansu_primitive_shell layout
  double exponent, coefficient, x, y, z;
  int shell_type;
  size_t basis_index;
  int atom_index;
};

// ---------------------------------------------------------------- Schwarz --
// Q = sqrt(max_comp c_a^2 c_b^2 N_a^2 N_b^2 [ab|ab]) per primitive pair in the
// reference's pair order; pair_b < 0 marks an auxiliary primitive (a|a).
// [ab|ab] = K^2 I with K = exp(-ab/p |AB|^2); the kernel returns sqrt(I) K.
// At P = Q the Hermite integrals are closed-form: R_{2i,2j,2k} =
// (2i-1)!! (2j-1)!! (2k-1)!! (-2 alpha)^n F_n(0), n = i + j + k, with
// F_n(0) = 1/(2n+1) and alpha = p/2, and R vanishes at odd indices.  So
// [ab|ab] = pref sum_n (-p)^n / (2n+1) [g_x g_y g_z]_n with one polynomial
// per direction, g[m] = (2m-1)!! sum_{t+tau=2m} (-1)^tau E_t E_tau.  Every
// loop bound is a template constant, so the tables stay in registers.
template <int LA, int LB>
__device__ __forceinline__ double schwarz_max_t(double p, const double (&PA)[3], const double (&PB)[3], double pref, double ea,
                                double eb, bool single, double ca2) {
  constexpr int LAB = LA + LB;
  // g[d][i][j][m] for every exponent pair (i, j) of direction d; entries
  // with m > i + j stay unused (zero)
  double g[3][LA + 1][LB + 1][LAB + 1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double E[LA + 1][LB + 1][LAB + 1];
    hermite_e_t<LA, LB, double>(p, PA[d], PB[d], E);
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
      for (int j = 0; j <= LB; ++j)
#pragma unroll
        for (int m = 0; m <= LAB; ++m) {
          double s = 0.0;
#pragma unroll
          for (int t = 0; t <= i + j; ++t) {
            const int tau = 2 * m - t;
            if (tau < 0 || tau > i + j) continue;
            s += (tau & 1) ? -E[i][j][t] * E[i][j][tau] : E[i][j][t] * E[i][j][tau];
          }
          g[d][i][j][m] = s * dfact2(2 * m - 1);
        }
  }
  // the component-independent part of the primitive norms, once
  const double pa_pow = pow(ea, (2.0 * LA + 3.0) / 4.0), pb_pow = single ? 1.0 : pow(eb, (2.0 * LB + 3.0) / 4.0);
  double best = 0.0;
#pragma unroll
  for (int ia = 0; ia < ncart(LA); ++ia)
#pragma unroll
    for (int ib = 0; ib < ncart(LB); ++ib) {
      const int ax = cart_exp(LA, ia, 0), ay = cart_exp(LA, ia, 1), az = cart_exp(LA, ia, 2);
      const int bx = cart_exp(LB, ib, 0), by = cart_exp(LB, ib, 1), bz = cart_exp(LB, ib, 2);
      double tot = 0.0, pw = 1.0;
#pragma unroll
      for (int n = 0; n <= LAB; ++n) {
        double cn = 0.0;
#pragma unroll
        for (int mx = 0; mx <= ax + bx; ++mx)
#pragma unroll
          for (int my = 0; my <= ay + by; ++my) {
            const int mz = n - mx - my;
            if (mz < 0 || mz > az + bz) continue;
            cn += g[0][ax][bx][mx] * g[1][ay][by][my] * g[2][az][bz][mz];
          }
        tot += cn * pw / double(2 * n + 1);
        pw *= -p;
      }
      double nrm = prim_norm_shape(ax, ay, az) * pa_pow;
      if (!single) nrm *= prim_norm_shape(bx, by, bz) * pb_pow;
      const double val = nrm * nrm * ca2 * pref * tot;
      if (fabs(best) < fabs(val)) best = val;
    }
  return best;
}

// The maximum over components does not depend on their order, so the
// auxiliary f order needs no separate table here.
__global__ __launch_bounds__(256) void k_schwarz_ref(const DPrim* __restrict__ prims, const int* __restrict__ pa,
                              const int* __restrict__ pb, size_t n, double* __restrict__ Q) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const DPrim p0 = prims[pa[id]];
  const bool single = pb[id] < 0;
  const DPrim p1 = single ? p0 : prims[pb[id]];
  const bool sw = !single && p1.shell_type > p0.shell_type;  // [ab|ab] = [ba|ba]
  const int la = sw ? p1.shell_type : p0.shell_type, lb = single ? 0 : (sw ? p0.shell_type : p1.shell_type);
  const double ea = sw ? p1.exponent : p0.exponent, eb = single ? 0.0 : (sw ? p0.exponent : p1.exponent);
  const double ca = sw ? p1.coefficient : p0.coefficient, cb = sw ? p0.coefficient : p1.coefficient;
  const double p = ea + eb;
  double PA[3], PB[3], r2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xa = d == 0 ? p0.x : d == 1 ? p0.y : p0.z, xb = d == 0 ? p1.x : d == 1 ? p1.y : p1.z;
    const double A = sw ? xb : xa, B = sw ? xa : xb;
    const double P = (ea * A + eb * B) / p;
    PA[d] = P - A;
    PB[d] = P - B;
    r2 += (A - B) * (A - B);
  }
  // K^2 enters as a common factor after the square root: K^2 alone can be
  // subnormal (|AB| large), K is not
  const double K = single ? 1.0 : exp(-ea * eb / p * r2);
  const double pref = kTwoPi25 / (p * p * sqrt(2.0 * p));
  const double ca2 = ca * ca * (single ? 1.0 : cb * cb);
  double best = 0.0;
  switch (la * 4 + lb) {
#define GANSU_SQ(x_, y_) \
  case x_ * 4 + y_: best = schwarz_max_t<x_, y_>(p, PA, PB, pref, ea, eb, single, ca2); break;
    GANSU_SQ(0, 0) GANSU_SQ(1, 0) GANSU_SQ(1, 1) GANSU_SQ(2, 0) GANSU_SQ(2, 1) GANSU_SQ(2, 2) GANSU_SQ(3, 0)
#undef GANSU_SQ
    default: best = __builtin_nan("");  // orbital shells are at most d, auxiliary ones at most f
  }
  Q[id] = sqrt(best) * K;
}

// --------------------------------------------------------- pair records ----
__global__ void k_pair_data(const DPrim* __restrict__ prims, const int* __restrict__ pa, const int* __restrict__ pb,
                            const double* __restrict__ Qg, size_t n, PairRec* __restrict__ out) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const DPrim a = prims[pa[id]], b = prims[pb[id]];
  const double p = a.exponent + b.exponent;
  const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z};
  PairRec r;
  double r2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double P = (a.exponent * A[d] + b.exponent * B[d]) / p;
    r.P[d] = P;
    r.PA[d] = P - A[d];
    r.PB[d] = P - B[d];
    r2 += (A[d] - B[d]) * (A[d] - B[d]);
  }
  r.p = p;
  const int la = a.shell_type, lb = b.shell_type;
  const double na = 0.71270547035499016035 * double(1 << la) * pow(a.exponent, (2.0 * la + 3.0) / 4.0);
  const double nb = 0.71270547035499016035 * double(1 << lb) * pow(b.exponent, (2.0 * lb + 3.0) / 4.0);
  r.Kc = exp(-a.exponent * b.exponent / p * r2) * a.coefficient * b.coefficient * na * nb;
  r.Q = Qg[id];
  out[id] = r;
}

// ------------------------------------------------------------- digestion --
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Digest one contracted block (compile-time shape).  `out` is [ab][cd].
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void digest_block(const double* out, bool active, const SPair& bra, const SPair& ket,
                                             bool pair_diff, const EriArgs& g) {
  constexpr int NAB = NA * NB, NCD = NC * ND;
  const int M = g.M;
  const bool ij_diff = bra.si != bra.sj, kl_diff = ket.si != ket.sj;
  const double fbk = (ij_diff ? 2.0 : 1.0) * (kl_diff ? 2.0 : 1.0);
  const double fo = pair_diff ? 2.0 : 1.0;
  const int bi = bra.bi, bj = bra.bj, bk = ket.bi, bl = ket.bj;
  const double* D = g.Dp;
  // J bra block: reduced over the wave (the bra is wave-uniform)
  {
    double dcd[NCD];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int d = 0; d < ND; ++d) dcd[c * ND + d] = active ? D[(size_t)(bk + c) * M + bl + d] : 0.0;
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) {
      double s = 0.0;
#pragma unroll
      for (int cd = 0; cd < NCD; ++cd) s += out[ab * NCD + cd] * dcd[cd];
      s = wave_sum(s * fbk);
      if (threadIdx.x == 0) unsafeAtomicAdd(&g.Jr[(size_t)(bi + ab / NB) * M + bj + ab % NB], s);
    }
  }
  if (!active) return;
  if (pair_diff) {
    double dab[NAB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) dab[a * NB + b] = D[(size_t)(bi + a) * M + bj + b];
#pragma unroll
    for (int cd = 0; cd < NCD; ++cd) {
      double s = 0.0;
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab) s += out[ab * NCD + cd] * dab[ab];
      unsafeAtomicAdd(&g.Jr[(size_t)(bk + cd / ND) * M + bl + cd % ND], s * fbk);
    }
  }
  // K: [a][c] with D[b][d]
  {
    double dbd[NB * ND];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int d = 0; d < ND; ++d) dbd[b * ND + d] = D[(size_t)(bj + b) * M + bl + d];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int d = 0; d < ND; ++d) s += out[(a * NB + b) * NCD + c * ND + d] * dbd[b * ND + d];
        unsafeAtomicAdd(&g.Kr[(size_t)(bi + a) * M + bk + c], s * fo);
      }
  }
  if (ij_diff) {  // [b][c] with D[a][d]
    double dad[NA * ND];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int d = 0; d < ND; ++d) dad[a * ND + d] = D[(size_t)(bi + a) * M + bl + d];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int d = 0; d < ND; ++d) s += out[(a * NB + b) * NCD + c * ND + d] * dad[a * ND + d];
        unsafeAtomicAdd(&g.Kr[(size_t)(bj + b) * M + bk + c], s * fo);
      }
  }
  if (kl_diff) {  // [a][d] with D[b][c]
    double dbc[NB * NC];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int c = 0; c < NC; ++c) dbc[b * NC + c] = D[(size_t)(bj + b) * M + bk + c];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int c = 0; c < NC; ++c) s += out[(a * NB + b) * NCD + c * ND + d] * dbc[b * NC + c];
        unsafeAtomicAdd(&g.Kr[(size_t)(bi + a) * M + bl + d], s * fo);
      }
  }
  if (ij_diff && kl_diff) {  // [b][d] with D[a][c]
    double dac[NA * NC];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) dac[a * NC + c] = D[(size_t)(bi + a) * M + bk + c];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int c = 0; c < NC; ++c) s += out[(a * NB + b) * NCD + c * ND + d] * dac[a * NC + c];
        unsafeAtomicAdd(&g.Kr[(size_t)(bj + b) * M + bl + d], s * fo);
      }
  }
}

// Contracted block (ab|cd), [ab][cd] order, accumulated into `out` over the
// screened primitive quartets of one bra and one ket shell pair.  nk is the
// number of ket primitive pairs to visit (0 for an idle lane).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void eri_block_t(const SPair& bra, const SPair& ket, int nk, const EriArgs& g,
                                            double* out) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  (void)NA; (void)NC;
    for (int ib = 0; ib < bra.nprim; ++ib) {
      const PairRec pb = g.pr[bra.prim_begin + ib];
      double Eb[3][LA + 1][LB + 1][LAB + 1];
#pragma unroll
      for (int d = 0; d < 3; ++d) hermite_e_t<LA, LB, double>(pb.p, pb.PA[d], pb.PB[d], Eb[d]);
      for (int ik = 0; ik < nk; ++ik) {
        const PairRec pk = g.pr[ket.prim_begin + ik];
        if (pb.Q * pk.Q < g.thr) continue;
        double Ek[3][LC + 1][LD + 1][LCD + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) hermite_e_t<LC, LD, double>(pk.p, pk.PA[d], pk.PB[d], Ek[d]);
        const double p = pb.p, q = pk.p;
        const double alpha = p * q / (p + q);
        const double pref = kTwoPi25 / (p * q * sqrt(p + q)) * pb.Kc * pk.Kc;
        double R[nherm(L)];
        hermite_r_t<L, double>(alpha, pb.P[0] - pk.P[0], pb.P[1] - pk.P[1], pb.P[2] - pk.P[2], pref, g.boys, R);
#pragma unroll
        for (int cd = 0; cd < NCD; ++cd) {
          const int ci = cd / ND, di = cd % ND;
          const int cx = cart_exp(LC, ci, 0), cy = cart_exp(LC, ci, 1), cz = cart_exp(LC, ci, 2);
          const int dx = cart_exp(LD, di, 0), dy = cart_exp(LD, di, 1), dz = cart_exp(LD, di, 2);
          // X[tuv] = sum_{tau nu phi} (-1)^(tau+nu+phi) E^cd_tau E^cd_nu E^cd_phi R[t+tau, u+nu, v+phi]
          double X[nherm(LAB)];
#pragma unroll
          for (int t = 0; t <= LAB; ++t)
#pragma unroll
            for (int u = 0; u <= LAB; ++u)
#pragma unroll
              for (int v = 0; v <= LAB; ++v) {
                if (t + u + v > LAB) continue;
                double s = 0.0;
#pragma unroll
                for (int tt = 0; tt <= LCD; ++tt) {
                  if (tt > cx + dx) continue;
#pragma unroll
                  for (int uu = 0; uu <= LCD; ++uu) {
                    if (uu > cy + dy) continue;
                    const double exy = Ek[0][cx][dx][tt] * Ek[1][cy][dy][uu];
#pragma unroll
                    for (int vv = 0; vv <= LCD; ++vv) {
                      if (vv > cz + dz) continue;
                      const double e = ((tt + uu + vv) & 1) ? -exy * Ek[2][cz][dz][vv] : exy * Ek[2][cz][dz][vv];
                      s += e * R[hidx(t + tt, u + uu, v + vv)];
                    }
                  }
                }
                X[hidx(t, u, v)] = s;
              }
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab) {
            const int ai = ab / NB, bi_ = ab % NB;
            const int ax = cart_exp(LA, ai, 0), ay = cart_exp(LA, ai, 1), az = cart_exp(LA, ai, 2);
            const int bx = cart_exp(LB, bi_, 0), by = cart_exp(LB, bi_, 1), bz = cart_exp(LB, bi_, 2);
            double s = 0.0;
#pragma unroll
            for (int t = 0; t <= LAB; ++t) {
              if (t > ax + bx) continue;
#pragma unroll
              for (int u = 0; u <= LAB; ++u) {
                if (u > ay + by) continue;
                const double exy = Eb[0][ax][bx][t] * Eb[1][ay][by][u];
#pragma unroll
                for (int v = 0; v <= LAB; ++v) {
                  if (v > az + bz) continue;
                  s += exy * Eb[2][az][bz][v] * X[hidx(t, u, v)];
                }
              }
            }
            out[ab * NCD + cd] += s;
          }
        }
      }
    }
  }

// Density-weighted screen of a contracted quartet on the direct path (see
// EriArgs::dmax): the bound of its integrals times the largest density
// element of the six shell blocks its J and K contributions use.
__device__ inline bool density_screened(const EriArgs& g, const SPair& bra, const SPair& ket) {
  if (!g.dmax) return false;
  const int n = g.nsh;
  const float* d = g.dmax;
  const float m = fmaxf(fmaxf(fmaxf(d[bra.si * n + bra.sj], d[ket.si * n + ket.sj]),
                              fmaxf(d[bra.si * n + ket.si], d[bra.si * n + ket.sj])),
                        fmaxf(d[bra.sj * n + ket.si], d[bra.sj * n + ket.sj]));
  return bra.qmax * ket.qmax * (double)m < g.dthr;
}

// Kets at positions >= the returned count of a bra's qmax-sorted ket list
// fall below the density screen for any shell block (bound times the largest
// density element): a binary search for the first such position in [0, n).
template <class QAt>
__device__ inline int density_ket_cut(const EriArgs& g, double qb, int n, QAt qat) {
  const double t = g.dthr / (qb * (double)g.dmax[g.nsh * g.nsh]);
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (qat(mid) >= t) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------- class-templated kernel --
// One lane = one contracted shell quartet; full register tiles.  Used for the
// classes whose block and Hermite tables fit in registers.
template <int LA, int LB, int LC, int LD, int MODE>
__global__ __launch_bounds__(64) void k_eri_class(ClassLaunch c, EriArgs g) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NCOMP = NAB * NCD;
  const BlockRec blk = c.blocks[blockIdx.x];
  const int lane = threadIdx.x;
  const int bra_id = c.bra_list[blk.bra_pos];
  const int ket_id = c.ket_list[blk.ket_start + (lane < blk.count ? lane : 0)];
  const SPair bra = g.sp[bra_id];
  const SPair ket = g.sp[ket_id];
  const bool active = lane < blk.count && !(MODE == kModeDirect && density_screened(g, bra, ket));

  double out[NCOMP];
#pragma unroll
  for (int k = 0; k < NCOMP; ++k) out[k] = 0.0;

  const bool f32 = blk.slot < 0;
  const size_t at = (size_t)(f32 ? -blk.slot - 1 : blk.slot) * NCOMP * 64 + lane;
  if constexpr (MODE == kModeStoredDigest) {
    if (f32) {
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) out[k] = active ? (double)c.store32[at + (size_t)k * 64] : 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) out[k] = active ? c.store[at + (size_t)k * 64] : 0.0;
    }
  } else {
    eri_block_t<LA, LB, LC, LD>(bra, ket, active ? ket.nprim : 0, g, out);
  }

  if constexpr (MODE == kModeStore) {
    if (f32) {
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) c.store32[at + (size_t)k * 64] = (float)out[k];
    } else {
#pragma unroll
      for (int k = 0; k < NCOMP; ++k) c.store[at + (size_t)k * 64] = out[k];
    }
  } else {
    digest_block<NA, NB, NC, ND>(out, active, bra, ket, bra_id != ket_id, g);
  }
}

// --------------------------------------------------- cooperative kernel ----
// The classes whose blocks exceed a lane's registers ((dp|pp) up to (dd|dd)):
// one workgroup of NT threads per contracted shell quartet.  Per group of up
// to NBG bra and KG ket primitive pairs, the pairs' data and Hermite E tables
// go to LDS once; the (bra, ket) primitive quartets that pass the screen are
// then taken KC at a time:
//   1. their Boys values, a thread each;
//   2. their Hermite R_tuv, level by level of the downward recursion, the
//      threads over (quartet, tuv) entries, ping-pong buffers in LDS;
//   3. X_b[cd][tuv] += sum E^cd (-1)^(tau+nu+phi) R[t+tau, u+nu, v+phi] for
//      the bra primitive b of each quartet, a thread per (cd, tuv) item.
// The bra contraction sum_b sum_tuv E^ab_b,tuv X_b[cd][tuv] then runs once
// per bra group, a thread per block component (tid, tid + NT, ...), into
// registers.  Nothing is held in private arrays, so the kernel needs no
// scratch memory.  blockIdx.y splits a block's kets over gridDim.y workgroups.
template <int LA, int LB, int LC, int LD>
struct CoopShape {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int NAB = NA * NB, NCD = NC * ND, NCOMP = NAB * NCD;
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  static constexpr int NH = nherm(L), NHB = nherm(LAB);
  static constexpr int NEB = (LA + 1) * (LB + 1) * (LAB + 1), NEK = (LC + 1) * (LD + 1) * (LCD + 1);
  static constexpr int NT = NCOMP > 640 ? 256 : 128;
  static constexpr int NOWN = (NCOMP + NT - 1) / NT;             // block components per thread
  static constexpr int NX = NCD * NHB;                            // X entries per bra primitive
  static constexpr int NXI = (NX + NT - 1) / NT;                  // X items per thread
  static constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
  // primitive quartets per chunk (C24H12 direct: 197 ms; 256 / NH 203 ms, 1024 / NH 218 ms)
  static constexpr int KC = clampi(512 / NH, 2, 32);
  static constexpr int NBG = clampi(768 / NX, 1, 4);             // bra primitive pairs per group
  static constexpr int KG = clampi(256 / NEK, 4, 64);            // ket primitive pairs per group
  static constexpr int NXO = NBG * NX > NCOMP ? NBG * NX : NCOMP;
};

// Packed Cartesian exponents (x | y << 2 | z << 4) of component c of an l <= 2 shell.
__device__ __forceinline__ int cart_pack(int l, int c) {
  if (l == 0) return 0;
  if (l == 1) return 1 << (2 * c);
  return c < 3 ? 2 << (2 * c) : c == 3 ? (1 | 1 << 2) : c == 4 ? (1 | 1 << 4) : (1 << 2 | 1 << 4);
}

template <int LA, int LB, int LC, int LD, int MODE>
__global__ __launch_bounds__((CoopShape<LA, LB, LC, LD>::NT)) void k_eri_coop(ClassLaunch c, EriArgs g) {
  using S = CoopShape<LA, LB, LC, LD>;
  constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NAB = S::NAB, NCD = S::NCD;
  constexpr int NCOMP = S::NCOMP, LAB = S::LAB, LCD = S::LCD, L = S::L, NEB = S::NEB, NEK = S::NEK;
  constexpr int NH = S::NH, NHB = S::NHB, NT = S::NT, NOWN = S::NOWN, NX = S::NX, NXI = S::NXI;
  constexpr int KC = S::KC, NBG = S::NBG, KG = S::KG;
  __shared__ double sR[2][KC * NH];       // R of two consecutive recursion levels, per quartet
  __shared__ double sXO[S::NXO];          // X_b[cd][tuv] of the bra group, then the block for digestion
  __shared__ double sEb[NBG][3][NEB], sEk[KG][3][NEK];
  __shared__ double sBp[NBG][6], sKp[KG][6];  // p, P[3], Kc, Q of the group's primitive pairs
  __shared__ double sF[KC][L + 1];        // scaled Boys values
  __shared__ double sPQ[KC][3];           // P - Q
  __shared__ int sPair[KC];               // surviving quartets of the chunk: bra << 8 | ket
  __shared__ int sTuv[NH];                // t | u << 4 | v << 8 of each level-ordered Hermite index
  __shared__ int sAb[NAB], sCd[NCD];      // packed exponents of a, c (bits 0-5) and b, d (bits 6-11)
  const int tid = threadIdx.x;
  for (int h = tid; h < NH; h += NT) {
    int n = 0;
    while (nherm(n) <= h) ++n;
    const int r = h - (n > 0 ? nherm(n - 1) : 0);
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= r) ++s;
    const int v = r - s * (s + 1) / 2;
    sTuv[h] = (n - s) | (s - v) << 4 | v << 8;
  }
  for (int ab = tid; ab < NAB; ab += NT) sAb[ab] = cart_pack(LA, ab / NB) | cart_pack(LB, ab % NB) << 6;
  for (int cd = tid; cd < NCD; cd += NT) sCd[cd] = cart_pack(LC, cd / ND) | cart_pack(LD, cd % ND) << 6;
  const BlockRec blk = c.blocks[blockIdx.x];
  const int bra_id = c.bra_list[blk.bra_pos];
  const SPair bra = g.sp[bra_id];
  const bool f32 = blk.slot < 0;
  const size_t slot_at = (size_t)(f32 ? -blk.slot - 1 : blk.slot) * NCOMP * 64;
  __syncthreads();

  for (int q = blockIdx.y; q < blk.count; q += gridDim.y) {
    const int ket_id = c.ket_list[blk.ket_start + q];
    const SPair ket = g.sp[ket_id];
    if (MODE == kModeDirect && density_screened(g, bra, ket)) continue;  // uniform over the workgroup
    double acc[NOWN];
#pragma unroll
    for (int j = 0; j < NOWN; ++j) acc[j] = 0.0;
    if constexpr (MODE == kModeStoredDigest) {
#pragma unroll
      for (int j = 0; j < NOWN; ++j) {
        const int k = tid + NT * j;
        if (k < NCOMP)
          acc[j] = f32 ? (double)c.store32[slot_at + (size_t)k * 64 + q] : c.store[slot_at + (size_t)k * 64 + q];
      }
    } else {
      for (int b0 = 0; b0 < bra.nprim; b0 += NBG) {
        const int nbl = min(NBG, bra.nprim - b0);
        for (int i = tid; i < nbl * NX; i += NT) sXO[i] = 0.0;
        bool any = false;
        for (int k0 = 0; k0 < ket.nprim; k0 += KG) {
          const int kn = min(KG, ket.nprim - k0);
          // primitive data and E tables of the group (the bra's only on its first ket group)
          const int nbi = k0 == 0 ? nbl : 0;
          for (int it = tid; it < 4 * (nbi + kn); it += NT) {
            const bool isb = it < 4 * nbi;
            const int i = isb ? it : it - 4 * nbi, n = isb ? nbi : kn;
            const int e = i / n, r = i - e * n;
            const PairRec* pp = g.pr + (isb ? bra.prim_begin + b0 + r : ket.prim_begin + k0 + r);
            if (e == 3) {
              double* dst = isb ? sBp[r] : sKp[r];
              dst[0] = pp->p;
              dst[1] = pp->P[0];
              dst[2] = pp->P[1];
              dst[3] = pp->P[2];
              dst[4] = pp->Kc;
              dst[5] = pp->Q;
            } else if (isb) {
              hermite_e_rt<double>(LA, LB, pp->p, pp->PA[e], pp->PB[e], sEb[r][e]);
            } else {
              hermite_e_rt<double>(LC, LD, pp->p, pp->PA[e], pp->PB[e], sEk[r][e]);
            }
          }
          __syncthreads();
          const int npair = nbl * kn;
          for (int p0 = 0; p0 < npair; p0 += KC) {
            // every wave takes the same ballot over the chunk's quartets, so
            // the survivor count needs no workgroup barrier; wave 0 writes
            // the compacted list, and only wave 0 (the Boys step) reads it
            // before the next barrier
            int nkc;
            {
              const int lane = tid & 63;
              const int pr = p0 + lane, ib = pr / kn, ik = pr - ib * kn;
              const bool keep = lane < KC && pr < npair && sBp[ib][5] * sKp[ik][5] >= g.thr;
              const unsigned long long m = __ballot(keep);
              if (tid < 64 && keep) sPair[__popcll(m & ((1ull << lane) - 1ull))] = ib << 8 | ik;
              nkc = (int)__popcll(m);
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (nkc == 0) continue;  // the same in every wave
            any = true;
            // 1. Boys values
            if (tid < nkc) {
              const int w = sPair[tid];
              const double* B = sBp[w >> 8];
              const double* K = sKp[w & 255];
              const double p = B[0], qe = K[0];
              const double alpha = p * qe / (p + qe);
              const double X = B[1] - K[1], Y = B[2] - K[2], Z = B[3] - K[3];
              double F[L + 1];
              boys_eval_t<L, double>(alpha * (X * X + Y * Y + Z * Z), g.boys, F);
              double sc = kTwoPi25 / (p * qe * sqrt(p + qe)) * B[4] * K[4];
              const double m2a = -2.0 * alpha;
#pragma unroll
              for (int n = 0; n <= L; ++n) {
                sF[tid][n] = F[n] * sc;
                sc *= m2a;
              }
              sPQ[tid][0] = X;
              sPQ[tid][1] = Y;
              sPQ[tid][2] = Z;
            }
            __syncthreads();
            // 2. R^(n) for n = L-1 .. 0; level n holds t + u + v <= L - n
#pragma unroll
            for (int n = L - 1; n >= 0; --n) {
              const int nh = nherm(L - n);
              const double* src = sR[(n + 1) & 1];
              double* dst = sR[n & 1];
              for (int it = tid; it < nkc * nh; it += NT) {
                const int t = it / nh, h = it - t * nh;
                double val;
                if (h == 0) {
                  val = sF[t][n];
                } else {
                  const int w = sTuv[h], tx = w & 15, u = (w >> 4) & 15, v = w >> 8;
                  const double* r = src + t * NH;
                  if (n == L - 1) {  // R^(L) holds only R_000 = F_L
                    val = sPQ[t][tx ? 0 : u ? 1 : 2] * sF[t][L];
                  } else if (tx > 0) {
                    val = sPQ[t][0] * r[hidx(tx - 1, u, v)];
                    if (tx > 1) val += double(tx - 1) * r[hidx(tx - 2, u, v)];
                  } else if (u > 0) {
                    val = sPQ[t][1] * r[hidx(tx, u - 1, v)];
                    if (u > 1) val += double(u - 1) * r[hidx(tx, u - 2, v)];
                  } else {
                    val = sPQ[t][2] * r[hidx(tx, u, v - 1)];
                    if (v > 1) val += double(v - 1) * r[hidx(tx, u, v - 2)];
                  }
                }
                dst[t * NH + h] = val;
              }
              __syncthreads();
            }
            // 3. X_b[cd][tuv]; the chunk's quartets are in bra-major order
#pragma unroll
            for (int j = 0; j < NXI; ++j) {
              const int it = tid + NT * j;
              if (it < NX) {
                const int cd = it / NHB, h = it - cd * NHB;
                const int w = sTuv[h], t = w & 15, u = (w >> 4) & 15, v = w >> 8;
                const int e = sCd[cd];
                const int cx = e & 3, cy = (e >> 2) & 3, cz = (e >> 4) & 3;
                const int dx = (e >> 6) & 3, dy = (e >> 8) & 3, dz = (e >> 10) & 3;
                const int ox = (cx * (LD + 1) + dx) * (LCD + 1), oy = (cy * (LD + 1) + dy) * (LCD + 1),
                          oz = (cz * (LD + 1) + dz) * (LCD + 1);
                double s = 0.0;
                int cur = sPair[0] >> 8;
                for (int k = 0; k < nkc; ++k) {
                  const int pw = sPair[k], ib = pw >> 8, ik = pw & 255;
                  if (ib != cur) {
                    sXO[cur * NX + it] += s;
                    s = 0.0;
                    cur = ib;
                  }
                  const double* ex = &sEk[ik][0][ox];
                  const double* ey = &sEk[ik][1][oy];
                  const double* ez = &sEk[ik][2][oz];
                  const double* R = sR[0] + k * NH;
                  if constexpr (LCD <= 2) {
                    // the quartet's ket E factors of this (c, d) into registers
                    // once, so that each multiply-add reads only R from LDS
                    // (the loops run to the class maximum, masked per lane).
                    // (dp|pp) and (dp|ds) 29 -> 24 ms at C24H12; with LCD = 3
                    // the registers cost a resident wave and (dp|dp) lost 15 %.
                    double fx[LCD + 1], fy[LCD + 1], fz[LCD + 1];
#pragma unroll
                    for (int i = 0; i <= LCD; ++i) {
                      fx[i] = i <= cx + dx ? ((i & 1) ? -ex[i] : ex[i]) : 0.0;
                      fy[i] = i <= cy + dy ? ((i & 1) ? -ey[i] : ey[i]) : 0.0;
                      fz[i] = i <= cz + dz ? ((i & 1) ? -ez[i] : ez[i]) : 0.0;
                    }
#pragma unroll
                    for (int tt = 0; tt <= LCD; ++tt) {
                      if (tt > cx + dx) continue;
#pragma unroll
                      for (int uu = 0; uu <= LCD - tt; ++uu) {
                        if (uu > cy + dy) continue;
                        const double exy = fx[tt] * fy[uu];
#pragma unroll
                        for (int vv = 0; vv <= LCD - tt - uu; ++vv) {
                          if (vv > cz + dz) continue;
                          s += exy * fz[vv] * R[hidx(t + tt, u + uu, v + vv)];
                        }
                      }
                    }
                  } else if constexpr (LCD == 3) {
                    // x and y factors in registers, z from LDS (the register
                    // budget of three resident waves)
                    double fx[LCD + 1], fy[LCD + 1];
#pragma unroll
                    for (int i = 0; i <= LCD; ++i) {
                      fx[i] = i <= cx + dx ? ((i & 1) ? -ex[i] : ex[i]) : 0.0;
                      fy[i] = i <= cy + dy ? ((i & 1) ? -ey[i] : ey[i]) : 0.0;
                    }
#pragma unroll
                    for (int tt = 0; tt <= LCD; ++tt) {
                      if (tt > cx + dx) continue;
#pragma unroll
                      for (int uu = 0; uu <= LCD - tt; ++uu) {
                        if (uu > cy + dy) continue;
                        const double exy = fx[tt] * fy[uu];
                        for (int vv = 0; vv <= cz + dz; ++vv) {
                          const double e3 = (vv & 1) ? -exy * ez[vv] : exy * ez[vv];
                          s += e3 * R[hidx(t + tt, u + uu, v + vv)];
                        }
                      }
                    }
                  } else {
                    for (int tt = 0; tt <= cx + dx; ++tt)
                      for (int uu = 0; uu <= cy + dy; ++uu) {
                        const double exy = ((tt + uu) & 1) ? -ex[tt] * ey[uu] : ex[tt] * ey[uu];
                        for (int vv = 0; vv <= cz + dz; ++vv) {
                          const double e3 = (vv & 1) ? -exy * ez[vv] : exy * ez[vv];
                          s += e3 * R[hidx(t + tt, u + uu, v + vv)];
                        }
                      }
                  }
                }
                sXO[cur * NX + it] += s;
              }
            }
            __syncthreads();  // the next chunk rewrites sPair, sF, sPQ and sR
          }
        }
        if (!any) continue;  // the same for every thread
        // bra contraction of the group
#pragma unroll
        for (int j = 0; j < NOWN; ++j) {
          const int k = tid + NT * j;
          if (k < NCOMP) {
            const int ab = k / NCD;
            const int e = sAb[ab];
            const int ax = e & 3, ay = (e >> 2) & 3, az = (e >> 4) & 3;
            const int bx = (e >> 6) & 3, by = (e >> 8) & 3, bz = (e >> 10) & 3;
            const int ox = (ax * (LB + 1) + bx) * (LAB + 1), oy = (ay * (LB + 1) + by) * (LAB + 1),
                      oz = (az * (LB + 1) + bz) * (LAB + 1);
            double s = 0.0;
            for (int ib = 0; ib < nbl; ++ib) {
              const double* ex = &sEb[ib][0][ox];
              const double* ey = &sEb[ib][1][oy];
              const double* ez = &sEb[ib][2][oz];
              const double* xr = &sXO[ib * NX + (k - ab * NCD) * NHB];
              for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u) {
                  const double exy = ex[t] * ey[u];
                  for (int v = 0; v <= az + bz; ++v) s += exy * ez[v] * xr[hidx(t, u, v)];
                }
            }
            acc[j] += s;
          }
        }
        __syncthreads();  // the next bra group rewrites sEb, sBp and sXO
      }
    }
    if constexpr (MODE == kModeStore) {
      // [slot][component][64 kets of the block]
#pragma unroll
      for (int j = 0; j < NOWN; ++j) {
        const int k = tid + NT * j;
        if (k < NCOMP) {
          if (f32) c.store32[slot_at + (size_t)k * 64 + q] = (float)acc[j];
          else c.store[slot_at + (size_t)k * 64 + q] = acc[j];
        }
      }
    } else {
      // digestion with the 8-fold weights: J of the bra and (for distinct
      // pairs) ket blocks, and the four K updates
#pragma unroll
      for (int j = 0; j < NOWN; ++j)
        if (tid + NT * j < NCOMP) sXO[tid + NT * j] = acc[j];
      __syncthreads();
      const int M = g.M;
      const bool ij_diff = bra.si != bra.sj, kl_diff = ket.si != ket.sj, pair_diff = bra_id != ket_id;
      const double fbk = (ij_diff ? 2.0 : 1.0) * (kl_diff ? 2.0 : 1.0);
      const double fo = pair_diff ? 2.0 : 1.0;
      const int bi = bra.bi, bj = bra.bj, bk = ket.bi, bl = ket.bj;
      const double* D = g.Dp;
      const double* O = sXO;
      for (int ab = tid; ab < NAB; ab += NT) {
        double s = 0.0;
        for (int cd = 0; cd < NCD; ++cd) s += O[ab * NCD + cd] * D[(size_t)(bk + cd / ND) * M + bl + cd % ND];
        unsafeAtomicAdd(&g.Jr[(size_t)(bi + ab / NB) * M + bj + ab % NB], s * fbk);
      }
      if (pair_diff)
        for (int cd = tid; cd < NCD; cd += NT) {
          double s = 0.0;
          for (int ab = 0; ab < NAB; ++ab) s += O[ab * NCD + cd] * D[(size_t)(bi + ab / NB) * M + bj + ab % NB];
          unsafeAtomicAdd(&g.Jr[(size_t)(bk + cd / ND) * M + bl + cd % ND], s * fbk);
        }
      constexpr int NT0 = NA * NC, NT1 = NB * NC, NT2 = NA * ND, NT3 = NB * ND;
      for (int it = tid; it < NT0 + NT1 + NT2 + NT3; it += NT) {
        double s = 0.0;
        if (it < NT0) {  // K[a][c] += (ab|cd) D[b][d]
          const int a = it / NC, cc = it % NC;
          for (int b = 0; b < NB; ++b)
            for (int d = 0; d < ND; ++d) s += O[(a * NB + b) * NCD + cc * ND + d] * D[(size_t)(bj + b) * M + bl + d];
          unsafeAtomicAdd(&g.Kr[(size_t)(bi + a) * M + bk + cc], s * fo);
        } else if (it < NT0 + NT1) {  // K[b][c] += (ab|cd) D[a][d]
          if (!ij_diff) continue;
          const int b = (it - NT0) / NC, cc = (it - NT0) % NC;
          for (int a = 0; a < NA; ++a)
            for (int d = 0; d < ND; ++d) s += O[(a * NB + b) * NCD + cc * ND + d] * D[(size_t)(bi + a) * M + bl + d];
          unsafeAtomicAdd(&g.Kr[(size_t)(bj + b) * M + bk + cc], s * fo);
        } else if (it < NT0 + NT1 + NT2) {  // K[a][d] += (ab|cd) D[b][c]
          if (!kl_diff) continue;
          const int a = (it - NT0 - NT1) / ND, d = (it - NT0 - NT1) % ND;
          for (int b = 0; b < NB; ++b)
            for (int cc = 0; cc < NC; ++cc) s += O[(a * NB + b) * NCD + cc * ND + d] * D[(size_t)(bj + b) * M + bk + cc];
          unsafeAtomicAdd(&g.Kr[(size_t)(bi + a) * M + bl + d], s * fo);
        } else {  // K[b][d] += (ab|cd) D[a][c]
          if (!(ij_diff && kl_diff)) continue;
          const int b = (it - NT0 - NT1 - NT2) / ND, d = (it - NT0 - NT1 - NT2) % ND;
          for (int a = 0; a < NA; ++a)
            for (int cc = 0; cc < NC; ++cc) s += O[(a * NB + b) * NCD + cc * ND + d] * D[(size_t)(bi + a) * M + bk + cc];
          unsafeAtomicAdd(&g.Kr[(size_t)(bj + b) * M + bl + d], s * fo);
        }
      }
      __syncthreads();  // the next ket rewrites sXO
    }
  }
}

// ---------------------------------------------------------- elementwise ----
__global__ void k_scale_density(const double* __restrict__ D, const double* __restrict__ w, int M,
                                double* __restrict__ Dp) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const int i = (int)(id / M), j = (int)(id % M);
  Dp[id] = D[id] * w[i] * w[j];
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// UHF (gpu_kernels.cu:289-340): F_a = H + J[D_a] + J[D_b] - K[D_a], likewise F_b.
__global__ void k_uhf_combine(const double* __restrict__ H, const double* __restrict__ Ja, const double* __restrict__ Ka,
                              const double* __restrict__ Jb, const double* __restrict__ Kb, size_t n,
                              double* __restrict__ Fa, double* __restrict__ Fb) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const double hj = H[id] + Ja[id] + Jb[id];
  Fa[id] = hj - Ka[id];
  Fb[id] = hj - Kb[id];
}

// ROHF (gpu_kernels.cu:351-418): F_c = H + J[D_c] + J[D_o] - (K[D_c] + K[D_o])/2,
// F_o = (H + J[D_c] + J[D_o] - K[D_c]/2 - K[D_o])/2.
__global__ void k_rohf_combine(const double* __restrict__ H, const double* __restrict__ Jc,
                               const double* __restrict__ Kc, const double* __restrict__ Jo,
                               const double* __restrict__ Ko, size_t n, double* __restrict__ Fc,
                               double* __restrict__ Fo) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const double hj = H[id] + Jc[id] + Jo[id];
  Fc[id] = hj - 0.5 * (Kc[id] + Ko[id]);
  Fo[id] = 0.5 * (hj - 0.5 * Kc[id] - Ko[id]);
}

// Unified ROHF Fock matrix in the MO basis (gpu_kernels.cu:421-497): orbital
// blocks closed (i < nc), open (nc <= i < nc + no), virtual; element (i, j)
// takes the value of the upper element (min, max) of the MO matrices.
__global__ void k_rohf_unified(const double* __restrict__ Fc, const double* __restrict__ Fo, RohfCoupling cp,
                               int nc, int no, int M, double* __restrict__ R) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const int r = (int)(id / M), c = (int)(id % M);
  const int i = min(r, c), j = max(r, c);
  const int ti = i < nc ? 0 : (i < nc + no ? 1 : 2), tj = j < nc ? 0 : (j < nc + no ? 1 : 2);
  const double fc = Fc[(size_t)i * M + j], fo = Fo[(size_t)i * M + j];
  double d;
  if (ti == 0 && tj == 0) d = 2.0 * (cp.Acc * fo + cp.Bcc * (fc - fo));
  else if (ti == 0 && tj == 1) d = 2.0 * (fc - fo);
  else if (ti == 0) d = fc;
  else if (ti == 1 && tj == 1) d = 2.0 * (cp.Aoo * fo + cp.Boo * (fc - fo));
  else if (ti == 1) d = 2.0 * fo;
  else d = 2.0 * (cp.Avv * fo + cp.Bvv * (fc - fo));
  R[id] = d;
}

__global__ void k_incr_step(int* st, int period) {
  const int n = st[0];
  st[1] = (period <= 0 || n % period == 0) ? 1 : 0;
  st[0] = n + 1;
}

__global__ void k_delta_density(const double* __restrict__ D, const double* __restrict__ w, int M,
                                const int* __restrict__ st, double* __restrict__ Dp, double* __restrict__ Dprev,
                                double* __restrict__ dDp) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const double v = D[id] * w[id / M] * w[id % M];
  const double prev = st[1] ? 0.0 : Dprev[id];
  Dp[id] = v;
  dDp[id] = v - prev;
  Dprev[id] = v;
}

__global__ void k_zero_if_full(double* __restrict__ Jr, double* __restrict__ Kr, size_t n, const int* __restrict__ st) {
  if (!st[1]) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    Jr[i] = 0.0;
    Kr[i] = 0.0;
  }
}

__global__ void k_shell_dmax(const double* __restrict__ dDp, int M, const int* __restrict__ bf,
                             const int* __restrict__ nf, int nsh, float* __restrict__ dmax) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nsh * nsh) return;
  const int i = id / nsh, j = id % nsh;
  double m = 0.0;
  for (int a = 0; a < nf[i]; ++a)
    for (int b = 0; b < nf[j]; ++b) m = fmax(m, fabs(dDp[(size_t)(bf[i] + a) * M + bf[j] + b]));
  dmax[id] = (float)m;
}

// dmax[nsh * nsh] = max over the nsh * nsh block maxima (one workgroup)
__global__ __launch_bounds__(1024) void k_dmax_reduce(float* dmax, int nsh) {
  __shared__ float red[1024];
  float m = 0.0f;
  for (int i = threadIdx.x; i < nsh * nsh; i += 1024) m = fmaxf(m, dmax[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) dmax[nsh * nsh] = red[0];
}

__global__ void k_finalize(const double* __restrict__ Jr, const double* __restrict__ Kr, const double* __restrict__ w,
                           int M, const double* __restrict__ H, double* J, double* K, double* F) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const int i = (int)(id / M), j = (int)(id % M);
  const size_t tr = (size_t)j * M + i;
  const double ww = 0.5 * w[i] * w[j];
  const double jv = ww * (Jr[id] + Jr[tr]);
  const double kv = ww * (Kr[id] + Kr[tr]);
  if (J) J[id] = jv;
  if (K) K[id] = kv;
  if (F) F[id] = H[id] + jv - 0.5 * kv;
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_t(const ClassLaunch& c, const EriArgs& a, int mode, hipStream_t s) {
  if (mode == kModeDirect)
    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, kModeDirect>), dim3(c.nblocks), dim3(64), 0, s, c, a);
  else if (mode == kModeStore)
    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, kModeStore>), dim3(c.nblocks), dim3(64), 0, s, c, a);
  else
    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD, kModeStoredDigest>), dim3(c.nblocks), dim3(64), 0, s, c, a);
  return hipGetLastError();
}


// Classes whose direct (integral-recomputing) slab tile fits a lane's
// registers: the integral block, R and both E tables.  The d-rich rest,
// (dd|ss) included (its d-d bra tables alone are 135 values), run the
// cooperative kernel.
constexpr bool slab_direct_class(int LA, int LB, int LC, int LD) {
  const int L = LA + LB + LC + LD;
  const int ncomp = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  return L <= 4 && ncomp <= 81 && !(LA == 2 && LB == 2);
}

// --------------------------------------------------- slab digestion kernel --
// A 256-thread workgroup owns BPW consecutive bra shell pairs (i j) of one
// quartet class and walks all of their kets (k l).  Every K update of a
// quartet lands in a row of shell i or j:
//   K[a][c] += v D[b][d], K[b][c] += v D[a][d], K[a][d] += v D[b][c],
//   K[b][d] += v D[a][c]      (a in i, b in j, c in k, d in l)
// so the workgroup accumulates K[i rows][:] and K[j rows][:] of each owned bra
// in LDS (with the matching D rows beside them) and flushes them once.  A
// thread reads the ket's metadata and D block once for all BPW bras, and its
// J contribution to the ket block leaves with one atomic per component per
// workgroup.  For the large bra blocks (d-p, d-d) the NA rows of a quartet are
// split over NA lanes instead (SPLIT), which keeps per-lane work short.
constexpr int kSlabThreads = 256;

template <int LA, int LB, int LC, int LD, int MODE>
struct SlabShape {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
  static constexpr bool SPLIT = MODE == kModeStoredDigest && NAB >= 18;
  static constexpr int BPW = MODE != kModeStoredDigest ? 1 : (NAB == 1 ? 8 : NAB <= 3 ? 4 : NAB <= 9 ? 2 : 1);
  static constexpr int LPQ = SPLIT ? NA : 1;          // lanes per quartet
  static constexpr int QPP = kSlabThreads / LPQ;      // quartets per pass
};

// Resident waves per SIMD the register allocation aims for in the stored
// digestion, which is latency-bound (dependent integral and D loads per ket):
// a third wave for (pp|ps) costs a 28-byte spill and cuts its time by 12 %
// (measured); forcing more waves on the 200+-register classes spills and
// loses (pp|pp) 15 %, (dd|dd) 75 %.
template <int LA, int LB, int LC, int LD, int MODE>
constexpr int slab_waves_per_eu() {
  return MODE == kModeStoredDigest && LA == 1 && LB == 1 && LC == 1 && LD == 0 ? 3 : 1;
}

template <int LA, int LB, int LC, int LD, int MODE>
__global__ __launch_bounds__(kSlabThreads) __attribute__((amdgpu_waves_per_eu(slab_waves_per_eu<LA, LB, LC, LD, MODE>())))
void k_fock_slab(SlabLaunch c, EriArgs g) {
  extern __shared__ double lds[];
  using S = SlabShape<LA, LB, LC, LD, MODE>;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NCOMP = NAB * NCD;
  constexpr int BPW = S::BPW, LPQ = S::LPQ, QPP = S::QPP;
  constexpr bool SPLIT = S::SPLIT;
  const int M = g.M;
  const int tid = threadIdx.x;
  const SlabWork wk = c.work[blockIdx.x];
  const int b0 = wk.bra0;
  const int nb = wk.nb;
  // per owned bra: [Ki NA*M][Kj NB*M][Di NA*M][Dj NB*M], then J of the bra block
  const int slab = 2 * (NA + NB) * M;
  double* jbl = lds + c.bpw * slab;  // [bpw][NAB]
  int bid[BPW], bbi[BPW], bbj[BPW], cnt[BPW], fblk[BPW], spl[BPW], fblk32[BPW];
  bool bdiff[BPW];
  int maxc = 0;
#pragma unroll
  for (int b = 0; b < BPW; ++b) {
    if (b < nb) {
      const BraRec br = c.bras[b0 + b];
      bid[b] = c.bra_list[br.bra_pos];
      const SPair bra = g.sp[bid[b]];
      bbi[b] = bra.bi; bbj[b] = bra.bj; bdiff[b] = bra.si != bra.sj;
      cnt[b] = min(br.count, wk.k_end); fblk[b] = br.first_block;
      if (MODE == kModeDirect && g.dmax)
        cnt[b] = min(cnt[b], density_ket_cut(g, bra.qmax, cnt[b], [&](int k) { return g.sp[c.ket_meta[k].w].qmax; }));
      spl[b] = br.split; fblk32[b] = br.first_block32;
      maxc = max(maxc, cnt[b]);
    } else {
      bid[b] = -1; bbi[b] = bbj[b] = 0; bdiff[b] = false; cnt[b] = 0; fblk[b] = 0; spl[b] = 0; fblk32[b] = 0;
    }
  }
  for (int b = 0; b < nb; ++b) {
    double* base = lds + b * slab;
    for (int x = tid; x < (NA + NB) * M; x += kSlabThreads) {
      base[x] = 0.0;
      const int r = x / M, col = x - r * M;
      const int row = r < NA ? bbi[b] + r : bbj[b] + (r - NA);
      base[(NA + NB) * M + x] = g.Dp[(size_t)row * M + col];
    }
  }
  for (int x = tid; x < BPW * NAB; x += kSlabThreads) jbl[x] = 0.0;
  __syncthreads();

  // SPLIT: lane -> (row a, ket); consecutive lanes read consecutive kets
  const int sub = SPLIT ? tid / QPP : 0;
  const int qi = SPLIT ? tid % QPP : tid;
  constexpr int NJB = SPLIT ? NB : NAB;  // J bra components a lane owns
  double jb[BPW][NJB];
#pragma unroll
  for (int b = 0; b < BPW; ++b)
#pragma unroll
    for (int k = 0; k < NJB; ++k) jb[b][k] = 0.0;

  for (int k0 = wk.k_begin; k0 < maxc; k0 += QPP) {
    const int kk = k0 + qi;
    if (sub < LPQ && kk < maxc) {
      const int4 km = c.ket_meta[kk];
      const int ket_id = km.w;
      const bool kl_diff = km.z != 0;
      const int bk = km.x, bl = km.y;
      double dcd[NCD], jk[NCD];
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          dcd[cc * ND + d] = g.Dp[(size_t)(bk + cc) * M + bl + d];
          jk[cc * ND + d] = 0.0;
        }
#pragma unroll
      for (int b = 0; b < BPW; ++b) {
        if (b >= nb || kk >= cnt[b]) continue;
        const double* Ki = lds + b * slab;
        const double* Kj_c = Ki + NA * M;
        double* KiW = lds + b * slab;
        double* KjW = KiW + NA * M;
        const double* Di = Kj_c + NB * M;
        const double* Dj = Di + NA * M;
        (void)Ki;
        const bool ij_diff = bdiff[b];
        const bool pair_diff = bid[b] != ket_id;
        const double fbk = (ij_diff ? 2.0 : 1.0) * (kl_diff ? 2.0 : 1.0);
        const double fo = pair_diff ? 2.0 : 1.0;
        const double fjk = pair_diff ? fbk : 0.0;
        const int bj = bbj[b];
        [[maybe_unused]] double out[MODE == kModeDirect ? NCOMP : 1];
        // digestion of one quartet; V(comp) yields its integrals.  The D
        // elements this ket needs are gathered from LDS into registers first
        // (the K rows share the LDS array, so reads cannot be hoisted across
        // the atomics otherwise) and K partial sums stay in registers until
        // one LDS atomic per (row, ket function).
        auto digest = [&](auto&& V) {
          constexpr int NAL = SPLIT ? 1 : NA;  // bra rows a of this lane
          double dal[NAL][ND], dak[NAL][NC], dbl[NB][ND], dbk[NB][NC];
#pragma unroll
          for (int aa = 0; aa < NAL; ++aa) {
            const int a = SPLIT ? sub : aa;
#pragma unroll
            for (int d = 0; d < ND; ++d) dal[aa][d] = Di[a * M + bl + d];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) dak[aa][cc] = Di[a * M + bk + cc];
          }
#pragma unroll
          for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
            for (int d = 0; d < ND; ++d) dbl[bb][d] = Dj[bb * M + bl + d];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) dbk[bb][cc] = Dj[bb * M + bk + cc];
          }
          double kac[NAL][NC], kbc[NB][NC], kad[NAL][ND], kbd[NB][ND];
#pragma unroll
          for (int aa = 0; aa < NAL; ++aa) {
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) kac[aa][cc] = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) kad[aa][d] = 0.0;
          }
#pragma unroll
          for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) kbc[bb][cc] = 0.0;
#pragma unroll
            for (int d = 0; d < ND; ++d) kbd[bb][d] = 0.0;
          }
#pragma unroll
          for (int aa = 0; aa < NAL; ++aa) {
            const int a = SPLIT ? sub : aa;
#pragma unroll
            for (int bb = 0; bb < NB; ++bb) {
              [[maybe_unused]] double dab = 0.0;
              if constexpr (MODE == kModeDirect) dab = Di[a * M + bj + bb];
              double jbs = 0.0;
#pragma unroll
              for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                  const double v = V(aa, bb, cc, d);
                  jbs += v * dcd[cc * ND + d];
                  if constexpr (MODE == kModeDirect) jk[cc * ND + d] += fjk * v * dab;
                  kac[aa][cc] += v * dbl[bb][d];
                  kbc[bb][cc] += v * dal[aa][d];
                  kad[aa][d] += v * dbk[bb][cc];
                  kbd[bb][d] += v * dak[aa][cc];
                }
              jb[b][SPLIT ? bb : a * NB + bb] += fbk * jbs;
            }
          }
#pragma unroll
          for (int aa = 0; aa < NAL; ++aa) {
            const int a = SPLIT ? sub : aa;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) unsafeAtomicAdd(&KiW[a * M + bk + cc], fo * kac[aa][cc]);
            if (kl_diff)
#pragma unroll
              for (int d = 0; d < ND; ++d) unsafeAtomicAdd(&KiW[a * M + bl + d], fo * kad[aa][d]);
          }
          if (ij_diff)
#pragma unroll
            for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
              for (int cc = 0; cc < NC; ++cc) unsafeAtomicAdd(&KjW[bb * M + bk + cc], fo * kbc[bb][cc]);
              if (kl_diff)
#pragma unroll
                for (int d = 0; d < ND; ++d) unsafeAtomicAdd(&KjW[bb * M + bl + d], fo * kbd[bb][d]);
            }
        };
        if constexpr (MODE == kModeDirect) {
          const SPair ket = g.sp[ket_id];
          const SPair bra = g.sp[bid[b]];
          if (density_screened(g, bra, ket)) continue;
#pragma unroll
          for (int k = 0; k < NCOMP; ++k) out[k] = 0.0;
          eri_block_t<LA, LB, LC, LD>(bra, ket, ket.nprim, g, out);
          digest([&](int aa, int bb, int cc, int d) -> double { return out[((aa * NB + bb) * NC + cc) * ND + d]; });
        } else {
          // the lane's integrals of this quartet: rows a of the lane, all (b, c, d)
          constexpr int NAL = SPLIT ? 1 : NA;
          constexpr int NV = NAL * NB * NCD;
          const int a0 = SPLIT ? sub : 0;
          const bool lo = kk < spl[b];
          const double* src = nullptr;
          const float* src32 = nullptr;
          if (lo) {
            src = c.store + ((size_t)(fblk[b] + kk / 64) * NCOMP + (size_t)a0 * NB * NCD) * 64 + (kk & 63);
          } else {
            const int kr = kk - spl[b];
            src32 = c.store32 + ((size_t)(fblk32[b] + kr / 64) * NCOMP + (size_t)a0 * NB * NCD) * 64 + (kr & 63);
          }
          if constexpr (NV <= 96) {
            // all loads in flight before the first use
            double vals[NV];
            if (lo) {
#pragma unroll
              for (int k = 0; k < NV; ++k) vals[k] = src[(size_t)k * 64];
            } else {
#pragma unroll
              for (int k = 0; k < NV; ++k) vals[k] = (double)src32[(size_t)k * 64];
            }
            digest([&](int aa, int bb, int cc, int d) -> double { return vals[((aa * NB + bb) * NC + cc) * ND + d]; });
          } else {
            // too many for registers: the NCD integrals of one (a, b) row are
            // loaded together at the row's first use (digest visits a row's
            // (c, d) in order, so the branch below resolves at compile time)
            double row[NCD];
            digest([&](int aa, int bb, int cc, int d) -> double {
              if (cc == 0 && d == 0) {
                const size_t r0 = (size_t)(aa * NB + bb) * NCD;
                if (lo) {
#pragma unroll
                  for (int k = 0; k < NCD; ++k) row[k] = src[(r0 + k) * 64];
                } else {
#pragma unroll
                  for (int k = 0; k < NCD; ++k) row[k] = (double)src32[(r0 + k) * 64];
                }
              }
              return row[cc * ND + d];
            });
          }
        }
      }
      // stored blocks: J of the ket block comes from the ket-major pass (k_jket)
      if constexpr (MODE == kModeDirect) {
#pragma unroll
        for (int cd = 0; cd < NCD; ++cd)
          if (jk[cd] != 0.0) unsafeAtomicAdd(&g.Jr[(size_t)(bk + cd / ND) * M + bl + cd % ND], jk[cd]);
      }
    }
  }
  // J of the bra blocks: LDS accumulation, then one atomic per component
#pragma unroll
  for (int b = 0; b < BPW; ++b) {
    if (b >= nb) break;
#pragma unroll
    for (int k = 0; k < NJB; ++k)
      if (jb[b][k] != 0.0) unsafeAtomicAdd(&jbl[b * NAB + (SPLIT ? sub * NB + k : k)], jb[b][k]);
  }
  __syncthreads();
  for (int b = 0; b < nb; ++b) {
    const double* Ki = lds + b * slab;
    const double* Kj = Ki + NA * M;
    for (int x = tid; x < NA * M; x += kSlabThreads) {
      const double v = Ki[x];
      if (v != 0.0) unsafeAtomicAdd(&g.Kr[(size_t)(bbi[b] + x / M) * M + x % M], v);
    }
    if (bdiff[b])
      for (int x = tid; x < NB * M; x += kSlabThreads) {
        const double v = Kj[x];
        if (v != 0.0) unsafeAtomicAdd(&g.Kr[(size_t)(bbj[b] + x / M) * M + x % M], v);
      }
    for (int ab = tid; ab < NAB; ab += kSlabThreads)
      unsafeAtomicAdd(&g.Jr[(size_t)(bbi[b] + ab / NB) * M + bbj[b] + ab % NB], jbl[b * NAB + ab]);
  }
}

template <int LA, int LB, int LC, int LD, int MODE>
size_t slab_lds_t(int M, int bpw) {
  using S = SlabShape<LA, LB, LC, LD, MODE>;
  return (size_t)bpw * (2ull * (S::NA + S::NB) * (size_t)M + S::NAB) * sizeof(double);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_slab_t(const SlabLaunch& c, const EriArgs& a, int mode, size_t max_lds, hipStream_t s) {
  if (c.nwork == 0) return hipSuccess;
  if (mode == kModeDirect) {
    if constexpr (slab_direct_class(LA, LB, LC, LD)) {
      if (c.bpw < 1 || c.bpw > SlabShape<LA, LB, LC, LD, kModeDirect>::BPW) return hipErrorInvalidValue;
      const size_t lds = slab_lds_t<LA, LB, LC, LD, kModeDirect>(a.M, c.bpw);
      if (lds > max_lds) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_fock_slab<LA, LB, LC, LD, kModeDirect>), dim3(c.nwork), dim3(kSlabThreads), lds, s, c,
                         a);
    } else {
      return hipErrorInvalidValue;
    }
  } else {
    if (c.bpw < 1 || c.bpw > SlabShape<LA, LB, LC, LD, kModeStoredDigest>::BPW) return hipErrorInvalidValue;
    const size_t lds = slab_lds_t<LA, LB, LC, LD, kModeStoredDigest>(a.M, c.bpw);
    if (lds > max_lds) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fock_slab<LA, LB, LC, LD, kModeStoredDigest>), dim3(c.nwork), dim3(kSlabThreads), lds, s,
                       c, a);
  }
  return hipGetLastError();
}


// ---------------------------------------------- ket-major J of stored blocks --
// J[k+c][l+d] += f sum_ab (ab|cd) D[i+a][j+b] for every stored quartet.  A
// 64-lane workgroup takes 64 consecutive ket positions, one ket component cd
// (blockIdx.y) and a range of bras (blockIdx.z): the bra's D block is uniform
// across the wave, the lanes read 64 consecutive kets of one stored chunk
// (coalesced), and each lane adds its sum to J once.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_jket(SlabLaunch c, EriArgs g, int bras_per_block) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NCOMP = NAB * NCD;
  const int tfirst = blockIdx.x * 64;
  const int t = tfirst + threadIdx.x;
  const int cd = blockIdx.y;
  const int s0 = blockIdx.z * bras_per_block;
  const int s1 = min(c.nbras, s0 + bras_per_block);
  const int M = g.M;
  double jk = 0.0;
  int ket_id = -1;
  bool kl_diff = false;
  int4 km = make_int4(0, 0, 0, -1);
  if (t < c.max_count) {
    km = c.ket_meta[t];
    ket_id = km.w;
    kl_diff = km.z != 0;
  }
  for (int s = s0; s < s1; ++s) {
    const BraRec br = c.bras[s];
    if (br.count <= tfirst) continue;  // uniform: no ket of this workgroup
    if (t >= br.count) continue;
    const int bra_id = c.bra_list[br.bra_pos];
    if (bra_id == ket_id) continue;  // diagonal quartet: its J is complete in the bra pass
    const SPair bra = g.sp[bra_id];
    const double f = (bra.si != bra.sj ? 2.0 : 1.0) * (kl_diff ? 2.0 : 1.0);
    double acc = 0.0;
    if (t < br.split) {
      const double* src = c.store + ((size_t)(br.first_block + t / 64) * NCOMP) * 64 + (t & 63) + (size_t)cd * 64;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          acc += g.Dp[(size_t)(bra.bi + a) * M + bra.bj + b] * src[(size_t)(a * NB + b) * NCD * 64];
    } else {
      const int tr = t - br.split;
      const float* src = c.store32 + ((size_t)(br.first_block32 + tr / 64) * NCOMP) * 64 + (tr & 63) + (size_t)cd * 64;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          acc += g.Dp[(size_t)(bra.bi + a) * M + bra.bj + b] * (double)src[(size_t)(a * NB + b) * NCD * 64];
    }
    jk += f * acc;
  }
  if (ket_id >= 0 && jk != 0.0) unsafeAtomicAdd(&g.Jr[(size_t)(km.x + cd / ND) * M + km.y + cd % ND], jk);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_jket_t(const SlabLaunch& c, const EriArgs& a, hipStream_t s) {
  constexpr int NCD = ncart(LC) * ncart(LD);
  // few long bra loops for light classes, more parallel ones for heavy ones
  const int bpb = ncart(LA) * ncart(LB) * NCD <= 9 ? 64 : 16;
  dim3 grid((c.max_count + 63) / 64, NCD, (c.nbras + bpb - 1) / bpb);
  hipLaunchKernelGGL((k_jket<LA, LB, LC, LD>), grid, dim3(64), 0, s, c, a, bpb);
  return hipGetLastError();
}

size_t g_max_lds = 0;

}  // namespace

// --------------------------------------------------------------- launchers --
hipError_t launch_schwarz_ref(const void* prims, const int* pair_a, const int* pair_b, size_t n, double* Q,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_schwarz_ref, dim3(nb), dim3(256), 0, s, (const DPrim*)prims, pair_a, pair_b, n, Q);
  return hipGetLastError();
}

hipError_t launch_pair_data(const void* prims, const int* pa, const int* pb, const double* Qg, size_t n, PairRec* out,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_pair_data, dim3(nb), dim3(256), 0, s, (const DPrim*)prims, pa, pb, Qg, n, out);
  return hipGetLastError();
}

// One wavefront per quartet; a block's kets are split over enough
// wavefronts (blockIdx.y) to fill the device.
template <int LA, int LB, int LC, int LD>
hipError_t launch_coop_t(const ClassLaunch& c, const EriArgs& a, int mode, hipStream_t s) {
  const unsigned split = (unsigned)std::min(64, std::max(1, (8192 + c.nblocks - 1) / c.nblocks));
  const dim3 grid((unsigned)c.nblocks, split), block(CoopShape<LA, LB, LC, LD>::NT);
  if (mode == kModeDirect)
    hipLaunchKernelGGL((k_eri_coop<LA, LB, LC, LD, kModeDirect>), grid, block, 0, s, c, a);
  else if (mode == kModeStore)
    hipLaunchKernelGGL((k_eri_coop<LA, LB, LC, LD, kModeStore>), grid, block, 0, s, c, a);
  else
    hipLaunchKernelGGL((k_eri_coop<LA, LB, LC, LD, kModeStoredDigest>), grid, block, 0, s, c, a);
  return hipGetLastError();
}

// Classes whose block and Hermite tables fit a lane's registers run the
// templated one-lane-per-quartet kernel; the d-rich rest the cooperative one.
hipError_t launch_eri_class(const ClassLaunch& c, const EriArgs& a, int mode, hipStream_t s) {
  if (c.nblocks == 0) return hipSuccess;
  const int key = ((c.LA * 3 + c.LB) * 3 + c.LC) * 3 + c.LD;
#define GANSU_CASE(a_, b_, c_, d_) \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_: return launch_t<a_, b_, c_, d_>(c, a, mode, s);
#define GANSU_COOP(a_, b_, c_, d_) \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_: return launch_coop_t<a_, b_, c_, d_>(c, a, mode, s);
  switch (key) {
    GANSU_CASE(0, 0, 0, 0)
    GANSU_CASE(1, 0, 0, 0)
    GANSU_CASE(1, 0, 1, 0)
    GANSU_CASE(1, 1, 0, 0)
    GANSU_CASE(1, 1, 1, 0)
    GANSU_CASE(1, 1, 1, 1)
    GANSU_CASE(2, 0, 0, 0)
    GANSU_CASE(2, 0, 1, 0)
    GANSU_CASE(2, 0, 1, 1)
    GANSU_CASE(2, 0, 2, 0)
    GANSU_CASE(2, 1, 0, 0)
    GANSU_CASE(2, 1, 1, 0)
    GANSU_COOP(2, 2, 0, 0)
    GANSU_COOP(2, 1, 1, 1)
    GANSU_COOP(2, 1, 2, 0)
    GANSU_COOP(2, 1, 2, 1)
    GANSU_COOP(2, 2, 1, 0)
    GANSU_COOP(2, 2, 1, 1)
    GANSU_COOP(2, 2, 2, 0)
    GANSU_COOP(2, 2, 2, 1)
    GANSU_COOP(2, 2, 2, 2)
    default: return hipErrorInvalidValue;
  }
#undef GANSU_CASE
#undef GANSU_COOP
}



void slab_geometry(int LA, int LB, int LC, int LD, int mode, int* bpw, int* qpp) {
  const int nab = ncart(LA) * ncart(LB);
  const bool stored = mode == kModeStoredDigest;
  const bool split = stored && nab >= 18;
  *bpw = !stored ? 1 : (nab == 1 ? 8 : nab <= 3 ? 4 : nab <= 9 ? 2 : 1);
  *qpp = split ? kSlabThreads / ncart(LA) : kSlabThreads;
  (void)LC; (void)LD;
}

hipError_t launch_jket(const SlabLaunch& c, const EriArgs& a, hipStream_t s) {
  if (c.nbras == 0 || c.max_count == 0) return hipSuccess;
  const int key = ((c.LA * 3 + c.LB) * 3 + c.LC) * 3 + c.LD;
#define GANSU_JK(a_, b_, c_, d_) \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_: return launch_jket_t<a_, b_, c_, d_>(c, a, s);
  switch (key) {
    GANSU_JK(0, 0, 0, 0) GANSU_JK(1, 0, 0, 0) GANSU_JK(1, 0, 1, 0) GANSU_JK(1, 1, 0, 0)
    GANSU_JK(1, 1, 1, 0) GANSU_JK(1, 1, 1, 1) GANSU_JK(2, 0, 0, 0) GANSU_JK(2, 0, 1, 0)
    GANSU_JK(2, 0, 1, 1) GANSU_JK(2, 0, 2, 0) GANSU_JK(2, 1, 0, 0) GANSU_JK(2, 1, 1, 0)
    GANSU_JK(2, 1, 1, 1) GANSU_JK(2, 1, 2, 0) GANSU_JK(2, 1, 2, 1) GANSU_JK(2, 2, 0, 0)
    GANSU_JK(2, 2, 1, 0) GANSU_JK(2, 2, 1, 1) GANSU_JK(2, 2, 2, 0) GANSU_JK(2, 2, 2, 1)
    GANSU_JK(2, 2, 2, 2)
    default: return hipErrorInvalidValue;
  }
#undef GANSU_JK
}

// Register-resident blocks (the classes of launch_eri_class's templated list)
// can also be computed inside the slab kernel; every class can be digested
// from stored blocks.
bool slab_available(int LA, int LB, int LC, int LD, int mode) {
  if (mode == kModeStoredDigest) return true;
  return mode == kModeDirect && slab_direct_class(LA, LB, LC, LD);
}

// LDS bytes of the slab kernel of a class at basis size M (0: no variant).
size_t slab_lds_bytes(int LA, int LB, int LC, int LD, int mode, int M) {
  if (!slab_available(LA, LB, LC, LD, mode)) return 0;
  const int NA = ncart(LA), NB = ncart(LB);
  int bpw = 1, qpp = 0;
  slab_geometry(LA, LB, LC, LD, mode, &bpw, &qpp);
  return (size_t)bpw * (2ull * (NA + NB) * (size_t)M + (size_t)NA * NB) * sizeof(double);
}

int slab_max_bpw(int LA, int LB, int LC, int LD, int mode) {
  int bpw = 1, qpp = 0;
  slab_geometry(LA, LB, LC, LD, mode, &bpw, &qpp);
  return bpw;
}

bool slab_fits(int LA, int LB, int LC, int LD, int mode, int M) {
  const size_t need = slab_lds_bytes(LA, LB, LC, LD, mode, M);
  if (need == 0) return false;
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return false;
  return need <= (size_t)v;
}

// Launches the slab kernel; hipErrorInvalidValue (nothing launched) when the
// class has no slab variant or its LDS footprint exceeds the device limit.
hipError_t launch_fock_slab(const SlabLaunch& c, const EriArgs& a, int mode, hipStream_t s) {
  if (c.nbras == 0) return hipSuccess;
  if (!slab_available(c.LA, c.LB, c.LC, c.LD, mode)) return hipErrorInvalidValue;
  if (g_max_lds == 0) {
    int dev = 0, v = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    g_max_lds = (size_t)v;
  }
  const int key = ((c.LA * 3 + c.LB) * 3 + c.LC) * 3 + c.LD;
#define GANSU_SLAB(a_, b_, c_, d_) \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_: return launch_slab_t<a_, b_, c_, d_>(c, a, mode, g_max_lds, s);
  if (mode == kModeDirect) {
    switch (key) {
      GANSU_SLAB(0, 0, 0, 0) GANSU_SLAB(1, 0, 0, 0) GANSU_SLAB(1, 0, 1, 0) GANSU_SLAB(1, 1, 0, 0)
      GANSU_SLAB(1, 1, 1, 0) GANSU_SLAB(1, 1, 1, 1) GANSU_SLAB(2, 0, 0, 0) GANSU_SLAB(2, 0, 1, 0)
      GANSU_SLAB(2, 0, 1, 1) GANSU_SLAB(2, 0, 2, 0) GANSU_SLAB(2, 1, 0, 0) GANSU_SLAB(2, 1, 1, 0)
      default: return hipErrorInvalidValue;
    }
  }
  switch (key) {
    GANSU_SLAB(0, 0, 0, 0) GANSU_SLAB(1, 0, 0, 0) GANSU_SLAB(1, 0, 1, 0) GANSU_SLAB(1, 1, 0, 0)
    GANSU_SLAB(1, 1, 1, 0) GANSU_SLAB(1, 1, 1, 1) GANSU_SLAB(2, 0, 0, 0) GANSU_SLAB(2, 0, 1, 0)
    GANSU_SLAB(2, 0, 1, 1) GANSU_SLAB(2, 0, 2, 0) GANSU_SLAB(2, 1, 0, 0) GANSU_SLAB(2, 1, 1, 0)
    GANSU_SLAB(2, 1, 1, 1) GANSU_SLAB(2, 1, 2, 0) GANSU_SLAB(2, 1, 2, 1) GANSU_SLAB(2, 2, 0, 0)
    GANSU_SLAB(2, 2, 1, 0) GANSU_SLAB(2, 2, 1, 1) GANSU_SLAB(2, 2, 2, 0) GANSU_SLAB(2, 2, 2, 1)
    GANSU_SLAB(2, 2, 2, 2)
    default: return hipErrorInvalidValue;
  }
#undef GANSU_SLAB
}

hipError_t launch_scale_density(const double* D, const double* w, int M, double* Dp, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_scale_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, w, M, Dp);
  return hipGetLastError();
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_incr_step(int* st, int period, hipStream_t s) {
  hipLaunchKernelGGL(k_incr_step, dim3(1), dim3(1), 0, s, st, period);
  return hipGetLastError();
}

hipError_t launch_delta_density(const double* D, const double* w, int M, const int* st, double* Dp, double* Dprev,
                                double* dDp, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_delta_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, w, M, st, Dp, Dprev, dDp);
  return hipGetLastError();
}

hipError_t launch_zero_if_full(double* Jr, double* Kr, size_t n, const int* st, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_zero_if_full, dim3(grid), dim3(256), 0, s, Jr, Kr, n, st);
  return hipGetLastError();
}

hipError_t launch_shell_dmax(const double* dDp, int M, const int* bf, const int* nf, int nsh, float* dmax,
                             hipStream_t s) {
  const int n = nsh * nsh;
  hipLaunchKernelGGL(k_shell_dmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dDp, M, bf, nf, nsh, dmax);
  hipLaunchKernelGGL(k_dmax_reduce, dim3(1), dim3(1024), 0, s, dmax, nsh);
  return hipGetLastError();
}

hipError_t launch_zero(double* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_uhf_combine(const double* H, const double* Ja, const double* Ka, const double* Jb, const double* Kb,
                              size_t n, double* Fa, double* Fb, hipStream_t s) {
  hipLaunchKernelGGL(k_uhf_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, Ja, Ka, Jb, Kb, n, Fa, Fb);
  return hipGetLastError();
}

hipError_t launch_rohf_combine(const double* H, const double* Jc, const double* Kc, const double* Jo, const double* Ko,
                               size_t n, double* Fc, double* Fo, hipStream_t s) {
  hipLaunchKernelGGL(k_rohf_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, Jc, Kc, Jo, Ko, n, Fc, Fo);
  return hipGetLastError();
}

hipError_t launch_rohf_unified(const double* Fc_mo, const double* Fo_mo, const RohfCoupling& cp, int nc, int no, int M,
                               double* R, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_rohf_unified, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Fc_mo, Fo_mo, cp, nc, no, M,
                     R);
  return hipGetLastError();
}

hipError_t launch_finalize(const double* Jr, const double* Kr, const double* w, int M, const double* H, double* J,
                           double* K, double* F, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Jr, Kr, w, M, H, J, K, F);
  return hipGetLastError();
}

}  // namespace gansu
