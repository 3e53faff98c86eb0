// 3-centre part of the RI-JK energy gradient, one kernel instance per class
// (orbital pair (LA, LB), auxiliary shell LC):
//   grad_X += sum_{P, mn} G^P_mn d(mn|P)/dX.
// The orbital centres take the E-shift 2a E(i+1) - i E(i-1) of their Hermite
// tables (built one order higher, R one order higher); the auxiliary centre
// follows by translational invariance.
//
// Work split: a wave takes one orbital shell pair and 64 auxiliary shells of
// angular momentum LC, one per lane.  The pair's Hermite tables (the largest
// private state) are built once per primitive pair into the wave's LDS slot and
// read back by every lane with compile-time offsets; each lane keeps its own
// R table and auxiliary Hermite intermediate in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "gradient_2e_impl.hpp"
#include "ri_kernels.hpp"

namespace gansu {
namespace {

using namespace dev;

constexpr int kRIG3Threads = 256;
constexpr int kRIG3Waves = kRIG3Threads / 64;

template <int LA, int LB>
constexpr int ri_grad_ne() {
  return (LA + 2) * (LB + 2) * (LA + LB + 3);
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(kRIG3Threads) void k_ri_grad_3c_t(RIGradArgs g, const int* __restrict__ plist,
                                                                int npairs, int aoff, int naux) {
  constexpr int NB = ncart(LB), NC = ncart(LC), NAB = ncart(LA) * NB;
  constexpr int LAB = LA + LB, L = LAB + LC + 1;
  constexpr int NE = ri_grad_ne<LA, LB>();
  // weights held in registers across the primitive loops where they are few
  // (loaded inside the loops instead, (pp|f) spilled 14 KB per lane: the
  // compiler hoists the loads anyway and keeps both halves live)
  constexpr bool PRELOAD = NAB * NC <= 96;
  extern __shared__ double smem[];
  const int n3 = 3 * g.natoms;
  double* s_grad = smem;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* s_E = smem + n3 + wave * 3 * NE;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const int ngroups = (naux + 63) >> 6;
  const long long nvirt = (long long)npairs * ngroups;
  const size_t M = g.M, MM = M * M;
  for (long long vb = (long long)blockIdx.x * kRIG3Waves + wave; vb < nvirt; vb += (long long)gridDim.x * kRIG3Waves) {
    const int pi = (int)(vb / ngroups), grp = (int)(vb % ngroups);
    const SPair sp = g.sp[plist[pi]];
    const int k = grp * 64 + lane;
    const bool in = k < naux;
    const int iaux = g.aux_list[aoff + (in ? k : 0)];
    const AuxShell C = g.shells[iaux];
    const bool active = in && sp.qmax * C.qmax >= g.thr;
    if (!__any(active)) continue;  // wave-uniform
    const bool diff = sp.si != sp.sj;
    auto weight = [&](int ab, int c) -> double {
      const size_t m = sp.bi + ab / NB, n = sp.bj + ab % NB;
      const double* Gp = g.G + (size_t)(C.basis + c - g.P0) * MM;
      double w = Gp[m * M + n];
      if (diff) w += Gp[n * M + m];
      return w * g.w[m] * g.w[n] * g.waux[C.basis + c];
    };
    double wts[PRELOAD ? NAB * NC : 1];
    if constexpr (PRELOAD) {
#pragma unroll
      for (int ab = 0; ab < NAB; ++ab)
#pragma unroll
        for (int c = 0; c < NC; ++c) wts[ab * NC + c] = active ? weight(ab, c) : 0.0;
    }
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    for (int ip = 0; ip < sp.nprim; ++ip) {  // wave-uniform trip count
      const PairRec pr = g.pr[sp.prim_begin + ip];
      const double exa = g.alpha[sp.prim_begin + ip], exb = pr.p - exa;
      // the previous primitive's table reads are complete before it is rebuilt
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < 3) {
        double E[LA + 2][LB + 2][LAB + 3];
        const double PA = lane == 0 ? pr.PA[0] : lane == 1 ? pr.PA[1] : pr.PA[2];
        const double PB = lane == 0 ? pr.PB[0] : lane == 1 ? pr.PB[1] : pr.PB[2];
        hermite_e_t<LA + 1, LB + 1, double>(pr.p, PA, PB, E);
        double* dst = s_E + lane * NE;
#pragma unroll
        for (int i = 0; i < LA + 2; ++i)
#pragma unroll
          for (int j = 0; j < LB + 2; ++j)
#pragma unroll
            for (int t = 0; t < LAB + 3; ++t) dst[(i * (LB + 2) + j) * (LAB + 3) + t] = E[i][j][t];
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (!active) continue;
      const auto& Eb = *reinterpret_cast<const double(*)[3][LA + 2][LB + 2][LAB + 3]>(s_E);
      for (int ic = 0; ic < C.nprim; ++ic) {
        const AuxPrim pc = g.prims[C.prim_begin + ic];
        if (pr.Q * pc.Q < g.thr) continue;
        double Ek[3][LC + 1][1][LC + 1];
#pragma unroll
        for (int d = 0; d < 3; ++d) hermite_e_t<LC, 0, double>(pc.alpha, 0.0, 0.0, Ek[d]);
        const double p = pr.p, q = pc.alpha;
        const double pref = kTwoPi25 / (p * q * sqrt(p + q)) * pr.Kc * pc.Kc;
        double R[nherm(L)];
        hermite_r_t<L, double>(p * q / (p + q), pr.P[0] - pc.C[0], pr.P[1] - pc.C[1], pr.P[2] - pc.C[2], pref,
                               g.boys, R);
        // one instance per auxiliary component (a fold rather than an unrolled
        // loop, whose full unrolling the compiler abandons for some classes,
        // leaving R runtime-indexed in scratch)
        auto component = [&]<int c>(std::integral_constant<int, c>) {
          constexpr int ce[3] = {cart_exp(LC, c, 0, true), cart_exp(LC, c, 1, true), cart_exp(LC, c, 2, true)};
          // Hermite intermediate of auxiliary function c, levels up to LAB+1
          double Gh[nherm(LAB + 1)];
#pragma unroll
          for (int t = 0; t <= LAB + 1; ++t)
#pragma unroll
            for (int u = 0; u <= LAB + 1; ++u)
#pragma unroll
              for (int v = 0; v <= LAB + 1; ++v) {
                if (t + u + v > LAB + 1) continue;
                double sacc = 0.0;
#pragma unroll
                for (int tt = 0; tt <= LC; ++tt) {
                  if (tt > ce[0]) continue;
#pragma unroll
                  for (int uu = 0; uu <= LC; ++uu) {
                    if (uu > ce[1]) continue;
                    const double exy = Ek[0][ce[0]][0][tt] * Ek[1][ce[1]][0][uu];
#pragma unroll
                    for (int vv = 0; vv <= LC; ++vv) {
                      if (vv > ce[2]) continue;
                      const double e = exy * Ek[2][ce[2]][0][vv];
                      sacc += (((tt + uu + vv) & 1) ? -e : e) * R[hidx(t + tt, u + uu, v + vv)];
                    }
                  }
                }
                Gh[hidx(t, u, v)] = sacc;
              }
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab) {
            double wt;
            if constexpr (PRELOAD) wt = wts[ab * NC + c];
            else wt = weight(ab, c);
            const int ae[3] = {cart_exp(LA, ab / NB, 0), cart_exp(LA, ab / NB, 1), cart_exp(LA, ab / NB, 2)};
            const int be[3] = {cart_exp(LB, ab % NB, 0), cart_exp(LB, ab % NB, 1), cart_exp(LB, ab % NB, 2)};
            double dA[3], dB[3];
            grad_detail::pair_grad_u<true, LA + 1, LB + 1, LAB + 3, LAB>(Eb, ae, be, exa, exb, Gh, dA, dB);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              gA[d] += wt * dA[d];
              gB[d] += wt * dB[d];
            }
          }
        };
        [&]<int... cs>(std::integer_sequence<int, cs...>) {
          (component(std::integral_constant<int, cs>{}), ...);
        }(std::make_integer_sequence<int, NC>{});
      }
    }
    if (active) {
      const int atA = g.shell_atom[sp.si], atB = g.shell_atom[sp.sj], atC = g.aux_atom[iaux];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        atomicAdd(&s_grad[3 * atA + d], gA[d]);
        atomicAdd(&s_grad[3 * atB + d], gB[d]);
        atomicAdd(&s_grad[3 * atC + d], -(gA[d] + gB[d]));
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&g.grad[i], s_grad[i]);
}

template <int LA, int LB, int LC>
hipError_t launch_t(const RIGradArgs& a, const int* plist, int npairs, int aoff, int naux, hipStream_t s) {
  const long long waves = (long long)npairs * ((naux + 63) / 64);
  const unsigned grid = (unsigned)std::min<long long>(8192, (waves + kRIG3Waves - 1) / kRIG3Waves);
  const size_t lds = (3 * (size_t)a.natoms + (size_t)kRIG3Waves * 3 * ri_grad_ne<LA, LB>()) * sizeof(double);
  hipLaunchKernelGGL((k_ri_grad_3c_t<LA, LB, LC>), dim3(grid), dim3(kRIG3Threads), lds, s, a, plist, npairs, aoff,
                     naux);
  return hipGetLastError();
}

template <int LA, int LB>
hipError_t launch_lc(int lc, const RIGradArgs& a, const int* plist, int npairs, int aoff, int naux, hipStream_t s) {
  switch (lc) {
    case 0: return launch_t<LA, LB, 0>(a, plist, npairs, aoff, naux, s);
    case 1: return launch_t<LA, LB, 1>(a, plist, npairs, aoff, naux, s);
    case 2: return launch_t<LA, LB, 2>(a, plist, npairs, aoff, naux, s);
    case 3: return launch_t<LA, LB, 3>(a, plist, npairs, aoff, naux, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_ri_grad_3c_class(int la, int lb, int lc, const RIGradArgs& a, const int* plist, int npairs,
                                   int aoff, int naux, hipStream_t s) {
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  if (npairs <= 0 || naux <= 0) return hipSuccess;
  switch (la * 3 + lb) {
    case 0: return launch_lc<0, 0>(lc, a, plist, npairs, aoff, naux, s);
    case 3: return launch_lc<1, 0>(lc, a, plist, npairs, aoff, naux, s);
    case 4: return launch_lc<1, 1>(lc, a, plist, npairs, aoff, naux, s);
    case 6: return launch_lc<2, 0>(lc, a, plist, npairs, aoff, naux, s);
    case 7: return launch_lc<2, 1>(lc, a, plist, npairs, aoff, naux, s);
    case 8: return launch_lc<2, 2>(lc, a, plist, npairs, aoff, naux, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gansu
