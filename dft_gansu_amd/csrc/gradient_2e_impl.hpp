// Two-electron backward kernel template, shared by the translation units that
// instantiate its quartet classes (gradient_2e_*.hip): see gradient_kernels.hip
// for the method.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gradient_kernels.hpp"
#include "md_device.hpp"

namespace gansu {
namespace grad_detail {

using namespace dev;


constexpr int kGradThreads = 256;

__device__ __forceinline__ double wave_sum_g(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sum_{t<=nt, u<=nu, v<=nv} Ex[t] Ey[u] Ez[v] G[hidx(t,u,v)]
__device__ inline double contract_e(const double* Ex, int nt, const double* Ey, int nu, const double* Ez, int nv,
                                    const double* G) {
  double s = 0.0;
  for (int t = 0; t <= nt; ++t)
    for (int u = 0; u <= nu; ++u) {
      const double exy = Ex[t] * Ey[u];
      for (int v = 0; v <= nv; ++v) s += exy * Ez[v] * G[hidx(t, u, v)];
    }
  return s;
}

// E tables of one primitive pair built to (la+1, lb+1): row (i, j) of
// direction d starts at E[d] + (i nj + j) nt.
struct ETab {
  const double* E[3];
  int nj, nt;
  __device__ const double* row(int d, int i, int j) const { return E[d] + (i * nj + j) * nt; }
};

// Derivatives along x, y, z of one component pair (ea, eb) with respect to
// both of its centres, against a Hermite intermediate G whose levels reach
// ea+eb+1 per direction.  With the one-direction partial sums
//   H_x[t] = sum_uv E_y[u] E_z[v] G[t,u,v]   (likewise H_y, H_z)
// d/dA_x = sum_t (2a E^(i+1,j)_t - i E^(i-1,j)_t) H_x[t] and
// d/dB_x = sum_t (2b E^(i,j+1)_t - j E^(i,j-1)_t) H_x[t], so the three partial
// sums serve all six derivatives.  want_b = false skips the second centre.
__device__ inline void pair_grad(const ETab& T, const int* ea, const int* eb, double exa, double exb, const double* G,
                                 bool want_b, double* dA, double* dB) {
  int n[3];
  const double* E[3];
  for (int d = 0; d < 3; ++d) { n[d] = ea[d] + eb[d]; E[d] = T.row(d, ea[d], eb[d]); }
  double H[3][6];
  for (int t = 0; t <= n[0] + 1; ++t) {
    double s = 0.0;
    for (int u = 0; u <= n[1]; ++u) {
      double su = 0.0;
      for (int v = 0; v <= n[2]; ++v) su += E[2][v] * G[hidx(t, u, v)];
      s += E[1][u] * su;
    }
    H[0][t] = s;
  }
  for (int u = 0; u <= n[1] + 1; ++u) {
    double s = 0.0;
    for (int t = 0; t <= n[0]; ++t) {
      double st = 0.0;
      for (int v = 0; v <= n[2]; ++v) st += E[2][v] * G[hidx(t, u, v)];
      s += E[0][t] * st;
    }
    H[1][u] = s;
  }
  for (int v = 0; v <= n[2] + 1; ++v) {
    double s = 0.0;
    for (int t = 0; t <= n[0]; ++t) {
      double st = 0.0;
      for (int u = 0; u <= n[1]; ++u) st += E[1][u] * G[hidx(t, u, v)];
      s += E[0][t] * st;
    }
    H[2][v] = s;
  }
  for (int d = 0; d < 3; ++d) {
    const double* up = T.row(d, ea[d] + 1, eb[d]);
    double a = 0.0;
    for (int t = 0; t <= n[d] + 1; ++t) a += up[t] * H[d][t];
    a *= 2.0 * exa;
    if (ea[d] > 0) {
      const double* dn = T.row(d, ea[d] - 1, eb[d]);
      double b = 0.0;
      for (int t = 0; t <= n[d] - 1; ++t) b += dn[t] * H[d][t];
      a -= ea[d] * b;
    }
    dA[d] = a;
    if (want_b) {
      const double* upb = T.row(d, ea[d], eb[d] + 1);
      double b = 0.0;
      for (int t = 0; t <= n[d] + 1; ++t) b += upb[t] * H[d][t];
      b *= 2.0 * exb;
      if (eb[d] > 0) {
        const double* dnb = T.row(d, ea[d], eb[d] - 1);
        double c2 = 0.0;
        for (int t = 0; t <= n[d] - 1; ++t) c2 += dnb[t] * H[d][t];
        b -= eb[d] * c2;
      }
      dB[d] = b;
    }
  }
}

// Register variant of pair_grad for the unrolled classes: E is the pair's
// table [d][i][j][t] built to (la+1, lb+1); ea, eb fold to constants once the
// component loops are unrolled, so every index below is compile-time.
template <bool WANT_B, int IA, int IB, int NT, int NMAX>
__device__ __forceinline__ void pair_grad_u(const double (&E)[3][IA + 1][IB + 1][NT], const int (&ea)[3],
                                            const int (&eb)[3], double exa, double exb, const double* G,
                                            double (&dA)[3], double (&dB)[3]) {
  const int n0 = ea[0] + eb[0], n1 = ea[1] + eb[1], n2 = ea[2] + eb[2];
  double H[3][NMAX + 2];
#pragma unroll
  for (int t = 0; t <= NMAX + 1; ++t) {
    double h0 = 0.0, h1 = 0.0, h2 = 0.0;
#pragma unroll
    for (int a = 0; a <= NMAX; ++a)
#pragma unroll
      for (int b = 0; b <= NMAX; ++b) {
        if (t <= n0 + 1 && a <= n1 && b <= n2) h0 += E[1][ea[1]][eb[1]][a] * E[2][ea[2]][eb[2]][b] * G[hidx(t, a, b)];
        if (t <= n1 + 1 && a <= n0 && b <= n2) h1 += E[0][ea[0]][eb[0]][a] * E[2][ea[2]][eb[2]][b] * G[hidx(a, t, b)];
        if (t <= n2 + 1 && a <= n0 && b <= n1) h2 += E[0][ea[0]][eb[0]][a] * E[1][ea[1]][eb[1]][b] * G[hidx(a, b, t)];
      }
    H[0][t] = h0;
    H[1][t] = h1;
    H[2][t] = h2;
  }
  const int n[3] = {n0, n1, n2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double up = 0.0, dn = 0.0, upb = 0.0, dnb = 0.0;
#pragma unroll
    for (int t = 0; t <= NMAX + 1; ++t) {
      if (t > n[d] + 1) continue;
      up += E[d][ea[d] + 1][eb[d]][t] * H[d][t];
      if (ea[d] > 0 && t <= n[d] - 1) dn += E[d][ea[d] > 0 ? ea[d] - 1 : 0][eb[d]][t] * H[d][t];
      if constexpr (WANT_B) {
        upb += E[d][ea[d]][eb[d] + 1][t] * H[d][t];
        if (eb[d] > 0 && t <= n[d] - 1) dnb += E[d][ea[d]][eb[d] > 0 ? eb[d] - 1 : 0][t] * H[d][t];
      }
    }
    dA[d] = 2.0 * exa * up - ea[d] * dn;
    if constexpr (WANT_B) dB[d] = 2.0 * exb * upb - eb[d] * dnb;
  }
}

// Classes that run the fully unrolled body (La+Lb+Lc+Ld <= 4 and not a d-d
// bra: measured faster even where it spills); the others run the
// runtime-indexed body with one wave per ket primitive pair as well.
__host__ __device__ constexpr bool grad_unrolled(int LA, int LB, int LC, int LD) {
  return LA + LB + LC + LD <= 4 && ncart(LA) * ncart(LB) < 36;
}
template <int LA, int LB, int LC, int LD>
constexpr bool kGradUnrolled = grad_unrolled(LA, LB, LC, LD);

// Lanes sharing one quartet: 8 for (dd|dd) and (dd|dp), whose few, long
// quartets leave most of the GPU idle otherwise; 1 elsewhere (measured: the
// redundant per-lane G intermediates cost more than they gain there).
template <int LA, int LB, int LC, int LD>
constexpr int kGradLanesPerQuartet = LA + LB + LC + LD >= 7 ? 8 : 1;

// Per-class instance: private arrays sized for the class (the largest,
// (dd|dd), needs about 10 KB per lane; the s/p classes stay in registers).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(kGradThreads) void k_grad_2e(ClassLaunch c, Grad2eArgs g, int maxnp, int maxnk) {
  extern __shared__ double s_grad[];
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  constexpr int NEB = (LA + 2) * (LB + 2) * (LAB + 3), NEK = (LC + 2) * (LD + 2) * (LCD + 3);
  const int n3 = 3 * g.natoms;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  const int M = g.M;
  // one wave = one work block (a bra and up to 64 kets, one per lane) and one
  // primitive pair of the bra; the loop bounds are wave-uniform
  // the heavy (runtime-indexed) classes also give each ket primitive pair its
  // own wave, for parallelism: their quartets are few and long
  constexpr bool KSPLIT = !kGradUnrolled<LA, LB, LC, LD>;
  // ... and for the heaviest classes QL lanes share one quartet (its bra
  // components ab = sub, sub + QL, ...): 64 / QL kets per wave
  constexpr int QL = kGradLanesPerQuartet<LA, LB, LC, LD>;
  const int nk = KSPLIT ? maxnk : 1;
  const long long nvirt = (long long)c.nblocks * maxnp * nk * QL;
  for (long long vb = (long long)blockIdx.x * nw + wave; vb < nvirt; vb += (long long)gridDim.x * nw) {
    const int chunk = (int)(vb % QL);
    const long long vq = vb / QL;
    const int ikp = (int)(vq % nk);
    const int b = (int)(vq / nk / maxnp), ibp = (int)((vq / nk) % maxnp);
    const int kq = chunk * (64 / QL) + lane / QL;  // this lane's ket in the block
    const int sub = lane % QL;
    const BlockRec blk = c.blocks[b];
    const int bra_id = c.bra_list[blk.bra_pos];
    const SPair bra = g.sp[bra_id];
    if (ibp >= bra.nprim || chunk * (64 / QL) >= blk.count) continue;  // wave-uniform
    const bool in_block = kq < blk.count;
    const int ket_id = c.ket_list[blk.ket_start + (in_block ? kq : 0)];
    const SPair ket = g.sp[ket_id];
    const bool active = in_block && (!KSPLIT || ikp < ket.nprim);
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
    // the bra primitive pair's E tables are wave-uniform, so three lanes
    // build them once into the wave's LDS slot and every lane reads them
    // there (instead of 64 private copies in scratch or registers)
    double* s_Eb = s_grad + 3 * g.natoms + wave * 3 * NEB;
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane < 3) {
        const PairRec pb = g.pr[bra.prim_begin + ibp];
        const double PA = lane == 0 ? pb.PA[0] : lane == 1 ? pb.PA[1] : pb.PA[2];
        const double PB = lane == 0 ? pb.PB[0] : lane == 1 ? pb.PB[1] : pb.PB[2];
        hermite_e_rt<double>(LA + 1, LB + 1, pb.p, PA, PB, s_Eb + lane * NEB);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (active) {
      // density blocks of Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)
      double dab[NAB], dcd[NCD], dac[NA * NC], dbd[NB * ND], dad[NA * ND], dbc[NB * NC];
      {
        const double* D = g.Dp;
        const size_t ra = bra.bi, rb = bra.bj, rc = ket.bi, rd = ket.bj;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
          for (int bb = 0; bb < NB; ++bb) dab[a * NB + bb] = D[(ra + a) * M + rb + bb];
#pragma unroll
          for (int cc = 0; cc < NC; ++cc) dac[a * NC + cc] = D[(ra + a) * M + rc + cc];
#pragma unroll
          for (int dd = 0; dd < ND; ++dd) dad[a * ND + dd] = D[(ra + a) * M + rd + dd];
        }
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
          for (int dd = 0; dd < ND; ++dd) dbd[bb * ND + dd] = D[(rb + bb) * M + rd + dd];
#pragma unroll
          for (int cc = 0; cc < NC; ++cc) dbc[bb * NC + cc] = D[(rb + bb) * M + rc + cc];
        }
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
          for (int dd = 0; dd < ND; ++dd) dcd[cc * ND + dd] = D[(rc + cc) * M + rd + dd];
      }
      auto gamma = [&](int ab, int cd) {
        const int a = ab / NB, bb = ab % NB, cc = cd / ND, dd = cd % ND;
        return 0.5 * dab[ab] * dcd[cd] - 0.125 * (dac[a * NC + cc] * dbd[bb * ND + dd] + dad[a * ND + dd] * dbc[bb * NC + cc]);
      };
      if constexpr (kGradUnrolled<LA, LB, LC, LD>) {
        const PairRec pb = g.pr[bra.prim_begin + ibp];
        const double exa = g.alpha[bra.prim_begin + ibp], exb = pb.p - exa;
        const auto& Eb = *reinterpret_cast<const double(*)[3][LA + 2][LB + 2][LAB + 3]>(s_Eb);
        for (int ik = 0; ik < ket.nprim; ++ik) {
          const PairRec pk = g.pr[ket.prim_begin + ik];
          if (pb.Q * pk.Q < g.thr) continue;
          const double exc = g.alpha[ket.prim_begin + ik];
          double Ek[3][LC + 2][LD + 2][LCD + 3];
#pragma unroll
          for (int d = 0; d < 3; ++d) hermite_e_t<LC + 1, LD + 1, double>(pk.p, pk.PA[d], pk.PB[d], Ek[d]);
          const double p = pb.p, q = pk.p;
          const double alpha = p * q / (p + q);
          const double pref = kTwoPi25 / (p * q * sqrt(p + q)) * pb.Kc * pk.Kc;
          double R[nherm(L + 1)];
          hermite_r_t<L + 1, double>(alpha, pb.P[0] - pk.P[0], pb.P[1] - pk.P[1], pb.P[2] - pk.P[2], pref, g.boys, R);
#pragma unroll
          for (int cd = 0; cd < NCD; ++cd) {
            const int ce[3] = {cart_exp(LC, cd / ND, 0), cart_exp(LC, cd / ND, 1), cart_exp(LC, cd / ND, 2)};
            const int de[3] = {cart_exp(LD, cd % ND, 0), cart_exp(LD, cd % ND, 1), cart_exp(LD, cd % ND, 2)};
            double G[nherm(LAB + 1)];
#pragma unroll
            for (int t = 0; t <= LAB + 1; ++t)
#pragma unroll
              for (int u = 0; u <= LAB + 1; ++u)
#pragma unroll
                for (int v = 0; v <= LAB + 1; ++v) {
                  if (t + u + v > LAB + 1) continue;
                  double sacc = 0.0;
#pragma unroll
                  for (int tt = 0; tt <= LCD; ++tt) {
                    if (tt > ce[0] + de[0]) continue;
#pragma unroll
                    for (int uu = 0; uu <= LCD; ++uu) {
                      if (uu > ce[1] + de[1]) continue;
                      const double exy = Ek[0][ce[0]][de[0]][tt] * Ek[1][ce[1]][de[1]][uu];
#pragma unroll
                      for (int vv = 0; vv <= LCD; ++vv) {
                        if (vv > ce[2] + de[2]) continue;
                        const double e = ((tt + uu + vv) & 1) ? -exy * Ek[2][ce[2]][de[2]][vv] : exy * Ek[2][ce[2]][de[2]][vv];
                        sacc += e * R[hidx(t + tt, u + uu, v + vv)];
                      }
                    }
                  }
                  G[hidx(t, u, v)] = sacc;
                }
#pragma unroll
            for (int ab = 0; ab < NAB; ++ab) {
              const double gm = gamma(ab, cd);
              const int ae[3] = {cart_exp(LA, ab / NB, 0), cart_exp(LA, ab / NB, 1), cart_exp(LA, ab / NB, 2)};
              const int be[3] = {cart_exp(LB, ab % NB, 0), cart_exp(LB, ab % NB, 1), cart_exp(LB, ab % NB, 2)};
              double dA[3], dB[3];
              pair_grad_u<true, LA + 1, LB + 1, LAB + 3, LAB>(Eb, ae, be, exa, exb, G, dA, dB);
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                gA[d] += gm * dA[d];
                gB[d] += gm * dB[d];
              }
            }
          }
#pragma unroll
          for (int ab = 0; ab < NAB; ++ab) {
            const int ae[3] = {cart_exp(LA, ab / NB, 0), cart_exp(LA, ab / NB, 1), cart_exp(LA, ab / NB, 2)};
            const int be[3] = {cart_exp(LB, ab % NB, 0), cart_exp(LB, ab % NB, 1), cart_exp(LB, ab % NB, 2)};
            double H[nherm(LCD + 1)];
#pragma unroll
            for (int t = 0; t <= LCD + 1; ++t)
#pragma unroll
              for (int u = 0; u <= LCD + 1; ++u)
#pragma unroll
                for (int v = 0; v <= LCD + 1; ++v) {
                  if (t + u + v > LCD + 1) continue;
                  double sacc = 0.0;
#pragma unroll
                  for (int tt = 0; tt <= LAB; ++tt) {
                    if (tt > ae[0] + be[0]) continue;
#pragma unroll
                    for (int uu = 0; uu <= LAB; ++uu) {
                      if (uu > ae[1] + be[1]) continue;
                      const double exy = Eb[0][ae[0]][be[0]][tt] * Eb[1][ae[1]][be[1]][uu];
#pragma unroll
                      for (int vv = 0; vv <= LAB; ++vv) {
                        if (vv > ae[2] + be[2]) continue;
                        sacc += exy * Eb[2][ae[2]][be[2]][vv] * R[hidx(t + tt, u + uu, v + vv)];
                      }
                    }
                  }
                  H[hidx(t, u, v)] = ((t + u + v) & 1) ? -sacc : sacc;
                }
#pragma unroll
            for (int cd = 0; cd < NCD; ++cd) {
              const double gm = gamma(ab, cd);
              const int ce[3] = {cart_exp(LC, cd / ND, 0), cart_exp(LC, cd / ND, 1), cart_exp(LC, cd / ND, 2)};
              const int de[3] = {cart_exp(LD, cd % ND, 0), cart_exp(LD, cd % ND, 1), cart_exp(LD, cd % ND, 2)};
              double dC[3], unused[3];
              pair_grad_u<false, LC + 1, LD + 1, LCD + 3, LCD>(Ek, ce, de, exc, 0.0, H, dC, unused);
#pragma unroll
              for (int d = 0; d < 3; ++d) gC[d] += gm * dC[d];
            }
          }
        }
      } else {
        const int ib = ibp;
        const PairRec pb = g.pr[bra.prim_begin + ib];
        const double exa = g.alpha[bra.prim_begin + ib], exb = pb.p - exa;
        const ETab TB{{s_Eb, s_Eb + NEB, s_Eb + 2 * NEB}, LB + 2, LAB + 3};
        for (int ik = KSPLIT ? ikp : 0; ik < (KSPLIT ? ikp + 1 : ket.nprim); ++ik) {
          const PairRec pk = g.pr[ket.prim_begin + ik];
          if (pb.Q * pk.Q < g.thr) continue;
          const double exc = g.alpha[ket.prim_begin + ik];
          double Ek[3][NEK];
          for (int d = 0; d < 3; ++d) hermite_e_rt<double>(LC + 1, LD + 1, pk.p, pk.PA[d], pk.PB[d], Ek[d]);
          const ETab TK{{Ek[0], Ek[1], Ek[2]}, LD + 2, LCD + 3};
          const double p = pb.p, q = pk.p;
          const double alpha = p * q / (p + q);
          const double pref = kTwoPi25 / (p * q * sqrt(p + q)) * pb.Kc * pk.Kc;
          double R[nherm(L + 1)];
          hermite_r_rt<double>(L + 1, alpha, pb.P[0] - pk.P[0], pb.P[1] - pk.P[1], pb.P[2] - pk.P[2], pref, g.boys, R);
          // bra centres
          for (int cd = 0; cd < NCD; ++cd) {
            int ce[3], de[3];
            for (int d = 0; d < 3; ++d) { ce[d] = cart_exp(LC, cd / ND, d); de[d] = cart_exp(LD, cd % ND, d); }
            const double* Kx = TK.row(0, ce[0], de[0]);
            const double* Ky = TK.row(1, ce[1], de[1]);
            const double* Kz = TK.row(2, ce[2], de[2]);
            double G[nherm(LAB + 1)];
            for (int t = 0; t <= LAB + 1; ++t)
              for (int u = 0; u + t <= LAB + 1; ++u)
                for (int v = 0; v + u + t <= LAB + 1; ++v) {
                  double sacc = 0.0;
                  for (int tt = 0; tt <= ce[0] + de[0]; ++tt)
                    for (int uu = 0; uu <= ce[1] + de[1]; ++uu) {
                      const double exy = Kx[tt] * Ky[uu];
                      for (int vv = 0; vv <= ce[2] + de[2]; ++vv) {
                        const double e = ((tt + uu + vv) & 1) ? -exy * Kz[vv] : exy * Kz[vv];
                        sacc += e * R[hidx(t + tt, u + uu, v + vv)];
                      }
                    }
                  G[hidx(t, u, v)] = sacc;
                }
            for (int ab = sub; ab < NAB; ab += QL) {
              const double gm = gamma(ab, cd);
              if (gm == 0.0) continue;
              int ae[3], be[3];
              for (int d = 0; d < 3; ++d) { ae[d] = cart_exp(LA, ab / NB, d); be[d] = cart_exp(LB, ab % NB, d); }
              double dA[3], dB[3];
              pair_grad(TB, ae, be, exa, exb, G, true, dA, dB);
              for (int d = 0; d < 3; ++d) {
                gA[d] += gm * dA[d];
                gB[d] += gm * dB[d];
              }
            }
          }
          // first ket centre
          for (int ab = sub; ab < NAB; ab += QL) {
            int ae[3], be[3];
            for (int d = 0; d < 3; ++d) { ae[d] = cart_exp(LA, ab / NB, d); be[d] = cart_exp(LB, ab % NB, d); }
            const double* Bx = TB.row(0, ae[0], be[0]);
            const double* By = TB.row(1, ae[1], be[1]);
            const double* Bz = TB.row(2, ae[2], be[2]);
            double H[nherm(LCD + 1)];
            for (int t = 0; t <= LCD + 1; ++t)
              for (int u = 0; u + t <= LCD + 1; ++u)
                for (int v = 0; v + u + t <= LCD + 1; ++v) {
                  double sacc = 0.0;
                  for (int tt = 0; tt <= ae[0] + be[0]; ++tt)
                    for (int uu = 0; uu <= ae[1] + be[1]; ++uu) {
                      const double exy = Bx[tt] * By[uu];
                      for (int vv = 0; vv <= ae[2] + be[2]; ++vv) sacc += exy * Bz[vv] * R[hidx(t + tt, u + uu, v + vv)];
                    }
                  H[hidx(t, u, v)] = ((t + u + v) & 1) ? -sacc : sacc;
                }
            for (int cd = 0; cd < NCD; ++cd) {
              const double gm = gamma(ab, cd);
              if (gm == 0.0) continue;
              int ce[3], de[3];
              for (int d = 0; d < 3; ++d) { ce[d] = cart_exp(LC, cd / ND, d); de[d] = cart_exp(LD, cd % ND, d); }
              double dC[3], unused[3];
              pair_grad(TK, ce, de, exc, 0.0, H, false, dC, unused);
              for (int d = 0; d < 3; ++d) gC[d] += gm * dC[d];
            }
          }
        }
      }
    }
    const double deg = (bra.si != bra.sj ? 2.0 : 1.0) * (ket.si != ket.sj ? 2.0 : 1.0) * (bra_id != ket_id ? 2.0 : 1.0);
    const int atA = g.shell_atom[bra.si], atB = g.shell_atom[bra.sj];
    for (int d = 0; d < 3; ++d) {
      const double sa = wave_sum_g(gA[d] * deg);  // the bra is wave-uniform
      const double sb = wave_sum_g(gB[d] * deg);
      if (lane == 0) {
        atomicAdd(&s_grad[3 * atA + d], sa);
        atomicAdd(&s_grad[3 * atB + d], sb);
      }
    }
    if (active) {
      const int atC = g.shell_atom[ket.si], atD = g.shell_atom[ket.sj];
      for (int d = 0; d < 3; ++d) {
        atomicAdd(&s_grad[3 * atC + d], gC[d] * deg);
        atomicAdd(&s_grad[3 * atD + d], -(gA[d] + gB[d] + gC[d]) * deg);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&g.grad[i], s_grad[i]);
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_grad_2e_t(const ClassLaunch& c, const Grad2eArgs& a, int maxnp, int maxnk, hipStream_t s) {
  const int waves = kGradThreads / 64;
  const long long nvirt = (long long)c.nblocks * maxnp * (kGradUnrolled<LA, LB, LC, LD> ? 1 : maxnk) *
                         kGradLanesPerQuartet<LA, LB, LC, LD>;
  const int grid = (int)std::min<long long>(2048, (nvirt + waves - 1) / waves);
  constexpr int NEB = (LA + 2) * (LB + 2) * (LA + LB + 3);
  const size_t lds = (3 * (size_t)a.natoms + (size_t)waves * 3 * NEB) * sizeof(double);
  hipLaunchKernelGGL((k_grad_2e<LA, LB, LC, LD>), dim3(grid), dim3(kGradThreads), lds, s, c, a, maxnp, maxnk);
  return hipGetLastError();
}


// Dispatch of one quartet class key ((LA*3+LB)*3+LC)*3+LD to its instance;
// false when this translation unit does not hold the class.
#define GANSU_GRAD_CASE(a_, b_, c_, d_)                                               \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_:                                            \
    *err = launch_grad_2e_t<a_, b_, c_, d_>(c, a, maxnp, maxnk, s);                  \
    return true;

}  // namespace grad_detail
}  // namespace gansu
