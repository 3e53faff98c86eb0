// Backward kernels of the RHF energy: the analytic nuclear gradient from
// McMurchie-Davidson derivative integrals (the reference has no gradient).
//
// A centre derivative of a Cartesian Gaussian is a two-term combination of
// Gaussians with one exponent raised or lowered,
//   d/dA_x [x_A^i exp(-a x_A^2)] = 2a x_A^(i+1) exp(..) - i x_A^(i-1) exp(..),
// so with Hermite E tables built one order higher than the energy needs and R
// one order higher, every derivative integral is a contraction of the same
// intermediates.
//
// Two-electron part: one lane = one contracted shell quartet, over the same
// class-sorted work blocks as the direct Fock build.  Per primitive quartet:
//   G_cd(tuv) = sum (-1)^(tau+nu+phi) E^cd_(tau nu phi) R_(t+tau, u+nu, v+phi)
// gives both bra-centre derivatives, the mirrored H_ab(tau nu phi) gives the
// first ket centre, and the last centre follows from translational invariance
// (the four centre derivatives of any primitive integral sum to zero).  The
// density enters through Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)
// and the 8-fold degeneracy of the quartet.  Four waves per workgroup walk the
// blocks; per-atom accumulators live in LDS and reach HBM once per workgroup.
#include <hip/hip_runtime.h>

#include "gradient_kernels.hpp"
#include "md_device.hpp"

namespace gansu {

using namespace dev;

namespace {

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

__device__ inline double pair_int(const ETab& T, const int* ea, const int* eb, const double* G) {
  return contract_e(T.row(0, ea[0], eb[0]), ea[0] + eb[0], T.row(1, ea[1], eb[1]), ea[1] + eb[1],
                    T.row(2, ea[2], eb[2]), ea[2] + eb[2], G);
}

// Derivative along direction d of the first (which = 0) or second centre of
// the pair, exponent ex of that primitive.
__device__ inline double pair_deriv(const ETab& T, int which, int d, double ex, int* ea, int* eb, const double* G) {
  int* e = which == 0 ? ea : eb;
  const int e0 = e[d];
  e[d] = e0 + 1;
  double v = 2.0 * ex * pair_int(T, ea, eb, G);
  if (e0 > 0) {
    e[d] = e0 - 1;
    v -= e0 * pair_int(T, ea, eb, G);
  }
  e[d] = e0;
  return v;
}

// Per-class instance: private arrays sized for the class (the largest,
// (dd|dd), needs about 10 KB per lane; the s/p classes stay in registers).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(kGradThreads) void k_grad_2e(ClassLaunch c, Grad2eArgs g) {
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
  for (int b = blockIdx.x * nw + wave; b < c.nblocks; b += gridDim.x * nw) {  // wave-uniform loop
    const BlockRec blk = c.blocks[b];
    const bool active = lane < blk.count;
    const int bra_id = c.bra_list[blk.bra_pos];
    const int ket_id = c.ket_list[blk.ket_start + (active ? lane : 0)];
    const SPair bra = g.sp[bra_id];
    const SPair ket = g.sp[ket_id];
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, gC[3] = {0.0, 0.0, 0.0};
    if (active) {
      // density blocks of Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)
      double dab[NAB], dcd[NCD], dac[NA * NC], dbd[NB * ND], dad[NA * ND], dbc[NB * NC];
      {
        const double* D = g.Dp;
        const size_t ra = bra.bi, rb = bra.bj, rc = ket.bi, rd = ket.bj;
        for (int a = 0; a < NA; ++a) {
          for (int bb = 0; bb < NB; ++bb) dab[a * NB + bb] = D[(ra + a) * M + rb + bb];
          for (int cc = 0; cc < NC; ++cc) dac[a * NC + cc] = D[(ra + a) * M + rc + cc];
          for (int dd = 0; dd < ND; ++dd) dad[a * ND + dd] = D[(ra + a) * M + rd + dd];
        }
        for (int bb = 0; bb < NB; ++bb) {
          for (int dd = 0; dd < ND; ++dd) dbd[bb * ND + dd] = D[(rb + bb) * M + rd + dd];
          for (int cc = 0; cc < NC; ++cc) dbc[bb * NC + cc] = D[(rb + bb) * M + rc + cc];
        }
        for (int cc = 0; cc < NC; ++cc)
          for (int dd = 0; dd < ND; ++dd) dcd[cc * ND + dd] = D[(rc + cc) * M + rd + dd];
      }
      auto gamma = [&](int ab, int cd) {
        const int a = ab / NB, bb = ab % NB, cc = cd / ND, dd = cd % ND;
        return 0.5 * dab[ab] * dcd[cd] - 0.125 * (dac[a * NC + cc] * dbd[bb * ND + dd] + dad[a * ND + dd] * dbc[bb * NC + cc]);
      };
      for (int ib = 0; ib < bra.nprim; ++ib) {
        const PairRec pb = g.pr[bra.prim_begin + ib];
        const double exa = g.alpha[bra.prim_begin + ib], exb = pb.p - exa;
        double Eb[3][NEB];
        for (int d = 0; d < 3; ++d) hermite_e_rt<double>(LA + 1, LB + 1, pb.p, pb.PA[d], pb.PB[d], Eb[d]);
        const ETab TB{{Eb[0], Eb[1], Eb[2]}, LB + 2, LAB + 3};
        for (int ik = 0; ik < ket.nprim; ++ik) {
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
            for (int ab = 0; ab < NAB; ++ab) {
              const double gm = gamma(ab, cd);
              if (gm == 0.0) continue;
              int ae[3], be[3];
              for (int d = 0; d < 3; ++d) { ae[d] = cart_exp(LA, ab / NB, d); be[d] = cart_exp(LB, ab % NB, d); }
              for (int d = 0; d < 3; ++d) {
                gA[d] += gm * pair_deriv(TB, 0, d, exa, ae, be, G);
                gB[d] += gm * pair_deriv(TB, 1, d, exb, ae, be, G);
              }
            }
          }
          // first ket centre
          for (int ab = 0; ab < NAB; ++ab) {
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
              for (int d = 0; d < 3; ++d) gC[d] += gm * pair_deriv(TK, 0, d, exc, ce, de, H);
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
hipError_t launch_grad_2e_t(const ClassLaunch& c, const Grad2eArgs& a, hipStream_t s) {
  const int waves = kGradThreads / 64;
  const int grid = std::min(2048, (c.nblocks + waves - 1) / waves);
  hipLaunchKernelGGL((k_grad_2e<LA, LB, LC, LD>), dim3(grid), dim3(kGradThreads), 3 * a.natoms * sizeof(double), s,
                     c, a);
  return hipGetLastError();
}

struct DPrimG {  // == gansu_primitive_shell layout
  double exponent, coefficient, x, y, z;
  int shell_type;
  size_t basis_index;
  int atom_index;
};

constexpr int kG1MaxE = 4 * 6 * 9;  // (la+2)(lb+4)(la+lb+5), la, lb <= 2

// One lane = one basis-function pair mu >= nu; derivatives of S, T and V
// with respect to both basis centres, and the Hellmann-Feynman term of every
// nucleus (-(d/dA + d/dB) of that nucleus' attraction integral).
__global__ __launch_bounds__(kGradThreads) void k_grad_1e(OneIntArgs a, const double* __restrict__ Dm,
                                                          const double* __restrict__ Wm, double* grad) {
  extern __shared__ double s_grad[];
  const int n3 = 3 * a.natoms;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = a.M;
  const int mu = id < (size_t)M * M ? (int)(id / M) : 0, nu = id < (size_t)M * M ? (int)(id % M) : 1;
  if (id < (size_t)M * M && nu <= mu) {
    const int sa = a.bf_shell[mu], sb = a.bf_shell[nu];
    const int ca = mu - a.shell_basis[sa], cb = nu - a.shell_basis[sb];
    const int la = a.shell_L[sa], lb = a.shell_L[sb];
    int ea[3], eb[3];
    for (int d = 0; d < 3; ++d) { ea[d] = cart_exp(la, ca, d); eb[d] = cart_exp(lb, cb, d); }
    const double ww = a.w[mu] * a.w[nu] * (mu == nu ? 1.0 : 2.0);
    const double dd = Dm[(size_t)mu * M + nu] * ww, wd = Wm[(size_t)mu * M + nu] * ww;
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    int atA = 0, atB = 0;
    const int nj = lb + 4, ntt = la + lb + 5;
    for (int ia = a.shell_off[sa]; ia < a.shell_off[sa + 1]; ++ia) {
      const DPrimG pa = ((const DPrimG*)a.prims)[a.shell_prims[ia]];
      atA = pa.atom_index;
      for (int ib = a.shell_off[sb]; ib < a.shell_off[sb + 1]; ++ib) {
        const DPrimG pb = ((const DPrimG*)a.prims)[a.shell_prims[ib]];
        atB = pb.atom_index;
        const double al = pa.exponent, be = pb.exponent, p = al + be;
        const double A[3] = {pa.x, pa.y, pa.z}, B[3] = {pb.x, pb.y, pb.z};
        double P[3], r2 = 0.0;
        double E[3][kG1MaxE];
        for (int d = 0; d < 3; ++d) {
          P[d] = (al * A[d] + be * B[d]) / p;
          r2 += (A[d] - B[d]) * (A[d] - B[d]);
          hermite_e_rt<double>(la + 1, lb + 3, p, P[d] - A[d], P[d] - B[d], E[d]);
        }
        auto row = [&](int d, int i, int j) -> const double* { return E[d] + (i * nj + j) * ntt; };
        auto s1 = [&](int d, int i, int j) { return row(d, i, j)[0]; };
        auto t1 = [&](int d, int i, int j) {
          return -2.0 * be * be * row(d, i, j + 2)[0] + be * (2 * j + 1) * row(d, i, j)[0] -
                 (j >= 2 ? 0.5 * j * (j - 1) * row(d, i, j - 2)[0] : 0.0);
        };
        const double K = exp(-al * be / p * r2);
        const double na = 0.71270547035499016035 * double(1 << la) * pow(al, (2.0 * la + 3.0) / 4.0);
        const double nb = 0.71270547035499016035 * double(1 << lb) * pow(be, (2.0 * lb + 3.0) / 4.0);
        const double pf = pa.coefficient * pb.coefficient * na * nb * K;
        const double ovl = pow(kPi / p, 1.5) * pf;
        double s0[3], t0[3];
        for (int d = 0; d < 3; ++d) { s0[d] = s1(d, ea[d], eb[d]); t0[d] = t1(d, ea[d], eb[d]); }
        // overlap and kinetic
        for (int k = 0; k < 2; ++k) {
          const double ex = k == 0 ? al : be;
          for (int d = 0; d < 3; ++d) {
            const int e0 = k == 0 ? ea[d] : eb[d];
            const int iu = k == 0 ? ea[d] + 1 : ea[d], ju = k == 0 ? eb[d] : eb[d] + 1;
            double ds = 2.0 * ex * s1(d, iu, ju), dt = 2.0 * ex * t1(d, iu, ju);
            if (e0 > 0) {
              const int il = k == 0 ? ea[d] - 1 : ea[d], jl = k == 0 ? eb[d] : eb[d] - 1;
              ds -= e0 * s1(d, il, jl);
              dt -= e0 * t1(d, il, jl);
            }
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            const double dS = ds * s0[d1] * s0[d2] * ovl;
            const double dT = (dt * s0[d1] * s0[d2] + ds * (t0[d1] * s0[d2] + s0[d1] * t0[d2])) * ovl;
            const double v = dd * dT - wd * dS;
            if (k == 0) gA[d] += v; else gB[d] += v;
          }
        }
        // nuclear attraction
        const int L = la + lb;
        for (int c = 0; c < a.natoms; ++c) {
          double R[nherm(5)];
          hermite_r_rt<double>(L + 1, p, P[0] - a.atom_xyz[3 * c], P[1] - a.atom_xyz[3 * c + 1],
                               P[2] - a.atom_xyz[3 * c + 2], 1.0, a.boys, R);
          const double vf = -a.Z[c] * 2.0 * kPi / p * pf * dd;
          double hf[3] = {0.0, 0.0, 0.0};
          for (int k = 0; k < 2; ++k) {
            const double ex = k == 0 ? al : be;
            for (int d = 0; d < 3; ++d) {
              int xa[3] = {ea[0], ea[1], ea[2]}, xb[3] = {eb[0], eb[1], eb[2]};
              int* e = k == 0 ? xa : xb;
              const int e0 = e[d];
              e[d] = e0 + 1;
              double v = 2.0 * ex *
                         contract_e(row(0, xa[0], xb[0]), xa[0] + xb[0], row(1, xa[1], xb[1]), xa[1] + xb[1],
                                    row(2, xa[2], xb[2]), xa[2] + xb[2], R);
              if (e0 > 0) {
                e[d] = e0 - 1;
                v -= e0 * contract_e(row(0, xa[0], xb[0]), xa[0] + xb[0], row(1, xa[1], xb[1]), xa[1] + xb[1],
                                     row(2, xa[2], xb[2]), xa[2] + xb[2], R);
              }
              v *= vf;
              if (k == 0) gA[d] += v; else gB[d] += v;
              hf[d] -= v;
            }
          }
          for (int d = 0; d < 3; ++d) atomicAdd(&s_grad[3 * c + d], hf[d]);
        }
      }
    }
    for (int d = 0; d < 3; ++d) {
      atomicAdd(&s_grad[3 * atA + d], gA[d]);
      atomicAdd(&s_grad[3 * atB + d], gB[d]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&grad[i], s_grad[i]);
}

}  // namespace

hipError_t launch_grad_2e(const ClassLaunch& c, const Grad2eArgs& a, hipStream_t s) {
  if (c.nblocks <= 0) return hipSuccess;
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  const int key = ((c.LA * 3 + c.LB) * 3 + c.LC) * 3 + c.LD;
#define GANSU_G(a_, b_, c_, d_) \
  case ((a_ * 3 + b_) * 3 + c_) * 3 + d_: return launch_grad_2e_t<a_, b_, c_, d_>(c, a, s);
  switch (key) {
    GANSU_G(0, 0, 0, 0) GANSU_G(1, 0, 0, 0) GANSU_G(1, 0, 1, 0) GANSU_G(1, 1, 0, 0)
    GANSU_G(1, 1, 1, 0) GANSU_G(1, 1, 1, 1) GANSU_G(2, 0, 0, 0) GANSU_G(2, 0, 1, 0)
    GANSU_G(2, 0, 1, 1) GANSU_G(2, 0, 2, 0) GANSU_G(2, 1, 0, 0) GANSU_G(2, 1, 1, 0)
    GANSU_G(2, 1, 1, 1) GANSU_G(2, 1, 2, 0) GANSU_G(2, 1, 2, 1) GANSU_G(2, 2, 0, 0)
    GANSU_G(2, 2, 1, 0) GANSU_G(2, 2, 1, 1) GANSU_G(2, 2, 2, 0) GANSU_G(2, 2, 2, 1)
    GANSU_G(2, 2, 2, 2)
    default: return hipErrorInvalidValue;
  }
#undef GANSU_G
}

hipError_t launch_grad_1e(const OneIntArgs& a, const double* D, const double* W, double* grad, hipStream_t s) {
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  const size_t n = (size_t)a.M * a.M;
  hipLaunchKernelGGL(k_grad_1e, dim3((unsigned)((n + kGradThreads - 1) / kGradThreads)), dim3(kGradThreads),
                     3 * a.natoms * sizeof(double), s, a, D, W, grad);
  return hipGetLastError();
}

}  // namespace gansu
