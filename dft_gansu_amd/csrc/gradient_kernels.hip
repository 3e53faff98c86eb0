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
// Two-electron part (template in gradient_2e_impl.hpp, instances in
// gradient_2e_*.hip): one wave = one work block of the direct Fock build (a
// bra shell pair and up to 64 kets, one per lane) and one primitive pair of
// the bra (and, for the heavy classes, one ket primitive pair per wave); the
// classes with La+Lb+Lc+Ld <= 4 other than (dd|ss) run a fully unrolled body
// whose tables live in registers.  Per primitive quartet:
//   G_cd(tuv) = sum (-1)^(tau+nu+phi) E^cd_(tau nu phi) R_(t+tau, u+nu, v+phi)
// gives both bra-centre derivatives, the mirrored H_ab(tau nu phi) gives the
// first ket centre, and the last centre follows from translational invariance
// (the four centre derivatives of any primitive integral sum to zero).  The
// density enters through Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)
// and the 8-fold degeneracy of the quartet.  Four waves per workgroup walk the
// work items; per-atom accumulators live in LDS and reach HBM once per
// workgroup.  This file holds the one-electron kernel and the dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gradient_2e_impl.hpp"
#include "gradient_kernels.hpp"
#include "md_device.hpp"

namespace gansu {

using namespace dev;

using namespace grad_detail;

namespace {

struct DPrimG {  // == gansu_primitive_shell layout
  double exponent, coefficient, x, y, z;
  int shell_type;
  size_t basis_index;
  int atom_index;
};

constexpr int kG1MaxE = 4 * 6 * 9;  // (la+2)(lb+4)(la+lb+5), la, lb <= 2
// A primitive pair whose prefactor times the density weights falls below this
// contributes nothing at FP64 precision to any component (the Gaussian
// product factor exp(-ab/p r^2) makes most pairs of a large molecule vanish).
constexpr double kG1Screen = 1e-17;

// Overlap and kinetic part: one lane = one basis-function pair mu >= nu;
// derivatives of S and T with respect to both basis centres.
__global__ __launch_bounds__(kGradThreads) void k_grad_1e(OneIntArgs a, const double* __restrict__ Dm,
                                                          const double* __restrict__ Wm, double* grad) {
  extern __shared__ double s_grad[];
  const int n3 = 3 * a.natoms;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = a.M;
  const int mu = id < (size_t)M * M ? (int)(id / M) : 0, nu = id < (size_t)M * M ? (int)(id % M) : 1;
  const bool act = id < (size_t)M * M && nu <= mu;
  const int sa = act ? a.bf_shell[mu] : 0, sb = act ? a.bf_shell[nu] : 0;
  const int ca = act ? mu - a.shell_basis[sa] : 0, cb = act ? nu - a.shell_basis[sb] : 0;
  const int la = act ? a.shell_L[sa] : 0, lb = act ? a.shell_L[sb] : 0;
  int ea[3], eb[3];
  for (int d = 0; d < 3; ++d) { ea[d] = cart_exp(la, ca, d); eb[d] = cart_exp(lb, cb, d); }
  const double ww = act ? a.w[mu] * a.w[nu] * (mu == nu ? 1.0 : 2.0) : 0.0;
  const double dd = act ? Dm[(size_t)mu * M + nu] * ww : 0.0, wd = act ? Wm[(size_t)mu * M + nu] * ww : 0.0;
  double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
  int atA = 0, atB = 0;
  const int pa0 = act ? a.shell_off[sa] : 0, pa1 = act ? a.shell_off[sa + 1] : 0;
  const int pb0 = act ? a.shell_off[sb] : 0, pb1 = act ? a.shell_off[sb + 1] : 0;
  auto prim_pair = [&](int ia, int ib, double* P, double& al, double& be, double& p, double& pf) {
    const DPrimG pa = ((const DPrimG*)a.prims)[a.shell_prims[ia]];
    const DPrimG pb = ((const DPrimG*)a.prims)[a.shell_prims[ib]];
    atA = pa.atom_index;
    atB = pb.atom_index;
    al = pa.exponent;
    be = pb.exponent;
    p = al + be;
    const double A[3] = {pa.x, pa.y, pa.z}, B[3] = {pb.x, pb.y, pb.z};
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      P[3 + d] = A[d];
      P[6 + d] = B[d];
      P[d] = (al * A[d] + be * B[d]) / p;
      r2 += (A[d] - B[d]) * (A[d] - B[d]);
    }
    const double K = exp(-al * be / p * r2);
    const double na = 0.71270547035499016035 * double(1 << la) * pow(al, (2.0 * la + 3.0) / 4.0);
    const double nb = 0.71270547035499016035 * double(1 << lb) * pow(be, (2.0 * lb + 3.0) / 4.0);
    pf = pa.coefficient * pb.coefficient * na * nb * K;
  };
  // overlap and kinetic
  {
    const int nj = lb + 4, ntt = la + lb + 5;
    for (int ia = pa0; ia < pa1; ++ia)
      for (int ib = pb0; ib < pb1; ++ib) {
        double P[9], al, be, p, pf;
        prim_pair(ia, ib, P, al, be, p, pf);
        const double ovl = pow(kPi / p, 1.5) * pf;
        const double p1 = 1.0 + p;  // the derivatives carry up to three powers of the exponents
        if (fabs(ovl) * (fabs(dd) + fabs(wd)) * p1 * p1 * p1 < kG1Screen) continue;
        double E[3][kG1MaxE];
        for (int d = 0; d < 3; ++d) hermite_e_rt<double>(la + 1, lb + 3, p, P[d] - P[3 + d], P[d] - P[6 + d], E[d]);
        auto row = [&](int d, int i, int j) -> const double* { return E[d] + (i * nj + j) * ntt; };
        auto s1 = [&](int d, int i, int j) { return row(d, i, j)[0]; };
        auto t1 = [&](int d, int i, int j) {
          return -2.0 * be * be * row(d, i, j + 2)[0] + be * (2 * j + 1) * row(d, i, j)[0] -
                 (j >= 2 ? 0.5 * j * (j - 1) * row(d, i, j - 2)[0] : 0.0);
        };
        double s0[3], t0[3];
        for (int d = 0; d < 3; ++d) { s0[d] = s1(d, ea[d], eb[d]); t0[d] = t1(d, ea[d], eb[d]); }
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
      }
  }
  if (act)
    for (int d = 0; d < 3; ++d) {
      atomicAdd(&s_grad[3 * atA + d], gA[d]);
      atomicAdd(&s_grad[3 * atB + d], gB[d]);
    }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&grad[i], s_grad[i]);
}

// Nuclear-attraction part, one instance per shell-pair class (LA >= LB): a
// wave takes one shell pair and 64 nuclei, one per lane.  Per primitive pair
// (wave-uniform) three lanes build the E tables (one order higher) into the
// wave's LDS slot; each lane evaluates R for its nucleus and contracts every
// component pair with compile-time offsets.  The basis-centre derivatives go
// to the pair's atoms after a wave reduction; the Hellmann-Feynman term
// -(dA + dB) goes to the lane's own nucleus (distinct LDS addresses).
constexpr int kG1Waves = kGradThreads / 64;

template <int LA, int LB>
__global__ __launch_bounds__(kGradThreads) void k_grad_1e_nuc(OneIntArgs a, const int2* __restrict__ pairs,
                                                              int npairs, const double* __restrict__ Dm,
                                                              double* grad) {
  constexpr int NB = ncart(LB), NAB = ncart(LA) * NB, LAB = LA + LB;
  constexpr int NE = (LA + 2) * (LB + 2) * (LAB + 3);
  extern __shared__ double smem[];
  const int n3 = 3 * a.natoms;
  double* s_grad = smem;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* s_E = smem + n3 + wave * 3 * NE;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const int ngroups = (a.natoms + 63) >> 6;
  const long long nvirt = (long long)npairs * ngroups;
  const int M = a.M;
  const DPrimG* prims = (const DPrimG*)a.prims;
  for (long long vb = (long long)blockIdx.x * kG1Waves + wave; vb < nvirt; vb += (long long)gridDim.x * kG1Waves) {
    const int2 sp = pairs[vb / ngroups];
    const int c = (int)(vb % ngroups) * 64 + lane;
    const bool act = c < a.natoms;
    const int s1 = sp.x, s2 = sp.y;
    const int b1 = a.shell_basis[s1], b2 = a.shell_basis[s2];
    const double f = s1 == s2 ? 1.0 : 2.0;
    double dd[NAB], ddmax = 0.0;
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) {
      const int mu = b1 + ab / NB, nu = b2 + ab % NB;
      dd[ab] = Dm[(size_t)mu * M + nu] * a.w[mu] * a.w[nu] * f;
      ddmax = fmax(ddmax, fabs(dd[ab]));
    }
    const double Cx = act ? a.atom_xyz[3 * c] : 0.0, Cy = act ? a.atom_xyz[3 * c + 1] : 0.0,
                 Cz = act ? a.atom_xyz[3 * c + 2] : 0.0;
    const double Zc = act ? (double)a.Z[c] : 0.0;
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0}, hf[3] = {0.0, 0.0, 0.0};
    int atA = 0, atB = 0;
    for (int ia = a.shell_off[s1]; ia < a.shell_off[s1 + 1]; ++ia)
      for (int ib = a.shell_off[s2]; ib < a.shell_off[s2 + 1]; ++ib) {  // wave-uniform
        const DPrimG pa = prims[a.shell_prims[ia]], pb = prims[a.shell_prims[ib]];
        atA = pa.atom_index;
        atB = pb.atom_index;
        const double al = pa.exponent, be = pb.exponent, p = al + be;
        const double A[3] = {pa.x, pa.y, pa.z}, B[3] = {pb.x, pb.y, pb.z};
        double P[3], r2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          P[d] = (al * A[d] + be * B[d]) / p;
          r2 += (A[d] - B[d]) * (A[d] - B[d]);
        }
        const double na = 0.71270547035499016035 * double(1 << LA) * pow(al, (2.0 * LA + 3.0) / 4.0);
        const double nb = 0.71270547035499016035 * double(1 << LB) * pow(be, (2.0 * LB + 3.0) / 4.0);
        const double vpf = 2.0 * kPi / p * pa.coefficient * pb.coefficient * na * nb * exp(-al * be / p * r2);
        if (fabs(vpf) * ddmax * a.zsum * (1.0 + p) * (1.0 + p) < kG1Screen) continue;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < 3) {
          double E[LA + 2][LB + 2][LAB + 3];
          const double Pd = lane == 0 ? P[0] : lane == 1 ? P[1] : P[2];
          const double Ad = lane == 0 ? A[0] : lane == 1 ? A[1] : A[2];
          const double Bd = lane == 0 ? B[0] : lane == 1 ? B[1] : B[2];
          hermite_e_t<LA + 1, LB + 1, double>(p, Pd - Ad, Pd - Bd, E);
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
        if (!act) continue;
        const auto& E = *reinterpret_cast<const double(*)[3][LA + 2][LB + 2][LAB + 3]>(s_E);
        double R[nherm(LAB + 1)];
        hermite_r_t<LAB + 1, double>(p, P[0] - Cx, P[1] - Cy, P[2] - Cz, 1.0, a.boys, R);
        const double vf = -Zc * vpf;
        // sum_tuv E_x[i0][j0][t] E_y[i1][j1][u] E_z[i2][j2][v] R[t,u,v]
        auto V = [&](int i0, int j0, int i1, int j1, int i2, int j2) {
          double sum = 0.0;
#pragma unroll
          for (int t = 0; t <= LAB + 1; ++t) {
            if (t > i0 + j0) continue;
#pragma unroll
            for (int u = 0; u <= LAB + 1; ++u) {
              if (u > i1 + j1) continue;
              const double exy = E[0][i0][j0][t] * E[1][i1][j1][u];
#pragma unroll
              for (int v = 0; v <= LAB + 1; ++v) {
                if (v > i2 + j2) continue;
                sum += exy * E[2][i2][j2][v] * R[hidx(t, u, v)];
              }
            }
          }
          return sum;
        };
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab) {
          const int e[6] = {cart_exp(LA, ab / NB, 0), cart_exp(LB, ab % NB, 0), cart_exp(LA, ab / NB, 1),
                            cart_exp(LB, ab % NB, 1), cart_exp(LA, ab / NB, 2), cart_exp(LB, ab % NB, 2)};
          const double w = vf * dd[ab];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            int x[6] = {e[0], e[1], e[2], e[3], e[4], e[5]};
            x[2 * d] = e[2 * d] + 1;
            double vA = 2.0 * al * V(x[0], x[1], x[2], x[3], x[4], x[5]);
            if (e[2 * d] > 0) {
              x[2 * d] = e[2 * d] - 1;
              vA -= e[2 * d] * V(x[0], x[1], x[2], x[3], x[4], x[5]);
            }
            x[2 * d] = e[2 * d];
            x[2 * d + 1] = e[2 * d + 1] + 1;
            double vB = 2.0 * be * V(x[0], x[1], x[2], x[3], x[4], x[5]);
            if (e[2 * d + 1] > 0) {
              x[2 * d + 1] = e[2 * d + 1] - 1;
              vB -= e[2 * d + 1] * V(x[0], x[1], x[2], x[3], x[4], x[5]);
            }
            gA[d] += w * vA;
            gB[d] += w * vB;
            hf[d] -= w * (vA + vB);
          }
        }
      }
    if (act)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (hf[d] != 0.0) atomicAdd(&s_grad[3 * c + d], hf[d]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double sa = wave_sum_g(gA[d]), sb = wave_sum_g(gB[d]);  // every lane reaches this point
      if (lane == 0 && (sa != 0.0 || sb != 0.0)) {
        atomicAdd(&s_grad[3 * atA + d], sa);
        atomicAdd(&s_grad[3 * atB + d], sb);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&grad[i], s_grad[i]);
}

template <int LA, int LB>
hipError_t launch_grad_1e_nuc_t(const OneIntArgs& a, const int2* pairs, int npairs, const double* D, double* grad,
                                hipStream_t s) {
  const long long waves = (long long)npairs * ((a.natoms + 63) / 64);
  const unsigned grid = (unsigned)std::min<long long>(8192, (waves + kG1Waves - 1) / kG1Waves);
  const size_t lds = (3 * (size_t)a.natoms + (size_t)kG1Waves * 3 * (LA + 2) * (LB + 2) * (LA + LB + 3)) *
                     sizeof(double);
  hipLaunchKernelGGL((k_grad_1e_nuc<LA, LB>), dim3(grid), dim3(kGradThreads), lds, s, a, pairs, npairs, D, grad);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_grad_1e_nuc(int la, int lb, const OneIntArgs& a, const int2* pairs, int npairs, const double* D,
                              double* grad, hipStream_t s) {
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  if (npairs <= 0) return hipSuccess;
  switch (la * 3 + lb) {
    case 0: return launch_grad_1e_nuc_t<0, 0>(a, pairs, npairs, D, grad, s);
    case 3: return launch_grad_1e_nuc_t<1, 0>(a, pairs, npairs, D, grad, s);
    case 4: return launch_grad_1e_nuc_t<1, 1>(a, pairs, npairs, D, grad, s);
    case 6: return launch_grad_1e_nuc_t<2, 0>(a, pairs, npairs, D, grad, s);
    case 7: return launch_grad_1e_nuc_t<2, 1>(a, pairs, npairs, D, grad, s);
    case 8: return launch_grad_1e_nuc_t<2, 2>(a, pairs, npairs, D, grad, s);
    default: return hipErrorInvalidValue;
  }
}

bool grad_2e_dispatch_0(int, const ClassLaunch&, const Grad2eArgs&, int, int, hipStream_t, hipError_t*);
bool grad_2e_dispatch_1(int, const ClassLaunch&, const Grad2eArgs&, int, int, hipStream_t, hipError_t*);
bool grad_2e_dispatch_2(int, const ClassLaunch&, const Grad2eArgs&, int, int, hipStream_t, hipError_t*);
bool grad_2e_dispatch_3(int, const ClassLaunch&, const Grad2eArgs&, int, int, hipStream_t, hipError_t*);

bool grad_2e_uses_scratch(int LA, int LB, int LC, int LD) { return !grad_unrolled(LA, LB, LC, LD); }

hipError_t launch_grad_2e(const ClassLaunch& c, const Grad2eArgs& a, int max_bra_prims, int max_ket_prims,
                          hipStream_t s) {
  if (c.nblocks <= 0) return hipSuccess;
  if (max_bra_prims <= 0 || max_ket_prims <= 0) return hipErrorInvalidValue;
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  const int key = ((c.LA * 3 + c.LB) * 3 + c.LC) * 3 + c.LD;
  hipError_t err = hipErrorInvalidValue;
  if (grad_2e_dispatch_0(key, c, a, max_bra_prims, max_ket_prims, s, &err) ||
      grad_2e_dispatch_1(key, c, a, max_bra_prims, max_ket_prims, s, &err) ||
      grad_2e_dispatch_2(key, c, a, max_bra_prims, max_ket_prims, s, &err) ||
      grad_2e_dispatch_3(key, c, a, max_bra_prims, max_ket_prims, s, &err))
    return err;
  return hipErrorInvalidValue;
}

hipError_t launch_grad_1e(const OneIntArgs& a, const double* D, const double* W, double* grad, hipStream_t s) {
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  const size_t n = (size_t)a.M * a.M;
  hipLaunchKernelGGL(k_grad_1e, dim3((unsigned)((n + kGradThreads - 1) / kGradThreads)), dim3(kGradThreads),
                     3 * a.natoms * sizeof(double), s, a, D, W, grad);
  return hipGetLastError();
}

}  // namespace gansu
