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
constexpr int kG1MaxEV = 4 * 4 * 7;  // (la+2)(lb+2)(la+lb+3): the attraction part
constexpr int kG1NucGroup = 8;       // nuclei per pass of the attraction part
// A primitive pair whose prefactor times the density weights falls below this
// contributes nothing at FP64 precision to any component (the Gaussian
// product factor exp(-ab/p r^2) makes most pairs of a large molecule vanish).
constexpr double kG1Screen = 1e-17;

// One lane = one basis-function pair mu >= nu; derivatives of S, T and V
// with respect to both basis centres, and the Hellmann-Feynman term of every
// nucleus (-(d/dA + d/dB) of that nucleus' attraction integral).  The nuclei
// are taken kG1NucGroup at a time in a wave-uniform loop: each lane keeps that
// group's Hellmann-Feynman sums over its primitive pairs in registers, and a
// wave reduction leaves one LDS atomic per nucleus and wave (an atomic per lane
// and nucleus serialised 64 lanes on one address).
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
  // nuclear attraction, kG1NucGroup nuclei per wave-uniform pass
  const int L = la + lb;
  const int nj = lb + 2, ntt = la + lb + 3;
  for (int c0 = 0; c0 < a.natoms; c0 += kG1NucGroup) {
    double hf[kG1NucGroup][3];
#pragma unroll
    for (int cc = 0; cc < kG1NucGroup; ++cc) hf[cc][0] = hf[cc][1] = hf[cc][2] = 0.0;
    for (int ia = pa0; ia < pa1; ++ia)
      for (int ib = pb0; ib < pb1; ++ib) {
        double P[9], al, be, p, pf;
        prim_pair(ia, ib, P, al, be, p, pf);
        const double vpf = 2.0 * kPi / p * pf * dd;
        if (fabs(vpf) * a.zsum * (1.0 + p) * (1.0 + p) < kG1Screen) continue;
        double E[3][kG1MaxEV];
        for (int d = 0; d < 3; ++d) hermite_e_rt<double>(la + 1, lb + 1, p, P[d] - P[3 + d], P[d] - P[6 + d], E[d]);
        auto row = [&](int d, int i, int j) -> const double* { return E[d] + (i * nj + j) * ntt; };
#pragma unroll
        for (int cc = 0; cc < kG1NucGroup; ++cc) {
          const int c = c0 + cc;
          if (c >= a.natoms) break;
          double R[nherm(5)];
          hermite_r_rt<double>(L + 1, p, P[0] - a.atom_xyz[3 * c], P[1] - a.atom_xyz[3 * c + 1],
                               P[2] - a.atom_xyz[3 * c + 2], 1.0, a.boys, R);
          const double vf = -a.Z[c] * vpf;
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
              hf[cc][d] -= v;
            }
          }
        }
      }
    // every lane of the wave reaches this point (the loop above is uniform)
#pragma unroll
    for (int cc = 0; cc < kG1NucGroup; ++cc)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double v = wave_sum_g(hf[cc][d]);
        if ((threadIdx.x & 63) == 0 && c0 + cc < a.natoms && v != 0.0) atomicAdd(&s_grad[3 * (c0 + cc) + d], v);
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

}  // namespace

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
