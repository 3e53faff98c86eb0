// RI (density fitting) kernels for gfx950.
//
//   (P|Q)    two-centre Coulomb metric over auxiliary shell pairs
//            (int2c2e.cu:468-687; screening Qaux_a Qaux_b < thr, int2c2e.cu:506)
//   (mn|P)   three-centre integrals, written straight into the B layout
//            B[m][P][n] (full m, n), screening Q_ab Qaux_c < thr (int3c2e.cu:1139)
//   W_P  = sum_mn B[m][P][n] D[m][n]           (gpu_manager.cu:1573, cublasDgemv)
//   J_mn = sum_P W_P B[m][P][n]                (weighted_sum_matrices_kernel)
// Both J passes stream B once from HBM.  K is a pair of large GEMMs on the
// same layout (ri.cpp).
#include <hip/hip_runtime.h>

#include "md_device.hpp"
#include "gradient_2e_impl.hpp"
#include "ri_kernels.hpp"

namespace gansu {

using namespace dev;

namespace {

// Raw integral block between a bra (pair or single) and a ket single, both
// expressed as Hermite pairs: out[(a*NB+b)*NC+c].
__device__ void block_pair_single(int la, int lb, int lc, double p, const double* P, const double* PA,
                                  const double* PB, double q, const double* Q, double pref, bool aux_a, bool aux_c,
                                  const double* boys, double* out) {
  const int nt = la + lb + 1;
  double Eb[3][4 * 3 * 6];  // (la+1)(lb+1)(la+lb+1), la <= 3 (aux bra), lb <= 2
  double Ek[3][4 * 1 * 4];
  for (int d = 0; d < 3; ++d) {
    hermite_e_rt<double>(la, lb, p, PA[d], PB[d], Eb[d]);
    hermite_e_rt<double>(lc, 0, q, 0.0, 0.0, Ek[d]);
  }
  const int L = la + lb + lc;
  double R[nherm(7)];
  hermite_r_rt<double>(L, p * q / (p + q), P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], pref, boys, R);
  const int NA = ncart(la), NB = ncart(lb), NC = ncart(lc), ntk = lc + 1;
  for (int c = 0; c < NC; ++c) {
    int ce[3];
    for (int d = 0; d < 3; ++d) ce[d] = cart_exp(lc, c, d, aux_c);
    for (int a = 0; a < NA; ++a)
      for (int b = 0; b < NB; ++b) {
        int ae[3], be[3];
        for (int d = 0; d < 3; ++d) { ae[d] = cart_exp(la, a, d, aux_a); be[d] = cart_exp(lb, b, d); }
        double s = 0.0;
        for (int t = 0; t <= ae[0] + be[0]; ++t)
          for (int u = 0; u <= ae[1] + be[1]; ++u)
            for (int v = 0; v <= ae[2] + be[2]; ++v) {
              const double eab = Eb[0][(ae[0] * (lb + 1) + be[0]) * nt + t] * Eb[1][(ae[1] * (lb + 1) + be[1]) * nt + u] *
                                 Eb[2][(ae[2] * (lb + 1) + be[2]) * nt + v];
              for (int tt = 0; tt <= ce[0]; ++tt)
                for (int uu = 0; uu <= ce[1]; ++uu)
                  for (int vv = 0; vv <= ce[2]; ++vv) {
                    const double ec = Ek[0][ce[0] * ntk + tt] * Ek[1][ce[1] * ntk + uu] * Ek[2][ce[2] * ntk + vv];
                    const double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                    s += eab * sg * ec * R[hidx(t + tt, u + uu, v + vv)];
                  }
            }
        out[(a * NB + b) * NC + c] += s;
      }
  }
}

__global__ void k_int2c2e(RI2cArgs g) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.npairs) return;
  // shell pair (P >= Q) of the triangular index
  size_t r = (size_t)((sqrt(8.0 * (double)id + 1.0) - 1.0) / 2.0);
  while ((r + 1) * (r + 2) / 2 <= id) ++r;
  while (r * (r + 1) / 2 > id) --r;
  const int sP = (int)r, sQ = (int)(id - r * (r + 1) / 2);
  const AuxShell A = g.shells[sP], B = g.shells[sQ];
  if (A.qmax * B.qmax < g.thr) return;
  const int NA = ncart(A.L), NB = ncart(B.L);
  double out[10 * 10];
  for (int k = 0; k < NA * NB; ++k) out[k] = 0.0;
  for (int ia = 0; ia < A.nprim; ++ia) {
    const AuxPrim pa = g.prims[A.prim_begin + ia];
    for (int ib = 0; ib < B.nprim; ++ib) {
      const AuxPrim pb = g.prims[B.prim_begin + ib];
      if (pa.Q * pb.Q < g.thr) continue;
      const double zero3[3] = {0.0, 0.0, 0.0};
      const double pref = kTwoPi25 / (pa.alpha * pb.alpha * sqrt(pa.alpha + pb.alpha)) * pa.Kc * pb.Kc;
      // (a| as a pair (a, s-dummy) with PA = PB = 0
      double tmp[10 * 10];
      for (int k = 0; k < NA * NB; ++k) tmp[k] = 0.0;
      block_pair_single(A.L, 0, B.L, pa.alpha, pa.C, zero3, zero3, pb.alpha, pb.C, pref, true, true, g.boys, tmp);
      for (int k = 0; k < NA * NB; ++k) out[k] += tmp[k];
    }
  }
  // aux Cartesian order on the bra side too
  for (int a = 0; a < NA; ++a)
    for (int b = 0; b < NB; ++b) {
      const int p = A.basis + a, q = B.basis + b;
      const double v = out[a * NB + b] * g.w[p] * g.w[q];
      g.V[(size_t)p * g.Maux + q] = v;
      g.V[(size_t)q * g.Maux + p] = v;
    }
}

__global__ void k_int3c2e(RI3cArgs g) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)g.nsp * g.naux_shells) return;
  const int isp = (int)(id / g.naux_shells), iaux = (int)(id % g.naux_shells);
  const SPair sp = g.sp[isp];
  const AuxShell C = g.shells[iaux];
  if (sp.qmax * C.qmax < g.thr) return;
  const int NA = ncart(sp.la), NB = ncart(sp.lb), NC = ncart(C.L);
  double out[6 * 6 * 10];
  for (int k = 0; k < NA * NB * NC; ++k) out[k] = 0.0;
  for (int ip = 0; ip < sp.nprim; ++ip) {
    const PairRec pr = g.pr[sp.prim_begin + ip];
    for (int ic = 0; ic < C.nprim; ++ic) {
      const AuxPrim pc = g.prims[C.prim_begin + ic];
      if (pr.Q * pc.Q < g.thr) continue;
      const double pref = kTwoPi25 / (pr.p * pc.alpha * sqrt(pr.p + pc.alpha)) * pr.Kc * pc.Kc;
      block_pair_single(sp.la, sp.lb, C.L, pr.p, pr.P, pr.PA, pr.PB, pc.alpha, pc.C, pref, false, true, g.boys, out);
    }
  }
  const size_t M = g.M, Maux = g.Maux;
  const bool diff = sp.si != sp.sj;
  for (int a = 0; a < NA; ++a)
    for (int b = 0; b < NB; ++b) {
      const int m = sp.bi + a, n = sp.bj + b;
      const double wmn = g.w[m] * g.w[n];
      for (int c = 0; c < NC; ++c) {
        const int P = C.basis + c;
        const double v = out[(a * NB + b) * NC + c] * wmn * g.waux[P];
        g.B[((size_t)m * Maux + P) * M + n] = v;
        if (diff) g.B[((size_t)n * Maux + P) * M + m] = v;
      }
    }
}

// ------------------------------------------------- RI-JK energy gradient --
// 2-centre term (the 3-centre term has per-class kernels of its own).  The
// centre derivatives take the E-shift 2a E(i+1) - i E(i-1) of the first
// function's Hermite table; the second centre follows by translational
// invariance.  Per-atom sums in LDS, flushed once per workgroup.
constexpr int kRIGradThreads = 256;

// sum_PQ Gam_PQ d(P|Q)/dX: one lane per auxiliary shell pair (P >= Q).
__global__ __launch_bounds__(kRIGradThreads) void k_ri_grad_2c(RIGradArgs g) {
  extern __shared__ double s_grad[];
  const int n3 = 3 * g.natoms;
  for (int i = threadIdx.x; i < n3; i += blockDim.x) s_grad[i] = 0.0;
  __syncthreads();
  const size_t ns = g.nsh;
  const size_t npairs = ns * (ns + 1) / 2;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < npairs; id += (size_t)gridDim.x * blockDim.x) {
    size_t r = (size_t)((sqrt(8.0 * (double)id + 1.0) - 1.0) / 2.0);
    while ((r + 1) * (r + 2) / 2 <= id) ++r;
    while (r * (r + 1) / 2 > id) --r;
    const int sP = (int)r, sQ = (int)(id - r * (r + 1) / 2);
    const AuxShell A = g.shells[sP], B = g.shells[sQ];
    if (A.qmax * B.qmax < g.thr) continue;
    const int la = A.L, lb = B.L, NA = ncart(la), NB = ncart(lb);
    double gP[3] = {0.0, 0.0, 0.0};
    for (int ia = 0; ia < A.nprim; ++ia) {
      const AuxPrim pa = g.prims[A.prim_begin + ia];
      double Ea[3][5 * 3 * 6];  // (la+2)(1+2)(la+1+3) for the (a, s-dummy) pair built to (la+1, 1)
      for (int d = 0; d < 3; ++d) hermite_e_rt<double>(la + 1, 1, pa.alpha, 0.0, 0.0, Ea[d]);
      const grad_detail::ETab TA{{Ea[0], Ea[1], Ea[2]}, 2, la + 3};
      for (int ib = 0; ib < B.nprim; ++ib) {
        const AuxPrim pb = g.prims[B.prim_begin + ib];
        if (pa.Q * pb.Q < g.thr) continue;
        double Eb[3][4 * 4];
        for (int d = 0; d < 3; ++d) hermite_e_rt<double>(lb, 0, pb.alpha, 0.0, 0.0, Eb[d]);
        const double p = pa.alpha, q = pb.alpha;
        const double pref = kTwoPi25 / (p * q * sqrt(p + q)) * pa.Kc * pb.Kc;
        double R[nherm(7)];
        hermite_r_rt<double>(la + lb + 1, p * q / (p + q), pa.C[0] - pb.C[0], pa.C[1] - pb.C[1], pa.C[2] - pb.C[2],
                             pref, g.boys, R);
        for (int b = 0; b < NB; ++b) {
          int ce[3];
          for (int d = 0; d < 3; ++d) ce[d] = cart_exp(lb, b, d, true);
          double Gh[nherm(4)];
          for (int t = 0; t <= la + 1; ++t)
            for (int u = 0; u + t <= la + 1; ++u)
              for (int v = 0; v + u + t <= la + 1; ++v) {
                double sacc = 0.0;
                for (int tt = 0; tt <= ce[0]; ++tt)
                  for (int uu = 0; uu <= ce[1]; ++uu) {
                    const double exy = Eb[0][ce[0] * (lb + 1) + tt] * Eb[1][ce[1] * (lb + 1) + uu];
                    for (int vv = 0; vv <= ce[2]; ++vv) {
                      const double e = exy * Eb[2][ce[2] * (lb + 1) + vv];
                      sacc += (((tt + uu + vv) & 1) ? -e : e) * R[hidx(t + tt, u + uu, v + vv)];
                    }
                  }
                Gh[hidx(t, u, v)] = sacc;
              }
          const int q_ = B.basis + b;
          for (int a = 0; a < NA; ++a) {
            const int p_ = A.basis + a;
            double wt = g.G[(size_t)p_ * g.Maux + q_];
            if (sP != sQ) wt += g.G[(size_t)q_ * g.Maux + p_];
            if (wt == 0.0) continue;
            wt *= g.waux[p_] * g.waux[q_];
            int ae[3];
            const int zero[3] = {0, 0, 0};
            for (int d = 0; d < 3; ++d) ae[d] = cart_exp(la, a, d, true);
            double dA[3], dB[3];
            grad_detail::pair_grad(TA, ae, zero, p, 0.0, Gh, false, dA, dB);
            for (int d = 0; d < 3; ++d) gP[d] += wt * dA[d];
          }
        }
      }
    }
    const int atP = g.aux_atom[sP], atQ = g.aux_atom[sQ];
    if (atP != atQ)
      for (int d = 0; d < 3; ++d) {
        atomicAdd(&s_grad[3 * atP + d], gP[d]);
        atomicAdd(&s_grad[3 * atQ + d], -gP[d]);
      }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x)
    if (s_grad[i] != 0.0) unsafeAtomicAdd(&g.grad[i], s_grad[i]);
}

__global__ void k_fill_cD(const double* __restrict__ c, const double* __restrict__ D, int P0, int pc, int M,
                          double* __restrict__ G) {
  const size_t n = (size_t)M * M, tot = n * pc;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < tot; id += (size_t)gridDim.x * blockDim.x)
    G[id] = c[P0 + id / n] * D[id % n];
}

// W_P = sum_{m,n} B[m][P][n] D[m][n]: one workgroup per P.
template <typename T>
__global__ void k_ri_w(const T* __restrict__ B, const double* __restrict__ D, int M, int Maux, double* __restrict__ W) {
  const int P = blockIdx.x;
  double s = 0.0;
  for (int m = 0; m < M; ++m) {
    const T* row = B + ((size_t)m * Maux + P) * M;
    const double* drow = D + (size_t)m * M;
    for (int n = threadIdx.x; n < M; n += blockDim.x) s += (double)row[n] * drow[n];
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) W[P] = sh[0];
}

// J[m][n] = sum_P W_P B[m][P][n] for n <= m (B is symmetric in m and n; the
// caller mirrors the rest).  A workgroup takes one m and 128 consecutive n
// (two per lane), the four waves a quarter of P each with four P in flight;
// the wave partial sums meet in LDS in wave order.
__global__ __launch_bounds__(256) void k_ri_j(const double* __restrict__ B, const double* __restrict__ W, int M,
                                              int Maux, double* __restrict__ J) {
  __shared__ double red[3][64][2];
  const int m = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if ((int)blockIdx.y * 128 > m) return;  // every column of the workgroup lies above the diagonal
  const int n0 = blockIdx.y * 128 + 2 * lane;
  const int P0 = (int)((long)w * Maux / 4), P1 = (int)((long)(w + 1) * Maux / 4);
  const double* col = B + (size_t)m * Maux * M + n0;
  double s0 = 0.0, s1 = 0.0;
  if (n0 <= m) {
    const bool two = n0 + 1 < M;
    int P = P0;
    for (; P + 4 <= P1; P += 4) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = col[(size_t)(P + u) * M];
        b[u] = two ? col[(size_t)(P + u) * M + 1] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s0 += W[P + u] * a[u];
        s1 += W[P + u] * b[u];
      }
    }
    for (; P < P1; ++P) {
      s0 += W[P] * col[(size_t)P * M];
      if (two) s1 += W[P] * col[(size_t)P * M + 1];
    }
  }
  if (w > 0) {
    red[w - 1][lane][0] = s0;
    red[w - 1][lane][1] = s1;
  }
  __syncthreads();
  if (w != 0 || n0 > m) return;
  J[(size_t)m * M + n0] = ((s0 + red[0][lane][0]) + red[1][lane][0]) + red[2][lane][0];
  if (n0 + 1 < M) J[(size_t)m * M + n0 + 1] = ((s1 + red[0][lane][1]) + red[1][lane][1]) + red[2][lane][1];
}

// W[p0 + p] += sum_m sum_i T[m][p][i] X[m][i] with X = sqrt(2) C_occ, i.e.
// W_P = sum_mn B[m][P][n] D[m][n] for D = X X^T, read from the K intermediate
// T (M x pc x nocc) instead of a pass over B.  One workgroup per p.
__global__ void k_ri_w_from_t(const double* __restrict__ T, const double* __restrict__ C, int M, int pc, int nocc,
                              int p0, double* __restrict__ W) {
  const int p = blockIdx.x;
  double s = 0.0;
  for (int m = 0; m < M; ++m) {
    const double* t = T + ((size_t)m * pc + p) * nocc;
    const double* c = C + (size_t)m * M;
    for (int i = threadIdx.x; i < nocc; i += blockDim.x) s += t[i] * c[i];
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) W[p0 + p] = 1.4142135623730951 * sh[0];
}

// mirror the upper triangle (column-major fill) of a symmetric matrix: row-major
// entries [i][j] with j < i are copied from [j][i]
__global__ void k_mirror_upper(double* __restrict__ K, int M) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const int i = (int)(id / M), j = (int)(id % M);
  // column-major upper (row <= col) is row-major lower: [j][i] with j >= i holds the data
  if (j > i) K[id] = K[(size_t)j * M + i];
}

// Sum of split-k partial products into a column-major nr x nc block of K
// (leading dimension ldk): K = (accumulate ? K : 0) + sum_b part[b], with
// part[b] column-major nr x nc.
__global__ void k_sum_parts(const double* __restrict__ part, int nparts, int nr, int nc, int accumulate, double* K,
                            int ldk) {
  const size_t n = (size_t)nr * nc;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int i = (int)(id % nr), j = (int)(id / nr);
  double* k = K + i + (size_t)j * ldk;
  double s = accumulate ? *k : 0.0;
  for (int b = 0; b < nparts; ++b) s += part[(size_t)b * n + id];
  *k = s;
}

__global__ void k_fock_ri(const double* __restrict__ H, const double* __restrict__ J, const double* __restrict__ K,
                          size_t n, double* __restrict__ F) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) F[id] = H[id] + J[id] - 0.5 * K[id];
}

}  // namespace

hipError_t launch_ri_grad_2c(const RIGradArgs& a, hipStream_t s) {
  if (a.natoms <= 0 || a.natoms > kGradMaxAtoms) return hipErrorInvalidValue;
  const size_t n = (size_t)a.nsh * (a.nsh + 1) / 2;
  if (n == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<size_t>(4096, (n + kRIGradThreads - 1) / kRIGradThreads);
  hipLaunchKernelGGL(k_ri_grad_2c, dim3(grid), dim3(kRIGradThreads), 3 * a.natoms * sizeof(double), s, a);
  return hipGetLastError();
}

hipError_t launch_fill_cD(const double* c, const double* D, int P0, int pc, int M, double* G, hipStream_t s) {
  const size_t tot = (size_t)M * M * pc;
  if (tot == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<size_t>(16384, (tot + 255) / 256);
  hipLaunchKernelGGL(k_fill_cD, dim3(grid), dim3(256), 0, s, c, D, P0, pc, M, G);
  return hipGetLastError();
}

hipError_t launch_int2c2e(const RI2cArgs& a, hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_int2c2e, dim3((unsigned)((a.npairs + 127) / 128)), dim3(128), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_int3c2e(const RI3cArgs& a, hipStream_t s) {
  const size_t n = (size_t)a.nsp * a.naux_shells;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_int3c2e, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, a);
  return hipGetLastError();
}

// B[m][P][n] = B[n][P][m] where n lies in an earlier block of bs (a multiple
// of 32) than m: 32 x 32 tiles of (m, n) through LDS, 64 P per workgroup.  A
// 32-tile lies inside one block, so a tile is copied whole or not at all;
// the copied entries are never read here (reads are at n-blocks above the
// m-block of their slab, writes below), so the workgroups do not race.
__global__ __launch_bounds__(256) void k_ri_mirror(double* __restrict__ B, int M, int Maux, int bs) {
  __shared__ double t[32][33];
  const int nt = (M + 31) / 32;
  const int tm = blockIdx.x / nt, tn = blockIdx.x % nt;
  if ((tn * 32) / bs >= (tm * 32) / bs) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int P1 = min(Maux, (int)blockIdx.y * 64 + 64);
  for (int P = blockIdx.y * 64; P < P1; ++P) {
    for (int y = ty; y < 32; y += 8) {
      const int n = tn * 32 + y, m = tm * 32 + tx;
      if (n < M && m < M) t[y][tx] = B[((size_t)n * Maux + P) * M + m];
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      const int m = tm * 32 + y, n = tn * 32 + tx;
      if (n < M && m < M) B[((size_t)m * Maux + P) * M + n] = t[tx][y];
    }
    __syncthreads();
  }
}

hipError_t launch_ri_mirror(double* B, int M, int Maux, int bs, hipStream_t s) {
  if (bs <= 0 || bs % 32 != 0) return hipErrorInvalidValue;
  const int nt = (M + 31) / 32;
  hipLaunchKernelGGL(k_ri_mirror, dim3((unsigned)(nt * nt), (unsigned)((Maux + 63) / 64)), dim3(256), 0, s, B, M,
                     Maux, bs);
  return hipGetLastError();
}

hipError_t launch_ri_w(const double* B, const double* D, int M, int Maux, double* W, hipStream_t s) {
  hipLaunchKernelGGL(k_ri_w<double>, dim3(Maux), dim3(256), 0, s, B, D, M, Maux, W);
  return hipGetLastError();
}

hipError_t launch_ri_j(const double* B, const double* W, int M, int Maux, double* J, hipStream_t s) {
  hipLaunchKernelGGL(k_ri_j, dim3((unsigned)M, (unsigned)((M + 127) / 128)), dim3(256), 0, s, B, W, M, Maux, J);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? launch_mirror_upper(J, M, s) : e;
}

hipError_t launch_ri_w_from_t(const double* T, const double* C, int M, int pc, int nocc, int p0, double* W,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_ri_w_from_t, dim3(pc), dim3(256), 0, s, T, C, M, pc, nocc, p0, W);
  return hipGetLastError();
}

hipError_t launch_mirror_upper(double* K, int M, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_mirror_upper, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, K, M);
  return hipGetLastError();
}

hipError_t launch_sum_parts(const double* part, int nparts, int nr, int nc, int accumulate, double* K, int ldk,
                            hipStream_t s) {
  const size_t n = (size_t)nr * nc;
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, nparts, nr, nc, accumulate,
                     K, ldk);
  return hipGetLastError();
}

hipError_t launch_fock_ri(const double* H, const double* J, const double* K, size_t n, double* F, hipStream_t s) {
  hipLaunchKernelGGL(k_fock_ri, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, J, K, n, F);
  return hipGetLastError();
}

}  // namespace gansu
