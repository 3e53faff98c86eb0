// One-electron integrals S, T, V (McMurchie-Davidson) and small dense helpers
// used by the RHF driver.  The reference computes S/T by MD and V by
// Obara-Saika per shell-type pair (int1e.cu, MD_kernel.txt, OS_kernel.txt);
// here one wave evaluates one shell pair, its lanes splitting the nuclei.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "gansu_internal.hpp"
#include "md_device.hpp"
#include "scf_kernels.hpp"

namespace gansu {

// Device array of the no-allocation entry point (gansu_oneint_noalloc).  In
// host code its name is only a stand-in: the device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kOneIntScratchBytes];

void hip_check(hipError_t e, const char* what);  // engine.cpp

using namespace dev;

namespace {

struct DPrim {
  double exponent, coefficient, x, y, z;
  int shell_type;
  size_t basis_index;
  int atom_index;
};

// S, T and V per shell-pair class (LA, LB): a wave takes one shell pair
// (s1 >= s2, decoded from the triangular index; the pairs of other classes
// are passed over) and every nucleus, 64 per round, one per lane.  Per
// primitive pair three lanes build the E tables (two orders higher in b, for
// the kinetic term) into the wave's LDS slot; lane ab < NAB forms S and T of
// component ab, every lane R for its nucleus and V of all components, summed
// over the wave at the end.  Every (mu, nu) of the pair is written once,
// with its mirror, so no output needs clearing and the sums are in a fixed
// order.  McMurchie-Davidson: S and T from the overlap E coefficients
// (T by the j +- 2 recurrence), V = -sum_C Z_C 2 pi / p sum_tuv E E E R_tuv.
constexpr int kOneIntWaves = 4;

__device__ __forceinline__ double oneint_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int LA, int LB>
__global__ __launch_bounds__(64 * kOneIntWaves) void k_oneint_t(OneIntArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB, LAB = LA + LB;
  constexpr int NJ = LB + 3, NTT = LAB + 3, NE = (LA + 1) * NJ * NTT;
  __shared__ double sE[kOneIntWaves][3][NE];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long npair = (long long)a.nshell * (a.nshell + 1) / 2;
  const DPrim* prims = (const DPrim*)a.prims;
  const int M = a.M;
  for (long long pid = (long long)blockIdx.x * kOneIntWaves + wave; pid < npair;
       pid += (long long)gridDim.x * kOneIntWaves) {
    long long r = (long long)((sqrt(8.0 * (double)pid + 1.0) - 1.0) / 2.0);
    while ((r + 1) * (r + 2) / 2 <= pid) ++r;
    while (r * (r + 1) / 2 > pid) --r;
    const int s1 = (int)r, s2 = (int)(pid - r * (r + 1) / 2);
    if (a.shell_L[s1] != LA || a.shell_L[s2] != LB) continue;  // wave-uniform
    const int ab = lane < NAB ? lane : 0;
    const int ea[3] = {cart_exp(LA, ab / NB, 0), cart_exp(LA, ab / NB, 1), cart_exp(LA, ab / NB, 2)};
    const int eb[3] = {cart_exp(LB, ab % NB, 0), cart_exp(LB, ab % NB, 1), cart_exp(LB, ab % NB, 2)};
    double S = 0.0, T = 0.0, vacc[NAB];
#pragma unroll
    for (int k = 0; k < NAB; ++k) vacc[k] = 0.0;
    for (int ia = a.shell_off[s1]; ia < a.shell_off[s1 + 1]; ++ia)
      for (int ib = a.shell_off[s2]; ib < a.shell_off[s2 + 1]; ++ib) {  // wave-uniform
        const DPrim pa = prims[a.shell_prims[ia]], pb = prims[a.shell_prims[ib]];
        const double al = pa.exponent, be = pb.exponent, p = al + be;
        const double A[3] = {pa.x, pa.y, pa.z}, B[3] = {pb.x, pb.y, pb.z};
        double P[3], r2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          P[d] = (al * A[d] + be * B[d]) / p;
          r2 += (A[d] - B[d]) * (A[d] - B[d]);
        }
        const double K = exp(-al * be / p * r2);
        const double na = 0.71270547035499016035 * double(1 << LA) * pow(al, (2.0 * LA + 3.0) / 4.0);
        const double nb = 0.71270547035499016035 * double(1 << LB) * pow(be, (2.0 * LB + 3.0) / 4.0);
        const double pf = pa.coefficient * pb.coefficient * na * nb * K;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // the previous primitive pair's readers are done
        if (lane < 3) {
          double E[LA + 1][NJ][NTT];
          const double Pd = lane == 0 ? P[0] : lane == 1 ? P[1] : P[2];
          const double Ad = lane == 0 ? A[0] : lane == 1 ? A[1] : A[2];
          const double Bd = lane == 0 ? B[0] : lane == 1 ? B[1] : B[2];
          hermite_e_t<LA, LB + 2, double>(p, Pd - Ad, Pd - Bd, E);
          double* dst = &sE[wave][lane][0];
#pragma unroll
          for (int i = 0; i <= LA; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
              for (int t = 0; t < NTT; ++t) dst[(i * NJ + j) * NTT + t] = E[i][j][t];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const auto& E = *reinterpret_cast<const double(*)[3][LA + 1][NJ][NTT]>(&sE[wave][0][0]);
        if (lane < NAB) {
          double s1d[3], t1d[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int i = ea[d], j = eb[d];
            const double sij = E[d][i][j][0], sp2 = E[d][i][j + 2][0];
            const double sm2 = j >= 2 ? E[d][i][j - 2][0] : 0.0;
            s1d[d] = sij;
            t1d[d] = -2.0 * be * be * sp2 + be * (2 * j + 1) * sij - 0.5 * j * (j - 1) * sm2;
          }
          const double ovl = pow(kPi / p, 1.5);
          S += pf * ovl * s1d[0] * s1d[1] * s1d[2];
          T += pf * ovl * (t1d[0] * s1d[1] * s1d[2] + s1d[0] * t1d[1] * s1d[2] + s1d[0] * s1d[1] * t1d[2]);
        }
        const double vpf = pf * 2.0 * kPi / p;
        for (int c0 = 0; c0 < a.natoms; c0 += 64) {
          const int c = c0 + lane;
          if (c >= a.natoms) continue;
          double R[nherm(LAB)];
          hermite_r_t<LAB, double>(p, P[0] - a.atom_xyz[3 * c], P[1] - a.atom_xyz[3 * c + 1],
                                   P[2] - a.atom_xyz[3 * c + 2], 1.0, a.boys, R);
          const double zf = -(double)a.Z[c] * vpf;
#pragma unroll
          for (int k = 0; k < NAB; ++k) {
            const int x0 = cart_exp(LA, k / NB, 0), x1 = cart_exp(LA, k / NB, 1), x2 = cart_exp(LA, k / NB, 2);
            const int y0 = cart_exp(LB, k % NB, 0), y1 = cart_exp(LB, k % NB, 1), y2 = cart_exp(LB, k % NB, 2);
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t <= LAB; ++t) {
              if (t > x0 + y0) continue;
#pragma unroll
              for (int u = 0; u <= LAB; ++u) {
                if (u > x1 + y1) continue;
                const double exy = E[0][x0][y0][t] * E[1][x1][y1][u];
#pragma unroll
                for (int v = 0; v <= LAB; ++v) {
                  if (v > x2 + y2) continue;
                  acc += exy * E[2][x2][y2][v] * R[hidx(t, u, v)];
                }
              }
            }
            vacc[k] += zf * acc;
          }
        }
      }
    const int b1 = a.shell_basis[s1], b2 = a.shell_basis[s2];
#pragma unroll
    for (int k = 0; k < NAB; ++k) {
      const double v = oneint_wave_sum(vacc[k]);  // every lane reaches this point
      if (lane == k) {
        const int mu = b1 + k / NB, nu = b2 + k % NB;
        const double ww = a.w[mu] * a.w[nu];
        a.S[(size_t)mu * M + nu] = a.S[(size_t)nu * M + mu] = S * ww;
        a.T[(size_t)mu * M + nu] = a.T[(size_t)nu * M + mu] = T * ww;
        a.V[(size_t)mu * M + nu] = a.V[(size_t)nu * M + mu] = v * ww;
      }
    }
  }
}

template <int LA, int LB>
void launch_oneint_class(const OneIntArgs& a, long long npair, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<long long>(16384, (npair + kOneIntWaves - 1) / kOneIntWaves);
  hipLaunchKernelGGL((k_oneint_t<LA, LB>), dim3(grid), dim3(64 * kOneIntWaves), 0, s, a);
}

// X[i][k] = U_col[i + k n] * (s_k < thr ? 0 : s_k^-1/2)   (hf.cu:171-242,
// gpu_kernels.cu:28-43: eigenvalues below 1e-6 are dropped, M columns kept)
__global__ void k_transform(const double* __restrict__ Ucol, const double* __restrict__ s, int n, double thr,
                            double* __restrict__ X) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n * n) return;
  const int i = (int)(id / n), k = (int)(id % n);
  const double sk = s[k];
  X[id] = Ucol[(size_t)k * n + i] * (sk < thr ? 0.0 : 1.0 / sqrt(sk));
}

// out = sum_i c_i in_i (DIIS extrapolation) or out = a*x + b*y
__global__ void k_lincomb(int nmat, const double* __restrict__ coef, const double* const* __restrict__ mats,
                          size_t n, double* __restrict__ out) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  double s = 0.0;
  for (int m = 0; m < nmat; ++m) s += coef[m] * mats[m][id];
  out[id] = s;
}

__global__ void k_axpby(double a, const double* __restrict__ x, double b, const double* __restrict__ y, size_t n,
                        double* __restrict__ out) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) out[id] = a * x[id] + b * y[id];
}

// Generalized Wolfsberg-Helmholz guess (gpu_kernels.cu:535-545, cx = 1.75 by
// gpu_manager.cu:1208): F_pq = cx S_pq (H_pp + H_qq) / 2 for every p, q, the
// diagonal included.
__global__ void k_gwh_guess(const double* __restrict__ H, const double* __restrict__ S, int M, double cx,
                            double* __restrict__ F) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * M) return;
  const size_t p = id / M, q = id % M;
  F[id] = cx * S[id] * (H[p * M + p] + H[q * M + q]) / 2.0;
}

// partial sums of sum_i x_i * y_i (deterministic two-pass reduction)
__global__ void k_dot_partial(const double* __restrict__ x, const double* __restrict__ y, size_t n,
                              double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_set_int(int* p, int v) { *p = v; }

__global__ void k_has_nan(const double* __restrict__ x, size_t n, int* flag) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n && !(x[id] == x[id])) *flag = 1;
}

}  // namespace

hipError_t launch_oneint(const OneIntArgs& a, hipStream_t s) {
  // one launch per shell-pair class; each writes the blocks of its pairs
  const long long npair = (long long)a.nshell * (a.nshell + 1) / 2;
  if (npair <= 0) return hipSuccess;
  launch_oneint_class<0, 0>(a, npair, s);
  launch_oneint_class<0, 1>(a, npair, s);
  launch_oneint_class<0, 2>(a, npair, s);
  launch_oneint_class<1, 0>(a, npair, s);
  launch_oneint_class<1, 1>(a, npair, s);
  launch_oneint_class<1, 2>(a, npair, s);
  launch_oneint_class<2, 0>(a, npair, s);
  launch_oneint_class<2, 1>(a, npair, s);
  launch_oneint_class<2, 2>(a, npair, s);
  return hipGetLastError();
}

hipError_t launch_transform(const double* Ucol, const double* evals, int n, double thr, double* X, hipStream_t s) {
  const size_t m = (size_t)n * n;
  hipLaunchKernelGGL(k_transform, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, Ucol, evals, n, thr, X);
  return hipGetLastError();
}

hipError_t launch_lincomb(int nmat, const double* d_coef, const double* const* d_mats, size_t n, double* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nmat, d_coef, d_mats, n, out);
  return hipGetLastError();
}

hipError_t launch_gwh_guess(const double* H, const double* S, int M, double* F, hipStream_t s) {
  const size_t n = (size_t)M * M;
  hipLaunchKernelGGL(k_gwh_guess, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, S, M, 1.75, F);
  return hipGetLastError();
}

hipError_t launch_axpby(double a, const double* x, double b, const double* y, size_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, x, b, y, n, out);
  return hipGetLastError();
}

hipError_t launch_dot_partial(const double* x, const double* y, size_t n, double* part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial, dim3(nblocks), dim3(256), 0, s, x, y, n, part);
  return hipGetLastError();
}

hipError_t launch_set_int(int* p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

namespace {
// layout of the tables in the array: each one 256-byte aligned, in this order
template <class F>
size_t oneint_layout(const OneIntTables& t, F&& place) {
  size_t at = 0;
  auto put = [&](const void* src, size_t bytes) {
    place(at, src, bytes);
    at += (bytes + 255) / 256 * 256;
  };
  put(t.prims.data(), t.prims.size() * sizeof(gansu_primitive_shell));
  put(t.off.data(), t.off.size() * sizeof(int));
  put(t.sprims.data(), t.sprims.size() * sizeof(int));
  put(t.sl.data(), t.sl.size() * sizeof(int));
  put(t.sb.data(), t.sb.size() * sizeof(int));
  put(t.bfs.data(), t.bfs.size() * sizeof(int));
  put(t.Z.data(), t.Z.size() * sizeof(int));
  put(t.w.data(), t.w.size() * sizeof(double));
  put(t.xyz.data(), t.xyz.size() * sizeof(double));
  put(t.boys.data(), t.boys.size() * sizeof(double));
  return at;
}

// the array's device address on the current device
unsigned char* scratch_ptr() {
  static std::mutex mu;
  static std::vector<unsigned char*> ptrs;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Error(GANSU_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
    ptrs[dev] = static_cast<unsigned char*>(p);
  }
  return ptrs[dev];
}
}  // namespace

size_t oneint_scratch_bytes(const OneIntTables& t) {
  return oneint_layout(t, [](size_t, const void*, size_t) {});
}

void oneint_noalloc(const OneIntTables& t, double* S, double* T, double* V, hipStream_t s) {
  const size_t need = oneint_scratch_bytes(t);
  if (need > kOneIntScratchBytes)
    throw Error(GANSU_E_ARG, "gansu_oneint_noalloc: the molecule's tables take " + std::to_string(need) +
                                 " bytes, beyond the " + std::to_string(kOneIntScratchBytes) +
                                 "-byte device array (use gansu_oneint)");
  unsigned char* base = scratch_ptr();
  const void* at[10];
  int k = 0;
  oneint_layout(t, [&](size_t off, const void* src, size_t bytes) {
    at[k++] = base + off;
    if (bytes) hip_check(hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync(scratch)");
  });
  const OneIntArgs a{at[0], (const int*)at[1], (const int*)at[2], (const int*)at[3], (const int*)at[4],
                     (const int*)at[5], (const double*)at[7], t.M, t.natoms, (const double*)at[8],
                     (const int*)at[6], (const double*)at[9], S, T, V, t.zsum, (int)t.sl.size()};
  hip_check(launch_oneint(a, s), "launch_oneint");
  // the host tables are released on return, and the next call reuses the array
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

hipError_t launch_has_nan(const double* x, size_t n, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_has_nan, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, flag);
  return hipGetLastError();
}

}  // namespace gansu
