// One-electron integrals S, T, V (McMurchie-Davidson) and small dense helpers
// used by the RHF driver.  The reference computes S/T by MD and V by
// Obara-Saika per shell-type pair (int1e.cu, MD_kernel.txt, OS_kernel.txt);
// here one lane evaluates one (mu, nu) basis-function pair.
#include <hip/hip_runtime.h>

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

__global__ void k_oneint(OneIntArgs a) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = a.M;
  if (id >= (size_t)M * M) return;
  const int mu = (int)(id / M), nu = (int)(id % M);
  if (nu > mu) return;  // lower triangle, mirrored at the end
  const int sa = a.bf_shell[mu], sb = a.bf_shell[nu];
  const int ca = mu - a.shell_basis[sa], cb = nu - a.shell_basis[sb];
  const int la = a.shell_L[sa], lb = a.shell_L[sb];
  int ea[3], eb[3];
  for (int d = 0; d < 3; ++d) { ea[d] = cart_exp(la, ca, d); eb[d] = cart_exp(lb, cb, d); }
  double S = 0.0, T = 0.0, V = 0.0;
  for (int ia = a.shell_off[sa]; ia < a.shell_off[sa + 1]; ++ia) {
    const DPrim pa = ((const DPrim*)a.prims)[a.shell_prims[ia]];
    for (int ib = a.shell_off[sb]; ib < a.shell_off[sb + 1]; ++ib) {
      const DPrim pb = ((const DPrim*)a.prims)[a.shell_prims[ib]];
      const double al = pa.exponent, be = pb.exponent, p = al + be;
      const double A[3] = {pa.x, pa.y, pa.z}, B[3] = {pb.x, pb.y, pb.z};
      double P[3], r2 = 0.0;
      double E[3][3 * 5 * 8];  // (la+1) x (lb+3) x (la+lb+3), la, lb <= 2
      const int nj = lb + 3, ntt = la + lb + 3;
      for (int d = 0; d < 3; ++d) {
        P[d] = (al * A[d] + be * B[d]) / p;
        r2 += (A[d] - B[d]) * (A[d] - B[d]);
        hermite_e_rt<double>(la, lb + 2, p, P[d] - A[d], P[d] - B[d], E[d]);
      }
      const double K = exp(-al * be / p * r2);
      const double na = 0.71270547035499016035 * double(1 << la) * pow(al, (2.0 * la + 3.0) / 4.0);
      const double nb = 0.71270547035499016035 * double(1 << lb) * pow(be, (2.0 * lb + 3.0) / 4.0);
      const double pf = pa.coefficient * pb.coefficient * na * nb * K;
      double s1[3], t1[3];
      for (int d = 0; d < 3; ++d) {
        const int i = ea[d], j = eb[d];
        const double sij = E[d][(i * nj + j) * ntt];
        const double sp2 = E[d][(i * nj + j + 2) * ntt];
        const double sm2 = j >= 2 ? E[d][(i * nj + j - 2) * ntt] : 0.0;
        s1[d] = sij;
        t1[d] = -2.0 * be * be * sp2 + be * (2 * j + 1) * sij - 0.5 * j * (j - 1) * sm2;
      }
      const double ovl = pow(kPi / p, 1.5);
      S += pf * ovl * s1[0] * s1[1] * s1[2];
      T += pf * ovl * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
      const int L = la + lb;
      const double* Ex = E[0] + (ea[0] * nj + eb[0]) * ntt;
      const double* Ey = E[1] + (ea[1] * nj + eb[1]) * ntt;
      const double* Ez = E[2] + (ea[2] * nj + eb[2]) * ntt;
      double vsum = 0.0;
      for (int c = 0; c < a.natoms; ++c) {
        double R[nherm(4)];
        hermite_r_rt<double>(L, p, P[0] - a.atom_xyz[3 * c], P[1] - a.atom_xyz[3 * c + 1],
                             P[2] - a.atom_xyz[3 * c + 2], 1.0, a.boys, R);
        double acc = 0.0;
        for (int t = 0; t <= ea[0] + eb[0]; ++t)
          for (int u = 0; u <= ea[1] + eb[1]; ++u)
            for (int v = 0; v <= ea[2] + eb[2]; ++v) acc += Ex[t] * Ey[u] * Ez[v] * R[hidx(t, u, v)];
        vsum -= a.Z[c] * acc;
      }
      V += pf * 2.0 * kPi / p * vsum;
    }
  }
  const double ww = a.w[mu] * a.w[nu];
  a.S[(size_t)mu * M + nu] = a.S[(size_t)nu * M + mu] = S * ww;
  a.T[(size_t)mu * M + nu] = a.T[(size_t)nu * M + mu] = T * ww;
  a.V[(size_t)mu * M + nu] = a.V[(size_t)nu * M + mu] = V * ww;
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
  const size_t n = (size_t)a.M * a.M;
  hipLaunchKernelGGL(k_oneint, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, a);
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
                     (const int*)at[6], (const double*)at[9], S, T, V, t.zsum};
  hip_check(launch_oneint(a, s), "launch_oneint");
  // the host tables are released on return, and the next call reuses the array
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

hipError_t launch_has_nan(const double* x, size_t n, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_has_nan, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, flag);
  return hipGetLastError();
}

}  // namespace gansu
