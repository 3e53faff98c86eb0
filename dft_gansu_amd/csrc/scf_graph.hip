// Kernels of the device-resident RHF iteration (declarations and the
// algorithms in scf_graph.hpp).  Every kernel takes its control state (the
// iteration's energy, convergence flag, DIIS counters, purification flags)
// from device memory, so the host captures an iteration once and replays it.
#include "scf_graph.hpp"

#include <cfloat>

namespace gansu {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kT = 16;  // output tile edge of one workgroup

// op(M)[r][c] with zero outside the n x n matrix
template <bool T>
__device__ __forceinline__ double ld_op(const double* __restrict__ M, int n, int r, int c) {
  if (r >= n || c >= n) return 0.0;
  return T ? M[(size_t)c * n + r] : M[(size_t)r * n + c];
}

// Partial product of one wave: the 16 x 16 block at (i0, j0) of op(A) op(B)
// over k = 4 q + (lane / 16), q in [q0, q1), by v_mfma_f64_16x16x4f64.
// Operand lanes: lane l holds A[row l%16][k l/16] and B[k l/16][col l%16];
// accumulator element r of lane l is C[l/16 + 4 r][l%16] (checked on gfx950).
// The operands are read straight from global memory (the matrices of an SCF
// iteration sit in L2): op(A) rows are contiguous when A is read transposed,
// op(B) columns when B is not.
template <bool TA, bool TB>
__device__ __forceinline__ d4 wave_product(int n, const double* __restrict__ A, const double* __restrict__ B, int i0,
                                           int j0, int q0, int q1) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  int q = q0;
  // eight k quads per round: sixteen independent loads in flight
  for (; q + 8 <= q1; q += 8) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = 4 * (q + u) + lk;
      a[u] = ld_op<TA>(A, n, i0 + li, k);
      b[u] = ld_op<TB>(B, n, k, j0 + li);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
  }
  for (; q < q1; ++q) {
    const int k = 4 * q + lk;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ld_op<TA>(A, n, i0 + li, k), ld_op<TB>(B, n, k, j0 + li), acc, 0, 0,
                                               0);
  }
  return acc;
}

// The workgroup's 16 x 16 tile of op(A) op(B): four waves split the k range
// (split-K), their partial blocks are summed through LDS in wave order.
// Thread t returns element (r, l) = (t / 64, t % 64) of the accumulator
// layout, i.e. C[i0 + l/16 + 4 r][j0 + l%16].
template <bool TA, bool TB>
__device__ __forceinline__ double tile_product(int n, const double* __restrict__ A, const double* __restrict__ B,
                                               int i0, int j0, double (*red)[4][64]) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int nq = (n + 3) / 4;
  const d4 acc = wave_product<TA, TB>(n, A, B, i0, j0, w * nq / 4, (w + 1) * nq / 4);
#pragma unroll
  for (int r = 0; r < 4; ++r) red[w][r][lane] = acc[r];
  __syncthreads();
  const int r = t >> 6;
  return ((red[0][r][lane] + red[1][r][lane]) + red[2][r][lane]) + red[3][r][lane];
}

__device__ __forceinline__ int tile_row(int i0) { return i0 + ((threadIdx.x & 63) >> 4) + 4 * (threadIdx.x >> 6); }
__device__ __forceinline__ int tile_col(int j0) { return j0 + (threadIdx.x & 15); }

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm64(int n, double alpha, const double* __restrict__ A,
                                                const double* __restrict__ B, double beta, double* __restrict__ C,
                                                const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[4][4][64];
  const int i0 = blockIdx.y * kT, j0 = blockIdx.x * kT;
  const double v = tile_product<TA, TB>(n, A, B, i0, j0, red);
  const int row = tile_row(i0), col = tile_col(j0);
  if (row >= n || col >= n) return;
  double* c = C + (size_t)row * n + col;
  *c = beta == 0.0 ? alpha * v : alpha * v + beta * *c;
}

// Sums over the 256 threads of a workgroup in a fixed order: a shuffle tree
// inside each wave, then the four wave sums in wave order.  sh: 8 doubles.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  return v;  // lane 0 holds the sum
}

__device__ __forceinline__ void block_sum256x2(double& a, double& b, double* sh) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane == 0) {
    sh[w] = a;
    sh[4 + w] = b;
  }
  __syncthreads();
  a = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  b = ((sh[4] + sh[5]) + sh[6]) + sh[7];
  __syncthreads();
}

__device__ __forceinline__ double block_sum256(double v, double* sh) {
  double z = 0.0;
  block_sum256x2(v, z, sh);
  return v;
}

// Gershgorin interval of each row of Fo: one wave per row.
__global__ __launch_bounds__(256) void k_sp2_rows(const double* __restrict__ Fo, int n, double* __restrict__ rows) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* row = Fo + (size_t)i * n;
  double r = 0.0;
  for (int j = lane; j < n; j += 64)
    if (j != i) r += fabs(row[j]);
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
  if (lane == 0) {
    rows[2 * i] = row[i] - r;
    rows[2 * i + 1] = row[i] + r;
  }
}

// P0 = (emax I - Fo) / (emax - emin) per 16 x 16 tile, with the tile's
// partial trace and sum of squares; every workgroup reduces the row
// intervals to (emin, emax) itself.
__global__ __launch_bounds__(256) void k_sp2_init(const double* __restrict__ Fo, int n,
                                                  const double* __restrict__ rows, ScfDevState* st,
                                                  double* __restrict__ P, double* __restrict__ part,
                                                  int* __restrict__ done) {
  __shared__ double lo[256], hi[256], sh[256];
  const int t = threadIdx.x;
  double l = DBL_MAX, h = -DBL_MAX;
  for (int i = t; i < n; i += 256) {
    l = fmin(l, rows[2 * i]);
    h = fmax(h, rows[2 * i + 1]);
  }
  lo[t] = l;
  hi[t] = h;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      lo[t] = fmin(lo[t], lo[t + o]);
      hi[t] = fmax(hi[t], hi[t + o]);
    }
    __syncthreads();
  }
  const double emin = lo[0], emax = hi[0];
  const double width = emax > emin ? emax - emin : 1.0;
  const int i0 = blockIdx.y * kT, j0 = blockIdx.x * kT;
  const int i = tile_row(i0), j = tile_col(j0);
  double tr = 0.0, fr = 0.0;
  if (i < n && j < n) {
    const double v = ((i == j ? emax : 0.0) - Fo[(size_t)i * n + j]) / width;
    P[(size_t)i * n + j] = v;
    if (i == j) tr = v;
    fr = v * v;
  }
  block_sum256x2(tr, fr, sh);
  const int wg = blockIdx.y * gridDim.x + blockIdx.x;
  if (t == 0) {
    part[2 * wg] = tr;
    part[2 * wg + 1] = fr;
    if (wg == 0) {
      done[0] = 0;
      st->sp2_used = -1;
      st->bounds[0] = emin;
      st->bounds[1] = emax;
    }
  }
}

// SP2 step k (1-based): P_k from P_(k-1).  Every workgroup sums the partial
// traces of P_(k-1) in the same order, so all take the same branch and the
// same stop decision; workgroup 0 records them.  Once the steps have
// stopped, later steps do nothing (k_sp2_final picks the buffer holding the
// last projector).  P is symmetric, so op(A) = P^T reads rows contiguously.
__global__ __launch_bounds__(256) void k_sp2_step(int n, double nocc, int k, const double* __restrict__ Pin,
                                                  double* __restrict__ Pout, double* __restrict__ part,
                                                  double* __restrict__ eps, int* __restrict__ done,
                                                  ScfDevState* st) {
  if (done[k - 1]) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) done[k] = 1;
    return;
  }
  __shared__ double red[4][4][64];
  __shared__ double sh[8];
  __shared__ int dec[2];
  const int t = threadIdx.x;
  const int nwg = gridDim.x * gridDim.y;
  const int wg = blockIdx.y * gridDim.x + blockIdx.x;
  const double* pin = part + (size_t)((k - 1) & 1) * 2 * nwg;
  double* pout = part + (size_t)(k & 1) * 2 * nwg;
  double tr = 0.0, fr = 0.0;
  for (int i = t; i < nwg; i += 256) {
    tr += pin[2 * i];
    fr += pin[2 * i + 1];
  }
  block_sum256x2(tr, fr, sh);
  if (t == 0) {
    const double e = tr - fr;  // idempotency error of P_(k-1)
    int d = 0;
    if (!(e > 0.0)) d = 1;
    else if (k >= 3 && e < 1e-4 && e >= eps[k - 3]) d = 1;
    dec[0] = d;
    dec[1] = tr > nocc;
    if (wg == 0) {
      eps[k - 1] = e;
      done[k] = d;
      if (d) st->sp2_used = k - 1;
    }
  }
  __syncthreads();
  if (dec[0]) return;  // P_(k-1) is the projector
  const int i0 = blockIdx.y * kT, j0 = blockIdx.x * kT;
  const bool square = dec[1];
  const double acc = tile_product<true, false>(n, Pin, Pin, i0, j0, red);
  const int row = tile_row(i0), col = tile_col(j0);
  double ptr = 0.0, pfr = 0.0;
  if (row < n && col < n) {
    const size_t id = (size_t)row * n + col;
    const double v = square ? acc : 2.0 * Pin[id] - acc;
    Pout[id] = v;
    if (row == col) ptr = v;
    pfr = v * v;
  }
  block_sum256x2(ptr, pfr, sh);
  if (t == 0) {
    pout[2 * wg] = ptr;
    pout[2 * wg + 1] = pfr;
  }
}

// Host-driven SP2 (the eager loop of the RI engines; the products run on
// rocBLAS).  Both kernels write (trace, sum of squares) partials per
// workgroup of a fixed grid, which k_sp2e_sum adds in a fixed order, so the
// stop decisions are reproducible.
constexpr int kSp2eBlocks = 512;

// P0 = (emax I - Fo) / (emax - emin) from the Gershgorin rows of Fo.
__global__ __launch_bounds__(256) void k_sp2e_init(const double* __restrict__ Fo, int n,
                                                   const double* __restrict__ rows, double* __restrict__ P,
                                                   double* __restrict__ part) {
  __shared__ double lo[256], hi[256], sh[8];
  const int t = threadIdx.x;
  double l = DBL_MAX, h = -DBL_MAX;
  for (int i = t; i < n; i += 256) {
    l = fmin(l, rows[2 * i]);
    h = fmax(h, rows[2 * i + 1]);
  }
  lo[t] = l;
  hi[t] = h;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      lo[t] = fmin(lo[t], lo[t + o]);
      hi[t] = fmax(hi[t], hi[t + o]);
    }
    __syncthreads();
  }
  const double emax = hi[0], width = hi[0] > lo[0] ? hi[0] - lo[0] : 1.0;
  double tr = 0.0, fr = 0.0;
  const size_t nn = (size_t)n * n;
  for (size_t id = (size_t)blockIdx.x * 256 + t; id < nn; id += (size_t)gridDim.x * 256) {
    const size_t i = id / n, j = id - i * n;
    const double v = ((i == j ? emax : 0.0) - Fo[id]) / width;
    P[id] = v;
    if (i == j) tr += v;
    fr += v * v;
  }
  block_sum256x2(tr, fr, sh);
  if (t == 0) {
    part[2 * blockIdx.x] = tr;
    part[2 * blockIdx.x + 1] = fr;
  }
}

// P <- P2 (square) or 2 P - P2, in place.
__global__ __launch_bounds__(256) void k_sp2e_apply(double* __restrict__ P, const double* __restrict__ P2, int n,
                                                    int square, double* __restrict__ part) {
  __shared__ double sh[8];
  double tr = 0.0, fr = 0.0;
  const size_t nn = (size_t)n * n;
  for (size_t id = (size_t)blockIdx.x * 256 + threadIdx.x; id < nn; id += (size_t)gridDim.x * 256) {
    const double v = square ? P2[id] : 2.0 * P[id] - P2[id];
    P[id] = v;
    if (id % (size_t)(n + 1) == 0) tr += v;
    fr += v * v;
  }
  block_sum256x2(tr, fr, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = tr;
    part[2 * blockIdx.x + 1] = fr;
  }
}

__global__ __launch_bounds__(256) void k_sp2e_sum(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[8];
  double tr = 0.0, fr = 0.0;
  for (int i = threadIdx.x; i < kSp2eBlocks; i += 256) {
    tr += part[2 * i];
    fr += part[2 * i + 1];
  }
  block_sum256x2(tr, fr, sh);
  if (threadIdx.x == 0) {
    out[0] = tr;
    out[1] = fr;
  }
}

// The projector of the stopped steps (P_used, in buffer used % 2) copied to
// P_out; the last step's buffer if the steps never stopped.
__global__ void k_sp2_final(int n, int steps, const double* __restrict__ b0, const double* __restrict__ b1,
                            double* __restrict__ P, const ScfDevState* st) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n * n) return;
  const int last = st->sp2_used >= 0 ? st->sp2_used : steps;
  P[id] = (last & 1) ? b1[id] : b0[id];
}

// partial sums of D.(H + F) per workgroup
__global__ __launch_bounds__(256) void k_scf_energy_part(const double* __restrict__ D, const double* __restrict__ H,
                                                         const double* __restrict__ F, size_t nn,
                                                         double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nn; i += (size_t)gridDim.x * 256)
    s += D[i] * (H[i] + F[i]);
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_scf_energy_final(const double* __restrict__ part, double conv, ScfDevState* st) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < kEnergyBlocks; ++b) s += part[b];
  const double E = 0.5 * s;
  st->E = E;
  if (fabs(E - st->Eprev) < conv) {
    st->converged = 1;
  } else {
    st->iter += 1;
    st->Eprev = E;
  }
}

__global__ void k_antisym(const double* __restrict__ A, int n, double* __restrict__ e, const ScfDevState* st) {
  if (st->converged) return;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)n * n) return;
  const size_t i = id / n, j = id % n;
  e[id] = A[id] - A[j * n + i];
}

// DIIS, part 1: store (e, F) in the ring slot and the partial dots of e
// with every held error.
__global__ __launch_bounds__(256) void k_diis_store(const double* __restrict__ e, const double* __restrict__ F,
                                                    double* __restrict__ errs, double* __restrict__ focks,
                                                    double* __restrict__ part, size_t nn, int size,
                                                    const ScfDevState* st) {
  if (st->converged) return;
  __shared__ double sh[256];
  const int count = st->diis_count;
  const int slot = count % size;
  const int np = count + 1 < size ? count + 1 : size;
  double acc[kGraphDiisMax];
#pragma unroll
  for (int i = 0; i < kGraphDiisMax; ++i) acc[i] = 0.0;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += (size_t)gridDim.x * blockDim.x) {
    const double ev = e[id];
    errs[(size_t)slot * nn + id] = ev;
    focks[(size_t)slot * nn + id] = F[id];
#pragma unroll
    for (int i = 0; i < kGraphDiisMax; ++i)
      if (i < np) acc[i] += (i == slot ? ev : errs[(size_t)i * nn + id]) * ev;
  }
#pragma unroll
  for (int i = 0; i < kGraphDiisMax; ++i) {
    if (i >= np) break;
    const double v = block_sum256(acc[i], sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * size + i] = v;
  }
}

// DIIS, part 2 (one wave): the Gram row of the new error (lane i sums slot
// i over the blocks), then the bordered system [G -1; -1 0] [c; l] = [0; -1]
// by Gaussian elimination with partial pivoting, lane r holding row r in
// registers (the pivot is the first row of largest magnitude, as in a
// sequential search).  A zero pivot leaves F as it is.
__global__ __launch_bounds__(64) void k_diis_solve(const double* __restrict__ part, int nblocks, int size,
                                                   ScfDevState* st) {
  if (st->converged) return;
  constexpr int K = kGraphDiisMax;
  constexpr int K1 = kGraphDiisMax + 1;
  const int lane = threadIdx.x;
  const int count = st->diis_count;
  const int slot = count % size;
  const int np = count + 1 < size ? count + 1 : size;
  for (int i = 0; i < np; ++i) {
    double r = 0.0;
    for (int b = lane; b < nblocks; b += 64) r += part[(size_t)b * size + i];
    r = wave_sum(r);
    if (lane == 0) {
      st->gram[slot * K + i] = r;
      st->gram[i * K + slot] = r;
    }
  }
  __syncthreads();
  if (lane == 0) {
    st->diis_count = count + 1;
    st->diis_ok = 0;
  }
  if (np < 2) return;
  const int ns = np + 1;
  double row[K1];
  double rhs = 0.0;
#pragma unroll
  for (int j = 0; j < K1; ++j) {
    double v = 0.0;
    if (lane < np && j < np) v = st->gram[lane * K + j];
    if ((lane < np && j == np) || (lane == np && j < np)) v = -1.0;
    row[j] = v;
  }
  if (lane == np) rhs = -1.0;
  bool singular = false;
#pragma unroll 17
  for (int c = 0; c < K1; ++c) {
    if (c >= ns || singular) continue;
    // pivot: largest |row[c]| among rows c..ns-1, the first one on ties
    double best = (lane >= c && lane < ns) ? fabs(row[c]) : -1.0;
    int piv = lane;
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o);
      const int op = __shfl_xor(piv, o);
      if (ob > best || (ob == best && op < piv)) {
        best = ob;
        piv = op;
      }
    }
    if (best == 0.0) {  // singular: F is kept (every lane sees the same best)
      singular = true;
      continue;
    }
    // swap rows c and piv
#pragma unroll
    for (int j = 0; j < K1; ++j) {
      const double vc = __shfl(row[j], c), vp = __shfl(row[j], piv);
      if (lane == c) row[j] = vp;
      else if (lane == piv) row[j] = vc;
    }
    {
      const double vc = __shfl(rhs, c), vp = __shfl(rhs, piv);
      if (lane == c) rhs = vp;
      else if (lane == piv) rhs = vc;
    }
    // eliminate below the pivot
    const double pc = __shfl(row[c], c);
    const double f = row[c] / pc;
#pragma unroll
    for (int j = 0; j < K1; ++j) {
      const double pj = __shfl(row[j], c);
      if (lane > c && lane < ns && j >= c) row[j] -= f * pj;
    }
    const double prhs = __shfl(rhs, c);
    if (lane > c && lane < ns) rhs -= f * prhs;
  }
  if (singular) return;
  // back substitution: x_r = (rhs_r - sum_{q>r} row_r[q] x_q) / row_r[r],
  // the sum taken as the x_q become known
  double x = 0.0;
#pragma unroll 17
  for (int r = K1 - 1; r >= 0; --r) {
    if (r >= ns) continue;
    double xr = 0.0;
    if (lane == r) {
      x = rhs / row[r];
      xr = x;
    }
    xr = __shfl(xr, r);
    if (lane < r) rhs -= row[r] * xr;
  }
  if (lane < np) st->coef[lane] = x;
  if (lane == 0) st->diis_ok = 1;
}

// DIIS, part 3: F = sum_i c_i F_i over the held slots.
__global__ void k_diis_combine(const double* __restrict__ focks, size_t nn, int size, double* __restrict__ F,
                               const ScfDevState* st) {
  if (st->converged || !st->diis_ok) return;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nn) return;
  const int np = st->diis_count < size ? st->diis_count : size;
  double s = 0.0;
  for (int i = 0; i < np; ++i) s += st->coef[i] * focks[(size_t)i * nn + id];
  F[id] = s;
}

}  // namespace

hipError_t launch_gemm64(bool ta, bool tb, int n, double alpha, const double* A, const double* B, double beta,
                         double* C, const int* skip, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int nb = (n + kT - 1) / kT;
  const dim3 grid(nb, nb), block(256);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm64<false, false>), grid, block, 0, s, n, alpha, A, B, beta, C, skip);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm64<true, false>), grid, block, 0, s, n, alpha, A, B, beta, C, skip);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm64<false, true>), grid, block, 0, s, n, alpha, A, B, beta, C, skip);
  else hipLaunchKernelGGL((k_gemm64<true, true>), grid, block, 0, s, n, alpha, A, B, beta, C, skip);
  return hipGetLastError();
}

size_t sp2_part_doubles(int n) {
  const size_t nb = (size_t)(n + kT - 1) / kT;
  return 2 * 2 * nb * nb;
}

int sp2_workgroups(int n) {
  const int nb = (n + kT - 1) / kT;
  return nb * nb;
}

hipError_t launch_sp2(const double* Fo, int n, int nocc, int steps, const Sp2Work& w, ScfDevState* st,
                      hipStream_t s) {
  const int nb = (n + kT - 1) / kT;
  if (w.nwg != nb * nb || steps < 1) return hipErrorInvalidValue;
  const dim3 grid(nb, nb), block(256);
  hipLaunchKernelGGL(k_sp2_rows, dim3((unsigned)((n + 3) / 4)), block, 0, s, Fo, n, w.rows);
  hipLaunchKernelGGL(k_sp2_init, grid, block, 0, s, Fo, n, w.rows, st, w.buf[0], w.part, w.done);
  for (int k = 1; k <= steps; ++k)
    hipLaunchKernelGGL(k_sp2_step, grid, block, 0, s, n, (double)nocc, k, w.buf[(k - 1) & 1], w.buf[k & 1], w.part,
                       w.eps, w.done, st);
  const size_t m = (size_t)n * n;
  hipLaunchKernelGGL(k_sp2_final, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, n, steps, w.buf[0], w.buf[1],
                     w.out, st);
  return hipGetLastError();
}

size_t sp2e_part_doubles() { return 2 * (size_t)kSp2eBlocks; }

hipError_t launch_sp2e_init(const double* Fo, int n, double* rows, double* P, double* part, double* trfr,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_sp2_rows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, Fo, n, rows);
  hipLaunchKernelGGL(k_sp2e_init, dim3(kSp2eBlocks), dim3(256), 0, s, Fo, n, rows, P, part);
  hipLaunchKernelGGL(k_sp2e_sum, dim3(1), dim3(256), 0, s, part, trfr);
  return hipGetLastError();
}

hipError_t launch_sp2e_apply(double* P, const double* P2, int n, bool square, double* part, double* trfr,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_sp2e_apply, dim3(kSp2eBlocks), dim3(256), 0, s, P, P2, n, square ? 1 : 0, part);
  hipLaunchKernelGGL(k_sp2e_sum, dim3(1), dim3(256), 0, s, part, trfr);
  return hipGetLastError();
}

hipError_t launch_scf_energy(const double* D, const double* H, const double* F, size_t nn, double conv,
                             double* part, ScfDevState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_scf_energy_part, dim3(kEnergyBlocks), dim3(256), 0, s, D, H, F, nn, part);
  hipLaunchKernelGGL(k_scf_energy_final, dim3(1), dim3(64), 0, s, part, conv, st);
  return hipGetLastError();
}

hipError_t launch_antisym(const double* A, int n, double* e, const ScfDevState* st, hipStream_t s) {
  const size_t m = (size_t)n * n;
  hipLaunchKernelGGL(k_antisym, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, A, n, e, st);
  return hipGetLastError();
}

hipError_t launch_diis(const double* e, double* F, double* errs, double* focks, double* part, size_t nn, int size,
                       ScfDevState* st, hipStream_t s) {
  if (size < 1 || size > kGraphDiisMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_diis_store, dim3(kDiisBlocks), dim3(256), 0, s, e, F, errs, focks, part, nn, size, st);
  hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(64), 0, s, part, kDiisBlocks, size, st);
  hipLaunchKernelGGL(k_diis_combine, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, focks, nn, size, F, st);
  return hipGetLastError();
}

}  // namespace gansu
