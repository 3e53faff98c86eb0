// Incremental RI exchange on the fp16 matrix cores of gfx950.
//
// For occupied factors X (D = X X^T, X = sqrt(2) C_occ) the RI exchange is
// K(X) = sum_P (B^P X)(B^P X)^T.  Between two builds the occupied space
// moves little: with X_n rotated onto the previous factor (orthogonal
// Procrustes, R = polar(X_n^T X_prev)), E = X_n R - X_prev is small and
//   D_n - D_prev = A E^T + E A^T,   A = (X_n R + X_prev) / 2,
//   K_n - K_prev = T_A T_E^T + T_E T_A^T,  T_A = B A, T_E = B E,
// with T_A = T_prev + T_E / 2 where T_prev = B X_prev is kept from the last
// build.  Two kernels carry the flops:
//   k_rik_tstep:  T_E = B E over all rows (m, P) of B (B as pre-split fp16
//                 hi/lo planes, E^T in fp16 hi/lo), updating T_A, T_E in
//                 place and the fitted density change
//                 dW_P = 2 sum_{n,i} T_A[n][P][i] E[n][i];
//   k_rik_dk:     Z = T_A T_E^T (K dimension Maux nocc), split-k partials;
//                 K_n = K_prev + Z + Z^T is summed in FP64 (k_rik_reduce).
// Each product runs as three fp16 MFMAs on a hi/lo split of both operands
// (x s = hi + lo, hi = fp16(x s), lo = fp16(x s - hi); hi*hi + hi*lo + lo*hi
// carries 22 bits), with FP32 accumulation; operands are scaled by powers of
// two so that their largest element sits at 2^14.  The rounding error of the
// increment scales with |E|, so the FP64 Fock matrix accumulates K_prev plus
// small, accurately formed increments; the caller refreshes K in FP64 from
// time to time.  With the previous state taken as zero (X_prev = 0, a
// "fresh" build) the same kernels give the whole of K = T T^T, T = B X,
// at the accuracy of the split products.
//
// MFMA v_mfma_f32_32x32x16_f16: lane l holds row (or column) l % 32 of an
// operand and 8 consecutive k of k-group l / 32; the 16 results of lane l
// are rows 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.  A and B share the
// k mapping, so any hardware permutation of k within a group cancels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ri_kernels.hpp"

namespace gansu {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16x __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ inline f16x mfma(const h8& a, const h8& b, const f16x& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// power-of-two scale taking max|x| to at most 2^14 (1 when max is 0)
__device__ inline float op_scale(float mx) {
  if (!(mx > 0.0f)) return 1.0f;
  int e;
  frexpf(mx, &e);  // mx = f 2^e, 0.5 <= f < 1
  return ldexpf(1.0f, 14 - e);
}

__device__ inline void split1(float x, float s, _Float16& hi, _Float16& lo) {
  const float y = x * s;
  hi = (_Float16)y;
  lo = (_Float16)(y - (float)hi);
}

// maximum of non-negative floats through their bit patterns
__device__ inline void atomic_max_pos(unsigned* w, float v) { atomicMax(w, __float_as_uint(v)); }

// ------------------------------------------------------------ E split -----
// One workgroup: max|E| -> scale sE (st->sE), then E^T in fp16 hi/lo,
// [c][k] with c < ncol_pad, k < kpad, zero outside nocc x M.
__global__ __launch_bounds__(1024) void k_rik_split_e(const double* __restrict__ E, int M, int nocc, int ncol_pad,
                                                      int kpad, RikState* st, uint16_t* __restrict__ Ehi_,
                                                      uint16_t* __restrict__ Elo_) {
  _Float16* Ehi = reinterpret_cast<_Float16*>(Ehi_);
  _Float16* Elo = reinterpret_cast<_Float16*>(Elo_);
  __shared__ float red[1024];
  float mx = 0.0f;
  const size_t n = (size_t)M * nocc;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) mx = fmaxf(mx, fabsf((float)E[i]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const float s = op_scale(red[0]);
  if (threadIdx.x == 0) {
    st->sE = s;
    st->emax = red[0];
  }
  const size_t tot = (size_t)ncol_pad * kpad;
  for (size_t i = threadIdx.x; i < tot; i += blockDim.x) {
    const int c = (int)(i / kpad), k = (int)(i % kpad);
    _Float16 hi = 0, lo = 0;
    if (c < nocc && k < M) split1((float)E[(size_t)k * nocc + c], s, hi, lo);
    Ehi[i] = hi;
    Elo[i] = lo;
  }
}

// ------------------------------------------------------------ B planes -----
// B (R x M, FP64, rows m Maux + P) as the fp16 hi/lo split of its FP32
// rounding scaled by sB, Bp[r][k / 8][hi 8 | lo 8] for k < kpad (zero for
// k >= M): the 8 consecutive k of both planes that one MFMA lane takes are
// one 32-byte piece.  Built once per geometry; the T update reads it as its
// A operand and the J increment as B = (hi + lo) / sB (22 bits).
__global__ void k_rik_planes(const double* __restrict__ B, long R, int M, int kpad, float sB,
                             uint16_t* __restrict__ Bp_) {
  _Float16* Bp = reinterpret_cast<_Float16*>(Bp_);
  const int ng = kpad / 8;
  const size_t n = (size_t)R * ng;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i / ng;
    const int k0 = (int)(i % ng) * 8;
    h8 hi, lo;
    if ((M & 1) == 0 && k0 + 8 <= M) {
      // a whole group inside the row: four 16-byte loads (rows start 16-byte aligned for even M)
      const double2* src = reinterpret_cast<const double2*>(B + r * M + k0);
      double x[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 v = src[j];
        x[2 * j] = v.x;
        x[2 * j + 1] = v.y;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        _Float16 h, l;
        split1((float)x[e], sB, h, l);
        hi[e] = h;
        lo[e] = l;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        _Float16 h = 0, l = 0;
        if (k0 + e < M) split1((float)B[r * M + k0 + e], sB, h, l);
        hi[e] = h;
        lo[e] = l;
      }
    }
    *(h8*)(Bp + i * 16) = hi;
    *(h8*)(Bp + i * 16 + 8) = lo;
  }
}

// ------------------------------------------------------------ T update -----
// Rows r = m Maux + P of B.  A workgroup of four waves takes 128 rows; each
// wave 32 rows and all NBW 32-column blocks of E (16 NBW accumulators per
// lane), so E^T is read once per 128 rows and each B row by one wave.  Per
// 32-k stage: the wave's A operand (hi/lo planes of its rows) is loaded from
// global memory two stages ahead into a two-slot register ring; E^T (fp16
// hi/lo, all columns) goes through registers into one of two LDS buffers one
// stage ahead, so one barrier per stage separates the writes of stage s + 1
// from the reads of stage s - 1.  The loop body is unrolled over the two
// buffer parities (nk is even), which keeps the ring in fixed registers.
// TAIL: the last, partial tile (per-element row checks in the epilogue).
constexpr int kTRows = 128;  // rows per workgroup

template <int NBW, bool TAIL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, NBW > 5 ? 1 : 8))) void k_rik_tstep(
    RikTArgs a) {
  constexpr int NCOL = NBW * 32;
  constexpr int NPIECE = 2 * NCOL * 4;            // 16-byte pieces of one E^T stage
  constexpr int NPT = (NPIECE + 255) / 256;        // per thread
  __shared__ __attribute__((aligned(16))) _Float16 lds[2][2][NCOL][40];  // [buffer][hi/lo][col][k] (+8 pad)
  const int t = threadIdx.x, lane = t & 63;
  const int g = lane >> 5, li = lane & 31;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const long rw = a.row0 + (long)blockIdx.x * kTRows + wv * 32;  // first row of the wave
  const int kq = a.kpad / 8;
  const int nk = a.kpad / 32;
  const long myrow = rw + li < a.R ? rw + li : a.R - 1;  // rows past R compute, never stored
  const u32x4* arow = reinterpret_cast<const u32x4*>(a.Bp) + (size_t)myrow * kq * 2;
  const float sB = a.sB, sE = a.st->sE;
  f16x acc[NBW];
#pragma unroll
  for (int b = 0; b < NBW; ++b)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[b][r] = 0.0f;
  // E^T stage pieces of this thread: piece p = t + 256 j -> (hi/lo, column, 8-k quarter)
  u32x4 ereg[NPT];
#define RIK_EPIECE(j) const int piece_ = t + 256 * (j), hl_ = piece_ / (NCOL * 4), rem_ = piece_ % (NCOL * 4)
#define RIK_ESRC(kc) (*(const u32x4*)((hl_ ? a.Elo : a.Ehi) + (size_t)(rem_ >> 2) * a.kpad + (kc) * 32 + (rem_ & 3) * 8))
#define RIK_EDST(buf) (*(u32x4*)&lds[buf][hl_][rem_ >> 2][(rem_ & 3) * 8])
  auto eload = [&](int kc) {
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      RIK_EPIECE(j);
      if (NPIECE % 256 == 0 || piece_ < NPIECE) ereg[j] = RIK_ESRC(kc);
    }
  };
  auto estore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
      RIK_EPIECE(j);
      if (NPIECE % 256 == 0 || piece_ < NPIECE) RIK_EDST(buf) = ereg[j];
    }
  };
  // A operand ring: [slot][k-half][hi/lo]
  u32x4 ap[2][2][2];
  auto load_a = [&](int slot, int kc) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const u32x4* p = arow + (size_t)(kc * 4 + ks * 2 + g) * 2;
      ap[slot][ks][0] = p[0];
      ap[slot][ks][1] = p[1];
    }
  };
  eload(0);
  load_a(0, 0);
  load_a(1, 1);
  estore(0);
  eload(1);
  __syncthreads();
  auto stage = [&]<int P>(int kc) {
    h8 ahi[2], alo[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      ahi[ks] = __builtin_bit_cast(h8, ap[P][ks][0]);
      alo[ks] = __builtin_bit_cast(h8, ap[P][ks][1]);
    }
    if (kc + 2 < nk) load_a(P, kc + 2);
    // the three products of a block are dependent: issue them for all
    // blocks in turn, so that NBW independent accumulators separate them
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      h8 bhi[NBW], blo[NBW];
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
        bhi[b] = *(const h8*)&lds[P][0][b * 32 + li][ks * 16 + g * 8];
        blo[b] = *(const h8*)&lds[P][1][b * 32 + li][ks * 16 + g * 8];
      }
#pragma unroll
      for (int b = 0; b < NBW; ++b) acc[b] = mfma(ahi[ks], bhi[b], acc[b]);
#pragma unroll
      for (int b = 0; b < NBW; ++b) acc[b] = mfma(ahi[ks], blo[b], acc[b]);
#pragma unroll
      for (int b = 0; b < NBW; ++b) acc[b] = mfma(alo[ks], bhi[b], acc[b]);
    }
    if (kc + 1 < nk) {
      estore(P ^ 1);
      if (kc + 2 < nk) eload(kc + 2);
    }
    __syncthreads();
  };
  for (int kc = 0; kc < nk; kc += 2) {
    stage.template operator()<0>(kc);
    stage.template operator()<1>(kc + 1);
  }
#undef RIK_EPIECE
#undef RIK_ESRC
#undef RIK_EDST

  // epilogue: T_E, T_A = T_prev + T_E / 2 (T_prev = T_A_old + T_E_old / 2),
  // dW rows = 2 sum_i T_A E (each row belongs to one wave: stored, not
  // accumulated), maxima for the next product's scales.  The rows of a wave
  // span at most two m (32 <= Maux), so E is read for m_A and m_A + 1 only.
  const float inv = 1.0f / (sB * sE);
  float mta = 0.0f, mte = 0.0f;
  const int nocc = a.nocc;
  const long nA = rw / a.Maux, nbnd = (nA + 1) * a.Maux;
  const long nB = nA + 1 < a.M ? nA + 1 : nA;
  float part[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) part[r] = 0.0f;
  if constexpr (!TAIL) {
    // wave-uniform row base, per-lane 32-bit byte offsets; all 32 loads of
    // a block are issued before its stores
    const unsigned nb4 = (unsigned)nocc * 4u;
    const unsigned lane_off = (unsigned)(4 * g) * nb4 + (unsigned)li * 4u;
    unsigned offr[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) offr[r] = lane_off + (unsigned)(8 * (r >> 2) + (r & 3)) * nb4;
    char* bA = (char*)(a.TA + (size_t)rw * nocc);
    char* bE = (char*)(a.TE + (size_t)rw * nocc);
    const double* eAp = a.E + (size_t)nA * nocc + li;
    const double* eBp = a.E + (size_t)nB * nocc + li;
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
      if (b * 32 + li < nocc) {
        const float eA = (float)eAp[b * 32], eB = (float)eBp[b * 32];
        float tao[16], teo[16];
        if (!a.fresh) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            tao[r] = *(const float*)(bA + offr[r] + b * 128);
            teo[r] = *(const float*)(bE + offr[r] + b * 128);
          }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const long row = rw + 4 * g + 8 * (r >> 2) + (r & 3);
          const float te = acc[b][r] * inv;
          const float tp = a.fresh ? 0.0f : tao[r] + 0.5f * teo[r];
          const float ta = tp + 0.5f * te;
          *(float*)(bA + offr[r] + b * 128) = ta;
          *(float*)(bE + offr[r] + b * 128) = te;
          mta = fmaxf(mta, fabsf(ta));
          mte = fmaxf(mte, fabsf(te));
          part[r] += ta * (row < nbnd ? eA : eB);
        }
      }
      __asm__ volatile("" ::: "memory");  // one block's loads in flight at a time
    }
  } else {
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
      const int c = b * 32 + li;
      if (c >= nocc) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long row = rw + 4 * g + 8 * (r >> 2) + (r & 3);
        if (row >= a.R) continue;
        const size_t idx = (size_t)row * nocc + c;
        const float te = acc[b][r] * inv;
        const float tp = a.fresh ? 0.0f : a.TA[idx] + 0.5f * a.TE[idx];
        const float ta = tp + 0.5f * te;
        a.TA[idx] = ta;
        a.TE[idx] = te;
        mta = fmaxf(mta, fabsf(ta));
        mte = fmaxf(mte, fabsf(te));
        part[r] += ta * (float)a.E[(size_t)(row < nbnd ? nA : nB) * nocc + c];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float v = part[r];
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const long row = rw + 8 * (r >> 2) + 4 * g + (r & 3);
    if (li == 0 && (!TAIL || row < a.R)) a.dWrow[row] = 2.0 * (double)v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mta = fmaxf(mta, __shfl_xor(mta, o, 64));
    mte = fmaxf(mte, __shfl_xor(mte, o, 64));
  }
  // the workgroup's maxima go to its own slot (k_rik_maxred reduces them)
  __shared__ float wmax[2][4];
  if (lane == 0) {
    wmax[0][wv] = mta;
    wmax[1][wv] = mte;
  }
  __syncthreads();
  if (t == 0) {
    const size_t w = (size_t)a.wg0 + blockIdx.x;
    a.wgmax[2 * w] = __float_as_uint(fmaxf(fmaxf(wmax[0][0], wmax[0][1]), fmaxf(wmax[0][2], wmax[0][3])));
    a.wgmax[2 * w + 1] = __float_as_uint(fmaxf(fmaxf(wmax[1][0], wmax[1][1]), fmaxf(wmax[1][2], wmax[1][3])));
  }
}

// max|T_A|, max|T_E| over the workgroups of k_rik_tstep (one workgroup)
__global__ __launch_bounds__(1024) void k_rik_maxred(const unsigned* __restrict__ wgmax, long nwg, RikState* st) {
  __shared__ unsigned m[2][1024];
  unsigned a0 = 0u, a1 = 0u;
  for (long i = threadIdx.x; i < nwg; i += 1024) {
    a0 = max(a0, wgmax[2 * i]);
    a1 = max(a1, wgmax[2 * i + 1]);
  }
  m[0][threadIdx.x] = a0;
  m[1][threadIdx.x] = a1;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      m[0][threadIdx.x] = max(m[0][threadIdx.x], m[0][threadIdx.x + o]);
      m[1][threadIdx.x] = max(m[1][threadIdx.x], m[1][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st->ta_max = m[0][0];
    st->te_max = m[1][0];
  }
}

// --------------------------------------------------------------- dK -------
// Zpart[chunk][m][n] = sum over the chunk's k of T_A[m][k] T_E[n][k]
// (K dimension Kd = Maux nocc).  Workgroup tile 256 x 128 (T_A rows x T_E
// rows), eight waves of 64 x 64 (2 x 2 MFMA blocks), 32-k stages: the T_A
// panel is re-read by 13 column tiles and the T_E panel by 7 row tiles at
// M = 1600 (26 panel reads per pair with 128 x 128 tiles, 20 here).  Stages
// are loaded into registers two ahead of the one the MFMAs use (two register
// sets) and split into fp16 hi/lo on their way into LDS, which holds two
// stages, so one barrier per stage separates the writes of stage s + 1 from
// the reads of stage s - 1.
constexpr int kDkWM = 4, kDkWN = 2;                 // waves along T_A and T_E rows
constexpr int kDkTM = 64 * kDkWM, kDkTN = 64 * kDkWN;  // tile rows of T_A, T_E
constexpr int kDkThreads = 64 * kDkWM * kDkWN;

// Instruction order of one loop half of k_rik_dk (per wave: 24 MFMAs, about
// 140 VALU of the hi/lo split, 16 LDS reads, 8 LDS writes, 6 global loads).
// Every wave of the workgroup runs the same phase between two barriers, so
// the split of the next stage must be threaded between the MFMAs of this
// one inside each wave: the first stage operands are read, then each MFMA
// is followed by six VALU and one LDS read (first eight) or write, and the
// second twelve also by a global load.
__device__ __forceinline__ void dk_interleave() {
  __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);  // LDS reads, first k half
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);  // VALU
    if (i < 8) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
    else __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);        // LDS write
  }
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);
    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load
  }
}

template <bool VEC>
__global__ __launch_bounds__(kDkThreads) void k_rik_dk(RikDkArgs a) {
  constexpr int NT = kDkThreads, NA = kDkTM * 8 / NT, NB = kDkTN * 8 / NT;  // float4 per thread and stage
  static_assert(NA * NT == kDkTM * 8 && NB * NT == kDkTN * 8, "staging");
  __shared__ __attribute__((aligned(16))) _Float16 sa[2][2][kDkTM][40];  // [buf][hi/lo][row][k]
  __shared__ __attribute__((aligned(16))) _Float16 sb[2][2][kDkTN][40];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int g = lane >> 5, li = lane & 31;
  const int wm = wave / kDkWN, wn = wave % kDkWN;
  const int ntm = (a.M + kDkTM - 1) / kDkTM, ntn = (a.M + kDkTN - 1) / kDkTN;
  // Workgroups reach the eight XCDs round-robin (XCD = linear id % 8): XCD x
  // walks the tiles of chunks x, x + 8, ... in row-major tile order, so the
  // workgroups resident on one XCD share the k stages of one chunk in its L2.
  const int nt2 = ntm * ntn;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int chunk = (slot / nt2) * 8 + xcd, tile = slot % nt2;
  if (chunk >= a.nchunk) return;
  const int tm = tile / ntn, tn = tile % ntn;
  const long k_begin = (long)chunk * a.kchunk;
  const long k_end = k_begin + a.kchunk < a.Kd ? k_begin + a.kchunk : a.Kd;
  const int nst = k_end > k_begin ? (int)((k_end - k_begin + 31) / 32) : 0;
  const float sA = op_scale(__uint_as_float(a.st->ta_max));
  const float sE = op_scale(__uint_as_float(a.st->te_max));

  // a stage: 32 k of the tile's T_A rows (A side) and T_E rows (B side) as
  // float4; piece idx -> row idx / 8, k offset 4 (idx % 8).  With VEC
  // (Kd % 4 == 0, so chunk ends are multiples of 4) the loads are
  // unconditional from clamped addresses and masked afterwards, which keeps
  // the loop free of divergent branches (and the load counters exact).
  struct Stage {
    float4 a[NA], b[NB];
    float fa[NA], fb[NB];  // 1 for loads inside the matrices, 0 for clamped ones
  };
  auto load1 = [&](const float* base, int row, long k, float4& v, float& f) {
    if constexpr (VEC) {
      const long kc = k + 4 <= k_end ? k : k_end - 4;
      const int rc = row < a.M ? row : a.M - 1;
      v = *(const float4*)(base + (size_t)rc * a.Kd + kc);
      f = (row < a.M && k < k_end) ? 1.0f : 0.0f;  // applied in stash, once the load has landed
    } else {
      float x[4] = {0.f, 0.f, 0.f, 0.f};
      for (int e = 0; e < 4; ++e)
        if (k + e < k_end && row < a.M) x[e] = base[(size_t)row * a.Kd + k + e];
      v = make_float4(x[0], x[1], x[2], x[3]);
      f = 1.0f;
    }
  };
  auto fetch = [&](int s, Stage& v) {
    const long k0 = k_begin + (long)s * 32;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int idx = t + NT * j;
      load1(a.TA, tm * kDkTM + (idx >> 3), k0 + (idx & 7) * 4, v.a[j], v.fa[j]);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int idx = t + NT * j;
      load1(a.TE, tn * kDkTN + (idx >> 3), k0 + (idx & 7) * 4, v.b[j], v.fb[j]);
    }
  };
  // split four floats into hi/lo halves and store each as one 8-byte word
  auto put4 = [&](const float4& v, float s, _Float16* hi, _Float16* lo) {
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    h4 h, l;
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      _Float16 p, q;
      split1(x[e], s, p, q);
      h[e] = p;
      l[e] = q;
    }
    *(h4*)hi = h;
    *(h4*)lo = l;
  };
  auto stash = [&](const Stage& v, int buf) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int idx = t + NT * j, r = idx >> 3, q = idx & 7;
      put4(v.a[j], sA * v.fa[j], &sa[buf][0][r][q * 4], &sa[buf][1][r][q * 4]);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int idx = t + NT * j, r = idx >> 3, q = idx & 7;
      put4(v.b[j], sE * v.fb[j], &sb[buf][0][r][q * 4], &sb[buf][1][r][q * 4]);
    }
  };

  f16x acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  auto compute = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      h8 ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm * 64 + i * 32 + li;
        ahi[i] = *(const h8*)&sa[buf][0][r][ks * 16 + g * 8];
        alo[i] = *(const h8*)&sa[buf][1][r][ks * 16 + g * 8];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = wn * 64 + j * 32 + li;
        bhi[j] = *(const h8*)&sb[buf][0][r][ks * 16 + g * 8];
        blo[j] = *(const h8*)&sb[buf][1][r][ks * 16 + g * 8];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma(ahi[i], bhi[j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma(ahi[i], blo[j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma(alo[i], bhi[j], acc[i][j]);
    }
  };

  // stage s sits in LDS buffer s % 2 and, before that, in register set s % 2;
  // the loop takes stages in pairs, and every fetch is issued (stages past
  // the end fetch the last one again), so the loop has no branches and each
  // stash waits only for its own stage's loads
  Stage r0, r1;
  const int last = nst - 1;
  if (nst > 0) {
    fetch(0, r0);
    stash(r0, 0);
    fetch(1 < last ? 1 : last, r1);
    fetch(2 < last ? 2 : last, r0);
  }
  __syncthreads();
  int s = 0;
  for (; s + 2 <= nst; s += 2) {
    // stage s in buffer 0, stage s + 1 in r1, stage s + 2 in r0
    compute(0);
    stash(r1, 1);
    fetch(s + 3 < last ? s + 3 : last, r1);
    dk_interleave();
    __syncthreads();
    // stage s + 1 in buffer 1, stage s + 2 in r0, stage s + 3 in r1
    compute(1);
    stash(r0, 0);
    fetch(s + 4 < last ? s + 4 : last, r0);
    dk_interleave();
    __syncthreads();
  }
  if (s < nst) compute(0);  // an odd last stage, already in buffer 0

  const float inv = 1.0f / (sA * sE);
  float* Z = a.Zpart + (size_t)chunk * a.M * a.M;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = tn * kDkTN + wn * 64 + j * 32 + li;
      if (col >= a.M) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = tm * kDkTM + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
        if (row < a.M) Z[(size_t)row * a.M + col] = acc[i][j][r] * inv;
      }
    }
}

// S = sum over chunks of Zpart in FP64, chunk by chunk (the order fixes the
// rounding): each thread four consecutive elements, four chunks' loads in
// flight.  Every partial is read once.
__global__ __launch_bounds__(256) void k_rik_sum(const float* __restrict__ Zpart, int nchunk, size_t mm,
                                                 double* __restrict__ S) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= mm) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (i0 + 4 <= mm && (mm & 3) == 0) {
    int ch = 0;
    for (; ch + 4 <= nchunk; ch += 4) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = *(const float4*)(Zpart + (size_t)(ch + u) * mm + i0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s[0] += (double)v[u].x;
        s[1] += (double)v[u].y;
        s[2] += (double)v[u].z;
        s[3] += (double)v[u].w;
      }
    }
    for (; ch < nchunk; ++ch) {
      const float4 v = *(const float4*)(Zpart + (size_t)ch * mm + i0);
      s[0] += (double)v.x;
      s[1] += (double)v.y;
      s[2] += (double)v.z;
      s[3] += (double)v.w;
    }
  } else {
    for (int ch = 0; ch < nchunk; ++ch)
      for (int e = 0; e < 4; ++e)
        if (i0 + e < mm) s[e] += (double)Zpart[(size_t)ch * mm + i0 + e];
  }
  for (int e = 0; e < 4; ++e)
    if (i0 + e < mm) S[i0 + e] = s[e];
}

// K = Kprev + S + S^T (Kprev null: 0); 32 x 32 tiles through LDS so that
// both S[m][n] and S[n][m] are read coalesced.
__global__ __launch_bounds__(256) void k_rik_symm(const double* __restrict__ S, int M, const double* __restrict__ Kprev,
                                                  double* __restrict__ K) {
  __shared__ double st[32][33];
  const int ntile = (M + 31) / 32;
  const int bi = blockIdx.x / ntile, bj = blockIdx.x % ntile;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int y = ty; y < 32; y += 8) {
    const int r = bj * 32 + y, c = bi * 32 + tx;
    st[tx][y] = (r < M && c < M) ? S[(size_t)r * M + c] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int r = bi * 32 + y, c = bj * 32 + tx;
    if (r >= M || c >= M) continue;
    const size_t id = (size_t)r * M + c;
    K[id] = (Kprev ? Kprev[id] : 0.0) + S[id] + st[y][tx];
  }
}

// dW[P] = sum_m dWrow[m Maux + P]; W = Wprev + dW; Wprev <- W
__global__ void k_rik_dw(const double* __restrict__ dWrow, int M, int Maux, double* __restrict__ dW,
                         double* __restrict__ W, double* __restrict__ Wprev) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= Maux) return;
  double s = 0.0;
  for (int m = 0; m < M; ++m) s += dWrow[(size_t)m * Maux + P];
  dW[P] = s;
  const double w = Wprev[P] + s;
  W[P] = w;
  Wprev[P] = w;
}

// J[m][n] = Jprev[m][n] + sum_P dW_P B[m][P][n]; Jprev <- J, with B from
// the planes as (hi + lo) / sB (exact sums in FP64; the 1 / sB power of two
// is applied to the total).  B and the planes are symmetric in m and n, so
// row m forms only its 8-column groups up to the one holding n = m (half of
// the planes' bytes); the caller mirrors J and Jprev.  A workgroup takes one
// m and 64 consecutive 8-column groups: each lane one group (two 16-byte
// loads per P), the four waves a quarter of P each, four P per round in
// flight; the wave partial sums meet in LDS in wave order.
__global__ __launch_bounds__(256) void k_rik_j(const uint16_t* __restrict__ Bp, float sB, int kpad,
                                               const double* __restrict__ dW, int M, int Maux,
                                               double* __restrict__ Jprev, double* __restrict__ J) {
  __shared__ double red[3][64][8];
  const int m = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ng = kpad / 8;
  if ((int)blockIdx.y * 64 > (m >> 3)) return;  // every group of the workgroup lies above the diagonal
  const int q = blockIdx.y * 64 + lane;          // 8-column group
  const int qmax = min(ng - 1, m >> 3);          // the groups of row m that are formed
  const int P0 = (int)((long)w * Maux / 4), P1 = (int)((long)(w + 1) * Maux / 4);
  const h8* slab = reinterpret_cast<const h8*>(Bp) + ((size_t)m * Maux * ng + q) * 2;
  const size_t pstride = (size_t)ng * 2;
  double s[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) s[e] = 0.0;
  if (q <= qmax) {
    int P = P0;
    for (; P + 4 <= P1; P += 4) {
      h8 hi[4], lo[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        hi[u] = slab[(size_t)(P + u) * pstride];
        lo[u] = slab[(size_t)(P + u) * pstride + 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double d = dW[P + u];
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += d * ((double)hi[u][e] + (double)lo[u][e]);
      }
    }
    for (; P < P1; ++P) {
      const h8 hi = slab[(size_t)P * pstride], lo = slab[(size_t)P * pstride + 1];
      const double d = dW[P];
#pragma unroll
      for (int e = 0; e < 8; ++e) s[e] += d * ((double)hi[e] + (double)lo[e]);
    }
  }
  if (w > 0)
#pragma unroll
    for (int e = 0; e < 8; ++e) red[w - 1][lane][e] = s[e];
  __syncthreads();
  if (w != 0 || q > qmax) return;
  const double inv = 1.0 / (double)sB;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int n = q * 8 + e;
    if (n >= M) continue;
    const double tot = (((s[e] + red[0][lane][e]) + red[1][lane][e]) + red[2][lane][e]) * inv;
    const size_t id = (size_t)m * M + n;
    const double v = Jprev[id] + tot;
    J[id] = v;
    Jprev[id] = v;
  }
}

// T chunk of an FP64 build (T[m][p][i], m < M, p < pc, i < nocc) into the
// incremental state: TA[m][p0 + p][i] = T, TE = 0.
__global__ void k_rik_seed(const double* __restrict__ T, int M, int pc, int nocc, int p0, int Maux,
                           float* __restrict__ TA, float* __restrict__ TE) {
  const size_t n = (size_t)M * pc * nocc;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
    const size_t m = id / ((size_t)pc * nocc), rem = id % ((size_t)pc * nocc);
    const size_t dst = (m * Maux + p0) * nocc + rem;
    TA[dst] = (float)T[id];
    TE[dst] = 0.0f;
  }
}

// X[m][i] = sqrt2 C[m][i] for i < nocc (C row-major M x M)
__global__ void k_rik_xocc(const double* __restrict__ C, int M, int nocc, double* __restrict__ X) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)M * nocc) return;
  X[id] = 1.4142135623730951 * C[(id / nocc) * M + id % nocc];
}

// out[0] = sum x^2, out[1] = sum y^2 (n entries each), out[2] = sum over the
// no x no matrix G of (G - I)^2 (one workgroup)
__global__ __launch_bounds__(1024) void k_rik_norms(const double* __restrict__ x, const double* __restrict__ y,
                                                    size_t n, const double* __restrict__ G, int no, double* out) {
  __shared__ double sx[1024], sy[1024], sg[1024];
  double a = 0.0, b = 0.0, c = 0.0;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
    a += x[i] * x[i];
    b += y[i] * y[i];
  }
  for (size_t i = threadIdx.x; i < (size_t)no * no; i += blockDim.x) {
    const double d = G[i] - ((i / no) == (i % no) ? 1.0 : 0.0);
    c += d * d;
  }
  sx[threadIdx.x] = a;
  sy[threadIdx.x] = b;
  sg[threadIdx.x] = c;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sx[threadIdx.x] += sx[threadIdx.x + o];
      sy[threadIdx.x] += sy[threadIdx.x + o];
      sg[threadIdx.x] += sg[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sx[0];
    out[1] = sy[0];
    out[2] = sg[0];
  }
}

// *w = max |(float)x| as float bits (w zeroed by the caller)
__global__ void k_rik_absmax(const double* __restrict__ x, size_t n, unsigned* w) {
  float m = 0.0f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)x[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_pos(w, m);
}

__global__ void k_set_u32(unsigned* p, unsigned v) { *p = v; }

__global__ void k_rik_reset(RikState* st) {
  st->ta_max = 0u;
  st->te_max = 0u;
}

template <int NBW>
hipError_t launch_tstep_nb(const RikTArgs& a, hipStream_t s) {
  const long full = a.R / kTRows;
  RikTArgs b = a;
  b.row0 = 0;
  b.wg0 = 0;
  if (full > 0) hipLaunchKernelGGL((k_rik_tstep<NBW, false>), dim3((unsigned)full), dim3(256), 0, s, b);
  if (a.R % kTRows) {
    b.row0 = full * kTRows;
    b.wg0 = full;
    hipLaunchKernelGGL((k_rik_tstep<NBW, true>), dim3(1), dim3(256), 0, s, b);
  }
  return hipGetLastError();
}

}  // namespace

int rik_blocks(int nocc) { return (nocc + 31) / 32; }
int rik_cols(int nocc) { return 64 * ((rik_blocks(nocc) + 1) / 2); }

hipError_t launch_rik_split_e(const double* E, int M, int nocc, int kpad, RikState* st, uint16_t* Ehi, uint16_t* Elo,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_rik_split_e, dim3(1), dim3(1024), 0, s, E, M, nocc, rik_cols(nocc), kpad, st, Ehi, Elo);
  return hipGetLastError();
}

size_t rik_tstep_workgroups(long R) { return (size_t)((R + kTRows - 1) / kTRows); }

int rik_kpad(int M) { return (M + 63) / 64 * 64; }

hipError_t launch_rik_planes(const double* B, long R, int M, int kpad, float sB, uint16_t* Bp, hipStream_t s) {
  const size_t n = (size_t)R * (kpad / 8);
  const unsigned grid = (unsigned)std::min<size_t>(65536, (n + 255) / 256);
  hipLaunchKernelGGL(k_rik_planes, dim3(grid), dim3(256), 0, s, B, R, M, kpad, sB, Bp);
  return hipGetLastError();
}

hipError_t launch_rik_tstep(const RikTArgs& a, hipStream_t s) {
  if (a.kpad % 64 != 0 || a.kpad < a.M || a.R <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rik_reset, dim3(1), dim3(1), 0, s, a.st);
  hipError_t e = hipErrorInvalidValue;
  switch (rik_blocks(a.nocc)) {  // 32-column blocks of E (all in each wave)
    case 1: e = launch_tstep_nb<1>(a, s); break;
    case 2: e = launch_tstep_nb<2>(a, s); break;
    case 3: e = launch_tstep_nb<3>(a, s); break;
    case 4: e = launch_tstep_nb<4>(a, s); break;
    case 5: e = launch_tstep_nb<5>(a, s); break;
    case 6: e = launch_tstep_nb<6>(a, s); break;
    case 7: e = launch_tstep_nb<7>(a, s); break;
    case 8: e = launch_tstep_nb<8>(a, s); break;
    case 9: e = launch_tstep_nb<9>(a, s); break;
    case 10: e = launch_tstep_nb<10>(a, s); break;
    default: break;
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rik_maxred, dim3(1), dim3(1024), 0, s, a.wgmax, (long)rik_tstep_workgroups(a.R), a.st);
  return hipGetLastError();
}

hipError_t launch_rik_dk(const RikDkArgs& a, int nchunk, hipStream_t s) {
  if (a.nchunk != nchunk) return hipErrorInvalidValue;
  const long nt2 = (long)((a.M + kDkTM - 1) / kDkTM) * ((a.M + kDkTN - 1) / kDkTN);
  const long nwg = nt2 * ((nchunk + 7) / 8) * 8;
  if ((a.Kd & 3) == 0) hipLaunchKernelGGL(k_rik_dk<true>, dim3((unsigned)nwg), dim3(kDkThreads), 0, s, a);
  else hipLaunchKernelGGL(k_rik_dk<false>, dim3((unsigned)nwg), dim3(kDkThreads), 0, s, a);
  return hipGetLastError();
}

int rik_dk_tiles(int M) { return ((M + kDkTM - 1) / kDkTM) * ((M + kDkTN - 1) / kDkTN); }

hipError_t launch_rik_reduce(const float* Zpart, int nchunk, int M, double* S, const double* Kprev, double* K,
                             hipStream_t s) {
  const size_t mm = (size_t)M * M;
  hipLaunchKernelGGL(k_rik_sum, dim3((unsigned)((mm / 4 + 256) / 256)), dim3(256), 0, s, Zpart, nchunk, mm, S);
  const int nt = (M + 31) / 32;
  hipLaunchKernelGGL(k_rik_symm, dim3((unsigned)(nt * nt)), dim3(256), 0, s, S, M, Kprev, K);
  return hipGetLastError();
}

hipError_t launch_rik_dw(const double* dWrow, int M, int Maux, double* dW, double* W, double* Wprev, hipStream_t s) {
  hipLaunchKernelGGL(k_rik_dw, dim3((unsigned)((Maux + 255) / 256)), dim3(256), 0, s, dWrow, M, Maux, dW, W, Wprev);
  return hipGetLastError();
}

hipError_t launch_rik_j(const uint16_t* Bp, float sB, int kpad, const double* dW, int M, int Maux, double* Jprev,
                        double* J, hipStream_t s) {
  const int ng = kpad / 8;
  hipLaunchKernelGGL(k_rik_j, dim3((unsigned)M, (unsigned)((ng + 63) / 64)), dim3(256), 0, s, Bp, sB, kpad, dW, M,
                     Maux, Jprev, J);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_mirror_upper(J, M, s);
  if (e == hipSuccess) e = launch_mirror_upper(Jprev, M, s);
  return e;
}

hipError_t launch_rik_seed(const double* T, int M, int pc, int nocc, int p0, int Maux, float* TA, float* TE,
                           hipStream_t s) {
  const size_t n = (size_t)M * pc * nocc;
  const unsigned grid = (unsigned)std::min<size_t>(65536, (n + 255) / 256);
  hipLaunchKernelGGL(k_rik_seed, dim3(grid), dim3(256), 0, s, T, M, pc, nocc, p0, Maux, TA, TE);
  return hipGetLastError();
}

hipError_t launch_rik_xocc(const double* C, int M, int nocc, double* X, hipStream_t s) {
  const size_t n = (size_t)M * nocc;
  hipLaunchKernelGGL(k_rik_xocc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C, M, nocc, X);
  return hipGetLastError();
}

hipError_t launch_rik_norms(const double* x, const double* y, size_t n, const double* G, int no, double* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_rik_norms, dim3(1), dim3(1024), 0, s, x, y, n, G, no, out);
  return hipGetLastError();
}

hipError_t launch_rik_absmax(const double* x, size_t n, unsigned* w, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, w, 0u);
  const unsigned grid = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(k_rik_absmax, dim3(grid), dim3(256), 0, s, x, n, w);
  return hipGetLastError();
}

}  // namespace gansu
