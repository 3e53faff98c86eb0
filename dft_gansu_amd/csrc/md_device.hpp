// McMurchie-Davidson building blocks for gfx950 (device code).
//
//  Boys      F_0..F_L(x): 7-term Taylor step from a 1/32-spaced grid of
//            F_n(x_i) for the top order, then downward recursion
//            F_{n-1} = (2x F_n + e^-x)/(2n-1); for x >= 31.98 the asymptotic
//            erf form with upward recursion.  (The reference evaluates every
//            order by a 6-term Taylor step, boys.hpp:104-146, or upward
//            recursion, boys.hpp:118-135.)
//  Hermite E E^{ij}_t per direction with E^{00}_0 = 1 (the pair factor
//            exp(-ab/p |AB|^2) is carried once per primitive pair), the
//            recursion E^{i+1,j}_t = E^{ij}_{t-1}/2p + X_PA E^{ij}_t + (t+1)E^{ij}_{t+1}.
//            (Et_functions.hpp:32-3002 holds 529 closed forms of the same.)
//  Hermite R R_{tuv} for t+u+v <= L by R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{tuv}
//            from R^n_{000} = (-2a)^n F_n, evaluated in place, level by level
//            (int2e.cu:1780-1840 computes the same with rolling buffers).
#pragma once
#include <hip/hip_runtime.h>

namespace gansu {
namespace dev {

constexpr int kBoysRows = 32;
constexpr double kPi = 3.14159265358979323846;
// 2 pi^(5/2)
constexpr double kTwoPi25 = 34.98683665524972497;

__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }
// level-ordered index of (t,u,v): all entries with t+u+v = n follow those of n-1
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

// Cartesian exponents of component c of an L shell.  Orbital order is
// types.hpp:197-228; aux==true selects the RI order for f (int2c2e.hpp:39-47).
__host__ __device__ constexpr int cart_exp(int L, int c, int axis, bool aux = false) {
  // s
  if (L == 0) return 0;
  if (L == 1) return c == axis ? 1 : 0;
  if (L == 2) {
    // xx yy zz xy xz yz
    constexpr int t[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    return t[c][axis];
  }
  if (L == 3) {
    constexpr int o[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0},
                              {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
    constexpr int r[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1},
                              {0, 2, 1}, {1, 2, 0}, {0, 1, 2}, {1, 0, 2}, {1, 1, 1}};
    return aux ? r[c][axis] : o[c][axis];
  }
  return 0;
}

// F_0..F_L(x) into F[0..L].  grid layout [i][n], i = round(32 x).
template <typename T>
__device__ __forceinline__ void boys_eval(int L, T x, const double* __restrict__ grid, T* F) {
  if (x < T(31.98)) {
    const int i = (int)(x * T(32) + T(0.5));
    const T dx = x - T(i) * T(0.03125);
    const double* g = grid + i * kBoysRows + L;
    // F_L(x) = sum_k F_{L+k}(x_i) (-dx)^k / k!, k = 0..6
    const T m = -dx;
    T f = T(g[6]);
    f = T(g[5]) + f * m * T(1.0 / 6.0);
    f = T(g[4]) + f * m * T(1.0 / 5.0);
    f = T(g[3]) + f * m * T(1.0 / 4.0);
    f = T(g[2]) + f * m * T(1.0 / 3.0);
    f = T(g[1]) + f * m * T(1.0 / 2.0);
    f = T(g[0]) + f * m;
    F[L] = f;
    const T ex = exp(-x);
    const T twox = x + x;
    for (int n = L; n > 0; --n) F[n - 1] = (twox * F[n] + ex) / T(2 * n - 1);
  } else {
    const T rx = T(1) / x;
    T f = T(0.5) * sqrt(T(kPi) * rx) * erf(sqrt(x));
    const T ex = exp(-x);
    F[0] = f;
    const T h = T(0.5) * rx;
    for (int n = 1; n <= L; ++n) {
      f = (T(2 * n - 1) * f - ex) * h;
      F[n] = f;
    }
  }
}

// Compile-time-L variant: keeps F in registers.
template <int L, typename T>
__device__ __forceinline__ void boys_eval_t(T x, const double* __restrict__ grid, T (&F)[L + 1]) {
  if (x < T(31.98)) {
    const int i = (int)(x * T(32) + T(0.5));
    const T dx = x - T(i) * T(0.03125);
    const double* g = grid + i * kBoysRows + L;
    const T m = -dx;
    T f = T(g[6]);
    f = T(g[5]) + f * m * T(1.0 / 6.0);
    f = T(g[4]) + f * m * T(1.0 / 5.0);
    f = T(g[3]) + f * m * T(1.0 / 4.0);
    f = T(g[2]) + f * m * T(1.0 / 3.0);
    f = T(g[1]) + f * m * T(1.0 / 2.0);
    f = T(g[0]) + f * m;
    F[L] = f;
    if constexpr (L > 0) {
      const T ex = exp(-x);
      const T twox = x + x;
#pragma unroll
      for (int n = L; n > 0; --n) F[n - 1] = (twox * F[n] + ex) * T(1.0 / (2 * n - 1));
    }
  } else {
    const T rx = T(1) / x;
    T f = T(0.5) * sqrt(T(kPi) * rx) * erf(sqrt(x));
    F[0] = f;
    if constexpr (L > 0) {
      const T ex = exp(-x);
      const T h = T(0.5) * rx;
#pragma unroll
      for (int n = 1; n <= L; ++n) {
        f = (T(2 * n - 1) * f - ex) * h;
        F[n] = f;
      }
    }
  }
}

// R_{tuv} (level-ordered, hidx) for t+u+v <= L, scaled by `pref`.
// In place: before step n, entries of levels < L-n hold R^{n+1}; step n
// rewrites levels L-n .. 1 (highest first) and then level 0.  All loop trip
// counts are template constants so the recursion unrolls into registers.
template <int L, typename T>
__device__ __forceinline__ void hermite_r_t(T alpha, T X, T Y, T Z, T pref, const double* __restrict__ grid,
                                            T (&R)[nherm(L)]) {
  T F[L + 1];
  boys_eval_t<L, T>(alpha * (X * X + Y * Y + Z * Z), grid, F);
  const T m2a = T(-2) * alpha;
  T sc = pref;
#pragma unroll
  for (int n = 0; n <= L; ++n) { F[n] *= sc; sc *= m2a; }
  R[0] = F[L];
#pragma unroll
  for (int n = L - 1; n >= 0; --n) {
#pragma unroll
    for (int k = L; k >= 1; --k) {
      if (k > L - n) continue;
#pragma unroll
      for (int t = L; t >= 0; --t) {
        if (t > k) continue;
#pragma unroll
        for (int u = L; u >= 0; --u) {
          if (u > k - t) continue;
          const int v = k - t - u;
          T val;
          if (t > 0) {
            val = X * R[hidx(t - 1, u, v)];
            if (t > 1) val += T(t - 1) * R[hidx(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * R[hidx(t, u - 1, v)];
            if (u > 1) val += T(u - 1) * R[hidx(t, u - 2, v)];
          } else {
            val = Z * R[hidx(t, u, v - 1)];
            if (v > 1) val += T(v - 1) * R[hidx(t, u, v - 2)];
          }
          R[hidx(t, u, v)] = val;
        }
      }
    }
    R[0] = F[n];
  }
}

// Runtime-L variant writing into a private array (generic kernels).
template <typename T>
__device__ inline void hermite_r_rt(int L, T alpha, T X, T Y, T Z, T pref, const double* __restrict__ grid, T* R) {
  T F[16];
  boys_eval<T>(L, alpha * (X * X + Y * Y + Z * Z), grid, F);
  const T m2a = T(-2) * alpha;
  T sc = pref;
  for (int n = 0; n <= L; ++n) { F[n] *= sc; sc *= m2a; }
  R[0] = F[L];
  for (int n = L - 1; n >= 0; --n) {
    for (int k = L - n; k >= 1; --k)
      for (int t = k; t >= 0; --t)
        for (int u = k - t; u >= 0; --u) {
          const int v = k - t - u;
          T val;
          if (t > 0) {
            val = X * R[hidx(t - 1, u, v)];
            if (t > 1) val += T(t - 1) * R[hidx(t - 2, u, v)];
          } else if (u > 0) {
            val = Y * R[hidx(t, u - 1, v)];
            if (u > 1) val += T(u - 1) * R[hidx(t, u - 2, v)];
          } else {
            val = Z * R[hidx(t, u, v - 1)];
            if (v > 1) val += T(v - 1) * R[hidx(t, u, v - 2)];
          }
          R[hidx(t, u, v)] = val;
        }
    R[0] = F[n];
  }
}

// E^{ij}_t for one direction, i <= LA, j <= LB: E[i][j][t] (t <= i + j).
template <int LA, int LB, typename T>
__device__ __forceinline__ void hermite_e_t(T p, T PA, T PB, T (&E)[LA + 1][LB + 1][LA + LB + 1]) {
  const T h = T(0.5) / p;
#pragma unroll
  for (int i = 0; i <= LA; ++i)
#pragma unroll
    for (int j = 0; j <= LB; ++j)
#pragma unroll
      for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = T(0);
  E[0][0][0] = T(1);
#pragma unroll
  for (int i = 0; i <= LA; ++i)
#pragma unroll
    for (int j = 0; j <= LB; ++j) {
      if (i == 0 && j == 0) continue;
#pragma unroll
      for (int t = 0; t <= LA + LB; ++t) {
        if (t > i + j) continue;
        T v;
        if (i > 0) {
          v = PA * E[i - 1][j][t];
          if (t + 1 <= i - 1 + j) v += T(t + 1) * E[i - 1][j][t + 1];
          if (t > 0) v += h * E[i - 1][j][t - 1];
        } else {
          v = PB * E[i][j - 1][t];
          if (t + 1 <= i + j - 1) v += T(t + 1) * E[i][j - 1][t + 1];
          if (t > 0) v += h * E[i][j - 1][t - 1];
        }
        E[i][j][t] = v;
      }
    }
}

// Runtime-L variant, E stored as E[(i*(lb+1)+j)*(la+lb+1)+t].
template <typename T>
__device__ inline void hermite_e_rt(int la, int lb, T p, T PA, T PB, T* E) {
  const int nt = la + lb + 1;
  const T h = T(0.5) / p;
  for (int k = 0; k < (la + 1) * (lb + 1) * nt; ++k) E[k] = T(0);
  E[0] = T(1);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        T v;
        if (i > 0) {
          const T* s = E + ((i - 1) * (lb + 1) + j) * nt;
          v = PA * s[t];
          if (t + 1 <= i - 1 + j) v += T(t + 1) * s[t + 1];
          if (t > 0) v += h * s[t - 1];
        } else {
          const T* s = E + (i * (lb + 1) + j - 1) * nt;
          v = PB * s[t];
          if (t + 1 <= i + j - 1) v += T(t + 1) * s[t + 1];
          if (t > 0) v += h * s[t - 1];
        }
        E[(i * (lb + 1) + j) * nt + t] = v;
      }
    }
}

}  // namespace dev
}  // namespace gansu
