// FP64 host oracle for the Fock-build path.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU in plain FP64 C++, what the reference
// (zhaokaifan02/DFT_GANSU) computes on its Fock-build path.  It is the checker
// for the HIP product library: only tests, the smoke check and the CPU-baseline
// leg of the benchmark may load it.  The product never links or calls it.
//
// Every function names the reference file:line it follows.  Parity status per
// function (what pins it) is listed in the table at the end of this header.
//
//   orc_parse_gbs/orc_build_molecule  basis_set.cpp:32-124, molecular.hpp:117-190,
//                                     parseXYZ.cpp:44-115, types.hpp:141-145
//   orc_cgto_norm                     basis_set.hpp:106-136 (odd-L quirk kept)
//   orc_boys_ref                      boys.hpp:104-182, parameters.h:20-37
//   orc_boys_exact                    series F_n(x)=e^-x sum (2x)^k/(2n+1)(2n+3)..
//   E (hermite_e)                     Et_functions.hpp:2999-3002 (E incl. K_AB)
//   R (hermite_r)                     int2e.cu:1790-1840 (R recursion)
//   orc_schwarz                       schwarz.txt:239-478, gpu_manager.cu:413-432
//   orc_eri_dense                     int2e.cu:1646-1947, int2e.hpp:127-161
//   orc_jk                            gpu_kernels.cu:230-287 (F = H + J - K/2)
//   orc_oneint                        int1e (S, T, V), standard MD formulas
//   orc_rhf                           hf.cu:171-314, rhf.hpp:298-327,
//                                     gpu_manager.cu:529-620, :827-836, :1020-1194
//   orc_rhf_guess                     orc_rhf from a guess density (rhf.hpp:551-584)
//   orc_int2c2e/orc_int3c2e/orc_ri_*  int2c2e.cu:468-687, int3c2e.cu:1102-1336,
//                                     gpu_manager.cu:1311-1632
//   orc_md_flops                      FP64 op count of prim_quartet per class
//                                     (int2e.cu:1646-1947 structure)
//
// Pinning (see tests/test_oracle.py):
//   Boys           closed forms F_n(0)=1/(2n+1), large-x asymptote, scipy 1F1,
//                  and the reference's own boys_30720.h digest (fixture).
//   E, R, ERI      (ss|ss) closed form; brute-force 3-D quadrature of (ps|ss)
//                  style integrals is replaced by an independent Obara-Saika
//                  recursion (orc_eri_os) written here; 8-fold symmetry,
//                  rotation invariance, Schwarz inequality, PSD of the ERI matrix.
//   one-electron   closed-form s/p overlaps; H2/STO-3G textbook energy.
//   SCF            H2 STO-3G at 1.4 bohr = -1.1167 Eh (Szabo-Ostlund 3.5.2);
//                  reference test known answers (Cholesky, 2x2 eigenvalues).
//   RI             B^T B reproduces (mn|P)(P|Q)^-1(Q|ls); Cholesky known answer.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>
#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <omp.h>

namespace {

struct Prim {  // == reference PrimitiveShell (types.hpp:116-125), 64 bytes
  double exponent, coefficient, x, y, z;
  int shell_type;
  size_t basis_index;
  int atom_index;
};
static_assert(sizeof(Prim) == 64, "PrimitiveShell layout");

struct TypeInfo { int count; size_t start_index; };

// Cartesian component order for the orbital basis (types.hpp:197-228).
const int kCartOrb[4][10][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
    {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}}};
// Cartesian order used for auxiliary f shells on the RI path (int2c2e.hpp:39-47).
const int kCartAuxF[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1},
                              {0, 2, 1}, {1, 2, 0}, {0, 1, 2}, {1, 0, 2}, {1, 1, 1}};

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline const int* cart(int l, int c, bool aux) {
  if (aux && l == 3) return kCartAuxF[c];
  return kCartOrb[l][c];
}

double dfact(int n) { double r = 1; for (int k = n; k > 1; k -= 2) r *= k; return r; }

// ------------------------------------------------------ FP64 op counting --
// Op: a double that counts the arithmetic performed on it (+, -, *, /; unary
// minus is a sign flip and is not counted) and, separately, the special
// functions (exp, sqrt, erf, pow, floor).  The MD routines below are templates
// over the number type: with double they are the oracle itself, with Op the
// same code counts its own floating-point work (orc_md_flops).
struct Op {
  double v;
  Op() : v(0.0) {}
  Op(double x) : v(x) {}  // constants enter free
  static thread_local long long n_arith, n_special;
};
thread_local long long Op::n_arith = 0, Op::n_special = 0;
inline Op operator+(Op a, Op b) { ++Op::n_arith; return Op(a.v + b.v); }
inline Op operator-(Op a, Op b) { ++Op::n_arith; return Op(a.v - b.v); }
inline Op operator*(Op a, Op b) { ++Op::n_arith; return Op(a.v * b.v); }
inline Op operator/(Op a, Op b) { ++Op::n_arith; return Op(a.v / b.v); }
inline Op operator-(Op a) { return Op(-a.v); }
inline Op& operator+=(Op& a, Op b) { ++Op::n_arith; a.v += b.v; return a; }
inline Op& operator*=(Op& a, Op b) { ++Op::n_arith; a.v *= b.v; return a; }
inline Op& operator-=(Op& a, Op b) { ++Op::n_arith; a.v -= b.v; return a; }
inline bool operator<(Op a, Op b) { return a.v < b.v; }
inline bool operator>(Op a, Op b) { return a.v > b.v; }
inline bool operator==(Op a, Op b) { return a.v == b.v; }
inline Op exp(Op a) { ++Op::n_special; return Op(std::exp(a.v)); }
inline Op sqrt(Op a) { ++Op::n_special; return Op(std::sqrt(a.v)); }
inline Op erf(Op a) { ++Op::n_special; return Op(std::erf(a.v)); }
inline Op pow(Op a, double b) { ++Op::n_special; return Op(std::pow(a.v, b)); }
inline double floor_of(double x) { return std::floor(x); }
inline double floor_of(Op x) { ++Op::n_special; return std::floor(x.v); }

// ---------------------------------------------------------------- Boys ----
// Exact Boys function by the convergent series
//   F_n(x) = exp(-x) * sum_k (2x)^k / ((2n+1)(2n+3)...(2n+2k+1))
// (all terms positive; no cancellation), valid for every x >= 0 used here.
double boys_series(int n, double x) {
  if (x > 200.0) {
    // asymptote (2n-1)!!/2^(n+1) sqrt(pi/x^(2n+1)); its relative error is the
    // Poisson tail e^-x x^k/k! summed over k <= n, below 1e-50 here for n <= 32
    return dfact(2 * n - 1) / std::pow(2.0, n + 1) * std::sqrt(M_PI / std::pow(x, 2 * n + 1));
  }
  double term = 1.0 / (2 * n + 1), sum = term;
  for (int k = 1; k < 2000; ++k) {
    term *= 2.0 * x / (2 * n + 2 * k + 1);
    sum += term;
    if (term < sum * 1e-18) break;
  }
  return std::exp(-x) * sum;
}

// Reference-style Boys, boys.hpp:149-182 with the constants of parameters.h:20-37.
constexpr double A_TR = 0.352905920120321, B_TR = 0.015532762923351;
constexpr double A_RS = 0.064048916778075, B_RS = 28.487431543672;
constexpr int LUT_NUM_XI = 1024, LUT_K_MAX = 5;
constexpr double LUT_XI_INTERVAL = 0.03125;

template <class T>
void boys_ref(int n, T x, const double* table, T* out) {
  using std::exp; using std::sqrt; using std::erf;
  if (x == T(0.0)) {
    for (int j = 0; j <= n; ++j) out[j] = T(1.0) / T(2 * j + 1);
  } else if (x < T(A_TR * n + B_TR)) {
    int xi = (int)floor_of(x / T(LUT_XI_INTERVAL) + T(0.5));
    T dx = x - T(LUT_XI_INTERVAL * xi);
    for (int j = 0; j <= n; ++j) {
      T num = 1.0; int fac = 1;
      T f = table[LUT_NUM_XI * j + xi];
      for (int k = 1; k <= LUT_K_MAX; ++k) {
        num *= -dx; fac *= k;
        f += T(table[LUT_NUM_XI * (j + k) + xi]) * num / T(fac);
      }
      out[j] = f;
    }
  } else {
    T em = 0.0, r2x = T(1.0) / (T(2) * x);
    T f = T(0.5) * sqrt(T(M_PI) / x);
    if (x < T(A_RS * n + B_RS)) { em = exp(-x); f *= erf(sqrt(x)); }
    out[0] = f;
    for (int j = 1; j <= n; ++j) { f = (T(2 * j - 1) * f - em) * r2x; out[j] = f; }
  }
}

struct BoysMode { int ref; const double* table; };
thread_local BoysMode g_boys = {0, nullptr};

void boys(int n, double x, double* out) {
  if (g_boys.ref && g_boys.table) { boys_ref(n, x, g_boys.table, out); return; }
  // exact: top value by the series, then stable downward recursion
  out[n] = boys_series(n, x);
  double ex = std::exp(-x);
  for (int j = n; j > 0; --j) out[j - 1] = (2 * x * out[j] + ex) / (2 * j - 1);
}
// the op-counted evaluation uses the reference-style Boys function
void boys(int n, Op x, Op* out) { boys_ref(n, x, g_boys.table, out); }

// ------------------------------------------------------ MD building blocks
// Hermite expansion coefficients E^{ij}_t for one Cartesian direction, with
// E^{00}_0 = 1 (the exp(-mu X^2) factor is applied once per pair by the caller).
// E[i][j][t], i<=la, j<=lb, t<=i+j.  Helgaker/Taylor recursion.
struct Etab { double e[7][7][13]; };
void hermite_e(int la, int lb, double p, double PA, double PB, Etab& E) {
  std::memset(&E, 0, sizeof(E));
  E.e[0][0][0] = 1.0;
  double h = 0.5 / p;
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          const double* s = E.e[i - 1][j];
          v = PA * s[t] + (t + 1 <= i - 1 + j ? (t + 1) * s[t + 1] : 0.0) + (t > 0 ? h * s[t - 1] : 0.0);
        } else {
          const double* s = E.e[i][j - 1];
          v = PB * s[t] + (t + 1 <= i + j - 1 ? (t + 1) * s[t + 1] : 0.0) + (t > 0 ? h * s[t - 1] : 0.0);
        }
        E.e[i][j][t] = v;
      }
    }
}

// Hermite Coulomb integrals R_{tuv} (n=0) for t+u+v <= L.
// R^n_{000} = (-2a)^n F_n(a |PQ|^2);  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}.
template <class T> struct RtabT { T r[13][13][13]; };
using Rtab = RtabT<double>;
template <class T>
void hermite_r(int L, T a, T X, T Y, T Z, RtabT<T>& R) {
  T F[32];
  boys(L, a * (X * X + Y * Y + Z * Z), F);
  // tmp[n][t][u][v]
  static thread_local std::vector<T> tmp;
  const int S = 13;
  tmp.assign((size_t)(L + 1) * S * S * S, T(0.0));
  auto at = [&](int n, int t, int u, int v) -> T& { return tmp[(((size_t)n * S + t) * S + u) * S + v]; };
  T m = 1.0;
  for (int n = 0; n <= L; ++n) { at(n, 0, 0, 0) = m * F[n]; m *= T(-2.0) * a; }
  for (int n = L - 1; n >= 0; --n) {
    int K = L - n;
    for (int t = 0; t <= K; ++t)
      for (int u = 0; u + t <= K; ++u)
        for (int v = 0; v + u + t <= K; ++v) {
          if (t + u + v == 0) continue;
          T val;
          if (t > 0) val = t > 1 ? X * at(n + 1, t - 1, u, v) + T(t - 1) * at(n + 1, t - 2, u, v) : X * at(n + 1, t - 1, u, v);
          else if (u > 0) val = u > 1 ? Y * at(n + 1, t, u - 1, v) + T(u - 1) * at(n + 1, t, u - 2, v) : Y * at(n + 1, t, u - 1, v);
          else val = v > 1 ? Z * at(n + 1, t, u, v - 1) + T(v - 1) * at(n + 1, t, u, v - 2) : Z * at(n + 1, t, u, v - 1);
          at(n, t, u, v) = val;
        }
  }
  for (int t = 0; t <= L; ++t)
    for (int u = 0; u + t <= L; ++u)
      for (int v = 0; v + u + t <= L; ++v) R.r[t][u][v] = at(0, t, u, v);
}

// Primitive normalisation, int2e.hpp:127-130.
double prim_norm(double a, int l, int m, int n) {
  return std::pow(2.0, l + m + n) / std::sqrt(dfact(2 * l - 1) * dfact(2 * m - 1) * dfact(2 * n - 1)) *
         std::pow(2.0 / M_PI, 0.75) * std::pow(a, (2.0 * (l + m + n) + 3.0) / 4.0);
}

struct PairData {
  double p, P[3], K;  // K = exp(-mu |AB|^2)
  Etab E[3];
};
void make_pair(const Prim& a, int la, const Prim& b, int lb, PairData& pd) {
  double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z};
  pd.p = a.exponent + b.exponent;
  double mu = a.exponent * b.exponent / pd.p, r2 = 0;
  for (int d = 0; d < 3; ++d) {
    pd.P[d] = (a.exponent * A[d] + b.exponent * B[d]) / pd.p;
    r2 += (A[d] - B[d]) * (A[d] - B[d]);
    hermite_e(la, lb, pd.p, pd.P[d] - A[d], pd.P[d] - B[d], pd.E[d]);
  }
  pd.K = std::exp(-mu * r2);
}

// Primitive [ab|cd] for all Cartesian components, raw (no coefficients, no
// norms).  out[((ia*nb+ib)*nc+ic)*nd+id].
template <class T>
void prim_quartet(const PairData& ab, int la, int lb, bool auxa, bool auxb, const PairData& cd, int lc, int ld,
                  bool auxc, bool auxd, T* out) {
  using std::sqrt;
  T p = ab.p, q = cd.p, alpha = p * q / (p + q);
  int L = la + lb + lc + ld;
  RtabT<T> R;
  hermite_r(L, alpha, T(ab.P[0]) - T(cd.P[0]), T(ab.P[1]) - T(cd.P[1]), T(ab.P[2]) - T(cd.P[2]), R);
  T pref = T(2.0 * std::pow(M_PI, 2.5)) / (p * q * sqrt(p + q)) * T(ab.K) * T(cd.K);
  int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  for (int ia = 0; ia < na; ++ia)
    for (int ib = 0; ib < nb; ++ib) {
      const int *A = cart(la, ia, auxa), *B = cart(lb, ib, auxb);
      for (int ic = 0; ic < nc; ++ic)
        for (int id = 0; id < nd; ++id) {
          const int *C = cart(lc, ic, auxc), *D = cart(ld, id, auxd);
          T s = 0;
          for (int t = 0; t <= A[0] + B[0]; ++t)
            for (int u = 0; u <= A[1] + B[1]; ++u)
              for (int v = 0; v <= A[2] + B[2]; ++v) {
                T eab = T(ab.E[0].e[A[0]][B[0]][t]) * T(ab.E[1].e[A[1]][B[1]][u]) * T(ab.E[2].e[A[2]][B[2]][v]);
                if (eab == T(0.0)) continue;
                for (int tt = 0; tt <= C[0] + D[0]; ++tt)
                  for (int uu = 0; uu <= C[1] + D[1]; ++uu)
                    for (int vv = 0; vv <= C[2] + D[2]; ++vv) {
                      T ecd = T(cd.E[0].e[C[0]][D[0]][tt]) * T(cd.E[1].e[C[1]][D[1]][uu]) * T(cd.E[2].e[C[2]][D[2]][vv]);
                      T sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                      s += eab * sg * ecd * R.r[t + tt][u + uu][v + vv];
                    }
              }
          out[((ia * nb + ib) * nc + ic) * nd + id] = pref * s;
        }
    }
}

// A primitive with an s function of exponent 0 at the same centre acts as the
// missing index for 2- and 3-centre integrals: p = a, P = A, E = E^{a0}, K = 1.
Prim s_dummy(const Prim& a) { Prim z = a; z.exponent = 0.0; z.shell_type = 0; z.coefficient = 1.0; return z; }

// Pair index enumeration of a pair type, gpu_manager.cu:413-432 and
// utils_cuda.hpp:35-44:  same type -> upper triangle (r1<=r2) from
// r2 = floor((sqrt(8i+1)-1)/2); different types -> (i / cnt1, i % cnt1).
void pair_of_index(size_t i, bool same, size_t cnt1, size_t& r1, size_t& r2) {
  if (same) {
    size_t y = (size_t)std::floor((std::sqrt(8.0 * (double)i + 1.0) - 1.0) / 2.0);
    while ((y + 1) * (y + 2) / 2 <= i) ++y;  // guard against sqrt rounding
    while (y * (y + 1) / 2 > i) --y;
    r1 = i - y * (y + 1) / 2; r2 = y;
  } else { r1 = i / cnt1; r2 = i % cnt1; }
}

struct Mol {
  std::vector<Prim> prims;
  std::vector<TypeInfo> types;
  std::vector<double> norm;
  int nbasis = 0;
};

std::string g_err;

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_boys_mode(int reference_style, const double* table) {
  g_boys.ref = reference_style; g_boys.table = table;
}

// F_0..F_n at x; mode 0 exact (series + downward), 1 reference-style.
void orc_boys(int n, double x, int reference_style, const double* table, double* out) {
  if (reference_style) boys_ref(n, x, table, out);
  else { out[n] = boys_series(n, x); double ex = std::exp(-x);
         for (int j = n; j > 0; --j) out[j - 1] = (2 * x * out[j] + ex) / (2 * j - 1); }
}

// The reference grid F_n(i/32), n=0..29, i=0..1023 (parameters.h:33-37,
// boys_30720.h), evaluated exactly; rounded to 15 decimals if round15.
void orc_boys_table(int round15, double* out) {
  for (int n = 0; n < 30; ++n)
    for (int i = 0; i < 1024; ++i) {
      double v = boys_series(n, i / 32.0);
      if (round15) { char buf[64]; std::snprintf(buf, sizeof buf, "%.15f", v); v = std::atof(buf); }
      out[n * 1024 + i] = v;
    }
}

// Contracted normalisation factor of one contracted shell, basis_set.hpp:106-136.
// quirk=1 keeps the reference's integer division L/2 in n_b.
double orc_cgto_norm(int L, int nprim, const double* exps, const double* coefs, int quirk) {
  double s = 0;
  for (int a = 0; a < nprim; ++a) {
    double na = std::pow(2.0 * M_PI, -0.75) * std::pow(4.0 * exps[a], L / 2.0 + 0.75);
    for (int b = 0; b < nprim; ++b) {
      double eb = quirk ? (double)(L / 2) + 0.75 : L / 2.0 + 0.75;
      double nb = std::pow(2.0 * M_PI, -0.75) * std::pow(4.0 * exps[b], eb);
      s += coefs[a] * na * coefs[b] * nb / std::pow(2.0 * (exps[a] + exps[b]), L + 1.5);
    }
  }
  s *= std::pow(2.0 * M_PI, 1.5);
  return std::pow(s, -0.5);
}

// ---- input path restated: .gbs + atoms -> primitive list (molecular.hpp:117-190)
static Mol g_mol;
int orc_build_molecule(const char* gbs_path, int natoms, const int* Z, const double* xyz_bohr, int quirk,
                       int* out_nprims, int* out_nbasis, int* out_ntypes) {
  static const char* names[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
                                "P", "S", "Cl", "Ar"};
  std::ifstream f(gbs_path);
  if (!f) { g_err = "cannot open gbs"; return 1; }
  struct Shell { int L; std::vector<double> e, c; };
  std::map<std::string, std::vector<Shell>> lib;
  std::string line, cur;
  bool in_el = false;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '!') continue;
    std::istringstream is(line);
    std::string tok; is >> tok;
    if (tok.empty()) continue;
    if (tok == "****") { in_el = false; continue; }
    if (!in_el) { cur = tok; in_el = true; lib[cur]; continue; }
    int np; is >> np;
    std::string lt = tok; for (auto& ch : lt) ch = (char)std::toupper(ch);
    auto Lof = [](char ch) { const char* s = "SPDFGHI"; return (int)(std::strchr(s, ch) - s); };
    std::vector<Shell> sh(lt.size());
    for (size_t k = 0; k < lt.size(); ++k) sh[k].L = Lof(lt[k]);
    for (int i = 0; i < np; ++i) {
      std::getline(f, line);
      for (auto& ch : line) if (ch == 'D' || ch == 'd') ch = 'E';
      std::istringstream ps(line);
      double e; ps >> e;
      for (size_t k = 0; k < lt.size(); ++k) { double c; ps >> c; sh[k].e.push_back(e); sh[k].c.push_back(c); }
    }
    for (auto& s : sh) lib[cur].push_back(s);
  }
  Mol m;
  size_t bi = 0;
  for (int a = 0; a < natoms; ++a) {
    auto it = lib.find(names[Z[a] - 1]);
    if (it == lib.end()) { g_err = "element missing"; return 2; }
    for (auto& s : it->second) {
      for (size_t k = 0; k < s.e.size(); ++k) {
        Prim p{};
        p.exponent = s.e[k]; p.coefficient = s.c[k];
        p.x = xyz_bohr[3 * a]; p.y = xyz_bohr[3 * a + 1]; p.z = xyz_bohr[3 * a + 2];
        p.shell_type = s.L; p.basis_index = bi; p.atom_index = a;
        m.prims.push_back(p);
      }
      double nf = orc_cgto_norm(s.L, (int)s.e.size(), s.e.data(), s.c.data(), quirk);
      for (int c = 0; c < ncart(s.L); ++c) m.norm.push_back(nf);
      bi += ncart(s.L);
    }
  }
  m.nbasis = (int)bi;
  std::sort(m.prims.begin(), m.prims.end(), [](const Prim& x, const Prim& y) { return x.shell_type < y.shell_type; });
  int maxL = m.prims.back().shell_type;
  m.types.assign(maxL + 1, {0, 0});
  for (auto& p : m.prims) m.types[p.shell_type].count++;
  for (int l = 1; l <= maxL; ++l) m.types[l].start_index = m.types[l - 1].start_index + m.types[l - 1].count;
  g_mol = m;
  *out_nprims = (int)m.prims.size(); *out_nbasis = m.nbasis; *out_ntypes = maxL + 1;
  return 0;
}
void orc_get_molecule(void* prims, void* types, double* norm) {
  std::memcpy(prims, g_mol.prims.data(), g_mol.prims.size() * sizeof(Prim));
  std::memcpy(types, g_mol.types.data(), g_mol.types.size() * sizeof(TypeInfo));
  std::memcpy(norm, g_mol.norm.data(), g_mol.norm.size() * sizeof(double));
}

// ---- Schwarz factors in reference pair order (schwarz.txt:239-478,
// gpu_manager.cu:2154-2179): Q = sqrt(max_comp c_a^2 c_b^2 N_a^2 N_b^2 [ab|ab]).
// Pair types s0<=s1 row-major over types.  Returns number of pairs written.
size_t orc_schwarz(const void* vprims, const void* vtypes, int ntypes, int aux_order, double* Q) {
  const Prim* P = (const Prim*)vprims;
  const TypeInfo* T = (const TypeInfo*)vtypes;
  std::vector<std::pair<size_t, size_t>> spans;  // (start, count) per pair type
  size_t off = 0;
  std::vector<std::pair<int, int>> pt;
  for (int s0 = 0; s0 < ntypes; ++s0)
    for (int s1 = s0; s1 < ntypes; ++s1) {
      size_t cnt = (s0 == s1) ? (size_t)T[s0].count * (T[s0].count + 1) / 2 : (size_t)T[s0].count * T[s1].count;
      spans.push_back({off, cnt}); pt.push_back({s0, s1}); off += cnt;
    }
  for (size_t k = 0; k < pt.size(); ++k) {
    int s0 = pt[k].first, s1 = pt[k].second;
    size_t cnt = spans[k].second;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < cnt; ++i) {
      size_t r1, r2;
      pair_of_index(i, s0 == s1, T[s1].count, r1, r2);
      const Prim& a = P[T[s0].start_index + r1];
      const Prim& b = P[T[s1].start_index + r2];
      int la = a.shell_type, lb = b.shell_type;
      PairData pd; make_pair(a, la, b, lb, pd);
      int na = ncart(la), nb = ncart(lb);
      std::vector<double> out((size_t)na * nb * na * nb);
      prim_quartet(pd, la, lb, aux_order, aux_order, pd, la, lb, aux_order, aux_order, out.data());
      double best = 0;
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
          const int *A = cart(la, ia, aux_order), *B = cart(lb, ib, aux_order);
          double nrm = prim_norm(a.exponent, A[0], A[1], A[2]) * prim_norm(b.exponent, B[0], B[1], B[2]);
          double v = nrm * nrm * a.coefficient * a.coefficient * b.coefficient * b.coefficient *
                     out[((ia * nb + ib) * na + ia) * nb + ib];
          if (std::fabs(best) < std::fabs(v)) best = v;
        }
      Q[spans[k].first + i] = std::sqrt(best);
    }
  }
  return off;
}

// Auxiliary per-primitive bound sqrt(max_comp c^2 N^2 (a|a)), schwarz.txt:479+.
void orc_schwarz_aux(const void* vprims, int nprims, double* Q) {
  const Prim* P = (const Prim*)vprims;
  for (int i = 0; i < nprims; ++i) {
    const Prim& a = P[i];
    int la = a.shell_type, na = ncart(la);
    Prim z = s_dummy(a);
    PairData pd; make_pair(a, la, z, 0, pd);
    std::vector<double> out((size_t)na * na);
    prim_quartet(pd, la, 0, true, false, pd, la, 0, true, false, out.data());
    double best = 0;
    for (int ia = 0; ia < na; ++ia) {
      const int* A = cart(la, ia, true);
      double nrm = prim_norm(a.exponent, A[0], A[1], A[2]);
      double v = nrm * nrm * a.coefficient * a.coefficient * out[ia * na + ia];
      if (std::fabs(best) < std::fabs(v)) best = v;
    }
    Q[i] = std::sqrt(best);
  }
}

}  // extern "C"

// ---- contracted integrals via per-primitive-quartet screening ------------
// Internal helper: the list of primitive pairs with their reference bound.
struct PPair { size_t a, b; double Q; };
static std::vector<PPair> pair_list(const Prim* P, const TypeInfo* T, int ntypes, const double* Q) {
  std::vector<PPair> out;
  size_t off = 0;
  for (int s0 = 0; s0 < ntypes; ++s0)
    for (int s1 = s0; s1 < ntypes; ++s1) {
      size_t cnt = (s0 == s1) ? (size_t)T[s0].count * (T[s0].count + 1) / 2 : (size_t)T[s0].count * T[s1].count;
      for (size_t i = 0; i < cnt; ++i) {
        size_t r1, r2; pair_of_index(i, s0 == s1, T[s1].count, r1, r2);
        out.push_back({T[s0].start_index + r1, T[s1].start_index + r2, Q[off + i]});
      }
      off += cnt;
    }
  return out;
}

// Contract one primitive quartet (pairs ab, cd) into a callback over basis
// indices with the value c N (cgto norms included), int2e.hpp:252-265.
template <class F>
static void contracted_contrib(const Prim* P, const double* norm, const PPair& x, const PPair& y, std::vector<double>& buf,
                               F&& sink) {
  const Prim &a = P[x.a], &b = P[x.b], &c = P[y.a], &d = P[y.b];
  int la = a.shell_type, lb = b.shell_type, lc = c.shell_type, ld = d.shell_type;
  PairData ab, cd; make_pair(a, la, b, lb, ab); make_pair(c, lc, d, ld, cd);
  int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  buf.resize((size_t)na * nb * nc * nd);
  prim_quartet(ab, la, lb, false, false, cd, lc, ld, false, false, buf.data());
  double cf = a.coefficient * b.coefficient * c.coefficient * d.coefficient;
  for (int ia = 0; ia < na; ++ia)
    for (int ib = 0; ib < nb; ++ib)
      for (int ic = 0; ic < nc; ++ic)
        for (int id = 0; id < nd; ++id) {
          const int *A = cart(la, ia, 0), *B = cart(lb, ib, 0), *C = cart(lc, ic, 0), *D = cart(ld, id, 0);
          size_t p = a.basis_index + ia, q = b.basis_index + ib, r = c.basis_index + ic, s = d.basis_index + id;
          double nrm = prim_norm(a.exponent, A[0], A[1], A[2]) * prim_norm(b.exponent, B[0], B[1], B[2]) *
                       prim_norm(c.exponent, C[0], C[1], C[2]) * prim_norm(d.exponent, D[0], D[1], D[2]);
          double v = buf[((ia * nb + ib) * nc + ic) * nd + id] * cf * nrm * norm[p] * norm[q] * norm[r] * norm[s];
          sink(p, q, r, s, v);
        }
}

extern "C" {

// Dense contracted ERI tensor out[M^4] (row-major [i][j][k][l]), the stored
// path of int2e.cu:1646-1947 + addToResult_case1 (int2e.hpp:252-265):
// quartets of primitive pairs x<=y, skipped when Q_x*Q_y < thr, scattered to
// the 8 index permutations according to primitive-level symmetry flags.
int orc_eri_dense(const void* vprims, const void* vtypes, int ntypes, const double* norm, int M, const double* Q,
                  double thr, double* out) {
  const Prim* P = (const Prim*)vprims;
  auto pl = pair_list(P, (const TypeInfo*)vtypes, ntypes, Q);
  size_t n = pl.size(), MM = M;
  std::memset(out, 0, sizeof(double) * MM * MM * MM * MM);
  std::vector<double> buf;
  auto idx = [&](size_t p, size_t q, size_t r, size_t s) { return ((p * MM + q) * MM + r) * MM + s; };
  for (size_t x = 0; x < n; ++x)
    for (size_t y = x; y < n; ++y) {
      if (pl[x].Q * pl[y].Q < thr) continue;
      bool sb = pl[x].a == pl[x].b, sk = pl[y].a == pl[y].b, sbk = (x == y);
      contracted_contrib(P, norm, pl[x], pl[y], buf, [&](size_t p, size_t q, size_t r, size_t s, double v) {
        out[idx(p, q, r, s)] += v;
        if (!sb) out[idx(q, p, r, s)] += v;
        if (!sk) out[idx(p, q, s, r)] += v;
        if (!sb && !sk) out[idx(q, p, s, r)] += v;
        if (!sbk) {
          out[idx(r, s, p, q)] += v;
          if (!sb) out[idx(r, s, q, p)] += v;
          if (!sk) out[idx(s, r, p, q)] += v;
          if (!sb && !sk) out[idx(s, r, q, p)] += v;
        }
      });
    }
  return 0;
}

// Direct J and K from the same screened primitive-quartet set, without a
// stored tensor.  J_ij = sum_kl (ij|kl) D_kl,  K_ij = sum_kl (ik|jl) D_kl.
// OpenMP over primitive-pair rows, per-thread J/K, deterministic reduction.
// sample_stride > 1 evaluates every sample_stride-th bra row only (used for a
// bounded CPU timing sample); the result is then partial.
int orc_jk(const void* vprims, const void* vtypes, int ntypes, const double* norm, int M, const double* Q, double thr,
           const double* D, double* J, double* K, int sample_stride) {
  const Prim* P = (const Prim*)vprims;
  auto pl = pair_list(P, (const TypeInfo*)vtypes, ntypes, Q);
  size_t n = pl.size(), MM = M;
  int nt = omp_get_max_threads();
  std::vector<std::vector<double>> Jt(nt, std::vector<double>(MM * MM, 0.0)), Kt(nt, std::vector<double>(MM * MM, 0.0));
  if (sample_stride < 1) sample_stride = 1;
#pragma omp parallel
  {
    int tid = omp_get_thread_num();
    double* Jl = Jt[tid].data(); double* Kl = Kt[tid].data();
    std::vector<double> buf;
#pragma omp for schedule(dynamic, 1)
    for (long long xx = 0; xx < (long long)n; xx += sample_stride) {
      size_t x = (size_t)xx;
      for (size_t y = x; y < n; ++y) {
        if (pl[x].Q * pl[y].Q < thr) continue;
        bool sb = pl[x].a == pl[x].b, sk = pl[y].a == pl[y].b, sbk = (x == y);
        contracted_contrib(P, norm, pl[x], pl[y], buf, [&](size_t p, size_t q, size_t r, size_t s, double v) {
          // the same 8 (or fewer) permutations as the dense scatter, digested
          auto put = [&](size_t i, size_t j, size_t k, size_t l) {
            Jl[i * MM + j] += v * D[k * MM + l];
            Kl[i * MM + k] += v * D[j * MM + l];
          };
          put(p, q, r, s);
          if (!sb) put(q, p, r, s);
          if (!sk) put(p, q, s, r);
          if (!sb && !sk) put(q, p, s, r);
          if (!sbk) {
            put(r, s, p, q);
            if (!sb) put(r, s, q, p);
            if (!sk) put(s, r, p, q);
            if (!sb && !sk) put(s, r, q, p);
          }
        });
      }
    }
  }
  for (size_t i = 0; i < MM * MM; ++i) {
    double sj = 0, sk = 0;
    for (int t = 0; t < nt; ++t) { sj += Jt[t][i]; sk += Kt[t][i]; }
    J[i] = sj; K[i] = sk;
  }
  return 0;
}

// F = H + J - K/2 from a dense tensor (gpu_kernels.cu:230-287).
void orc_fock_from_dense(const double* eri, const double* D, const double* H, int M, double* F) {
  size_t MM = M;
  for (size_t i = 0; i < MM; ++i)
    for (size_t j = 0; j < MM; ++j) {
      double s = 0;
      for (size_t k = 0; k < MM; ++k)
        for (size_t l = 0; l < MM; ++l)
          s += (eri[((i * MM + j) * MM + k) * MM + l] - 0.5 * eri[((i * MM + k) * MM + j) * MM + l]) * D[k * MM + l];
      F[i * MM + j] = H[i * MM + j] + s;
    }
}

// ---- one-electron integrals S, T, V (contracted, cgto norms applied) -----
int orc_oneint(const void* vprims, int nprims, const double* norm, int M, int natoms, const int* Z, const double* xyz,
               double* S, double* T, double* V) {
  const Prim* P = (const Prim*)vprims;
  size_t MM = M;
  std::memset(S, 0, sizeof(double) * MM * MM);
  std::memset(T, 0, sizeof(double) * MM * MM);
  std::memset(V, 0, sizeof(double) * MM * MM);
  for (int ia = 0; ia < nprims; ++ia)
    for (int ib = 0; ib < nprims; ++ib) {
      const Prim &a = P[ia], &b = P[ib];
      int la = a.shell_type, lb = b.shell_type;
      PairData pd; make_pair(a, la, b, lb + 2, pd);  // +2 for the kinetic term
      double p = pd.p;
      // nuclear attraction needs R tables per nucleus
      std::vector<Rtab> Rn(natoms);
      for (int c = 0; c < natoms; ++c)
        hermite_r(la + lb, p, pd.P[0] - xyz[3 * c], pd.P[1] - xyz[3 * c + 1], pd.P[2] - xyz[3 * c + 2], Rn[c]);
      double sp = std::pow(M_PI / p, 1.5) * pd.K;
      double bb = b.exponent;
      for (int ca = 0; ca < ncart(la); ++ca)
        for (int cb = 0; cb < ncart(lb); ++cb) {
          const int *A = cart(la, ca, 0), *B = cart(lb, cb, 0);
          double s1[3], t1[3];
          for (int d = 0; d < 3; ++d) {
            int i = A[d], j = B[d];
            double sij = pd.E[d].e[i][j][0];
            double sp2 = pd.E[d].e[i][j + 2][0];
            double sm2 = j >= 2 ? pd.E[d].e[i][j - 2][0] : 0.0;
            s1[d] = sij;
            t1[d] = -2.0 * bb * bb * sp2 + bb * (2 * j + 1) * sij - 0.5 * j * (j - 1) * sm2;
          }
          double s = s1[0] * s1[1] * s1[2];
          double t = t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2];
          double v = 0;
          for (int c = 0; c < natoms; ++c) {
            double acc = 0;
            for (int tt = 0; tt <= A[0] + B[0]; ++tt)
              for (int u = 0; u <= A[1] + B[1]; ++u)
                for (int w = 0; w <= A[2] + B[2]; ++w)
                  acc += pd.E[0].e[A[0]][B[0]][tt] * pd.E[1].e[A[1]][B[1]][u] * pd.E[2].e[A[2]][B[2]][w] *
                         Rn[c].r[tt][u][w];
            v -= Z[c] * acc;
          }
          double nrm = prim_norm(a.exponent, A[0], A[1], A[2]) * prim_norm(b.exponent, B[0], B[1], B[2]) *
                       a.coefficient * b.coefficient;
          size_t p_ = a.basis_index + ca, q_ = b.basis_index + cb;
          double nn = nrm * norm[p_] * norm[q_];
          S[p_ * MM + q_] += nn * s * sp;
          T[p_ * MM + q_] += nn * t * sp;
          V[p_ * MM + q_] += nn * v * 2.0 * M_PI / p * pd.K;
        }
    }
  return 0;
}

// ---- dense linear algebra (host, FP64) ------------------------------------
// Cyclic Jacobi eigensolver: A (n x n symmetric, row-major) -> ascending
// eigenvalues w and eigenvectors in the columns of V (row-major V[i][k]).
void orc_eigh(int n, const double* Ain, double* w, double* V) {
  std::vector<double> A(Ain, Ain + (size_t)n * n);
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) V[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) { double a = A[i * n + j] * A[i * n + j]; tot += a; if (i != j) off += a; }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq; A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk; A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq; V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
  std::vector<double> Vs((size_t)n * n);
  for (int k = 0; k < n; ++k) { w[k] = A[ord[k] * n + ord[k]]; for (int i = 0; i < n; ++i) Vs[i * n + k] = V[i * n + ord[k]]; }
  std::memcpy(V, Vs.data(), sizeof(double) * n * n);
}

// Lower Cholesky L (row-major, upper zeroed) of an SPD matrix: the convention
// of choleskyDecomposition (gpu_manager.cu:1311-1349).  Returns 0 on success.
int orc_cholesky(int n, const double* A, double* L) {
  std::memset(L, 0, sizeof(double) * n * n);
  for (int j = 0; j < n; ++j) {
    double s = A[j * n + j];
    for (int k = 0; k < j; ++k) s -= L[j * n + k] * L[j * n + k];
    if (s <= 0) return 1;
    L[j * n + j] = std::sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      double t = A[i * n + j];
      for (int k = 0; k < j; ++k) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / L[j * n + j];
    }
  }
  return 0;
}

static void matmul(int n, const double* A, bool ta, const double* B, bool tb, double* C) {
  std::vector<double> tmp((size_t)n * n, 0.0);
#pragma omp parallel for
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double a = ta ? A[k * n + i] : A[i * n + k];
      if (a == 0.0) continue;
      for (int j = 0; j < n; ++j) tmp[(size_t)i * n + j] += a * (tb ? B[j * n + k] : B[k * n + j]);
    }
  std::memcpy(C, tmp.data(), sizeof(double) * n * n);
}

// Gaussian elimination with partial pivoting (getrf/getrs semantics).
static int solve_lin(int n, std::vector<double> A, std::vector<double>& b) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r) if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) return 1;
    if (piv != c) { for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]); std::swap(b[c], b[piv]); }
    for (int r = c + 1; r < n; ++r) {
      double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r * n + k] * b[k];
    b[r] = s / A[r * n + r];
  }
  return 0;
}

// Optimal damping factor for RHF (the known answer 0.613207547 of the
// reference's test_gpu_manager.cu:468-492 pins this formula).
// s = tr(Fp dD), c = tr(dF dD) (gpu_manager.cu:923-958).
double orc_optimal_damping(int n, const double* F, const double* Fp, const double* D, const double* Dp) {
  double s = 0, c = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double dF = F[i * n + j] - Fp[i * n + j], dD = D[j * n + i] - Dp[j * n + i];
      s += Fp[i * n + j] * dD; c += dF * dD;
    }
  if (c <= -s / 2.0) return 1.0;
  return -s / (2.0 * c);
}

// ---- RHF SCF with DIIS (hf.cu:245-314, rhf.hpp:298-327) ------------------
// fock(D, F) is provided as a dense ERI tensor (M small) or computed by orc_jk.
// Returns the number of iterations; energies in out[0]=E_elec, out[1]=E_total.
int orc_ri_jk(int Maux, int M, const double* B, const double* D, double* J, double* K);

// riB (Maux x M x M, B^P_mn) selects RI-J/K Fock builds (ERI_RI_RHF, rhf.hpp:650).
// D0 (M x M, may be NULL) replaces the core guess: F(D0), then C, D and F
// (InitialGuess_RHF_SAD::guess, rhf.hpp:551-584, with D0 = D_alpha + D_beta).
static int rhf_impl(int M, int nocc, const double* S, const double* H, const double* eri, const void* vprims,
                    const void* vtypes, int ntypes, const double* norm, const double* Q, double thr, double enuc,
                    int maxiter, double conv, int diis_size, int diis_include_transform, double* out_energy,
                    double* out_D, double* out_F, const double* riB, int Maux, const double* D0) {
  size_t MM = (size_t)M * M;
  std::vector<double> w(M), U(MM), X(MM), tmp(MM), Fp(MM), Cp(MM), C(MM), D(MM), F(MM), J(MM), K(MM);
  orc_eigh(M, S, w.data(), U.data());
  for (int i = 0; i < M; ++i)
    for (int k = 0; k < M; ++k) X[i * M + k] = U[i * M + k] * (w[k] < 1e-6 ? 0.0 : 1.0 / std::sqrt(w[k]));
  auto coeffs = [&](const double* Fm) {
    matmul(M, X.data(), true, Fm, false, tmp.data());
    matmul(M, tmp.data(), false, X.data(), false, Fp.data());
    std::vector<double> e(M);
    orc_eigh(M, Fp.data(), e.data(), Cp.data());
    matmul(M, X.data(), false, Cp.data(), false, C.data());
  };
  auto density = [&]() {
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) {
        double s = 0;
        for (int k = 0; k < nocc; ++k) s += C[i * M + k] * C[j * M + k];
        D[i * M + j] = 2.0 * s;
      }
  };
  auto fock = [&]() {
    if (riB) {
      orc_ri_jk(Maux, M, riB, D.data(), J.data(), K.data());
      for (size_t i = 0; i < MM; ++i) F[i] = H[i] + J[i] - 0.5 * K[i];
      return;
    }
    if (eri) { orc_fock_from_dense(eri, D.data(), H, M, F.data()); return; }
    orc_jk(vprims, vtypes, ntypes, norm, M, Q, thr, D.data(), J.data(), K.data(), 1);
    for (size_t i = 0; i < MM; ++i) F[i] = H[i] + J[i] - 0.5 * K[i];
  };
  auto energy = [&]() {
    double e = 0;
    for (size_t i = 0; i < MM; ++i) e += D[i] * H[i];
    double f = 0;
    for (size_t i = 0; i < MM; ++i) f += D[i] * F[i];
    return 0.5 * (e + f);
  };
  if (D0) {
    std::copy(D0, D0 + MM, D.begin());
    fock();
    coeffs(F.data());
  } else {
    coeffs(H);  // core guess (rhf.hpp:384-396)
  }
  density(); fock();
  std::vector<double> errs((size_t)diis_size * MM), focks((size_t)diis_size * MM);
  int diis_iter = 0;
  double prev = 1.7976931348623157e308, E = 0;
  int it = 0;
  while (it < maxiter) {
    coeffs(F.data()); density(); fock();
    E = energy();
    if (std::fabs(E - prev) < conv) break;
    ++it; prev = E;
    // DIIS: e = FDS - SDF (gpu_manager.cu:1020-1058)
    std::vector<double> a(MM), b(MM), e(MM);
    matmul(M, F.data(), false, D.data(), false, a.data()); matmul(M, a.data(), false, S, false, b.data());
    matmul(M, S, false, D.data(), false, a.data()); matmul(M, a.data(), false, F.data(), false, e.data());
    for (size_t i = 0; i < MM; ++i) e[i] = b[i] - e[i];
    if (diis_include_transform) {
      matmul(M, X.data(), false, e.data(), false, a.data()); matmul(M, a.data(), false, X.data(), true, e.data());
    }
    int slot = diis_iter % diis_size;
    std::memcpy(&errs[slot * MM], e.data(), sizeof(double) * MM);
    std::memcpy(&focks[slot * MM], F.data(), sizeof(double) * MM);
    int np = std::min(diis_iter + 1, diis_size);
    if (np > 1) {
      int ns = np + 1;
      std::vector<double> B((size_t)ns * ns, 0.0), rhs(ns, 0.0);
      for (int i = 0; i < np; ++i) {
        for (int j = i; j < np; ++j) {
          double s = 0;
          for (size_t k = 0; k < MM; ++k) s += errs[i * MM + k] * errs[j * MM + k];
          B[i * ns + j] = B[j * ns + i] = s;
        }
        B[i * ns + np] = B[np * ns + i] = -1.0;
      }
      rhs[np] = -1.0;
      if (solve_lin(ns, B, rhs) == 0) {  // a singular system keeps F as it is
        for (size_t k = 0; k < MM; ++k) {
          double s = 0;
          for (int i = 0; i < np; ++i) s += rhs[i] * focks[i * MM + k];
          F[k] = s;
        }
      }
    }
    ++diis_iter;
  }
  out_energy[0] = E; out_energy[1] = E + enuc;
  if (out_D) std::memcpy(out_D, D.data(), sizeof(double) * MM);
  if (out_F) std::memcpy(out_F, F.data(), sizeof(double) * MM);
  return it;
}

int orc_rhf(int M, int nocc, const double* S, const double* H, const double* eri, const void* vprims,
            const void* vtypes, int ntypes, const double* norm, const double* Q, double thr, double enuc, int maxiter,
            double conv, int diis_size, int diis_include_transform, double* out_energy, double* out_D,
            double* out_F, const double* riB, int Maux) {
  return rhf_impl(M, nocc, S, H, eri, vprims, vtypes, ntypes, norm, Q, thr, enuc, maxiter, conv, diis_size,
                  diis_include_transform, out_energy, out_D, out_F, riB, Maux, nullptr);
}

int orc_rhf_guess(int M, int nocc, const double* S, const double* H, const double* eri, const void* vprims,
                  const void* vtypes, int ntypes, const double* norm, const double* Q, double thr, double enuc,
                  int maxiter, double conv, int diis_size, int diis_include_transform, double* out_energy,
                  double* out_D, double* out_F, const double* riB, int Maux, const double* D0) {
  return rhf_impl(M, nocc, S, H, eri, vprims, vtypes, ntypes, norm, Q, thr, enuc, maxiter, conv, diis_size,
                  diis_include_transform, out_energy, out_D, out_F, riB, Maux, D0);
}

// ---- RI: (P|Q), (mn|P), B = L^-1 (P|mn), RI-J/K (gpu_manager.cu:1448-1632)
// Aux primitives use the RI Cartesian order.  thr screens primitive pairs /
// aux primitives exactly as int2c2e.cu:506 and int3c2e.cu:1139.
int orc_int2c2e(const void* vaux, int naux_prims, const double* aux_norm, int Maux, const double* Qaux, double thr,
                double* out) {
  const Prim* P = (const Prim*)vaux;
  std::memset(out, 0, sizeof(double) * (size_t)Maux * Maux);
  std::vector<double> buf;
  for (int i = 0; i < naux_prims; ++i)
    for (int j = i; j < naux_prims; ++j) {
      if (Qaux[i] * Qaux[j] < thr) continue;
      const Prim &a = P[i], &b = P[j];
      int la = a.shell_type, lb = b.shell_type;
      Prim za = s_dummy(a), zb = s_dummy(b);
      PairData pa, pb; make_pair(a, la, za, 0, pa); make_pair(b, lb, zb, 0, pb);
      int na = ncart(la), nb = ncart(lb);
      buf.resize((size_t)na * nb);
      prim_quartet(pa, la, 0, true, false, pb, lb, 0, true, false, buf.data());
      for (int ca = 0; ca < na; ++ca)
        for (int cb = 0; cb < nb; ++cb) {
          const int *A = cart(la, ca, true), *B = cart(lb, cb, true);
          size_t p = a.basis_index + ca, r = b.basis_index + cb;
          double v = buf[ca * nb + cb] * a.coefficient * b.coefficient * prim_norm(a.exponent, A[0], A[1], A[2]) *
                     prim_norm(b.exponent, B[0], B[1], B[2]) * aux_norm[p] * aux_norm[r];
          out[p * Maux + r] += v;
          if (i != j) out[r * Maux + p] += v;
        }
    }
  return 0;
}

// out[P][m][n] (aux-major, full m,n), int3c2e.cu:1102-1336 + addToResult_3center.
int orc_int3c2e(const void* vprims, const void* vtypes, int ntypes, const double* norm, int M, const double* Q,
                const void* vaux, int naux_prims, const double* aux_norm, int Maux, const double* Qaux, double thr,
                double* out) {
  const Prim* P = (const Prim*)vprims;
  const Prim* A = (const Prim*)vaux;
  auto pl = pair_list(P, (const TypeInfo*)vtypes, ntypes, Q);
  size_t MM = M;
  std::memset(out, 0, sizeof(double) * (size_t)Maux * MM * MM);
  // primitives grouped by their contracted auxiliary function, so that one
  // thread owns all writes to that function's rows
  std::vector<std::vector<int>> groups;
  {
    std::map<size_t, int> gid;
    for (int c = 0; c < naux_prims; ++c) {
      auto it = gid.find(A[c].basis_index);
      if (it == gid.end()) { gid[A[c].basis_index] = (int)groups.size(); groups.push_back({c}); }
      else groups[it->second].push_back(c);
    }
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (long long gi = 0; gi < (long long)groups.size(); ++gi)
  for (int c : groups[gi]) {
    std::vector<double> buf;
    const Prim& pc = A[c];
    int lc = pc.shell_type, nc = ncart(lc);
    Prim zc = s_dummy(pc);
    PairData pcd; make_pair(pc, lc, zc, 0, pcd);
    for (size_t x = 0; x < pl.size(); ++x) {
      if (pl[x].Q * Qaux[c] < thr) continue;
      const Prim &a = P[pl[x].a], &b = P[pl[x].b];
      int la = a.shell_type, lb = b.shell_type, na = ncart(la), nb = ncart(lb);
      PairData ab; make_pair(a, la, b, lb, ab);
      buf.resize((size_t)na * nb * nc);
      prim_quartet(ab, la, lb, false, false, pcd, lc, 0, true, false, buf.data());
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
          for (int ic = 0; ic < nc; ++ic) {
            const int *CA = cart(la, ia, 0), *CB = cart(lb, ib, 0), *CC = cart(lc, ic, true);
            size_t p = a.basis_index + ia, q = b.basis_index + ib, r = pc.basis_index + ic;
            double v = buf[(ia * nb + ib) * nc + ic] * a.coefficient * b.coefficient * pc.coefficient *
                       prim_norm(a.exponent, CA[0], CA[1], CA[2]) * prim_norm(b.exponent, CB[0], CB[1], CB[2]) *
                       prim_norm(pc.exponent, CC[0], CC[1], CC[2]) * norm[p] * norm[q] * aux_norm[r];
            // every primitive of this auxiliary function is handled by this thread
            out[(r * MM + p) * MM + q] += v;
            if (pl[x].a != pl[x].b) out[(r * MM + q) * MM + p] += v;
          }
    }
  }
  return 0;
}

// B = L^-1 T with T = (P|mn) viewed as Maux x M^2 (solve_lower_triangular,
// gpu_manager.cu:1359-1416).  In place on T.
int orc_ri_B(int Maux, const double* L, size_t ncols, double* T) {
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < (long long)ncols; ++c)
    for (int i = 0; i < Maux; ++i) {
      double s = T[(size_t)i * ncols + c];
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * Maux + k] * T[(size_t)k * ncols + c];
      T[(size_t)i * ncols + c] = s / L[(size_t)i * Maux + i];
    }
  return 0;
}

// RI-J and RI-K from B (Maux x M x M): J = sum_P B^P (B^P . D), K = sum_P B^P D B^P.
int orc_ri_jk(int Maux, int M, const double* B, const double* D, double* J, double* K) {
  size_t MM = (size_t)M * M;
  std::vector<double> W(Maux);
  for (int P = 0; P < Maux; ++P) { double s = 0; for (size_t k = 0; k < MM; ++k) s += B[P * MM + k] * D[k]; W[P] = s; }
  for (size_t k = 0; k < MM; ++k) { double s = 0; for (int P = 0; P < Maux; ++P) s += W[P] * B[P * MM + k]; J[k] = s; }
  std::memset(K, 0, sizeof(double) * MM);
  int nt = omp_get_max_threads();
  std::vector<std::vector<double>> Kt(nt, std::vector<double>(MM, 0.0));
#pragma omp parallel
  {
    std::vector<double> T1(MM);
    double* Kl = Kt[omp_get_thread_num()].data();
#pragma omp for
    for (int P = 0; P < Maux; ++P) {
      const double* Bp = B + P * MM;
      for (int i = 0; i < M; ++i)
        for (int l = 0; l < M; ++l) {
          double s = 0;
          for (int k = 0; k < M; ++k) s += Bp[i * M + k] * D[k * M + l];
          T1[i * M + l] = s;
        }
      for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
          double s = 0;
          for (int l = 0; l < M; ++l) s += T1[i * M + l] * Bp[j * M + l];
          Kl[i * M + j] += s;
        }
    }
  }
  for (size_t k = 0; k < MM; ++k) { double s = 0; for (int t = 0; t < nt; ++t) s += Kt[t][k]; K[k] = s; }
  return 0;
}
}  // extern "C"


// ---- independent check: Obara-Saika / Head-Gordon-Pople primitive ERI ----
// A second, unrelated recursion for one raw primitive Cartesian integral
// [ab|cd] (no coefficients, no norms): VRR on a and c, then HRR to b and d.
// Exponential-time; for small angular momenta in tests only.
namespace {
struct OSCtx { double p, q, rho, PA[3], QC[3], WP[3], WQ[3], AB[3], CD[3], base; double F[40]; };
double os_vrr(const OSCtx& c, int a[3], int cc[3], int m) {
  int i = -1;
  for (int d = 0; d < 3; ++d) if (a[d] > 0) { i = d; break; }
  if (i >= 0) {
    int am[3] = {a[0], a[1], a[2]}; am[i]--;
    double v = c.PA[i] * os_vrr(c, am, cc, m) + c.WP[i] * os_vrr(c, am, cc, m + 1);
    if (am[i] > 0) {
      int amm[3] = {am[0], am[1], am[2]}; amm[i]--;
      v += am[i] / (2 * c.p) * (os_vrr(c, amm, cc, m) - c.rho / c.p * os_vrr(c, amm, cc, m + 1));
    }
    if (cc[i] > 0) { int cm[3] = {cc[0], cc[1], cc[2]}; cm[i]--; v += cc[i] / (2 * (c.p + c.q)) * os_vrr(c, am, cm, m + 1); }
    return v;
  }
  for (int d = 0; d < 3; ++d) if (cc[d] > 0) { i = d; break; }
  if (i >= 0) {
    int cm[3] = {cc[0], cc[1], cc[2]}; cm[i]--;
    double v = c.QC[i] * os_vrr(c, a, cm, m) + c.WQ[i] * os_vrr(c, a, cm, m + 1);
    if (cm[i] > 0) {
      int cmm[3] = {cm[0], cm[1], cm[2]}; cmm[i]--;
      v += cm[i] / (2 * c.q) * (os_vrr(c, a, cmm, m) - c.rho / c.q * os_vrr(c, a, cmm, m + 1));
    }
    return v;
  }
  return c.base * c.F[m];
}
double os_hrr(const OSCtx& c, int a[3], int b[3], int cc[3], int d[3]) {
  for (int i = 0; i < 3; ++i)
    if (b[i] > 0) {
      int a1[3] = {a[0], a[1], a[2]}, b1[3] = {b[0], b[1], b[2]};
      a1[i]++; b1[i]--;
      return os_hrr(c, a1, b1, cc, d) + c.AB[i] * os_hrr(c, a, b1, cc, d);
    }
  for (int i = 0; i < 3; ++i)
    if (d[i] > 0) {
      int c1[3] = {cc[0], cc[1], cc[2]}, d1[3] = {d[0], d[1], d[2]};
      c1[i]++; d1[i]--;
      return os_hrr(c, a, b, c1, d1) + c.CD[i] * os_hrr(c, a, b, cc, d1);
    }
  return os_vrr(c, a, cc, 0);
}
}  // namespace

extern "C" {
// One raw primitive integral by Obara-Saika.  ang = {ax,ay,az,bx,..,dz}.
// centres/exps: A,B,C,D (xyz each) and alpha,beta,gamma,delta.
double orc_eri_prim_os(const double* ctr, const double* ex, const int* ang) {
  OSCtx c;
  double a = ex[0], b = ex[1], g = ex[2], d = ex[3];
  c.p = a + b; c.q = g + d; c.rho = c.p * c.q / (c.p + c.q);
  double P[3], Q[3], W[3], rab = 0, rcd = 0, rpq = 0;
  for (int i = 0; i < 3; ++i) {
    P[i] = (a * ctr[i] + b * ctr[3 + i]) / c.p;
    Q[i] = (g * ctr[6 + i] + d * ctr[9 + i]) / c.q;
    W[i] = (c.p * P[i] + c.q * Q[i]) / (c.p + c.q);
    c.PA[i] = P[i] - ctr[i]; c.QC[i] = Q[i] - ctr[6 + i];
    c.WP[i] = W[i] - P[i]; c.WQ[i] = W[i] - Q[i];
    c.AB[i] = ctr[i] - ctr[3 + i]; c.CD[i] = ctr[6 + i] - ctr[9 + i];
    rab += c.AB[i] * c.AB[i]; rcd += c.CD[i] * c.CD[i]; rpq += (P[i] - Q[i]) * (P[i] - Q[i]);
  }
  int L = 0; for (int i = 0; i < 12; ++i) L += ang[i];
  c.base = 2.0 * std::pow(M_PI, 2.5) / (c.p * c.q * std::sqrt(c.p + c.q)) * std::exp(-a * b / c.p * rab) *
           std::exp(-g * d / c.q * rcd);
  c.F[L] = boys_series(L, c.rho * rpq);
  double ex_ = std::exp(-c.rho * rpq);
  for (int j = L; j > 0; --j) c.F[j - 1] = (2 * c.rho * rpq * c.F[j] + ex_) / (2 * j - 1);
  int A_[3] = {ang[0], ang[1], ang[2]}, B_[3] = {ang[3], ang[4], ang[5]};
  int C_[3] = {ang[6], ang[7], ang[8]}, D_[3] = {ang[9], ang[10], ang[11]};
  return os_hrr(c, A_, B_, C_, D_);
}

// The same raw primitive integral through this file's McMurchie-Davidson path.
double orc_eri_prim_md(const double* ctr, const double* ex, const int* ang) {
  Prim p[4];
  int L[4];
  for (int k = 0; k < 4; ++k) {
    p[k] = Prim{};
    p[k].exponent = ex[k]; p[k].x = ctr[3 * k]; p[k].y = ctr[3 * k + 1]; p[k].z = ctr[3 * k + 2];
    L[k] = ang[3 * k] + ang[3 * k + 1] + ang[3 * k + 2];
  }
  PairData ab, cd;
  make_pair(p[0], L[0], p[1], L[1], ab); make_pair(p[2], L[2], p[3], L[3], cd);
  std::vector<double> out((size_t)ncart(L[0]) * ncart(L[1]) * ncart(L[2]) * ncart(L[3]));
  prim_quartet(ab, L[0], L[1], false, false, cd, L[2], L[3], false, false, out.data());
  int idx[4];
  for (int k = 0; k < 4; ++k) {
    idx[k] = -1;
    for (int c = 0; c < ncart(L[k]); ++c) {
      const int* t = cart(L[k], c, false);
      if (t[0] == ang[3 * k] && t[1] == ang[3 * k + 1] && t[2] == ang[3 * k + 2]) idx[k] = c;
    }
  }
  return out[((idx[0] * ncart(L[1]) + idx[1]) * ncart(L[2]) + idx[2]) * ncart(L[3]) + idx[3]];
}

// FP64 work of one primitive quartet of class (la lb | lc ld) in the MD
// routine above (prim_quartet: Boys, R recursion, E products, contraction
// sum, prefactor), counted by running it on Op numbers, plus the
// accumulation of the primitive block into the contracted one (one multiply
// by the coefficient product and one add per component).  The reference-style
// Boys function takes the branch of the argument x_boys (table Taylor
// expansion when x < 0.3529 n + 0.0155, else the closed form and upward
// recursion).  out[0] = arithmetic operations, out[1] = special-function
// calls (exp, sqrt, erf, pow, floor).  A generic geometry: no E coefficient
// vanishes, so nothing is skipped.
int orc_md_flops(int la, int lb, int lc, int ld, double x_boys, const double* boys_table, double* out) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la + lb + lc + ld > 12 || !boys_table) return 1;
  Prim q[4];
  const double ex[4] = {1.3, 0.7, 0.9, 0.45};
  const double c[4][3] = {{0.1, 0.2, 0.3}, {1.1, -0.4, 0.6}, {-0.5, 0.8, 1.7}, {0.7, 1.3, -0.9}};
  for (int k = 0; k < 4; ++k) {
    q[k] = Prim{};
    q[k].exponent = ex[k]; q[k].x = c[k][0]; q[k].y = c[k][1]; q[k].z = c[k][2];
  }
  PairData ab, cd;
  make_pair(q[0], la, q[1], lb, ab);
  make_pair(q[2], lc, q[3], ld, cd);
  // place the pair centres so that alpha |PQ|^2 = x_boys
  const double p = ab.p, qq = cd.p, alpha = p * qq / (p + qq);
  const double d[3] = {ab.P[0] - cd.P[0], ab.P[1] - cd.P[1], ab.P[2] - cd.P[2]};
  const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double want = std::sqrt(x_boys / alpha);
  for (int k = 0; k < 3; ++k) cd.P[k] = ab.P[k] - d[k] * (want / r);
  const int ncomp = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  std::vector<Op> blk(ncomp);
  const BoysMode saved = g_boys;
  g_boys = BoysMode{1, boys_table};
  Op::n_arith = 0;
  Op::n_special = 0;
  prim_quartet(ab, la, lb, false, false, cd, lc, ld, false, false, blk.data());
  g_boys = saved;
  out[0] = (double)Op::n_arith + 2.0 * ncomp;
  out[1] = (double)Op::n_special;
  return 0;
}

// The same count for the nuclear-derivative routine (grad_prim_quartet, the
// body of orc_grad_2e): the R table one order higher, Gamma(D) per component
// from a density of ones, and the 12 centre derivatives of every component
// (one or two shifted Hermite contractions each), accumulated.  Norms enter
// as per-component products (the primitive norm itself is a per-primitive
// constant and is not counted).
int orc_grad_md_flops(int la, int lb, int lc, int ld, double x_boys, const double* boys_table, double* out);

int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"

// ---- analytic RHF nuclear gradient (FP64 host) ---------------------------
// The reference has no gradient; this is the oracle of the engine's backward
// kernels.  Every centre derivative of a Cartesian Gaussian is a two-term
// combination of Gaussians with one exponent raised or lowered:
//   d/dA_x [x_A^i exp(-a x_A^2)] = 2a x_A^(i+1) exp(..) - i x_A^(i-1) exp(..)
// with x_A = x - A_x, so each derivative integral is formed from integrals of
// the same primitive pair with shifted angular momentum (E tables built one
// order higher, R one order higher).  Energy (RHF, D = 2 C_occ C_occ^T):
//   E = sum D h + 1/2 sum D_ab D_cd (ab|cd) - 1/4 sum D_ac D_bd (ab|cd) + V_nn
//   dE/dX = sum D dh/dX + sum Gamma_abcd d(ab|cd)/dX - sum W dS/dX + dV_nn/dX
// with Gamma_abcd = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc) summed over all
// ordered index quadruples, W = 2 C_occ eps C_occ^T (= D F D / 2 at convergence).

// Raw primitive (ab|cd) for explicit exponent vectors (E tables of ab/cd must
// reach the requested orders; R must reach sum of all orders).
template <class T>
static T md_raw4(const PairData& ab, const PairData& cd, const RtabT<T>& R, const int* A, const int* B, const int* C,
                 const int* D) {
  T s = 0;
  for (int t = 0; t <= A[0] + B[0]; ++t)
    for (int u = 0; u <= A[1] + B[1]; ++u)
      for (int v = 0; v <= A[2] + B[2]; ++v) {
        const T eab = T(ab.E[0].e[A[0]][B[0]][t]) * T(ab.E[1].e[A[1]][B[1]][u]) * T(ab.E[2].e[A[2]][B[2]][v]);
        if (eab == T(0.0)) continue;
        for (int tt = 0; tt <= C[0] + D[0]; ++tt)
          for (int uu = 0; uu <= C[1] + D[1]; ++uu)
            for (int vv = 0; vv <= C[2] + D[2]; ++vv) {
              const T ecd = T(cd.E[0].e[C[0]][D[0]][tt]) * T(cd.E[1].e[C[1]][D[1]][uu]) * T(cd.E[2].e[C[2]][D[2]][vv]);
              const T sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
              s += eab * sg * ecd * R.r[t + tt][u + uu][v + vv];
            }
      }
  return s;
}

// Gradient contribution of one primitive quartet f[0..3] (shell types l):
// Gamma(D) times the centre derivatives of every component, added to gl.
// deg: the permutational weight of the quartet.
template <class T>
static void grad_prim_quartet(const Prim* const f[4], const int l[4], const double* norm, size_t MM, const double* D,
                              double deg, T* gl) {
  using std::sqrt;
  PairData ab, cd;
  make_pair(*f[0], l[0] + 1, *f[1], l[1] + 1, ab);
  make_pair(*f[2], l[2] + 1, *f[3], l[3] + 1, cd);
  const T p = ab.p, q = cd.p;
  RtabT<T> R;
  hermite_r(l[0] + l[1] + l[2] + l[3] + 1, p * q / (p + q), T(ab.P[0]) - T(cd.P[0]), T(ab.P[1]) - T(cd.P[1]),
            T(ab.P[2]) - T(cd.P[2]), R);
  const T pref = T(2.0 * std::pow(M_PI, 2.5)) / (p * q * sqrt(p + q)) * T(ab.K) * T(cd.K);
  const T cf = T(f[0]->coefficient) * T(f[1]->coefficient) * T(f[2]->coefficient) * T(f[3]->coefficient);
  for (int i0 = 0; i0 < ncart(l[0]); ++i0)
    for (int i1 = 0; i1 < ncart(l[1]); ++i1)
      for (int i2 = 0; i2 < ncart(l[2]); ++i2)
        for (int i3 = 0; i3 < ncart(l[3]); ++i3) {
          const int ic[4] = {i0, i1, i2, i3};
          size_t bf[4];
          int e[4][3];
          T nrm = 1.0;
          for (int k = 0; k < 4; ++k) {
            const int* c = cart(l[k], ic[k], false);
            for (int d = 0; d < 3; ++d) e[k][d] = c[d];
            bf[k] = f[k]->basis_index + ic[k];
            nrm *= T(prim_norm(f[k]->exponent, c[0], c[1], c[2])) * T(norm[bf[k]]);
          }
          const size_t pp = bf[0], qq = bf[1], rr = bf[2], ss = bf[3];
          const T gam = T(0.5) * T(D[pp * MM + qq]) * T(D[rr * MM + ss]) -
                        T(0.125) * (T(D[pp * MM + rr]) * T(D[qq * MM + ss]) + T(D[pp * MM + ss]) * T(D[qq * MM + rr]));
          if (gam == T(0.0)) continue;
          const T fac = T(deg) * cf * nrm * gam * pref;
          for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) {
              const int e0 = e[k][d];
              e[k][d] = e0 + 1;
              T v = T(2.0 * f[k]->exponent) * md_raw4(ab, cd, R, e[0], e[1], e[2], e[3]);
              if (e0 > 0) {
                e[k][d] = e0 - 1;
                v -= T(e0) * md_raw4(ab, cd, R, e[0], e[1], e[2], e[3]);
              }
              e[k][d] = e0;
              gl[3 * f[k]->atom_index + d] += fac * v;
            }
        }
}

extern "C" {

// Two-electron part: grad[3*natoms] = sum Gamma(D) d(ab|cd)/dX over the same
// screened primitive quartets as orc_jk (thr applies to Q_x Q_y).
int orc_grad_2e(const void* vprims, const void* vtypes, int ntypes, const double* norm, int M, const double* Q,
                double thr, const double* D, int natoms, double* grad) {
  const Prim* P = (const Prim*)vprims;
  auto pl = pair_list(P, (const TypeInfo*)vtypes, ntypes, Q);
  const size_t n = pl.size(), MM = M;
  const int nt = omp_get_max_threads();
  std::vector<std::vector<double>> gt(nt, std::vector<double>((size_t)3 * natoms, 0.0));
#pragma omp parallel
  {
    double* gl = gt[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 1)
    for (long long xx = 0; xx < (long long)n; ++xx) {
      const size_t x = (size_t)xx;
      for (size_t y = x; y < n; ++y) {
        if (pl[x].Q * pl[y].Q < thr) continue;
        const Prim* f[4] = {&P[pl[x].a], &P[pl[x].b], &P[pl[y].a], &P[pl[y].b]};
        int l[4];
        for (int k = 0; k < 4; ++k) l[k] = f[k]->shell_type;
        const bool sb = pl[x].a == pl[x].b, sk = pl[y].a == pl[y].b, sbk = x == y;
        const double deg = (sb ? 1.0 : 2.0) * (sk ? 1.0 : 2.0) * (sbk ? 1.0 : 2.0);
        grad_prim_quartet(f, l, norm, MM, D, deg, gl);
      }
    }
  }
  for (int i = 0; i < 3 * natoms; ++i) {
    double s = 0;
    for (int t = 0; t < nt; ++t) s += gt[t][i];
    grad[i] = s;
  }
  return 0;
}

// One-electron part: grad = sum D d(T+V)/dX - sum W dS/dX, including the
// Hellmann-Feynman term of each nucleus (by translational invariance of every
// single-nucleus attraction integral, d/dC = -(d/dA + d/dB)).
int orc_grad_1e(const void* vprims, int nprims, const double* norm, int M, int natoms, const int* Z, const double* xyz,
                const double* D, const double* W, double* grad) {
  const Prim* P = (const Prim*)vprims;
  const size_t MM = M;
  std::fill(grad, grad + 3 * natoms, 0.0);
  std::vector<Rtab> Rn(natoms);
  std::vector<double> vc(natoms);
  for (int ia = 0; ia < nprims; ++ia)
    for (int ib = 0; ib < nprims; ++ib) {
      const Prim &a = P[ia], &b = P[ib];
      const int la = a.shell_type, lb = b.shell_type;
      PairData pd;
      make_pair(a, la + 1, b, lb + 3, pd);  // shifted orders; +2 more on b for the kinetic term
      const double p = pd.p, bb = b.exponent;
      for (int c = 0; c < natoms; ++c)
        hermite_r(la + lb + 1, p, pd.P[0] - xyz[3 * c], pd.P[1] - xyz[3 * c + 1], pd.P[2] - xyz[3 * c + 2], Rn[c]);
      const double sp = std::pow(M_PI / p, 1.5) * pd.K, vp = 2.0 * M_PI / p * pd.K;
      // S, T and the per-nucleus V of one (possibly shifted) component pair
      auto stv = [&](const int* A, const int* B, double& s, double& t) {
        double s1[3], t1[3];
        for (int d = 0; d < 3; ++d) {
          const int i = A[d], j = B[d];
          const double sij = pd.E[d].e[i][j][0];
          const double sp2 = pd.E[d].e[i][j + 2][0];
          const double sm2 = j >= 2 ? pd.E[d].e[i][j - 2][0] : 0.0;
          s1[d] = sij;
          t1[d] = -2.0 * bb * bb * sp2 + bb * (2 * j + 1) * sij - 0.5 * j * (j - 1) * sm2;
        }
        s = s1[0] * s1[1] * s1[2] * sp;
        t = (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]) * sp;
        for (int c = 0; c < natoms; ++c) {
          double acc = 0;
          for (int tt = 0; tt <= A[0] + B[0]; ++tt)
            for (int u = 0; u <= A[1] + B[1]; ++u)
              for (int w = 0; w <= A[2] + B[2]; ++w)
                acc += pd.E[0].e[A[0]][B[0]][tt] * pd.E[1].e[A[1]][B[1]][u] * pd.E[2].e[A[2]][B[2]][w] *
                       Rn[c].r[tt][u][w];
          vc[c] = -Z[c] * acc * vp;
        }
      };
      for (int ca = 0; ca < ncart(la); ++ca)
        for (int cb = 0; cb < ncart(lb); ++cb) {
          int e[2][3];
          const int *A = cart(la, ca, 0), *B = cart(lb, cb, 0);
          for (int d = 0; d < 3; ++d) { e[0][d] = A[d]; e[1][d] = B[d]; }
          const size_t p_ = a.basis_index + ca, q_ = b.basis_index + cb;
          const double nn = prim_norm(a.exponent, A[0], A[1], A[2]) * prim_norm(b.exponent, B[0], B[1], B[2]) *
                            a.coefficient * b.coefficient * norm[p_] * norm[q_];
          const double dpq = D[p_ * MM + q_] * nn, wpq = W[p_ * MM + q_] * nn;
          const Prim* fk[2] = {&a, &b};
          for (int k = 0; k < 2; ++k)
            for (int d = 0; d < 3; ++d) {
              const int e0 = e[k][d];
              double s, t, ds, dt;
              e[k][d] = e0 + 1;
              stv(e[0], e[1], s, t);
              const double up = 2.0 * fk[k]->exponent;
              ds = up * s; dt = up * t;
              std::vector<double> dv(natoms);
              for (int c = 0; c < natoms; ++c) dv[c] = up * vc[c];
              if (e0 > 0) {
                e[k][d] = e0 - 1;
                stv(e[0], e[1], s, t);
                ds -= e0 * s; dt -= e0 * t;
                for (int c = 0; c < natoms; ++c) dv[c] -= e0 * vc[c];
              }
              e[k][d] = e0;
              double vsum = 0;
              for (int c = 0; c < natoms; ++c) {
                vsum += dv[c];
                grad[3 * c + d] -= dpq * dv[c];  // nucleus c moves opposite to the basis centres
              }
              grad[3 * fk[k]->atom_index + d] += dpq * (dt + vsum) - wpq * ds;
            }
        }
    }
  return 0;
}

// Nuclear repulsion gradient.
void orc_grad_nuc(int natoms, const int* Z, const double* xyz, double* grad) {
  for (int a = 0; a < natoms; ++a)
    for (int d = 0; d < 3; ++d) grad[3 * a + d] = 0.0;
  for (int a = 0; a < natoms; ++a)
    for (int b = 0; b < natoms; ++b) {
      if (a == b) continue;
      double r[3], r2 = 0;
      for (int d = 0; d < 3; ++d) { r[d] = xyz[3 * a + d] - xyz[3 * b + d]; r2 += r[d] * r[d]; }
      const double f = Z[a] * Z[b] / (r2 * std::sqrt(r2));
      for (int d = 0; d < 3; ++d) grad[3 * a + d] -= f * r[d];
    }
}

}  // extern "C"

// ---- RI-JK energy gradient: derivative 3-centre and 2-centre integrals ---
// With A^P_mn = (mn|P), V_PQ = (P|Q), gamma = A.D, c = V^-1 gamma,
// At^P = sum_Q V^-1_PQ A^Q and M_PQ = tr(A^P D A^Q D), the RI two-electron
// energy 1/2 gamma^T V^-1 gamma - 1/4 sum V^-1_PQ M_PQ has the derivative
//   sum_P sum_mn G^P_mn d(mn|P)/dX + sum_PQ Gam_PQ d(P|Q)/dX,
//   G^P = c_P D - 1/2 D At^P D,   Gam = -1/2 c c^T + 1/4 V^-1 M V^-1,
// which the two functions below contract (the caller forms G and Gam).
// Centre derivatives use the E-shift of the orbital (or first auxiliary)
// function; the remaining centre follows by translational invariance.
extern "C" {

// grad[3*natoms] += sum_P sum_mn G[P][m][n] d(mn|P)/dX over the screened set
// of orc_int3c2e (G full in m, n).
int orc_grad_3c(const void* vprims, const void* vtypes, int ntypes, const double* norm, int M, const double* Q,
                const void* vaux, int naux_prims, const double* aux_norm, int Maux, const double* Qaux, double thr,
                const double* G, int natoms, double* grad) {
  const Prim* P = (const Prim*)vprims;
  const Prim* A = (const Prim*)vaux;
  auto pl = pair_list(P, (const TypeInfo*)vtypes, ntypes, Q);
  const size_t MM = M;
  (void)Maux;
  const int nt = omp_get_max_threads();
  std::vector<std::vector<double>> gt(nt, std::vector<double>((size_t)3 * natoms, 0.0));
#pragma omp parallel for schedule(dynamic, 16)
  for (long long ci = 0; ci < naux_prims; ++ci) {
    double* gl = gt[omp_get_thread_num()].data();
    const Prim& pc = A[ci];
    const int lc = pc.shell_type;
    Prim zc = s_dummy(pc);
    PairData pcd;
    make_pair(pc, lc, zc, 0, pcd);
    for (size_t x = 0; x < pl.size(); ++x) {
      if (pl[x].Q * Qaux[ci] < thr) continue;
      const Prim &a = P[pl[x].a], &b = P[pl[x].b];
      const int la = a.shell_type, lb = b.shell_type;
      PairData ab;
      make_pair(a, la + 1, b, lb + 1, ab);
      const double p = ab.p, q = pcd.p;
      Rtab R;
      hermite_r(la + lb + lc + 1, p * q / (p + q), ab.P[0] - pcd.P[0], ab.P[1] - pcd.P[1], ab.P[2] - pcd.P[2], R);
      const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q)) * ab.K * pcd.K;
      const int zero3[3] = {0, 0, 0};
      for (int ia = 0; ia < ncart(la); ++ia)
        for (int ib = 0; ib < ncart(lb); ++ib)
          for (int ic = 0; ic < ncart(lc); ++ic) {
            const int *CA = cart(la, ia, 0), *CB = cart(lb, ib, 0), *CC = cart(lc, ic, true);
            const size_t pp = a.basis_index + ia, qq = b.basis_index + ib, r = pc.basis_index + ic;
            double w = G[(r * MM + pp) * MM + qq];
            if (pl[x].a != pl[x].b) w += G[(r * MM + qq) * MM + pp];
            if (w == 0.0) continue;
            const double fac = w * pref * a.coefficient * b.coefficient * pc.coefficient *
                               prim_norm(a.exponent, CA[0], CA[1], CA[2]) * prim_norm(b.exponent, CB[0], CB[1], CB[2]) *
                               prim_norm(pc.exponent, CC[0], CC[1], CC[2]) * norm[pp] * norm[qq] * aux_norm[r];
            int e[2][3];
            for (int d = 0; d < 3; ++d) { e[0][d] = CA[d]; e[1][d] = CB[d]; }
            const Prim* f[2] = {&a, &b};
            double dsum[3] = {0, 0, 0};
            for (int k = 0; k < 2; ++k)
              for (int d = 0; d < 3; ++d) {
                const int e0 = e[k][d];
                e[k][d] = e0 + 1;
                double v = 2.0 * f[k]->exponent * md_raw4(ab, pcd, R, e[0], e[1], CC, zero3);
                if (e0 > 0) {
                  e[k][d] = e0 - 1;
                  v -= e0 * md_raw4(ab, pcd, R, e[0], e[1], CC, zero3);
                }
                e[k][d] = e0;
                gl[3 * f[k]->atom_index + d] += fac * v;
                dsum[d] += fac * v;
              }
            for (int d = 0; d < 3; ++d) gl[3 * pc.atom_index + d] -= dsum[d];
          }
    }
  }
  for (int i = 0; i < 3 * natoms; ++i)
    for (int t = 0; t < nt; ++t) grad[i] += gt[t][i];
  return 0;
}

// grad[3*natoms] += sum_PQ Gam[P][Q] d(P|Q)/dX over the screened set of
// orc_int2c2e.
int orc_grad_2c(const void* vaux, int naux_prims, const double* aux_norm, int Maux, const double* Qaux, double thr,
                const double* Gam, int natoms, double* grad) {
  const Prim* P = (const Prim*)vaux;
  (void)natoms;
  const int zero3[3] = {0, 0, 0};
  for (int i = 0; i < naux_prims; ++i)
    for (int j = i; j < naux_prims; ++j) {
      if (Qaux[i] * Qaux[j] < thr) continue;
      const Prim &a = P[i], &b = P[j];
      const int la = a.shell_type, lb = b.shell_type;
      Prim za = s_dummy(a), zb = s_dummy(b);
      PairData pa, pb;
      make_pair(a, la + 1, za, 0, pa);
      make_pair(b, lb, zb, 0, pb);
      const double p = pa.p, q = pb.p;
      Rtab R;
      hermite_r(la + lb + 1, p * q / (p + q), pa.P[0] - pb.P[0], pa.P[1] - pb.P[1], pa.P[2] - pb.P[2], R);
      const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q)) * pa.K * pb.K;
      for (int ca = 0; ca < ncart(la); ++ca)
        for (int cb = 0; cb < ncart(lb); ++cb) {
          const int *CA = cart(la, ca, true), *CB = cart(lb, cb, true);
          const size_t pp = a.basis_index + ca, r = b.basis_index + cb;
          double w = Gam[pp * Maux + r];
          if (i != j) w += Gam[r * Maux + pp];
          if (w == 0.0) continue;
          const double fac = w * pref * a.coefficient * b.coefficient * prim_norm(a.exponent, CA[0], CA[1], CA[2]) *
                             prim_norm(b.exponent, CB[0], CB[1], CB[2]) * aux_norm[pp] * aux_norm[r];
          int e[3] = {CA[0], CA[1], CA[2]};
          for (int d = 0; d < 3; ++d) {
            const int e0 = e[d];
            e[d] = e0 + 1;
            double v = 2.0 * a.exponent * md_raw4(pa, pb, R, e, zero3, CB, zero3);
            if (e0 > 0) {
              e[d] = e0 - 1;
              v -= e0 * md_raw4(pa, pb, R, e, zero3, CB, zero3);
            }
            e[d] = e0;
            grad[3 * a.atom_index + d] += fac * v;
            grad[3 * b.atom_index + d] -= fac * v;
          }
        }
    }
  return 0;
}

}  // extern "C"

extern "C" int orc_grad_md_flops(int la, int lb, int lc, int ld, double x_boys, const double* boys_table, double* out) {
  if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la + lb + lc + ld > 11 || !boys_table) return 1;
  const int L[4] = {la, lb, lc, ld};
  Prim q[4];
  const double ex[4] = {1.3, 0.7, 0.9, 0.45};
  const double c[4][3] = {{0.1, 0.2, 0.3}, {1.1, -0.4, 0.6}, {-0.5, 0.8, 1.7}, {0.7, 1.3, -0.9}};
  size_t off = 0;
  for (int k = 0; k < 4; ++k) {
    q[k] = Prim{};
    q[k].exponent = ex[k]; q[k].coefficient = 1.0; q[k].shell_type = L[k];
    q[k].x = c[k][0]; q[k].y = c[k][1]; q[k].z = c[k][2];
    q[k].basis_index = off; q[k].atom_index = k;
    off += ncart(L[k]);
  }
  // pair centres moved along P - Q so that alpha |PQ|^2 = x_boys (as orc_md_flops)
  PairData ab, cd;
  make_pair(q[0], la + 1, q[1], lb + 1, ab);
  make_pair(q[2], lc + 1, q[3], ld + 1, cd);
  const double alpha = ab.p * cd.p / (ab.p + cd.p);
  const double d[3] = {ab.P[0] - cd.P[0], ab.P[1] - cd.P[1], ab.P[2] - cd.P[2]};
  const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double scale = std::sqrt(x_boys / alpha) / r;
  // shift the ket primitives rigidly: P_cd moves by (scale - 1) (Q - P)
  for (int k = 2; k < 4; ++k) {
    q[k].x += (scale - 1.0) * -d[0];
    q[k].y += (scale - 1.0) * -d[1];
    q[k].z += (scale - 1.0) * -d[2];
  }
  const size_t M = off;
  std::vector<double> D(M * M, 1.0), nrm(M, 1.0);
  const Prim* f[4] = {&q[0], &q[1], &q[2], &q[3]};
  std::vector<Op> gl(12);
  const BoysMode saved = g_boys;
  g_boys = BoysMode{1, boys_table};
  Op::n_arith = 0;
  Op::n_special = 0;
  grad_prim_quartet(f, L, nrm.data(), M, D.data(), 1.0, gl.data());
  g_boys = saved;
  out[0] = (double)Op::n_arith;
  out[1] = (double)Op::n_special;
  return 0;
}
