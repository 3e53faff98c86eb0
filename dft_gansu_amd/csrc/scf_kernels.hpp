// Launchers for one-electron integrals and the dense helpers of the RHF driver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <vector>

#include "gansu_fock.h"

namespace gansu {

struct OneIntArgs {
  const void* prims;       // gansu_primitive_shell[]
  const int* shell_off;    // [nshell+1] ranges into shell_prims
  const int* shell_prims;  // primitive indices grouped by contracted shell
  const int* shell_L;
  const int* shell_basis;
  const int* bf_shell;     // basis function -> shell
  const double* w;         // cgto norm x component factor per basis function
  int M;
  int natoms;
  const double* atom_xyz;
  const int* Z;
  const double* boys;
  double *S, *T, *V;
  double zsum;             // sum of the nuclear charges
  int nshell;              // contracted shells (entries of shell_L)
};

hipError_t launch_oneint(const OneIntArgs& a, hipStream_t s);

// Host copies of the tables k_oneint reads (the device side of OneIntArgs).
struct OneIntTables {
  std::vector<gansu_primitive_shell> prims;
  std::vector<int> off, sprims, sl, sb, bfs, Z;
  std::vector<double> w, xyz, boys;
  int M = 0, natoms = 0;
  double zsum = 0.0;
};
// No-allocation one-electron integrals: the tables are copied into a
// file-scope device array of kOneIntScratchBytes (one per GPU, found with
// hipGetSymbolAddress on the first call on each device) and k_oneint runs
// on them; no hipMalloc or hipFree.  Tables beyond the array are an error
// (GANSU_E_ARG).  Calls that overlap on one GPU share the array and race.
constexpr size_t kOneIntScratchBytes = size_t(4) << 20;
size_t oneint_scratch_bytes(const OneIntTables& t);  // bytes the tables take in the array
void oneint_noalloc(const OneIntTables& t, double* S, double* T, double* V, hipStream_t s);
hipError_t launch_transform(const double* Ucol, const double* evals, int n, double thr, double* X, hipStream_t s);
hipError_t launch_lincomb(int nmat, const double* d_coef, const double* const* d_mats, size_t n, double* out,
                          hipStream_t s);
hipError_t launch_axpby(double a, const double* x, double b, const double* y, size_t n, double* out, hipStream_t s);
// Initial Fock matrix of the GWH guess: F_pq = 1.75 S_pq (H_pp + H_qq) / 2.
hipError_t launch_gwh_guess(const double* H, const double* S, int M, double* F, hipStream_t s);
hipError_t launch_dot_partial(const double* x, const double* y, size_t n, double* part, int nblocks, hipStream_t s);
// *flag = 1 if x holds a NaN (flag left as it is otherwise).
hipError_t launch_has_nan(const double* x, size_t n, int* flag, hipStream_t s);
hipError_t launch_set_int(int* p, int v, hipStream_t s);

}  // namespace gansu
