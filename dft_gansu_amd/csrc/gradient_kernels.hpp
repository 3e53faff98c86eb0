// Backward (nuclear-gradient) kernels: device layouts and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

#include "fock_kernels.hpp"
#include "scf_kernels.hpp"

namespace gansu {

// Two-electron part: grad_X += sum_quartets deg * sum_abcd Gamma_abcd d(ab|cd)/dX
// with Gamma = 1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc) on the raw-basis
// density.
struct Grad2eArgs {
  const SPair* sp;
  const PairRec* pr;
  const double* alpha;  // exponent of the first primitive of each pair record
  const double* boys;
  double thr;
  const double* Dp;  // raw-basis density (M x M)
  int M;
  const int* shell_atom;
  int natoms;
  double* grad;  // 3 natoms, accumulated
};

// Largest atom count whose gradient accumulator fits the LDS of one workgroup.
constexpr int kGradMaxAtoms = 2048;  // 48 KiB of LDS

// max_bra_prims / max_ket_prims: largest primitive-pair count of a bra / ket
// of the class (the kernel gives each (work block, bra primitive pair) its own
// wave, and for the classes without a register-resident body each ket
// primitive pair too).
hipError_t launch_grad_2e(const ClassLaunch& c, const Grad2eArgs& a, int max_bra_prims, int max_ket_prims,
                          hipStream_t s);

// The class's kernel runs the runtime-indexed body (private arrays in scratch)
// rather than the register-resident unrolled one.
bool grad_2e_uses_scratch(int LA, int LB, int LC, int LD);

// One-electron part, grad += sum D d(T+V)/dX - sum W dS/dX, both matrices in
// the caller's basis (the kernels apply the raw-basis factors w):
// launch_grad_1e adds the overlap and kinetic terms, launch_grad_1e_nuc the
// nuclear-attraction terms (basis centres and Hellmann-Feynman) of the shell
// pairs pairs[0..npairs) = {s1, s2} of class (la, lb), la >= lb, each
// unordered pair once.
hipError_t launch_grad_1e(const OneIntArgs& a, const double* D, const double* W, double* grad, hipStream_t s);
hipError_t launch_grad_1e_nuc(int la, int lb, const OneIntArgs& a, const int2* pairs, int npairs, const double* D,
                              double* grad, hipStream_t s);

}  // namespace gansu
