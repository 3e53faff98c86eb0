// Device-resident pieces of the RHF iteration: FP64 matrix products on the
// matrix cores, SP2 density purification, the energy and convergence test,
// and DIIS, all as kernels that read their control state from device
// memory, so that a whole SCF iteration can be captured into one HIP graph.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace gansu {

constexpr int kGraphDiisMax = 16;  // largest diis_size of the device DIIS

// Control state of the captured iteration (one per SCF run, device memory).
struct ScfDevState {
  double E, Eprev;                 // energy of this and of the previous iteration
  int iter;                        // completed iterations (the driver's `it`)
  int converged;                   // |E - Eprev| < threshold was reached
  int diis_count;                  // DIIS steps taken
  int diis_ok;                     // the last DIIS solve produced coefficients
  int sp2_used;                    // purification steps of the last density (-1: not idempotent)
  int pad_;
  double bounds[2];                // Gershgorin bounds of the last orthogonal-basis Fock matrix
  double gram[kGraphDiisMax * kGraphDiisMax];  // <e_i, e_j> by slot
  double coef[kGraphDiisMax];
};

// C = alpha op(A) op(B) + beta C, all n x n row-major FP64 (op = transpose if
// ta / tb).  With skip != nullptr nothing is done when *skip != 0.
hipError_t launch_gemm64(bool ta, bool tb, int n, double alpha, const double* A, const double* B, double beta,
                         double* C, const int* skip, hipStream_t s);

// SP2 purification (Niklasson, Phys. Rev. B 66, 155115 (2002)) of the
// orthogonal-basis Fock matrix Fo into the projector on its nocc lowest
// eigenvectors.  P0 = (emax I - Fo) / (emax - emin) from Gershgorin bounds;
// step k: P <- P^2 if tr(P) > nocc, else P <- 2P - P^2.  The steps stop once
// the idempotency error tr(P) - tr(P^2) has entered the quadratic regime
// (< 1e-4) and no longer falls between every second step (the rounding
// floor); later steps do nothing.  Sp2Work holds two n x n ping-pong buffers
// and the per-step control words; the projector is copied to `out`, and
// state->sp2_used gives the steps used (-1 if the floor was not reached
// within `steps`, e.g. no gap at the Fermi level; `out` then holds the last
// step's matrix).
struct Sp2Work {
  double* buf[2];
  double* out;
  double* rows;   // [n][2] Gershgorin interval of each row
  double* part;   // [2][nwg][2] partial (trace, sum of squares) per step parity
  double* eps;    // [steps + 1] idempotency error of P_k
  int* done;      // [steps + 1]
  int nwg;        // sp2_workgroups(n)
};
size_t sp2_part_doubles(int n);  // size of Sp2Work::part
int sp2_workgroups(int n);
hipError_t launch_sp2(const double* Fo, int n, int nocc, int steps, const Sp2Work& w, ScfDevState* st,
                      hipStream_t s);

// Host-driven SP2 for the eager loop (the products P P on rocBLAS by the
// caller): init forms P0 from the Gershgorin interval of Fo (rows: 2n
// doubles) and apply replaces P by P2 (square) or 2 P - P2; both leave
// (tr P, sum of squares of P) in trfr[2].  part: sp2e_part_doubles().
size_t sp2e_part_doubles();
hipError_t launch_sp2e_init(const double* Fo, int n, double* rows, double* P, double* part, double* trfr,
                            hipStream_t s);
hipError_t launch_sp2e_apply(double* P, const double* P2, int n, bool square, double* part, double* trfr,
                             hipStream_t s);

// E = 1/2 sum D.(H + F); converged = |E - Eprev| < conv, else iter += 1 and
// Eprev = E (the loop tail of hf.cu:278-307).  part: kEnergyBlocks doubles.
constexpr int kEnergyBlocks = 32;
hipError_t launch_scf_energy(const double* D, const double* H, const double* F, size_t nn, double conv,
                             double* part, ScfDevState* st, hipStream_t s);

// e = A - A^T (the DIIS error F D S - S D F from A = F D S); skipped once converged.
hipError_t launch_antisym(const double* A, int n, double* e, const ScfDevState* st, hipStream_t s);

// One DIIS step on the device (rhf.hpp:298-327 semantics, as the host
// driver's): store (e, F) in slot count % size, the new row of the Gram
// matrix, the bordered system solved by Gaussian elimination with partial
// pivoting, F <- sum c_i F_i once two pairs are held (kept as it is when the
// system is singular).  Nothing is done once converged.  part holds
// kDiisBlocks x size doubles.
constexpr int kDiisBlocks = 128;
hipError_t launch_diis(const double* e, double* F, double* errs, double* focks, double* part, size_t nn, int size,
                       ScfDevState* st, hipStream_t s);

}  // namespace gansu
