// RHF driver, one-electron integrals and the nuclear gradient on the host side.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "engine.hpp"
#include "scf_kernels.hpp"

namespace gansu {

struct ScfResult {
  double e_elec = 0, e_total = 0;
  int iterations = 0, fock_builds = 0, converged = 0;
  int graph_launches = 0;    // scf_solver=purify: captured iterations launched
  std::vector<double> D;     // final density (M x M)
  std::vector<double> grad;  // dE/dX (3 natoms), when requested
  std::vector<double> C;     // final MO coefficients (M x M row-major, MOs in columns)
  std::vector<double> eps;   // their orbital energies (ascending)
};

// HF::single_point_energy (hf.cu:245-314) for RHF; with gradient = true the
// analytic nuclear gradient of the converged energy follows.
// Da0/Db0 (host, M x M) replace the guess when force_density is set or
// initial_guess=density: F(Da0 + Db0) starts the loop (HF::solve with
// densities, rhf.cu:122-134); without both the core guess runs then.
ScfResult rhf_run(const Molecule& mol, const Molecule* aux, const Params& prm, bool gradient = false,
                  const double* Da0 = nullptr, const double* Db0 = nullptr, bool force_density = false);

struct UhfResult {
  double e_elec = 0, e_total = 0;
  int iterations = 0, fock_builds = 0, converged = 0;
  std::vector<double> Da, Db;  // final alpha and beta densities (M x M)
};
// HF::single_point_energy for UHF (uhf.cu): core, GWH or SAD guess with the
// symmetry break, per-spin DIIS or damping.  Da0/Db0 (host, M x M) replace
// the guess as in rhf_run: F(Da0, Db0), then C per spin, the symmetry break
// and D (HF::solve with densities).
UhfResult uhf_run(const Molecule& mol, const Molecule* aux, const Params& prm, const double* Da0 = nullptr,
                  const double* Db0 = nullptr, bool force_density = false);
// HF::single_point_energy for ROHF (rohf.cu): core or GWH guess, DIIS,
// damping or optimal damping on the unified Fock matrix; Da / Db of the
// result hold the closed / open densities.
UhfResult rohf_run(const Molecule& mol, const Molecule* aux, const Params& prm);

// Molden file of the orbitals C (M x n_mo row-major, MOs in columns) with
// energies eps and occupations occ (rhf.cu:363-425 layout).
void write_molden(const Molecule& mol, const double* C, const double* eps, const double* occ, int nmo,
                  const std::string& path);

// Device copies of a molecule's shell tables for the one-electron kernels.
struct OneIntHost {
  DevBuf<gansu_primitive_shell> prims;
  DevBuf<int> off, sprims, sl, sb, bfs, Z;
  DevBuf<double> w, xyz, boys;
  int M = 0, natoms = 0;
  double zsum = 0.0;
  std::vector<int> shell_L;  // angular momentum of each contracted shell (host copy)
  explicit OneIntHost(const Molecule& mol);
  OneIntArgs args(double* S, double* T, double* V) const;
};

void compute_oneint(const Molecule& mol, double* dS, double* dT, double* dV, hipStream_t s);
OneIntTables make_oneint_tables(const Molecule& mol);
// compute_oneint without device allocation (oneint_noalloc)
void compute_oneint_noalloc(const Molecule& mol, double* dS, double* dT, double* dV, hipStream_t s);
// d_grad (device, 3 natoms) = sum D d(T+V)/dX - sum W dS/dX, overwritten.
void compute_oneint_gradient(const Molecule& mol, const double* dD, const double* dW, double* dgrad, hipStream_t s);
// dV_nn/dX (host, 3 natoms)
std::vector<double> nuclear_repulsion_gradient(const Molecule& mol);

}  // namespace gansu
