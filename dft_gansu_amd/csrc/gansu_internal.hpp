// Internal host-side types shared by the engine sources (not part of the C ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "gansu_fock.h"

namespace gansu {

// Error carried to the C boundary as (code, message).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& msg);

// Highest angular momentum the compiled kernels handle.
constexpr int kMaxL = 2;     // orbital basis: s, p, d
constexpr int kMaxLAux = 3;  // auxiliary basis: s, p, d, f

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

struct Atom { int Z; double xyz[3]; };

struct Molecule {
  std::vector<Atom> atoms;
  std::vector<gansu_primitive_shell> prims;   // sorted by shell type (molecular.hpp:173)
  std::vector<gansu_shell_type_info> types;
  std::vector<double> norm;                   // per basis function
  int nbasis = 0;
  int n_electrons = 0, n_alpha = 0, n_beta = 0;
  std::string gbs_path;                       // basis file the molecule was built from
  std::vector<int> atom_basis;                // first basis function of each atom, natoms + 1 entries
  double nuclear_repulsion() const;
};

// Superposition-of-atomic-densities guess (utils.cpp:21-58, rhf.hpp:551-584):
// per-spin atomic densities read from the .sad file beside the basis file,
// placed on the diagonal blocks of row-major M x M matrices Da and Db.
void sad_density(const Molecule& m, std::vector<double>& Da, std::vector<double>& Db);
// The solved densities in the cache layout of HF::generate_sad_cache
// (hf.hpp:372-402): "M", then the M x M alpha and beta rows.
void write_sad_cache(const std::string& path, int M, const std::vector<double>& Da, const std::vector<double>& Db);

Molecule build_molecule(const std::vector<Atom>& atoms, const std::string& gbs, int charge, int beta_to_alpha,
                        int normalization);
std::vector<Atom> read_xyz(const std::string& path);

struct Params {
  std::map<std::string, std::string> kv;
  Params();
  void set(const std::string& key, const std::string& value, bool from_recipe);
  void load_recipe(const std::string& path);
  std::string get(const std::string& key) const;
  double getd(const std::string& key) const;
  int geti(const std::string& key) const;
};

// Contracted shell rebuilt from the primitive list: primitives that share a
// basis_index form one contracted Gaussian (molecular.hpp:135-150).
struct CShell {
  int L;
  int atom;
  int basis;            // first basis function
  double xyz[3];
  std::vector<int> prims;  // indices into the (type-sorted) primitive array
};
std::vector<CShell> contracted_shells(const gansu_primitive_shell* prims, size_t n, int max_l);

// Boys grid generated from the series F_n(x) = e^-x sum_k (2x)^k/((2n+1)..(2n+2k+1)),
// layout [x_index][n] with kBoysRows values of n per grid point x_i = i/32.
constexpr int kBoysRows = 32;
constexpr int kBoysGrid = 1024;
std::vector<double> boys_grid_xn();

// Option strings "a=b;c=d" for gansu_fock_create.
std::map<std::string, std::string> parse_options(const char* s);

}  // namespace gansu
