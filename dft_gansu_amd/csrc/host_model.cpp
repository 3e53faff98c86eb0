// Host input path: .gbs basis files, .xyz geometries, molecule assembly,
// contracted normalisation, parameter recipes, Boys grid generation.
//
// Semantics follow the reference input path:
//   .gbs   basis_set.cpp:32-124 (SP split, Fortran D exponents, "****" ends an element)
//   .xyz   parseXYZ.cpp:44-115  (Angstrom -> bohr 1.8897259886, types.hpp:141-145;
//                                overlap check 1e-4 bohr; atom-count check)
//   molecule  molecular.hpp:66-190 (primitive list, std::sort by shell type,
//                                ShellTypeInfo, electron counts)
//   norm   basis_set.hpp:106-136 (per contracted shell; odd-L integer-division quirk)
//   recipe parameter_manager.cpp:42-88, :164-225
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "gansu_internal.hpp"

namespace gansu {

namespace {

const char* const kElements[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs",
    "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};
constexpr int kNumElements = sizeof(kElements) / sizeof(kElements[0]);

int element_z(const std::string& sym) {
  for (int i = 0; i < kNumElements; ++i)
    if (sym == kElements[i]) return i + 1;
  throw Error(GANSU_E_IO, "unknown element symbol: " + sym);
}

int shell_letter_l(char c) {
  switch (c) {
    case 'S': case 's': return 0;
    case 'P': case 'p': return 1;
    case 'D': case 'd': return 2;
    case 'F': case 'f': return 3;
    case 'G': case 'g': return 4;
    case 'H': case 'h': return 5;
    case 'I': case 'i': return 6;
  }
  return -1;
}

struct GbsShell { int L; std::vector<double> exps, coefs; };
using GbsLibrary = std::map<std::string, std::vector<GbsShell>>;

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

GbsLibrary read_gbs(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(GANSU_E_IO, "Cannot open basis set file: " + path);
  std::vector<std::string> lines;
  for (std::string l; std::getline(in, l);) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  GbsLibrary lib;
  size_t i = 0;
  // header: everything before the first line that starts with a letter
  while (i < lines.size() && !(lines[i].size() && std::isalpha((unsigned char)lines[i][0]))) ++i;
  while (i < lines.size()) {
    std::string head = trim(lines[i++]);
    if (head.empty()) continue;
    std::istringstream hs(head);
    std::string elem;
    hs >> elem;
    auto& shells = lib[elem];
    while (i < lines.size()) {
      std::string l = trim(lines[i++]);
      if (l == "****") break;
      if (l.empty()) continue;
      std::istringstream ls(l);
      std::string label;
      size_t nprim = 0;
      ls >> label >> nprim;
      if (label.empty() || label.size() > 2) throw Error(GANSU_E_BASIS, "Invalid basis function name: " + label);
      std::vector<GbsShell> sh(label.size());
      for (size_t k = 0; k < label.size(); ++k) {
        sh[k].L = shell_letter_l(label[k]);
        if (sh[k].L < 0) throw Error(GANSU_E_BASIS, "Invalid basis function name: " + label);
      }
      for (size_t p = 0; p < nprim; ++p) {
        if (i >= lines.size()) throw Error(GANSU_E_IO, "truncated basis file: " + path);
        std::string pl = lines[i++];
        for (char& c : pl) if (c == 'D' || c == 'd') c = 'E';
        std::istringstream ps(pl);
        double e = 0;
        ps >> e;
        for (size_t k = 0; k < label.size(); ++k) {
          double c = 0;
          ps >> c;
          sh[k].exps.push_back(e);
          sh[k].coefs.push_back(c);
        }
        if (!ps) throw Error(GANSU_E_IO, "bad primitive line in " + path + ": " + pl);
      }
      for (auto& s : sh) shells.push_back(std::move(s));
    }
  }
  return lib;
}

// Contracted normalisation of one shell, basis_set.hpp:106-136.  The factor
// depends on L only (not on the Cartesian component).  normalization == 0
// keeps the reference's integer division L/2 for the second primitive's
// exponent, which mis-normalises odd L (energies are invariant to it).
double contracted_norm(int L, const std::vector<double>& e, const std::vector<double>& c, int normalization) {
  const double pre = std::pow(2.0 * M_PI, -0.75);
  const double ea = L / 2.0 + 0.75;
  const double eb = normalization == 0 ? double(L / 2) + 0.75 : ea;
  double s = 0.0;
  for (size_t a = 0; a < e.size(); ++a) {
    const double na = pre * std::pow(4.0 * e[a], ea);
    for (size_t b = 0; b < e.size(); ++b) {
      const double nb = pre * std::pow(4.0 * e[b], eb);
      s += c[a] * na * c[b] * nb / std::pow(2.0 * (e[a] + e[b]), L + 1.5);
    }
  }
  s *= std::pow(2.0 * M_PI, 1.5);
  return std::pow(s, -0.5);
}

}  // namespace

std::string element_symbol(int Z) {
  if (Z < 1 || Z > kNumElements) throw Error(GANSU_E_ARG, "atomic number out of range");
  return kElements[Z - 1];
}

double Molecule::nuclear_repulsion() const {
  double e = 0.0;
  for (size_t a = 0; a < atoms.size(); ++a)
    for (size_t b = 0; b < a; ++b) {
      double d2 = 0;
      for (int k = 0; k < 3; ++k) d2 += (atoms[a].xyz[k] - atoms[b].xyz[k]) * (atoms[a].xyz[k] - atoms[b].xyz[k]);
      e += double(atoms[a].Z) * atoms[b].Z / std::sqrt(d2);
    }
  return e;
}

std::vector<Atom> read_xyz(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(GANSU_E_IO, "Failed to open xyz file: " + path);
  std::string line;
  if (!std::getline(in, line)) throw Error(GANSU_E_IO, "Failed to read atom count from file: " + path);
  int count = 0;
  try { count = std::stoi(line); } catch (...) { throw Error(GANSU_E_IO, "Invalid atom count in file: " + path); }
  if (!std::getline(in, line)) throw Error(GANSU_E_IO, "Failed to read comment line from file: " + path);
  const double a2b = 1.8897259886;   // types.hpp:141-145
  const double dup = 1e-4;           // bohr, parseXYZ.cpp:69
  std::vector<Atom> atoms;
  int lineno = 2;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.empty()) continue;
    std::istringstream ls(line);
    std::string sym;
    double x, y, z;
    if (!(ls >> sym >> x >> y >> z))
      throw Error(GANSU_E_IO, "Invalid line format in file at line " + std::to_string(lineno) + ": " + line);
    Atom a{element_z(sym), {x * a2b, y * a2b, z * a2b}};
    for (const auto& b : atoms) {
      double d2 = 0;
      for (int k = 0; k < 3; ++k) d2 += (a.xyz[k] - b.xyz[k]) * (a.xyz[k] - b.xyz[k]);
      if (d2 < dup * dup)
        throw Error(GANSU_E_IO, "Detected overlapping atoms at line " + std::to_string(lineno));
    }
    atoms.push_back(a);
  }
  if ((int)atoms.size() != count) throw Error(GANSU_E_IO, "Atom count mismatch in file: " + path);
  return atoms;
}

Molecule build_molecule(const std::vector<Atom>& atoms, const std::string& gbs, int charge, int beta_to_alpha,
                        int normalization) {
  if (atoms.empty()) throw Error(GANSU_E_ARG, "No atoms are given.");
  if (normalization != 0 && normalization != 1) throw Error(GANSU_E_ARG, "normalization must be 0 or 1");
  GbsLibrary lib = read_gbs(gbs);
  Molecule m;
  m.atoms = atoms;
  m.gbs_path = gbs;
  size_t basis = 0;
  for (size_t ai = 0; ai < atoms.size(); ++ai) {
    const Atom& at = atoms[ai];
    m.atom_basis.push_back((int)basis);
    if (at.Z < 1 || at.Z > kNumElements) throw Error(GANSU_E_ARG, "bad atomic number");
    auto it = lib.find(kElements[at.Z - 1]);
    if (it == lib.end()) throw Error(GANSU_E_BASIS, std::string("Element not in basis set: ") + kElements[at.Z - 1]);
    for (const GbsShell& s : it->second) {
      for (size_t k = 0; k < s.exps.size(); ++k) {
        gansu_primitive_shell p;
        std::memset(&p, 0, sizeof p);
        p.exponent = s.exps[k];
        p.coefficient = s.coefs[k];
        p.x = at.xyz[0]; p.y = at.xyz[1]; p.z = at.xyz[2];
        p.shell_type = s.L;
        p.basis_index = basis;
        p.atom_index = (int)ai;
        m.prims.push_back(p);
      }
      const double nf = contracted_norm(s.L, s.exps, s.coefs, normalization);
      for (int c = 0; c < ncart(s.L); ++c) m.norm.push_back(nf);
      basis += ncart(s.L);
    }
  }
  m.nbasis = (int)basis;
  m.atom_basis.push_back((int)basis);
  // Same comparator and algorithm as molecular.hpp:173 (std::sort, not stable):
  // under libstdc++ the resulting order is the reference's order.
  std::sort(m.prims.begin(), m.prims.end(),
            [](const gansu_primitive_shell& a, const gansu_primitive_shell& b) { return a.shell_type < b.shell_type; });
  const int maxl = m.prims.back().shell_type;
  m.types.assign(maxl + 1, gansu_shell_type_info{0, 0});
  for (const auto& p : m.prims) m.types[p.shell_type].count++;
  for (int l = 1; l <= maxl; ++l) m.types[l].start_index = m.types[l - 1].start_index + m.types[l - 1].count;

  int ne = -charge;
  for (const auto& a : atoms) ne += a.Z;
  if (ne < 1) throw Error(GANSU_E_ELECTRONS, "The number of electrons is less than one.");
  m.n_electrons = ne;
  m.n_alpha = (ne + 1) / 2 + beta_to_alpha;
  m.n_beta = ne / 2 - beta_to_alpha;
  if (m.n_beta < 0) throw Error(GANSU_E_ELECTRONS, "The number of beta-spin electrons is less than zero.");
  return m;
}

namespace {

// One element's block of a .sad file: a header line "<symbol> <n>" followed by
// the n x n alpha and then the n x n beta density, row-major, whitespace-separated.
bool read_sad_element(const std::string& path, const std::string& symbol, int& n, std::vector<double>& a,
                      std::vector<double>& b) {
  std::ifstream in(path);
  if (!in) throw Error(GANSU_E_IO, "Failed to open SAD file: " + path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream hs(line);
    std::string tag;
    int cnt = 0;
    if (!(hs >> tag >> cnt) || cnt <= 0) continue;
    if (tag != symbol) {
      for (int i = 0; i < 2 * cnt; ++i) std::getline(in, line);
      continue;
    }
    n = cnt;
    a.resize((size_t)n * n);
    b.resize((size_t)n * n);
    for (double& v : a)
      if (!(in >> v)) throw Error(GANSU_E_IO, "truncated SAD block for " + symbol + " in " + path);
    for (double& v : b)
      if (!(in >> v)) throw Error(GANSU_E_IO, "truncated SAD block for " + symbol + " in " + path);
    return true;
  }
  return false;
}

}  // namespace

void sad_density(const Molecule& m, std::vector<double>& Da, std::vector<double>& Db) {
  const size_t M = (size_t)m.nbasis;
  if (m.atom_basis.size() != m.atoms.size() + 1) throw Error(GANSU_E_STATE, "molecule has no per-atom basis ranges");
  // the .sad file sits beside the basis file, with the extension replaced
  std::string path = m.gbs_path;
  const size_t slash = path.find_last_of('/');
  const size_t dot = path.find_last_of('.');
  if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) path.resize(dot);
  path += ".sad";
  Da.assign(M * M, 0.0);
  Db.assign(M * M, 0.0);
  struct Block { int n; std::vector<double> a, b; };
  std::map<int, Block> cache;  // per element
  for (size_t i = 0; i < m.atoms.size(); ++i) {
    const int Z = m.atoms[i].Z;
    auto it = cache.find(Z);
    if (it == cache.end()) {
      Block blk{0, {}, {}};
      if (!read_sad_element(path, kElements[Z - 1], blk.n, blk.a, blk.b))
        throw Error(GANSU_E_BASIS, std::string("Element not found in SAD file: ") + kElements[Z - 1]);
      it = cache.emplace(Z, std::move(blk)).first;
    }
    const Block& blk = it->second;
    const int start = m.atom_basis[i], nat = m.atom_basis[i + 1] - start;
    if (blk.n != nat)
      throw Error(GANSU_E_BASIS, std::string("SAD block of ") + kElements[Z - 1] + " has " + std::to_string(blk.n) +
                                     " functions, the basis gives the atom " + std::to_string(nat));
    for (int p = 0; p < nat; ++p)
      for (int q = 0; q < nat; ++q) {
        Da[(size_t)(start + p) * M + start + q] = blk.a[(size_t)p * nat + q];
        Db[(size_t)(start + p) * M + start + q] = blk.b[(size_t)p * nat + q];
      }
  }
}

void write_sad_cache(const std::string& path, int M, const std::vector<double>& Da, const std::vector<double>& Db) {
  std::ofstream out(path);
  if (!out) throw Error(GANSU_E_IO, "Failed to open the SAD cache file: " + path);
  out << M << '\n';
  char buf[32];
  for (const std::vector<double>* mat : {&Da, &Db})
    for (int i = 0; i < M; ++i) {
      for (int j = 0; j < M; ++j) {
        std::snprintf(buf, sizeof buf, "%.10e ", (*mat)[(size_t)i * M + j]);
        out << buf;
      }
      out << '\n';
    }
  if (!out) throw Error(GANSU_E_IO, "failed writing " + path);
}

std::vector<CShell> contracted_shells(const gansu_primitive_shell* prims, size_t n, int max_l) {
  std::map<size_t, size_t> by_basis;  // basis index -> shell slot
  std::vector<CShell> shells;
  for (size_t i = 0; i < n; ++i) {
    const auto& p = prims[i];
    if (p.shell_type < 0 || p.shell_type > max_l)
      throw Error(GANSU_E_UNSUPPORTED, "angular momentum " + std::to_string(p.shell_type) + " beyond compiled classes");
    auto it = by_basis.find(p.basis_index);
    if (it == by_basis.end()) {
      CShell s;
      s.L = p.shell_type; s.atom = p.atom_index; s.basis = (int)p.basis_index;
      s.xyz[0] = p.x; s.xyz[1] = p.y; s.xyz[2] = p.z;
      by_basis[p.basis_index] = shells.size();
      shells.push_back(s);
      shells.back().prims.push_back((int)i);
    } else {
      CShell& s = shells[it->second];
      if (s.L != p.shell_type || s.xyz[0] != p.x || s.xyz[1] != p.y || s.xyz[2] != p.z)
        throw Error(GANSU_E_ARG, "primitives sharing a basis_index differ in type or centre");
      s.prims.push_back((int)i);
    }
  }
  // order by L, then basis index (deterministic)
  std::stable_sort(shells.begin(), shells.end(), [](const CShell& a, const CShell& b) {
    return a.L != b.L ? a.L < b.L : a.basis < b.basis;
  });
  return shells;
}

// ---------------------------------------------------------------- params --
Params::Params() {
  // parameter_manager.cpp:42-66 defaults, plus this engine's keys.
  kv = {{"parameter_file", ""},
        {"xyzfilename", ""},
        {"gbsfilename", ""},
        {"verbose", "0"},
        {"method", "rhf"},
        {"charge", "0"},
        {"beta_to_alpha", "0"},
        {"eri_method", "stored"},
        {"auxiliary_gbsfilename", ""},
        {"maxiter", "100"},
        {"convergence_energy_threshold", "1e-6"},
        {"schwarz_screening_threshold", "1e-12"},
        {"initial_guess", "core"},
        {"convergence_method", "diis"},
        {"damping_factor", "0.9"},
        {"diis_size", "8"},
        {"diis_include_transform", "0"},
        {"rohf_parameter_name", "roothaan"},
        {"geometry_optimization", "0"},
        {"geometry_optimization_method", "powell"},
        {"export_sad_cache", "0"},
        // engine keys (no reference counterpart)
        {"precision", "fp64"},
        {"fp32_bound_threshold", "1e-6"},
        {"normalization", "reference"},
        {"ri_k", "density"},
        {"scf_solver", "diagonalize"},
        {"incremental_fock", "0"},
        {"incremental_threshold", "1e-14"}};
}

static std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

void Params::set(const std::string& key_in, const std::string& value_in, bool /*from_recipe*/) {
  std::string key = lower(trim(key_in));
  std::string value = trim(value_in);
  if (value == "true" || value == "True" || value == "TRUE") value = "1";
  if (value == "false" || value == "False" || value == "FALSE") value = "0";
  // The reference accepts any key (is_valid_key compares against the wrong
  // map, parameter_manager.cpp:92-95); unknown keys are rejected here.
  if (!kv.count(key)) throw Error(GANSU_E_PARAM, "Unknown parameter: " + key);
  const bool is_path = key == "xyzfilename" || key == "gbsfilename" || key == "auxiliary_gbsfilename" ||
                       key == "parameter_file";
  kv[key] = is_path ? value : lower(value);
}

void Params::load_recipe(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error(GANSU_E_IO, "Failed to open parameter file: " + path);
  for (std::string line; std::getline(in, line);) {
    size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (trim(line).empty()) continue;
    size_t eq = line.find('=');
    if (eq == std::string::npos) throw Error(GANSU_E_PARAM, "Invalid line format: " + line);
    set(line.substr(0, eq), line.substr(eq + 1), true);
  }
}

std::string Params::get(const std::string& key) const {
  auto it = kv.find(lower(key));
  if (it == kv.end()) throw Error(GANSU_E_PARAM, "Unknown parameter: " + key);
  return it->second;
}
double Params::getd(const std::string& key) const {
  try { return std::stod(get(key)); } catch (const Error&) { throw; } catch (...) {
    throw Error(GANSU_E_PARAM, "bad numeric value for " + key);
  }
}
int Params::geti(const std::string& key) const {
  try { return std::stoi(get(key)); } catch (const Error&) { throw; } catch (...) {
    throw Error(GANSU_E_PARAM, "bad integer value for " + key);
  }
}

std::map<std::string, std::string> parse_options(const char* s) {
  std::map<std::string, std::string> out;
  if (!s) return out;
  std::string str(s);
  size_t pos = 0;
  while (pos <= str.size()) {
    size_t semi = str.find(';', pos);
    std::string item = str.substr(pos, semi == std::string::npos ? std::string::npos : semi - pos);
    item = trim(item);
    if (!item.empty()) {
      size_t eq = item.find('=');
      if (eq == std::string::npos) throw Error(GANSU_E_PARAM, "option without '=': " + item);
      out[lower(trim(item.substr(0, eq)))] = lower(trim(item.substr(eq + 1)));
    }
    if (semi == std::string::npos) break;
    pos = semi + 1;
  }
  return out;
}

// ------------------------------------------------------------ Boys grid --
// F_n(x_i), x_i = i/32, from the positive series; exp(-x) applied last.  The
// reference ships the same grid rounded to 15 decimals (boys_30720.h); this
// one keeps full double precision and two extra rows for the 7-term Taylor
// step used by the kernels.
std::vector<double> boys_grid_xn() {
  std::vector<double> g((size_t)kBoysGrid * kBoysRows);
  for (int i = 0; i < kBoysGrid; ++i) {
    const double x = i / 32.0;
    for (int n = 0; n < kBoysRows; ++n) {
      double term = 1.0 / (2 * n + 1), sum = term;
      for (int k = 1; k < 4000; ++k) {
        term *= 2.0 * x / (2 * n + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-19) break;
      }
      g[(size_t)i * kBoysRows + n] = std::exp(-x) * sum;
    }
  }
  return g;
}

}  // namespace gansu
