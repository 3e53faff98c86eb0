// Molden output of a molecule's orbitals, in the layout of the reference's
// RHF::export_molden_file (rhf.cu:363-425): [Atoms] in bohr, [GTO] per atom
// with one entry per contracted shell (its primitives' exponents and raw
// contraction coefficients), [MO] with energy, spin, occupation and the
// coefficients of every basis function (17 significant digits from the
// first coefficient on, as the reference's stream formatting leaves it).
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "gansu_internal.hpp"

namespace gansu {

std::string element_symbol(int Z);

void write_molden(const Molecule& mol, const double* C, const double* eps, const double* occ, int nmo,
                  const std::string& path) {
  const int M = mol.nbasis;
  if (nmo <= 0 || nmo > M) throw Error(GANSU_E_ARG, "n_mo out of range");
  std::ofstream out(path);
  if (!out) throw Error(GANSU_E_IO, "cannot open " + path);
  out << "[Molden Format]\n" << "generated by GANSU\n" << "[Atoms] (AU)\n";
  for (size_t a = 0; a < mol.atoms.size(); ++a)
    out << element_symbol(mol.atoms[a].Z) << " " << a + 1 << " " << mol.atoms[a].Z << " " << mol.atoms[a].xyz[0]
        << " " << mol.atoms[a].xyz[1] << " " << mol.atoms[a].xyz[2] << "\n";
  out << "[GTO]\n";
  // primitives by the first basis function of their shell
  std::vector<std::vector<const gansu_primitive_shell*>> by_basis(M);
  std::vector<int> atom_of(M, -1);
  for (const auto& p : mol.prims) {
    if (p.basis_index >= (size_t)M) throw Error(GANSU_E_ARG, "basis_index out of range");
    by_basis[p.basis_index].push_back(&p);
    atom_of[p.basis_index] = p.atom_index;
  }
  static const char* const kShell[] = {"s", "p", "d", "f", "g"};
  for (size_t a = 0; a < mol.atoms.size(); ++a) {
    out << a + 1 << " " << 0 << "\n";
    for (int j = 0; j < M; ++j) {
      if (by_basis[j].empty() || atom_of[j] != (int)a) continue;
      const int L = by_basis[j].front()->shell_type;
      out << " " << (L >= 0 && L < 5 ? kShell[L] : "?") << " " << by_basis[j].size() << " " << "1.00" << "\n";
      for (const auto* p : by_basis[j]) out << "\t" << p->exponent << " " << p->coefficient << "\n";
    }
    out << "\n";
  }
  out << "\n" << "[MO]\n";
  for (int i = 0; i < nmo; ++i) {
    out << "Sym= A\n" << "Ene= " << eps[i] << "\n" << "Spin= Alpha\n" << "Occup= " << occ[i] << "\n";
    for (int j = 0; j < M; ++j) out << " " << j + 1 << " " << std::setprecision(17) << C[(size_t)j * nmo + i] << "\n";
  }
  if (!out) throw Error(GANSU_E_IO, "write failed: " + path);
}

}  // namespace gansu
