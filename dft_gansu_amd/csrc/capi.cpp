// extern "C" boundary (include/gansu_fock.h): every entry point converts C++
// exceptions into an error code plus a thread-local message.
#include <cstring>
#include <string>

#include "engine.hpp"
#include "gansu_fock.h"
#include "scf.hpp"

namespace gansu {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace gansu

struct gansu_molecule { gansu::Molecule m; };
struct gansu_params { gansu::Params p; };
struct gansu_fock_engine { std::unique_ptr<gansu::FockEngine> e; };

namespace {
template <class F>
int guarded(F&& f) {
  try {
    f();
    return GANSU_OK;
  } catch (const gansu::Error& e) {
    gansu::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    gansu::set_last_error(e.what());
    return GANSU_E_ARG;
  } catch (...) {
    gansu::set_last_error("unknown error");
    return GANSU_E_ARG;
  }
}
}  // namespace

extern "C" {

const char* gansu_last_error(void) { return gansu::g_last_error.c_str(); }
const char* gansu_version(void) { return "gansu-mi355x 0.1 (gfx950)"; }

int gansu_molecule_from_xyz(const char* xyz_path, const char* gbs_path, int charge, int beta_to_alpha,
                            int normalization, gansu_molecule** out) {
  return guarded([&] {
    if (!xyz_path || !gbs_path || !out) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto atoms = gansu::read_xyz(xyz_path);
    *out = new gansu_molecule{gansu::build_molecule(atoms, gbs_path, charge, beta_to_alpha, normalization)};
  });
}

int gansu_molecule_from_atoms(int n_atoms, const int* Z, const double* xyz_bohr, const char* gbs_path, int charge,
                              int beta_to_alpha, int normalization, gansu_molecule** out) {
  return guarded([&] {
    if (n_atoms <= 0 || !Z || !xyz_bohr || !gbs_path || !out) throw gansu::Error(GANSU_E_ARG, "null argument");
    std::vector<gansu::Atom> atoms(n_atoms);
    for (int i = 0; i < n_atoms; ++i) {
      atoms[i].Z = Z[i];
      for (int k = 0; k < 3; ++k) atoms[i].xyz[k] = xyz_bohr[3 * i + k];
    }
    *out = new gansu_molecule{gansu::build_molecule(atoms, gbs_path, charge, beta_to_alpha, normalization)};
  });
}

int gansu_molecule_sizes(const gansu_molecule* mol, int* sizes) {
  return guarded([&] {
    if (!mol || !sizes) throw gansu::Error(GANSU_E_ARG, "null argument");
    const auto& m = mol->m;
    sizes[0] = (int)m.prims.size(); sizes[1] = m.nbasis; sizes[2] = (int)m.types.size();
    sizes[3] = (int)m.atoms.size(); sizes[4] = m.n_electrons; sizes[5] = m.n_alpha; sizes[6] = m.n_beta;
  });
}

int gansu_molecule_get(const gansu_molecule* mol, gansu_primitive_shell* prims, gansu_shell_type_info* types,
                       double* norm, int* Z, double* xyz_bohr) {
  return guarded([&] {
    if (!mol) throw gansu::Error(GANSU_E_ARG, "null molecule");
    const auto& m = mol->m;
    if (prims) std::memcpy(prims, m.prims.data(), m.prims.size() * sizeof(gansu_primitive_shell));
    if (types) std::memcpy(types, m.types.data(), m.types.size() * sizeof(gansu_shell_type_info));
    if (norm) std::memcpy(norm, m.norm.data(), m.norm.size() * sizeof(double));
    for (size_t i = 0; i < m.atoms.size(); ++i) {
      if (Z) Z[i] = m.atoms[i].Z;
      if (xyz_bohr) for (int k = 0; k < 3; ++k) xyz_bohr[3 * i + k] = m.atoms[i].xyz[k];
    }
  });
}

double gansu_molecule_nuclear_repulsion(const gansu_molecule* mol) { return mol ? mol->m.nuclear_repulsion() : 0.0; }
void gansu_molecule_destroy(gansu_molecule* mol) { delete mol; }

int gansu_sad_density(const gansu_molecule* mol, double* Da, double* Db) {
  return guarded([&] {
    if (!mol || !Da || !Db) throw gansu::Error(GANSU_E_ARG, "null argument");
    std::vector<double> a, b;
    gansu::sad_density(mol->m, a, b);
    std::memcpy(Da, a.data(), a.size() * sizeof(double));
    std::memcpy(Db, b.data(), b.size() * sizeof(double));
  });
}

int gansu_params_create(gansu_params** out) {
  return guarded([&] {
    if (!out) throw gansu::Error(GANSU_E_ARG, "null argument");
    *out = new gansu_params{};
  });
}
int gansu_params_load_recipe(gansu_params* p, const char* path) {
  return guarded([&] {
    if (!p || !path) throw gansu::Error(GANSU_E_ARG, "null argument");
    p->p.load_recipe(path);
  });
}
int gansu_params_set(gansu_params* p, const char* key, const char* value) {
  return guarded([&] {
    if (!p || !key || !value) throw gansu::Error(GANSU_E_ARG, "null argument");
    p->p.set(key, value, false);
  });
}
int gansu_params_get(const gansu_params* p, const char* key, char* buf, int buflen) {
  return guarded([&] {
    if (!p || !key || !buf || buflen <= 0) throw gansu::Error(GANSU_E_ARG, "null argument");
    const std::string v = p->p.get(key);
    if ((int)v.size() + 1 > buflen) throw gansu::Error(GANSU_E_ARG, "buffer too small");
    std::memcpy(buf, v.c_str(), v.size() + 1);
  });
}
void gansu_params_destroy(gansu_params* p) { delete p; }

int gansu_fock_create(const gansu_primitive_shell* prims, size_t n_prims, const gansu_shell_type_info* types,
                      int n_types, const double* cgto_norm, int num_basis, const gansu_primitive_shell* aux,
                      size_t n_aux, const gansu_shell_type_info* aux_types, int n_aux_types, const double* aux_norm,
                      int num_aux, const char* options, gansu_fock_engine** out) {
  return guarded([&] {
    if (!out) throw gansu::Error(GANSU_E_ARG, "null output");
    auto opts = gansu::parse_options(options);
    auto e = std::make_unique<gansu::FockEngine>(prims, n_prims, types, n_types, cgto_norm, num_basis, aux, n_aux,
                                                 aux_types, n_aux_types, aux_norm, num_aux, opts);
    *out = new gansu_fock_engine{std::move(e)};
  });
}

int gansu_fock_precompute(gansu_fock_engine* e) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->precompute();
  });
}

int gansu_fock_build_rhf(gansu_fock_engine* e, const double* d_D, const double* d_H, double* d_F, void* stream) {
  return guarded([&] {
    if (!e || !d_D || !d_H || !d_F) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    e->e->build_jk(d_D, nullptr, nullptr, d_H, d_F, (hipStream_t)stream);
  });
}

int gansu_fock_build_jk(gansu_fock_engine* e, const double* d_D, double* d_J, double* d_K, void* stream) {
  return guarded([&] {
    if (!e || !d_D || !d_J || !d_K) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    e->e->build_jk(d_D, d_J, d_K, nullptr, nullptr, (hipStream_t)stream);
  });
}

int gansu_fock_build_rhf_occ(gansu_fock_engine* e, const double* d_D, const double* d_C, int nocc, const double* d_H,
                             double* d_F, void* stream) {
  return guarded([&] {
    if (!e || !d_D || !d_C || !d_H || !d_F) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    e->e->build_jk(d_D, nullptr, nullptr, d_H, d_F, (hipStream_t)stream, d_C, nocc);
  });
}

int gansu_fock_build_uhf(gansu_fock_engine* e, const double* d_Da, const double* d_Db, const double* d_H, double* d_Fa,
                         double* d_Fb, void* stream) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->build_uhf(d_Da, d_Db, d_H, d_Fa, d_Fb, (hipStream_t)stream);
  });
}

int gansu_fock_build_rohf(gansu_fock_engine* e, const double* d_Dc, const double* d_Do, const double* d_H,
                          const double* d_C, const double* d_S, const char* coupling, int num_closed, int num_open,
                          double* d_Fc, double* d_Fo, double* d_F, void* stream) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    const std::string name = coupling ? coupling : "roothaan";
    gansu::RohfCoupling cp{};
    if (!gansu::rohf_coupling(name, &cp)) throw gansu::Error(GANSU_E_PARAM, "unknown ROHF coupling set: " + name);
    gansu::DeviceScope ds(e->e->device());
    e->e->build_rohf(d_Dc, d_Do, d_H, d_C, d_S, cp, num_closed, num_open, d_Fc, d_Fo, d_F, (hipStream_t)stream);
  });
}

int gansu_fock_update_geometry(gansu_fock_engine* e, const double* xyz_bohr, int n_atoms) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->update_geometry(xyz_bohr, n_atoms);
  });
}

const char* gansu_fock_algorithm_name(const gansu_fock_engine* e) { return e ? e->e->algorithm_name() : ""; }

int gansu_fock_schwarz(gansu_fock_engine* e, double* host_Q, size_t n) {
  return guarded([&] {
    if (!e || !host_Q) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    auto q = e->e->schwarz_ref();
    if (n != q.size()) throw gansu::Error(GANSU_E_ARG, "size mismatch: engine has " + std::to_string(q.size()));
    std::memcpy(host_Q, q.data(), n * sizeof(double));
  });
}

size_t gansu_fock_num_primitive_pairs(const gansu_fock_engine* e) { return e ? e->e->num_primitive_pairs() : 0; }

int gansu_fock_stats(const gansu_fock_engine* e, double* stats) {
  return guarded([&] {
    if (!e || !stats) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    e->e->stats(stats);
  });
}

int gansu_fock_class_stats(gansu_fock_engine* e, int max_classes, int* n_classes, int* lab, double* n_contracted,
                           double* n_primitive) {
  return guarded([&] {
    if (!e || !n_classes) throw gansu::Error(GANSU_E_ARG, "null argument");
    if (max_classes > 0 && (!lab || !n_contracted || !n_primitive)) throw gansu::Error(GANSU_E_ARG, "null output");
    gansu::DeviceScope ds(e->e->device());
    const auto& cc = e->e->class_counts();
    *n_classes = (int)cc.size();
    for (int c = 0; c < (int)cc.size() && c < max_classes; ++c) {
      for (int k = 0; k < 4; ++k) lab[4 * c + k] = cc[c].L[k];
      n_contracted[c] = cc[c].contracted;
      n_primitive[c] = cc[c].primitive;
    }
  });
}

int gansu_fock_set_profiling(gansu_fock_engine* e, int enable) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->set_profiling(enable != 0);
  });
}

int gansu_fock_last_profile(gansu_fock_engine* e, double* rows, int max_rows, int* n_rows) {
  return guarded([&] {
    if (!e || !n_rows) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::DeviceScope ds(e->e->device());
    auto v = e->e->last_profile();
    const int n = (int)(v.size() / 5);
    *n_rows = n;
    if (rows) std::memcpy(rows, v.data(), sizeof(double) * 5 * std::min(n, max_rows));
  });
}

void gansu_fock_destroy(gansu_fock_engine* e) {
  if (!e) return;
  gansu::DeviceScope ds(e->e->device());  // buffers and streams live on the engine's device
  delete e;
}

int gansu_oneint(const gansu_molecule* mol, double* d_S, double* d_T, double* d_V, void* stream) {
  return guarded([&] {
    if (!mol || !d_S || !d_T || !d_V) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::compute_oneint(mol->m, d_S, d_T, d_V, (hipStream_t)stream);
  });
}

int gansu_oneint_noalloc(const gansu_molecule* mol, double* d_S, double* d_T, double* d_V, void* stream) {
  return guarded([&] {
    if (!mol || !d_S || !d_T || !d_V) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::compute_oneint_noalloc(mol->m, d_S, d_T, d_V, (hipStream_t)stream);
  });
}

int gansu_rhf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                  double* D_out, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto r = gansu::rhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (D_out) std::memcpy(D_out, r.D.data(), r.D.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = r.graph_launches; }
  });
}

int gansu_rhf_gradient_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                             const double* Da0, const double* Db0, int force_density, double* energies, double* grad,
                             double* D_out, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies || !grad) throw gansu::Error(GANSU_E_ARG, "null argument");
    if ((Da0 == nullptr) != (Db0 == nullptr)) throw gansu::Error(GANSU_E_ARG, "give both densities or neither");
    auto r = gansu::rhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p, true, Da0, Db0, force_density != 0);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    std::memcpy(grad, r.grad.data(), r.grad.size() * sizeof(double));
    if (D_out) std::memcpy(D_out, r.D.data(), r.D.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = r.graph_launches; }
  });
}

int gansu_rhf_run_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                        const double* Da0, const double* Db0, int force_density, double* energies, double* D_out,
                        int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    if ((Da0 == nullptr) != (Db0 == nullptr)) throw gansu::Error(GANSU_E_ARG, "give both densities or neither");
    auto r = gansu::rhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p, false, Da0, Db0, force_density != 0);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (D_out) std::memcpy(D_out, r.D.data(), r.D.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = r.graph_launches; }
  });
}

int gansu_fock_gradient_rhf(gansu_fock_engine* e, const double* d_D, int n_atoms, double* d_grad, void* stream) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->gradient_2e(d_D, n_atoms, d_grad, (hipStream_t)stream);
  });
}

int gansu_fock_gradient_rhf_occ(gansu_fock_engine* e, const double* d_D, const double* d_C, int nocc, int n_atoms,
                                double* d_grad, void* stream) {
  return guarded([&] {
    if (!e) throw gansu::Error(GANSU_E_ARG, "null engine");
    gansu::DeviceScope ds(e->e->device());
    e->e->gradient_2e(d_D, n_atoms, d_grad, (hipStream_t)stream, d_C, nocc);
  });
}

int gansu_oneint_gradient(const gansu_molecule* mol, const double* d_D, const double* d_W, double* d_grad,
                          void* stream) {
  return guarded([&] {
    if (!mol || !d_D || !d_W || !d_grad) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::compute_oneint_gradient(mol->m, d_D, d_W, d_grad, (hipStream_t)stream);
  });
}

int gansu_molecule_nuclear_repulsion_gradient(const gansu_molecule* mol, double* grad) {
  return guarded([&] {
    if (!mol || !grad) throw gansu::Error(GANSU_E_ARG, "null argument");
    const auto g = gansu::nuclear_repulsion_gradient(mol->m);
    std::memcpy(grad, g.data(), g.size() * sizeof(double));
  });
}

int gansu_rhf_run_mo(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                     double* energies, double* C_out, double* eps_out, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto r = gansu::rhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (C_out) std::memcpy(C_out, r.C.data(), r.C.size() * sizeof(double));
    if (eps_out) std::memcpy(eps_out, r.eps.data(), r.eps.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = r.graph_launches; }
  });
}

int gansu_write_molden(const gansu_molecule* mol, const double* C, const double* eps, const double* occ, int n_mo,
                       const char* path) {
  return guarded([&] {
    if (!mol || !C || !eps || !occ || !path) throw gansu::Error(GANSU_E_ARG, "null argument");
    gansu::write_molden(mol->m, C, eps, occ, n_mo, path);
  });
}

int gansu_uhf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                  double* Da, double* Db, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto r = gansu::uhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (Da) std::memcpy(Da, r.Da.data(), r.Da.size() * sizeof(double));
    if (Db) std::memcpy(Db, r.Db.data(), r.Db.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = 0; }
  });
}

int gansu_uhf_run_guess(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                        const double* Da0, const double* Db0, int force_density, double* energies, double* Da,
                        double* Db, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    if ((Da0 == nullptr) != (Db0 == nullptr)) throw gansu::Error(GANSU_E_ARG, "give both densities or neither");
    auto r = gansu::uhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p, Da0, Db0, force_density != 0);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (Da) std::memcpy(Da, r.Da.data(), r.Da.size() * sizeof(double));
    if (Db) std::memcpy(Db, r.Db.data(), r.Db.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = 0; }
  });
}

int gansu_rohf_run(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p, double* energies,
                   double* Dc, double* Do, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto r = gansu::rohf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    if (Dc) std::memcpy(Dc, r.Da.data(), r.Da.size() * sizeof(double));
    if (Do) std::memcpy(Do, r.Db.data(), r.Db.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = 0; }
  });
}

int gansu_export_sad_cache(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                           const char* path) {
  return guarded([&] {
    if (!mol || !p || !path) throw gansu::Error(GANSU_E_ARG, "null argument");
    const std::string method = p->p.get("method");
    const gansu::Molecule* aux = aux_mol ? &aux_mol->m : nullptr;
    const size_t M = (size_t)mol->m.nbasis, nn = M * M;
    std::vector<double> a(nn), b(nn);
    // the per-spin densities of each method's export_density_matrix
    // (rhf.cu:312-327, uhf.cu:331-343, rohf.cu:288-310)
    if (method == "rhf") {
      auto r = gansu::rhf_run(mol->m, aux, p->p);
      for (size_t i = 0; i < nn; ++i) a[i] = b[i] = 0.5 * r.D[i];
    } else if (method == "uhf") {
      auto r = gansu::uhf_run(mol->m, aux, p->p);
      a = r.Da;
      b = r.Db;
    } else if (method == "rohf") {
      auto r = gansu::rohf_run(mol->m, aux, p->p);
      for (size_t i = 0; i < nn; ++i) {
        a[i] = 0.5 * r.Da[i] + r.Db[i];
        b[i] = 0.5 * r.Da[i];
      }
    } else {
      throw gansu::Error(GANSU_E_PARAM, "Unknown method: " + method);
    }
    gansu::write_sad_cache(path, (int)M, a, b);
  });
}

int gansu_rhf_gradient(const gansu_molecule* mol, const gansu_molecule* aux_mol, const gansu_params* p,
                       double* energies, double* grad, int* info) {
  return guarded([&] {
    if (!mol || !p || !energies || !grad) throw gansu::Error(GANSU_E_ARG, "null argument");
    auto r = gansu::rhf_run(mol->m, aux_mol ? &aux_mol->m : nullptr, p->p, true);
    energies[0] = r.e_elec;
    energies[1] = r.e_total;
    std::memcpy(grad, r.grad.data(), r.grad.size() * sizeof(double));
    if (info) { info[0] = r.iterations; info[1] = r.fock_builds; info[2] = r.converged; info[3] = r.graph_launches; }
  });
}

}  // extern "C"
