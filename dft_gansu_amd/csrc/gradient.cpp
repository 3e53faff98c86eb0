// Two-electron part of the RHF nuclear gradient on the engine's four-centre
// work lists (the direct Fock build's class-sorted blocks and Schwarz bounds).
#include <algorithm>
#include <string>

#include "engine.hpp"
#include "gradient_kernels.hpp"

namespace gansu {

void FockEngine::gradient_2e(const double* d_D, int natoms, double* d_grad, hipStream_t s, const double* d_C,
                             int nocc) {
  if (!precomputed_) throw Error(GANSU_E_STATE, "gradient before precompute");
  if (!d_D || !d_grad) throw Error(GANSU_E_ARG, "null device pointer");
  int max_atom = -1;
  for (const auto& sh : shells_) max_atom = std::max(max_atom, sh.atom);
  if (natoms <= max_atom || natoms <= 0) throw Error(GANSU_E_ARG, "n_atoms smaller than the atoms the basis sits on");
  if (natoms > kGradMaxAtoms)
    throw Error(GANSU_E_UNSUPPORTED, "gradient limited to " + std::to_string(kGradMaxAtoms) + " atoms");
  if (method_ == "ri") {
    gradient_ri(d_D, d_C, nocc, natoms, d_grad, s);
    return;
  }
  if (!d_alpha_.ptr) d_alpha_.upload(pair_alpha_h_);
  if (!d_shell_atom_.ptr) {
    std::vector<int> sa(shells_.size());
    for (size_t i = 0; i < shells_.size(); ++i) sa[i] = shells_[i].atom;
    d_shell_atom_.upload(sa);
  }
  prof_clear();
  GANSU_HIP(launch_zero(d_grad, 3 * (size_t)natoms, s));
  GANSU_HIP(launch_scale_density(d_D, d_w_.ptr, M_, d_Dp_.ptr, s));
  const Grad2eArgs g{d_sp_.ptr, d_pr_.ptr, d_alpha_.ptr, boys_.ptr, thr_, d_Dp_.ptr, M_,
                     d_shell_atom_.ptr, natoms, d_grad};
  // Classes run concurrently on the build stream and the three side streams
  // (forked and joined with events).  The runtime-indexed classes, whose
  // private segment reaches ~9 KB per lane, are kept to the build stream and
  // the first side stream, so that at most two hardware queues hold such
  // scratch; the register-resident classes use the other two.
  const bool concurrent = !profiling_ && ev_fork_ != nullptr;
  if (concurrent) {
    GANSU_HIP(hipEventRecord(ev_fork_, s));
    for (int i = 0; i < kSideStreams; ++i) GANSU_HIP(hipStreamWaitEvent(side_[i], ev_fork_, 0));
  }
  struct Item { QuartetClass* qc; double cost; bool heavy; };
  std::vector<Item> items;
  for (auto& qc : qclass_) {
    if (qc->blocks_h.empty()) continue;
    const double ncomp = double(ncart(qc->LA) * ncart(qc->LB) * ncart(qc->LC) * ncart(qc->LD));
    items.push_back({qc.get(), qc->n_quartets * ncomp * (1 + qc->LA + qc->LB + qc->LC + qc->LD),
                     grad_2e_uses_scratch(qc->LA, qc->LB, qc->LC, qc->LD)});
  }
  std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
  double load[4] = {0, 0, 0, 0};
  for (const Item& it : items) {
    QuartetClass* qc = it.qc;
    int k = 0;
    if (concurrent) {
      const int lo = it.heavy ? 0 : 2, hi = it.heavy ? 2 : 4;
      k = lo;
      for (int j = lo; j < hi; ++j)
        if (load[j] < load[k]) k = j;
      load[k] += it.cost;
    }
    hipStream_t q = k == 0 ? s : side_[k - 1];
    prof_begin(q, 400000.0 + 1000 * qc->LA + 100 * qc->LB + 10 * qc->LC + qc->LD, 0.0, 0.0, qc->n_quartets);
    const ClassLaunch c{qc->LA, qc->LB, qc->LC, qc->LD, qc->blocks.ptr, (int)qc->blocks_h.size(),
                        pclass_[qc->A].list.ptr, pclass_[qc->B].list.ptr, nullptr};
    int maxnp = 1, maxnk = 1;
    for (const auto& br : qc->bras_h) maxnp = std::max(maxnp, sp_h_[pclass_[qc->A].sorted[br.bra_pos]].nprim);
    for (int kk = 0; kk < qc->max_count; ++kk) maxnk = std::max(maxnk, sp_h_[pclass_[qc->B].sorted[kk]].nprim);
    GANSU_HIP(launch_grad_2e(c, g, maxnp, maxnk, q));
    prof_end(q);
  }
  if (concurrent)
    for (int i = 0; i < kSideStreams; ++i) {
      GANSU_HIP(hipEventRecord(ev_join_[i], side_[i]));
      GANSU_HIP(hipStreamWaitEvent(s, ev_join_[i], 0));
    }
}

}  // namespace gansu
