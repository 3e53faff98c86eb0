// Fock engine host side: screening, work decomposition and build orchestration
// for the four-centre paths (stored / direct).  The RI path lives in ri.cpp.
//
// Reference counterparts: ERI_Stored::precomputation (eri.cu:60-116),
// computeSchwarzUpperBounds (gpu_manager.cu:2154-2179), makeShellPairTypeInfo
// (gpu_manager.cu:413-432), computeERIMatrix (gpu_manager.cu:434-517) and
// computeFockMatrix_RHF (gpu_manager.cu:678-691).
#include "engine.hpp"
#include "scf_kernels.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdlib>

namespace gansu {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(GANSU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// 1/sqrt((2l-1)!!(2m-1)!!(2n-1)!!): the component-dependent part of the
// primitive norm (int2e.hpp:127-130), folded into the raw-basis scaling.
double component_factor(int L, int c, bool aux) {
  int e[3];
  // orbital order (types.hpp:197-228); aux f uses the RI order (int2c2e.hpp:39-47)
  static const int d2[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  static const int f3o[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0},
                                 {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
  static const int f3r[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1},
                                 {0, 2, 1}, {1, 2, 0}, {0, 1, 2}, {1, 0, 2}, {1, 1, 1}};
  for (int d = 0; d < 3; ++d) {
    if (L == 0) e[d] = 0;
    else if (L == 1) e[d] = (c == d);
    else if (L == 2) e[d] = d2[c][d];
    else e[d] = aux ? f3r[c][d] : f3o[c][d];
  }
  return 1.0 / std::sqrt(dfact(2 * e[0] - 1) * dfact(2 * e[1] - 1) * dfact(2 * e[2] - 1));
}

int pair_class_id(int la, int lb) { return la * (la + 1) / 2 + lb; }

}  // namespace

FockEngine::FockEngine(const gansu_primitive_shell* prims, size_t n_prims, const gansu_shell_type_info* types,
                       int n_types, const double* norm, int num_basis, const gansu_primitive_shell* aux, size_t n_aux,
                       const gansu_shell_type_info* aux_types, int n_aux_types, const double* aux_norm, int num_aux,
                       const std::map<std::string, std::string>& opts) {
  if (!prims || n_prims == 0 || !types || n_types <= 0 || !norm || num_basis <= 0)
    throw Error(GANSU_E_ARG, "empty primitive list or basis");
  // the engine's device: the caller's current one (no device yet: 0, and the
  // first call that needs one reports the HIP error)
  if (hipGetDevice(&dev_) != hipSuccess) dev_ = 0;
  {
    const char* tr = std::getenv("GANSU_RI_TRACE");
    rik_trace_ = tr && tr[0] == '1';
  }
  method_ = "stored";
  for (const auto& kv : opts) {
    if (kv.first == "eri_method") method_ = kv.second;
    else if (kv.first == "schwarz_screening_threshold") thr_ = std::stod(kv.second);
    else if (kv.first == "precision") {
      if (kv.second != "fp64" && kv.second != "mixed") throw Error(GANSU_E_PARAM, "precision must be fp64|mixed");
      mixed_ = kv.second == "mixed";
    } else if (kv.first == "fp32_bound_threshold") fp32_thr_ = std::stod(kv.second);
    else if (kv.first == "autotune") autotune_ = kv.second != "0";
    else if (kv.first == "graph") use_graph_ = kv.second != "0";
    else if (kv.first == "nan_check") nan_check_ = kv.second != "0";
    else if (kv.first == "ri_refresh") {
      rik_refresh_ = std::stoi(kv.second);
      if (rik_refresh_ < 0) throw Error(GANSU_E_PARAM, "ri_refresh must be >= 0");
    } else if (kv.first == "ri_tol") {
      rik_tol_ = std::stod(kv.second);
      if (!(rik_tol_ >= 0.0)) throw Error(GANSU_E_PARAM, "ri_tol must be >= 0");
    }
    else if (kv.first == "incremental_fock") {
      incr_period_ = std::stoi(kv.second);
      if (incr_period_ < 0) throw Error(GANSU_E_PARAM, "incremental_fock must be >= 0");
    } else if (kv.first == "incremental_threshold") {
      incr_thr_ = std::stod(kv.second);
      if (!(incr_thr_ > 0.0)) throw Error(GANSU_E_PARAM, "incremental_threshold must be > 0");
    }
    else if (kv.first == "ri_k") {
      if (kv.second != "density" && kv.second != "occupied") throw Error(GANSU_E_PARAM, "ri_k must be density|occupied");
      ri_k_occ_ = kv.second == "occupied";
    }
    else throw Error(GANSU_E_PARAM, "Unknown engine option: " + kv.first);
  }
  if (method_ != "stored" && method_ != "direct" && method_ != "ri")
    throw Error(GANSU_E_PARAM, "Unknown eri_method: " + method_ + " (stored|direct|ri)");
  if (incr_period_ > 0 && method_ != "direct")
    throw Error(GANSU_E_PARAM, "incremental_fock needs eri_method=direct");
  prims_.assign(prims, prims + n_prims);
  types_.assign(types, types + n_types);
  norm_.assign(norm, norm + num_basis);
  M_ = num_basis;
  size_t tot = 0;
  for (int t = 0; t < n_types; ++t) {
    if ((size_t)types[t].start_index != tot) throw Error(GANSU_E_ARG, "shell type infos are not contiguous");
    tot += types[t].count;
    for (int k = 0; k < types[t].count; ++k)
      if (prims[types[t].start_index + k].shell_type != t) throw Error(GANSU_E_ARG, "primitives not sorted by type");
  }
  if (tot != n_prims) throw Error(GANSU_E_ARG, "shell type counts do not cover the primitive list");
  shells_ = contracted_shells(prims, n_prims, kMaxL);
  w_h_.assign(M_, 0.0);
  for (const auto& s : shells_) {
    if (s.basis + ncart(s.L) > M_) throw Error(GANSU_E_ARG, "basis_index out of range");
    for (int c = 0; c < ncart(s.L); ++c) w_h_[s.basis + c] = norm_[s.basis + c] * component_factor(s.L, c, false);
  }
  if (method_ == "ri") {
    if (!aux || n_aux == 0 || !aux_types || n_aux_types <= 0 || !aux_norm || num_aux <= 0)
      throw Error(GANSU_E_ARG, "eri_method=ri needs an auxiliary basis");
    aux_prims_.assign(aux, aux + n_aux);
    aux_types_.assign(aux_types, aux_types + n_aux_types);
    aux_norm_.assign(aux_norm, aux_norm + num_aux);
    Maux_ = num_aux;
  }
}

FockEngine::~FockEngine() {
  prof_clear();
  rohf_release();
  graph_release();
  if (cap_) (void)hipStreamDestroy(cap_);
  streams_destroy();
  ri_destroy(ri_);
}

void FockEngine::streams_create() {
  if (ev_fork_) return;
  for (int i = 0; i < kSideStreams; ++i) {
    GANSU_HIP(hipStreamCreateWithFlags(&side_[i], hipStreamNonBlocking));
    GANSU_HIP(hipEventCreateWithFlags(&ev_join_[i], hipEventDisableTiming));
  }
  GANSU_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
}

void FockEngine::streams_destroy() {
  for (int i = 0; i < kSideStreams; ++i) {
    if (side_[i]) (void)hipStreamDestroy(side_[i]);
    if (ev_join_[i]) (void)hipEventDestroy(ev_join_[i]);
    side_[i] = nullptr;
    ev_join_[i] = nullptr;
  }
  if (ev_fork_) (void)hipEventDestroy(ev_fork_);
  ev_fork_ = nullptr;
}

void FockEngine::set_profiling(bool on) {
  prof_clear();
  profiling_ = on;
}

int FockEngine::graph_captures() const { return n_graph_captures_; }

void FockEngine::prof_clear() {
  for (auto& r : prof_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  prof_.clear();
}

void FockEngine::prof_begin(hipStream_t s, double id, double bytes, double flops, double units) {
  if (!profiling_) return;
  ProfRow r{id, bytes, flops, units, nullptr, nullptr};
  GANSU_HIP(hipEventCreate(&r.a));
  GANSU_HIP(hipEventCreate(&r.b));
  GANSU_HIP(hipEventRecord(r.a, s));
  prof_.push_back(r);
}

void FockEngine::prof_end(hipStream_t s) {
  if (!profiling_) return;
  GANSU_HIP(hipEventRecord(prof_.back().b, s));
}

std::vector<double> FockEngine::last_profile() {
  std::vector<double> out;
  for (auto& r : prof_) {
    GANSU_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    GANSU_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    out.insert(out.end(), {r.id, (double)ms, r.bytes, r.flops, r.units});
  }
  return out;
}

const char* FockEngine::algorithm_name() const {
  if (method_ == "stored") return "Stored";
  if (method_ == "direct") return "Direct";
  return "RI";
}

void FockEngine::precompute() {
  graph_release();  // a captured build refers to the buffers rebuilt below
  if (!boys_.ptr) boys_.upload(boys_grid_xn());
  d_w_.upload(w_h_);
  d_Dp_.alloc((size_t)M_ * M_);
  d_Jr_.alloc((size_t)M_ * M_);
  d_Kr_.alloc((size_t)M_ * M_);
  precompute_pairs();
  if (method_ == "ri") precompute_ri();
  else precompute_quartets();
  if (incr_period_ > 0) {
    const int nsh = (int)shells_.size();
    std::vector<int> bf(nsh), nf(nsh);
    for (int i = 0; i < nsh; ++i) {
      bf[i] = shells_[i].basis;
      nf[i] = ncart(shells_[i].L);
    }
    d_shell_bf_.upload(bf);
    d_shell_nf_.upload(nf);
    d_dmax_.alloc((size_t)nsh * nsh + 1);
    d_Dprev_.alloc((size_t)M_ * M_);
    d_dDp_.alloc((size_t)M_ * M_);
    d_incr_st_.alloc(2);
    reset_incremental(0);
  }
  streams_create();
  GANSU_HIP(hipDeviceSynchronize());
  precomputed_ = true;
}

void FockEngine::precompute_pairs() {
  d_prims_.upload(prims_);
  const int nt = (int)types_.size();
  // reference primitive pair list (gpu_manager.cu:413-432 order)
  std::vector<int> pa, pb;
  std::vector<size_t> type_off((size_t)nt * nt, 0);
  for (int s0 = 0; s0 < nt; ++s0)
    for (int s1 = s0; s1 < nt; ++s1) {
      type_off[(size_t)s0 * nt + s1] = pa.size();
      const int c0 = types_[s0].count, c1 = types_[s1].count;
      const int b0 = (int)types_[s0].start_index, b1 = (int)types_[s1].start_index;
      if (s0 == s1) {
        for (int r2 = 0; r2 < c0; ++r2)
          for (int r1 = 0; r1 <= r2; ++r1) { pa.push_back(b0 + r1); pb.push_back(b0 + r2); }
      } else {
        for (int r1 = 0; r1 < c0; ++r1)
          for (int r2 = 0; r2 < c1; ++r2) { pa.push_back(b0 + r1); pb.push_back(b1 + r2); }
      }
    }
  ref_pairs_ = pa.size();
  {
    DevBuf<int> dpa, dpb;
    dpa.upload(pa);
    dpb.upload(pb);
    d_Qref_.alloc(ref_pairs_);
    GANSU_HIP(launch_schwarz_ref(d_prims_.ptr, dpa.ptr, dpb.ptr, ref_pairs_, d_Qref_.ptr, 0));
    GANSU_HIP(hipDeviceSynchronize());
  }
  Qref_h_ = d_Qref_.download();
  auto refpos = [&](int x, int y) -> size_t {
    int tx = prims_[x].shell_type, ty = prims_[y].shell_type;
    if (tx > ty || (tx == ty && x > y)) { std::swap(x, y); std::swap(tx, ty); }
    const size_t off = type_off[(size_t)tx * nt + ty];
    const int rx = x - (int)types_[tx].start_index, ry = y - (int)types_[ty].start_index;
    if (tx == ty) return off + (size_t)ry * (ry + 1) / 2 + rx;
    return off + (size_t)rx * types_[ty].count + ry;
  };
  double qglob = 0.0;
  for (double q : Qref_h_) qglob = std::max(qglob, q);

  // contracted shell pairs, L_i >= L_j (i >= j for equal L)
  std::vector<int> ppa, ppb;
  std::vector<double> pq;
  sp_h_.clear();
  const int ns = (int)shells_.size();
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j <= i; ++j) {
      // shells_ is sorted by L, so L_i >= L_j here
      const CShell &si = shells_[i], &sj = shells_[j];
      SPair sp;
      std::memset(&sp, 0, sizeof sp);
      sp.prim_begin = (int)ppa.size();
      sp.bi = si.basis; sp.bj = sj.basis; sp.si = i; sp.sj = j; sp.la = si.L; sp.lb = sj.L;
      double qmax = 0.0;
      for (int a : si.prims)
        for (int b : sj.prims) {
          const double q = Qref_h_[refpos(a, b)];
          ppa.push_back(a); ppb.push_back(b); pq.push_back(q);
          qmax = std::max(qmax, q);
        }
      sp.nprim = (int)ppa.size() - sp.prim_begin;
      sp.qmax = qmax;
      if (qmax * qglob < thr_) {  // no ket can rescue this pair
        ppa.resize(sp.prim_begin); ppb.resize(sp.prim_begin); pq.resize(sp.prim_begin);
        continue;
      }
      sp_h_.push_back(sp);
    }
  {
    DevBuf<int> dpa, dpb;
    DevBuf<double> dq;
    dpa.upload(ppa); dpb.upload(ppb); dq.upload(pq);
    d_pr_.alloc(ppa.size());
    GANSU_HIP(launch_pair_data(d_prims_.ptr, dpa.ptr, dpb.ptr, dq.ptr, ppa.size(), d_pr_.ptr, 0));
    GANSU_HIP(hipDeviceSynchronize());
  }
  pair_q_h_ = pq;
  class_counts_.clear();
  pair_alpha_h_.resize(ppa.size());
  for (size_t i = 0; i < ppa.size(); ++i) pair_alpha_h_[i] = prims_[ppa[i]].exponent;
  d_sp_.upload(sp_h_);
}

void FockEngine::precompute_quartets() {
  // pair classes sorted by qmax (descending; ties by id)
  const int npc = (kMaxL + 1) * (kMaxL + 2) / 2;
  pclass_.clear();
  pclass_.resize(npc);
  for (int la = 0; la <= kMaxL; ++la)
    for (int lb = 0; lb <= la; ++lb) { pclass_[pair_class_id(la, lb)].la = la; pclass_[pair_class_id(la, lb)].lb = lb; }
  for (int id = 0; id < (int)sp_h_.size(); ++id) pclass_[pair_class_id(sp_h_[id].la, sp_h_[id].lb)].sorted.push_back(id);
  for (auto& pc : pclass_) {
    std::stable_sort(pc.sorted.begin(), pc.sorted.end(),
                     [&](int x, int y) { return sp_h_[x].qmax > sp_h_[y].qmax; });
    pc.list.upload(pc.sorted);
    std::vector<int4> meta(pc.sorted.size());
    for (size_t k = 0; k < pc.sorted.size(); ++k) {
      const SPair& p = sp_h_[pc.sorted[k]];
      meta[k] = make_int4(p.bi, p.bj, p.si != p.sj ? 1 : 0, pc.sorted[k]);
    }
    pc.meta.upload(meta);
  }
  // quartet classes (bra class A >= ket class B) and their work blocks; with
  // precision=mixed on the stored path each bra's kets split into an FP64
  // prefix (qmax_bra * qmax_ket >= fp32_bound_threshold; the ket list is
  // sorted by qmax) and an FP32 suffix
  qclass_.clear();
  n_quartets_ = 0;
  n_fp32_quartets_ = 0;
  const bool split32 = mixed_ && method_ == "stored";
  for (int A = 0; A < npc; ++A)
    for (int B = 0; B <= A; ++B) {
      auto qc = std::make_unique<QuartetClass>();
      qc->A = A; qc->B = B;
      qc->LA = pclass_[A].la; qc->LB = pclass_[A].lb; qc->LC = pclass_[B].la; qc->LD = pclass_[B].lb;
      const auto& bl = pclass_[A].sorted;
      const auto& kl = pclass_[B].sorted;
      // kets with qb * qk >= t form a prefix of the descending list
      auto prefix = [&](double qb, double t) {
        int lo = 0, hi = (int)kl.size();
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (qb * sp_h_[kl[mid]].qmax >= t) lo = mid + 1; else hi = mid;
        }
        return lo;
      };
      for (int s = 0; s < (int)bl.size(); ++s) {
        const double qb = sp_h_[bl[s]].qmax;
        int count = prefix(qb, thr_);
        if (A == B) count = std::min(count, s + 1);
        if (count <= 0) continue;
        const int split = split32 ? std::min(count, prefix(qb, fp32_thr_)) : count;
        qc->bras_h.push_back(BraRec{s, qc->n_blocks64, count, split, qc->n_blocks32, 0, 0, 0});
        for (int k0 = 0; k0 < split; k0 += 64)
          qc->blocks_h.push_back(BlockRec{s, k0, std::min(64, split - k0), qc->n_blocks64++});
        for (int k0 = split; k0 < count; k0 += 64)
          qc->blocks_h.push_back(BlockRec{s, k0, std::min(64, count - k0), -1 - qc->n_blocks32++});
        qc->n_quartets += count;
        qc->n_fp32 += count - split;
      }
      n_quartets_ += qc->n_quartets;
      n_fp32_quartets_ += qc->n_fp32;
      qc->blocks.upload(qc->blocks_h);
      qc->bras.upload(qc->bras_h);
      for (int mode : {(int)kModeDirect, (int)kModeStoredDigest})
        qc->slab_ok[mode] = slab_fits(qc->LA, qc->LB, qc->LC, qc->LD, mode, M_);
      for (const auto& br : qc->bras_h) qc->max_count = std::max(qc->max_count, br.count);
      // default slab work records: the largest bra tile, 4 passes of kets
      for (int mode : {(int)kModeDirect, (int)kModeStoredDigest}) {
        int bpw = 1, qpp = 256;
        slab_geometry(qc->LA, qc->LB, qc->LC, qc->LD, mode, &bpw, &qpp);
        make_slab_work(*qc, mode, bpw, 4 * qpp);
      }
      qclass_.push_back(std::move(qc));
    }

  stored_bytes_ = 0;
  if (method_ == "stored") {
    EriArgs g{d_sp_.ptr, d_pr_.ptr, boys_.ptr, thr_, d_Dp_.ptr, M_, d_Jr_.ptr, d_Kr_.ptr};
    for (auto& qc : qclass_) {
      const size_t ncomp = (size_t)ncart(qc->LA) * ncart(qc->LB) * ncart(qc->LC) * ncart(qc->LD);
      qc->store.alloc((size_t)qc->n_blocks64 * 64 * ncomp);
      qc->store32.alloc((size_t)qc->n_blocks32 * 64 * ncomp);
      stored_bytes_ += double(qc->store.n) * sizeof(double) + double(qc->store32.n) * sizeof(float);
      ClassLaunch c{qc->LA, qc->LB, qc->LC, qc->LD, qc->blocks.ptr, (int)qc->blocks_h.size(),
                    pclass_[qc->A].list.ptr, pclass_[qc->B].list.ptr, qc->store.ptr, qc->store32.ptr};
      GANSU_HIP(launch_eri_class(c, g, kModeStore, 0));
    }
    GANSU_HIP(hipDeviceSynchronize());
  }
  // autotune the digestion tiles on this geometry (the stored path's
  // per-build kernel; the direct path's cost is the integrals themselves)
  if (method_ == "stored" && autotune_) {
    GANSU_HIP(launch_fill(d_Dp_.ptr, (size_t)M_ * M_, 1e-3, 0));
    for (auto& qc : qclass_)
      if (!qc->blocks_h.empty() && qc->slab_ok[kModeStoredDigest]) tune_slab(*qc, kModeStoredDigest);
    GANSU_HIP(launch_zero(d_Jr_.ptr, (size_t)M_ * M_, 0));
    GANSU_HIP(launch_zero(d_Kr_.ptr, (size_t)M_ * M_, 0));
    GANSU_HIP(hipDeviceSynchronize());
  }
}

void FockEngine::make_slab_work(QuartetClass& qc, int mode, int bpw, int kch) {
  auto& w = qc.work_h[mode];
  w.clear();
  for (int b0 = 0; b0 < (int)qc.bras_h.size(); b0 += bpw) {
    const int nb = std::min(bpw, (int)qc.bras_h.size() - b0);
    int maxc = 0;
    for (int b = 0; b < nb; ++b) maxc = std::max(maxc, qc.bras_h[b0 + b].count);
    for (int k0 = 0; k0 < maxc; k0 += kch) w.push_back(SlabWork{b0, nb, k0, std::min(maxc, k0 + kch)});
  }
  qc.work[mode].upload(w);
  qc.bpw[mode] = bpw;
  qc.kch[mode] = kch;
}

SlabLaunch FockEngine::slab_launch(QuartetClass& qc, int mode) {
  return SlabLaunch{qc.LA, qc.LB, qc.LC, qc.LD, qc.bras.ptr, (int)qc.bras_h.size(), qc.work[mode].ptr,
                    (int)qc.work_h[mode].size(), pclass_[qc.A].list.ptr, pclass_[qc.B].list.ptr, qc.store.ptr,
                    qc.max_count, qc.store32.ptr, pclass_[qc.B].meta.ptr, qc.bpw[mode]};
}

void FockEngine::tune_slab(QuartetClass& qc, int mode) {
  int bpw_max = 1, qpp = 256;
  slab_geometry(qc.LA, qc.LB, qc.LC, qc.LD, mode, &bpw_max, &qpp);
  const EriArgs g{d_sp_.ptr, d_pr_.ptr, boys_.ptr, thr_, d_Dp_.ptr, M_, d_Jr_.ptr, d_Kr_.ptr};
  hipEvent_t e0, e1;
  GANSU_HIP(hipEventCreate(&e0));
  GANSU_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_bpw = qc.bpw[mode], best_kch = qc.kch[mode];
  for (int bpw = bpw_max; bpw >= 1; bpw /= 2)
    for (int mult : {1, 2, 4, 8}) {
      make_slab_work(qc, mode, bpw, mult * qpp);
      const SlabLaunch c = slab_launch(qc, mode);
      // a warm-up and one timed launch per candidate (the precompute, tuning
      // included, is part of every SCF's wall time)
      GANSU_HIP(launch_fock_slab(c, g, mode, 0));
      GANSU_HIP(hipEventRecord(e0, 0));
      GANSU_HIP(launch_fock_slab(c, g, mode, 0));
      GANSU_HIP(hipEventRecord(e1, 0));
      GANSU_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      GANSU_HIP(hipEventElapsedTime(&t, e0, e1));
      if (t < best) { best = t; best_bpw = bpw; best_kch = mult * qpp; }
    }
  make_slab_work(qc, mode, best_bpw, best_kch);
  qc.t_slab = best;
  const SlabLaunch c = slab_launch(qc, mode);
  GANSU_HIP(launch_jket(c, g, 0));
  GANSU_HIP(hipEventRecord(e0, 0));
  GANSU_HIP(launch_jket(c, g, 0));
  GANSU_HIP(hipEventRecord(e1, 0));
  GANSU_HIP(hipEventSynchronize(e1));
  GANSU_HIP(hipEventElapsedTime(&qc.t_jket, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void FockEngine::reset_incremental(hipStream_t s) {
  if (incr_period_ > 0 && d_incr_st_.ptr) GANSU_HIP(launch_set_int(d_incr_st_.ptr, 0, s));
}

void FockEngine::build_four_center(const double* d_D, hipStream_t s, bool incremental) {
  prof_clear();
  EriArgs g{d_sp_.ptr, d_pr_.ptr, boys_.ptr, thr_, d_Dp_.ptr, M_, d_Jr_.ptr, d_Kr_.ptr};
  if (incremental) {
    // G(D) = G(D_prev) + G(D - D_prev): Jr, Kr keep the previous build's sums
    // (cleared on the full builds), the kernels digest the density change and
    // skip the quartets whose bound times the change is below the threshold
    const int nsh = (int)shells_.size();
    GANSU_HIP(launch_incr_step(d_incr_st_.ptr, incr_period_, s));
    GANSU_HIP(launch_delta_density(d_D, d_w_.ptr, M_, d_incr_st_.ptr, d_Dp_.ptr, d_Dprev_.ptr, d_dDp_.ptr, s));
    GANSU_HIP(launch_zero_if_full(d_Jr_.ptr, d_Kr_.ptr, (size_t)M_ * M_, d_incr_st_.ptr, s));
    GANSU_HIP(launch_shell_dmax(d_dDp_.ptr, M_, d_shell_bf_.ptr, d_shell_nf_.ptr, nsh, d_dmax_.ptr, s));
    g.Dp = d_dDp_.ptr;
    g.dmax = d_dmax_.ptr;
    g.nsh = nsh;
    g.dthr = incr_thr_;
  } else {
    GANSU_HIP(launch_scale_density(d_D, d_w_.ptr, M_, d_Dp_.ptr, s));
    GANSU_HIP(launch_zero(d_Jr_.ptr, (size_t)M_ * M_, s));
    GANSU_HIP(launch_zero(d_Kr_.ptr, (size_t)M_ * M_, s));
  }
  const int mode = method_ == "stored" ? kModeStoredDigest : kModeDirect;
  // launch items: (class, 0 = digestion, 1 = ket-J pass), costed by the
  // integrals they visit; profiling runs them in order on the build stream
  struct Item { QuartetClass* qc; int kind; double cost; };
  std::vector<Item> items;
  for (auto& qc : qclass_) {
    if (qc->blocks_h.empty()) continue;
    const double ncomp = double(ncart(qc->LA) * ncart(qc->LB) * ncart(qc->LC) * ncart(qc->LD));
    // tuned kernel times where the autotuner ran, else the integral count
    const bool timed = mode == kModeStoredDigest && qc->t_slab > 0.f;
    const double w = timed ? 1e9 * qc->t_slab : qc->n_quartets * ncomp * (mode == kModeDirect ? 10.0 : 1.0);
    items.push_back({qc.get(), 0, w});
    if (mode == kModeStoredDigest && qc->slab_ok[mode])
      items.push_back({qc.get(), 1, timed ? 1e9 * qc->t_jket : 0.5 * w});
  }
  const bool concurrent = !profiling_ && ev_fork_ != nullptr;
  std::vector<hipStream_t> st{s};
  if (concurrent) {
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    GANSU_HIP(hipEventRecord(ev_fork_, s));
    for (int i = 0; i < kSideStreams; ++i) {
      GANSU_HIP(hipStreamWaitEvent(side_[i], ev_fork_, 0));
      st.push_back(side_[i]);
    }
  }
  // longest first onto the least-loaded stream (no class kernel needs a
  // private segment, so any queue can take any of them)
  std::vector<double> load(st.size(), 0.0);
  for (const Item& it : items) {
    QuartetClass* qc = it.qc;
    const bool slab = qc->slab_ok[mode];
    const size_t k = std::min_element(load.begin(), load.end()) - load.begin();
    load[k] += it.cost;
    hipStream_t q = st[k];
    const double ncomp = double(ncart(qc->LA) * ncart(qc->LB) * ncart(qc->LC) * ncart(qc->LD));
    // algorithmic traffic of the stored digestion: each surviving integral read once
    const double bytes =
        mode == kModeStoredDigest ? ((qc->n_quartets - qc->n_fp32) * 8.0 + qc->n_fp32 * 4.0) * ncomp : 0.0;
    const double f32id = qc->n_fp32 > 0 && mode == kModeStoredDigest ? 1000000.0 : 0.0;
    const double cls = 1000 * qc->LA + 100 * qc->LB + 10 * qc->LC + qc->LD;
    const SlabLaunch c = slab_launch(*qc, mode);
    if (it.kind == 1) {  // ket-major J of the stored blocks
      prof_begin(q, f32id + 300000.0 + 10000.0 * mode + cls, bytes, 0.0, qc->n_quartets);
      GANSU_HIP(launch_jket(c, g, q));
      prof_end(q);
      continue;
    }
    prof_begin(q, f32id + 10000.0 * mode + cls + (slab ? 100000.0 : 0.0), bytes, 0.0, qc->n_quartets);
    if (slab) {
      GANSU_HIP(launch_fock_slab(c, g, mode, q));
    } else {  // per-lane kernel: digests J of bra and ket completely
      ClassLaunch cl{qc->LA, qc->LB, qc->LC, qc->LD, qc->blocks.ptr, (int)qc->blocks_h.size(),
                     pclass_[qc->A].list.ptr, pclass_[qc->B].list.ptr, qc->store.ptr, qc->store32.ptr};
      GANSU_HIP(launch_eri_class(cl, g, mode, q));
    }
    prof_end(q);
  }
  if (concurrent)
    for (int i = 0; i < kSideStreams; ++i) {
      GANSU_HIP(hipEventRecord(ev_join_[i], side_[i]));
      GANSU_HIP(hipStreamWaitEvent(s, ev_join_[i], 0));
    }
}

void FockEngine::build_jk(const double* d_D, double* d_J, double* d_K, const double* d_H, double* d_F,
                          hipStream_t s, const double* d_C, int nocc) {
  if (!precomputed_) throw Error(GANSU_E_STATE, "build before precompute");
  if (!d_D) throw Error(GANSU_E_ARG, "null density");
  if (d_C && (nocc <= 0 || nocc > M_)) throw Error(GANSU_E_ARG, "nocc out of range");
  if (method_ == "ri") { build_ri(d_D, d_J, d_K, d_H, d_F, s, d_C, nocc); return; }
  const bool incr = incr_period_ > 0 && !multi_density_;
  if (!use_graph_ || profiling_ || incr) {
    build_four_center(d_D, s, incr);
    GANSU_HIP(launch_finalize(d_Jr_.ptr, d_Kr_.ptr, d_w_.ptr, M_, d_H, d_J, d_K, d_F, s));
    return;
  }
  // Graph path: the build's launches (four streams joined by events) are
  // captured once per set of caller buffers and replayed as one graph launch
  // on the caller's stream; an SCF loop over fixed D/H/F buffers replays.
  GANSU_HIP(hipGraphLaunch(graph_for(GraphKey{d_D, d_J, d_K, d_H, d_F}), s));
}

void FockEngine::record_build(const double* d_D, const double* d_H, double* d_F, hipStream_t s) {
  if (!precomputed_) throw Error(GANSU_E_STATE, "build before precompute");
  if (method_ == "ri") throw Error(GANSU_E_UNSUPPORTED, "a captured build needs eri_method stored or direct");
  if (!d_D || !d_H || !d_F) throw Error(GANSU_E_ARG, "null device pointer");
  build_four_center(d_D, s, incr_period_ > 0);
  GANSU_HIP(launch_finalize(d_Jr_.ptr, d_Kr_.ptr, d_w_.ptr, M_, d_H, nullptr, nullptr, d_F, s));
}

void FockEngine::graph_release() {
  for (auto& e : gcache_) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e = GraphEntry{};
  }
}

hipGraphExec_t FockEngine::graph_for(const GraphKey& k) {
  ++graph_clock_;
  for (auto& e : gcache_)
    if (e.exec && e.key == k) {
      e.used = graph_clock_;
      return e.exec;
    }
  GraphEntry& e = gcache_[gcache_[0].used <= gcache_[1].used ? 0 : 1];  // least recently used
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  if (e.graph) (void)hipGraphDestroy(e.graph);
  e = GraphEntry{};
  if (!cap_) GANSU_HIP(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
  GANSU_HIP(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
  try {
    build_four_center(k.D, cap_);
    GANSU_HIP(launch_finalize(d_Jr_.ptr, d_Kr_.ptr, d_w_.ptr, M_, k.H, k.J, k.K, k.F, cap_));
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(cap_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  GANSU_HIP(hipStreamEndCapture(cap_, &e.graph));
  GANSU_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
  e.key = k;
  e.used = graph_clock_;
  ++n_graph_captures_;
  return e.exec;
}

void FockEngine::build_uhf(const double* d_Da, const double* d_Db, const double* d_H, double* d_Fa, double* d_Fb,
                           hipStream_t s) {
  if (!precomputed_) throw Error(GANSU_E_STATE, "build before precompute");
  if (!d_Da || !d_Db || !d_H || !d_Fa || !d_Fb) throw Error(GANSU_E_ARG, "null device pointer");
  const size_t nn = (size_t)M_ * M_;
  if (uhf_.n < 4 * nn) uhf_.alloc(4 * nn);
  double *Ja = uhf_.ptr, *Ka = Ja + nn, *Jb = Ka + nn, *Kb = Jb + nn;
  multi_density_ = true;  // two densities per build: full builds
  try {
    build_jk(d_Da, Ja, Ka, nullptr, nullptr, s);
    build_jk(d_Db, Jb, Kb, nullptr, nullptr, s);
  } catch (...) {
    multi_density_ = false;
    throw;
  }
  multi_density_ = false;
  GANSU_HIP(launch_uhf_combine(d_H, Ja, Ka, Jb, Kb, nn, d_Fa, d_Fb, s));
}

void FockEngine::update_geometry(const double* xyz, int n_atoms) {
  if (!xyz || n_atoms <= 0) throw Error(GANSU_E_ARG, "empty geometry");
  auto move = [&](std::vector<gansu_primitive_shell>& ps) {
    for (auto& p : ps) {
      if (p.atom_index < 0 || p.atom_index >= n_atoms) throw Error(GANSU_E_ARG, "atom_index outside n_atoms");
      p.x = xyz[3 * p.atom_index];
      p.y = xyz[3 * p.atom_index + 1];
      p.z = xyz[3 * p.atom_index + 2];
    }
  };
  move(prims_);
  move(aux_prims_);
  shells_ = contracted_shells(prims_.data(), prims_.size(), kMaxL);
  graph_release();
  if (precomputed_) {
    precomputed_ = false;
    // the RI state is replaced by precompute_ri, which keeps its large buffers
    d_alpha_.release();
    d_shell_atom_.release();
    precompute();
  }
}

std::vector<double> FockEngine::schwarz_ref() const {
  if (!precomputed_) throw Error(GANSU_E_STATE, "schwarz factors before precompute");
  return Qref_h_;
}

const std::vector<FockEngine::ClassCount>& FockEngine::class_counts() {
  if (!precomputed_) throw Error(GANSU_E_STATE, "class counts before precompute");
  if (method_ == "ri") throw Error(GANSU_E_UNSUPPORTED, "class counts of the four-centre paths (stored, direct)");
  if (!class_counts_.empty()) return class_counts_;
  // each shell pair's primitive-pair Q, descending
  std::vector<std::vector<double>> qs(sp_h_.size());
  for (size_t i = 0; i < sp_h_.size(); ++i) {
    const SPair& p = sp_h_[i];
    qs[i].assign(pair_q_h_.begin() + p.prim_begin, pair_q_h_.begin() + p.prim_begin + p.nprim);
    std::sort(qs[i].begin(), qs[i].end(), std::greater<double>());
  }
  double total = 0.0;
  for (const auto& qc : qclass_) {
    ClassCount cc{{qc->LA, qc->LB, qc->LC, qc->LD}, qc->n_quartets, 0.0};
    const auto& bl = pclass_[qc->A].sorted;
    const auto& kl = pclass_[qc->B].sorted;
    for (const BraRec& br : qc->bras_h) {
      const auto& X = qs[bl[br.bra_pos]];
      for (int k = 0; k < br.count; ++k) {
        const auto& Y = qs[kl[k]];
        // pairs with X_i Y_j >= thr: for descending X the admissible prefix of
        // Y shrinks, so one pointer walks Y once
        size_t j = Y.size();
        double n = 0.0;
        for (double x : X) {
          while (j > 0 && x * Y[j - 1] < thr_) --j;
          if (j == 0) break;
          n += (double)j;
        }
        cc.primitive += n;
      }
    }
    total += cc.primitive;
    class_counts_.push_back(cc);
  }
  n_prim_quartets_ = total;
  return class_counts_;
}

void FockEngine::stats(double* out) const {
  for (int k = 0; k < 8; ++k) out[k] = 0.0;
  out[0] = n_quartets_;
  out[1] = n_prim_quartets_;
  out[2] = stored_bytes_;
  out[3] = n_fp32_quartets_;
  out[6] = n_graph_captures_;
  const double P = double(M_) * (M_ + 1) / 2;
  out[4] = P * (P + 1) / 2;
  if (method_ == "ri") out[5] = double(M_) * Maux_ * M_ * sizeof(double);
  out[7] = ri_incremental_builds();
}

}  // namespace gansu
