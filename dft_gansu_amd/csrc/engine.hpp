// Fock engine: owns integral-side state (pair data, work blocks, stored
// integrals, RI tensors) and runs precompute / Fock builds on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "fock_kernels.hpp"
#include "gansu_internal.hpp"

namespace gansu {

void hip_check(hipError_t e, const char* what);
#define GANSU_HIP(x) ::gansu::hip_check((x), #x)

// Makes `dev` the current device for a scope and restores the caller's
// device afterwards, so that an engine created on device k runs there
// whatever the caller's current device is at a later call.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != dev && hipSetDevice(dev) == hipSuccess) prev = cur;
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
};

// RAII device buffer.
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) GANSU_HIP(hipMalloc((void**)&ptr, count * sizeof(T)));
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void upload(const T* h, size_t count) {
    alloc(count);
    if (count) GANSU_HIP(hipMemcpy(ptr, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
  std::vector<T> download() const {
    std::vector<T> h(n);
    if (n) GANSU_HIP(hipMemcpy(h.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

struct QuartetClass {
  int A, B;  // pair class ids (bra >= ket)
  int LA, LB, LC, LD;
  std::vector<BlockRec> blocks_h;
  DevBuf<BlockRec> blocks;
  std::vector<BraRec> bras_h;  // one entry per bra with kets (slab digestion)
  DevBuf<BraRec> bras;
  bool slab_ok[3] = {true, true, true};  // per mode: the slab kernel has a variant whose LDS fits
  std::vector<SlabWork> work_h[3];  // per mode (direct, -, stored digest)
  DevBuf<SlabWork> work[3];
  int bpw[3] = {1, 1, 1};  // bras per workgroup of work[mode]
  int kch[3] = {0, 0, 0};  // kets per workgroup of work[mode]
  float t_slab = 0.f, t_jket = 0.f;  // tuned kernel times (ms), launch-order costs
  int max_count = 0;
  DevBuf<double> store;  // stored integrals of this class (eri_method=stored)
  // precision=mixed: the quartets whose bound qmax_bra * qmax_ket is below
  // fp32_bound_threshold (a suffix of each bra's kets) are stored in FP32
  DevBuf<float> store32;
  int n_blocks64 = 0, n_blocks32 = 0;
  double n_quartets = 0, n_fp32 = 0;
};

struct PairClass {
  int la, lb;
  std::vector<int> sorted;  // shell-pair ids sorted by qmax desc
  DevBuf<int> list;
  DevBuf<int4> meta;  // per sorted position: {bi, bj, si != sj, id}
};

// Coupling set of the unified ROHF Fock matrix by its parameter name
// (rohf_parameter_name, rohf.cu:30-39); false for an unknown name.
bool rohf_coupling(const std::string& name, RohfCoupling* out);

class FockEngine {
 public:
  FockEngine(const gansu_primitive_shell* prims, size_t n_prims, const gansu_shell_type_info* types, int n_types,
             const double* norm, int num_basis, const gansu_primitive_shell* aux, size_t n_aux,
             const gansu_shell_type_info* aux_types, int n_aux_types, const double* aux_norm, int num_aux,
             const std::map<std::string, std::string>& opts);
  ~FockEngine();

  void precompute();
  // d_C/nocc (optional): occupied orbitals, D = 2 C_occ C_occ^T
  void build_jk(const double* d_D, double* d_J, double* d_K, const double* d_H, double* d_F, hipStream_t s,
                const double* d_C = nullptr, int nocc = 0);
  // Four-centre RHF build F = H + J - K/2 launched on s without the engine's
  // own graph, so that a caller capturing s records it into its own graph
  // (stored and direct engines; RI builds call rocBLAS and are refused).
  void record_build(const double* d_D, const double* d_H, double* d_F, hipStream_t s);
  // UHF (gpu_kernels.cu:289-340): F_a = H + J[D_a + D_b] - K[D_a], F_b likewise.
  void build_uhf(const double* d_Da, const double* d_Db, const double* d_H, double* d_Fa, double* d_Fb, hipStream_t s);
  // ROHF (gpu_manager.cu:733-817): F_c, F_o from J/K of the closed and open
  // densities, the unified MO-basis matrix R with coupling cp over nc closed
  // and no open orbitals (columns of the row-major C), and F = S C R C^T S.
  void build_rohf(const double* d_Dc, const double* d_Do, const double* d_H, const double* d_C, const double* d_S,
                  const RohfCoupling& cp, int nc, int no, double* d_Fc, double* d_Fo, double* d_F, hipStream_t s);
  // Moves every primitive (and auxiliary primitive) to the position of its
  // atom and repeats the precompute if it had been done.
  void update_geometry(const double* xyz_bohr, int n_atoms);
  // incremental_fock: the next build is a full one (a new SCF starts)
  void reset_incremental(hipStream_t s);
  const char* algorithm_name() const;
  std::vector<double> schwarz_ref() const;  // host copy, reference pair order
  size_t num_primitive_pairs() const { return ref_pairs_; }
  void stats(double* out) const;
  // Per quartet class of the stored/direct paths: (LA LB LC LD), the
  // contracted quartets the builds visit and the primitive quartets among
  // them that pass the primitive screen Q_ab Q_cd >= thr (the MD unit of
  // work).  Counted on the host once, on first use.
  struct ClassCount { int L[4]; double contracted, primitive; };
  const std::vector<ClassCount>& class_counts();
  int num_basis() const { return M_; }
  int device() const { return dev_; }  // the device current when the engine was created
  void set_profiling(bool on);
  int graph_captures() const;  // graphs captured so far (the build replays the last one)
  int ri_incremental_builds() const;  // precision=mixed RI builds formed as increments
  std::vector<double> last_profile();  // 5 values per launch
  // Two-electron part of the RHF nuclear gradient for density D (device,
  // M x M): d_grad (device, 3 natoms) = sum Gamma(D) d(ab|cd)/dX, overwritten.
  // RI engines need the occupied orbitals: d_C (M x M row-major, first nocc
  // columns occupied, D = 2 C_occ C_occ^T); the four-centre paths ignore them.
  void gradient_2e(const double* d_D, int natoms, double* d_grad, hipStream_t s, const double* d_C = nullptr,
                   int nocc = 0);

 private:
  void precompute_pairs();     // Schwarz factors, shell pairs, primitive-pair records
  void precompute_quartets();  // quartet classes, work blocks, stored integrals
  // Slab work records of a class for bpw bras and kch kets per workgroup.
  void make_slab_work(QuartetClass& qc, int mode, int bpw, int kch);
  // Autotuner: times the class's slab kernel over candidate (bpw, kch) and
  // keeps the fastest.
  void tune_slab(QuartetClass& qc, int mode);
  SlabLaunch slab_launch(QuartetClass& qc, int mode);
  void precompute_ri();
  // seed: also write the T chunks into the incremental state (precision=mixed)
  void build_ri_k_occ(const double* d_C, int nocc, double* K, hipStream_t s, bool seed = false);
  // precision=mixed incremental exchange (ri_mfma.hip); false: a full build is due
  bool build_ri_incremental(const double* d_C, int nocc, double* J, double* K, hipStream_t s);
  void rik_prepare(int nocc);
  static int rik_chunks(int M, long Kd);  // split-k depth of the dK product
  void gradient_ri(const double* d_D, const double* d_C, int nocc, int natoms, double* d_grad, hipStream_t s);
  void ri_k_product(const double* A, int kdim, bool accumulate, double* K, hipStream_t s);
  // incremental: a delta-density build (incremental_fock), Jr and Kr
  // accumulate on top of the previous build's
  void build_four_center(const double* d_D, hipStream_t s, bool incremental = false);
  void build_ri(const double* d_D, double* d_J, double* d_K, const double* d_H, double* d_F, hipStream_t s,
                const double* d_C, int nocc);

  int dev_ = 0;
  std::string method_;  // stored | direct | ri
  bool nan_check_ = true;  // eri_method=ri: a build whose F holds a NaN fails (rhf.hpp:665-674)
  double thr_ = 1e-12;
  bool mixed_ = false;
  double fp32_thr_ = 1e-6;
  bool ri_k_occ_ = false;  // RI-K from occupied orbitals when they are given
  // precision=mixed RI: a full FP64 build after rik_refresh_ increments,
  // for a change |E| / |X| above rik_accept_, and for the first small change
  // after increments whose |E| / |X| summed to more than rik_tol_
  int rik_refresh_ = 16;
  double rik_tol_ = 0.01;
  double rik_accept_ = 0.5;
  bool rik_trace_ = false;  // GANSU_RI_TRACE=1: one stderr line per mixed RI build
  int rik_ns_iters_ = 40;  // Newton-Schulz steps of the orbital alignment
  double last_rik_rel_ = 0.0;
  static constexpr int kRikMaxBlocks = 10;  // nocc <= 320 for the incremental path
  bool autotune_ = true;   // time candidate digestion tiles at precompute
  // eri_method=direct, incremental_fock=N > 0: builds from the change of the
  // density with density-weighted screening, a full build every N builds
  int incr_period_ = 0;
  double incr_thr_ = 1e-14;  // density-weighted screen: bound x |density element| below this is skipped
  bool multi_density_ = false;  // inside a UHF/ROHF build (two densities): no increments
  DevBuf<double> d_Dprev_, d_dDp_;
  DevBuf<float> d_dmax_;
  DevBuf<int> d_shell_bf_, d_shell_nf_, d_incr_st_;
  int M_ = 0;
  std::vector<gansu_primitive_shell> prims_;
  std::vector<gansu_shell_type_info> types_;
  std::vector<double> norm_;
  std::vector<double> w_h_;  // raw-basis scaling per basis function
  std::vector<CShell> shells_;
  bool precomputed_ = false;

  // device state
  DevBuf<double> boys_;
  DevBuf<gansu_primitive_shell> d_prims_;
  DevBuf<double> d_Qref_;
  size_t ref_pairs_ = 0;
  std::vector<double> Qref_h_;
  DevBuf<PairRec> d_pr_;
  std::vector<double> pair_alpha_h_;  // exponent of the first primitive of each pair record
  std::vector<double> pair_q_h_;      // Q of each pair record (host copy)
  std::vector<ClassCount> class_counts_;
  DevBuf<double> d_alpha_;
  DevBuf<int> d_shell_atom_;
  DevBuf<SPair> d_sp_;
  std::vector<SPair> sp_h_;
  std::vector<PairClass> pclass_;
  std::vector<std::unique_ptr<QuartetClass>> qclass_;
  DevBuf<double> d_w_, d_Dp_, d_Jr_, d_Kr_;
  double n_quartets_ = 0, n_prim_quartets_ = 0, stored_bytes_ = 0, n_fp32_quartets_ = 0;

  // Concurrent class launches: the build stream forks to kSideStreams more
  // streams (default priority) through an event and joins them at the end.
  // Each launch item (a class's digestion or ket-J pass) goes to the stream
  // with the least assigned work, largest items first.
  static constexpr int kSideStreams = 3;
  hipStream_t side_[kSideStreams] = {};
  hipEvent_t ev_fork_ = nullptr, ev_join_[kSideStreams] = {};
  void streams_create();
  void streams_destroy();

  // Graph of a four-centre build for one set of caller buffers.
  struct GraphKey {
    const double* D;
    double *J, *K;
    const double* H;
    double* F;
    bool operator==(const GraphKey& o) const { return D == o.D && J == o.J && K == o.K && H == o.H && F == o.F; }
  };
  bool use_graph_ = true;
  hipStream_t cap_ = nullptr;
  // two cached graphs (an RHF loop uses one, a UHF loop alternates two)
  struct GraphEntry {
    GraphKey key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
  };
  GraphEntry gcache_[2];
  unsigned long long graph_clock_ = 0;
  int n_graph_captures_ = 0;
  hipGraphExec_t graph_for(const GraphKey& k);
  void graph_release();
  DevBuf<double> uhf_;  // Ja, Ka, Jb, Kb of a UHF build (Jc, Kc, Jo, Ko of a ROHF build)
  DevBuf<double> rohf_;  // ROHF MO-basis work matrices
  void* rohf_blas_ = nullptr;  // rocBLAS handle of the ROHF transforms
  void rohf_release();

  // profiling: events around each launch of a build
  struct ProfRow { double id, bytes, flops, units; hipEvent_t a, b; };
  bool profiling_ = false;
  std::vector<ProfRow> prof_;
  void prof_begin(hipStream_t s, double id, double bytes, double flops, double units);
  void prof_end(hipStream_t s);
  void prof_clear();

  // RI state
  std::vector<gansu_primitive_shell> aux_prims_;
  std::vector<gansu_shell_type_info> aux_types_;
  std::vector<double> aux_norm_;
  int Maux_ = 0;

 public:
  struct RIState;

 private:
  RIState* ri_ = nullptr;  // owned; released by ri_destroy
};

void ri_destroy(FockEngine::RIState*);

}  // namespace gansu
