// Device-resident tree engine (see engine.cpp).
#pragma once
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "gemm_ops.hpp"
#include "model.hpp"
#include "smallmat.hpp"
#include "tree_ops.hpp"

namespace qtg {

class Cmf;

// out[i] = (float)in[i] for n doubles
int launch_to_f32(const double* in, float* out, int64_t n, hipStream_t s);

class Engine {
  friend class Cmf;

 public:
  ~Engine();
  int create(const std::string& tree_text, int precision, bool host_only = false);
  int set_hamiltonian(int64_t nterms, const double* coeff, const int64_t* term_off, const int64_t* fmode,
                      const int64_t* fkind, const int64_t* fpool, const double* pool);
  int set_psi(const double* host, hipStream_t s);
  int set_psi_device(const double* dev, hipStream_t s);
  // HamiltonianRepresentation::build: rho tree, rho^-1, h-matrices, holes
  int build(hipStream_t s);
  // LayerDerivative at every node with the current representation
  int layer_derivative_all(hipStream_t s, double ph_re, double ph_im);
  // Derivative (build + LayerDerivative everywhere)
  int derivative(hipStream_t s, double ph_re, double ph_im);
  int energy(hipStream_t s, double* out);
  int get_matrix(int which, int term, int node, double* out, hipStream_t s);

  const HTree& tree() const { return tree_; }
  const std::string& tree_text() const { return treeText_; }
  const HSop& sop() const { return sop_; }
  double sweep_flops() const { return flops_; }        // executed by this plan
  double reference_flops() const { return refFlops_; }  // of the reference algorithm
  int64_t work_elems() const { return workSize_; }       // device workspace (complex elements)
  int64_t pool_elems() const { return poolSize_; }       // matrix pool (complex elements)
  double* psi64() { return d_psi64_; }
  double* dpsi64() { return d_dpsi64_; }
  double* apply64() { return d_apply64_; }
  bool complex64() const { return c64_; }
  bool planned() const { return planned_; }
  bool host_only() const { return host_only_; }
  int64_t sandwich_items() const { return nSw_; }
  int64_t apply_terms_chunks() const { return nChunks_; }
  int64_t apply_terms() const { return (int64_t)h_terms_.size(); }
  double eps_ = 1e-10;
  int chunkTerms_ = 32;
  int chunkTermsCompressed_ = 4;
  bool compress_ = true;  // share equal h/hole matrices, aggregate Apply summands
  bool graph_ = true;     // replay Derivative() from a captured HIP graph
  // 0: LDS-resident node kernels (node tensors up to ~150 KB);
  // 1: grouped MFMA contractions through a global workspace (any node size)
  int path_ = 1;
  int64_t splitWaves_ = 2048;  // wavefronts a grouped hole-contraction launch is split up to
  int splitMinK_ = 16;         // shortest K chunk of such a split
  int gmSplitK_ = 32;    // mode products longer than 2 gmSplitK_ split along K, chunks >= gmSplitK_
  int gcLongK_ = 512;    // hole contractions longer than 2 gcLongK_ always split along K
  // Plan parameters from the tuning table (tuning.cpp) for this plan's
  // class, unless the caller set them as options; tuning_ false skips it.
  bool tuning_ = true;
  unsigned userSet_ = 0;  // bits: options set explicitly (see tuning.cpp)
  std::string tunedFrom_;  // the table line that was applied (empty: none)
  void apply_tuning();
  const std::string& tuned_from() const { return tunedFrom_; }
  // Apply + projector + rho^-1 on an arbitrary FP64 state (complex128 engines):
  // dpsi = -i phase rho^-1 (1 - P) Apply(psi) at every node whose active flag
  // is set (all when active is null); energy (4 doubles: <A|HA>, <A|A> of the
  // top node) is written when not null.  Uses the last build's matrices.
  int apply_finish(const double* psi64, double* dpsi64, const uint8_t* active, double ph_re, double ph_im,
                   double* energy_dev, hipStream_t s);
  // dE/dA* at every node (complex128 engines, grouped path) of the state set
  // by set_psi, whose non-top nodes are orthonormal; grad_host receives the
  // gradient in the engine's layout, energy the expectation value.
  int gradient(hipStream_t s, double* grad_host, double* energy);
  // The energy functional E(A) = <Psi|H|Psi> / <Psi|Psi> as a forward /
  // backward pair (complex128 engines, grouped path, one state at the top):
  // forward runs the sweep at phase 1 and leaves E and 1/<Psi|Psi> on the
  // device (energy_host, when not null, receives E); backward writes dE/dA*
  // into the engine's gradient buffer (grad_host, when not null, receives
  // it).  With gradRecompute_ backward first recomputes the sweep's
  // matrices from the current state, so nothing of the forward pass has to
  // be kept in between (other sweeps may overwrite the matrix pool);
  // otherwise it uses those of the last forward.
  int forward(hipStream_t s, double* energy_host);
  int backward(hipStream_t s, double* grad_host);
  // steps of A <- A - eta dE/dA* at every node, each followed by the QR
  // gauge (every non-top node orthonormalised, R pushed into its parent),
  // all on the device.  energies_host (steps + 1) receives E before every
  // step and after the last; psi_host (may be null) the final state.
  int gradient_descent(hipStream_t s, int64_t steps, double eta, double* energies_host, double* psi_host);
  // QR gauge of the state on the device (see gradient_descent)
  int canonicalize(hipStream_t s);
  bool gradRecompute_ = false;
  // the grouped sweep's eigensolver on the side stream (fork after the
  // build, join before the solve) instead of in line on the engine stream;
  // measured equal for replayed sweeps and faster for CMF, whose first tick
  // after a build overlaps the eigensolver
  bool heevSide_ = true;
  // hole layers of complex128 binary-tree sweeps as one launch each
  // (hole_fused_kernel: kets built in LDS, bras of all layers in one launch
  // before them); false keeps a mode-product and a hole-contraction launch
  // per layer
  bool fuseHoles_ = true;
  // device bytes this engine holds
  int64_t device_bytes() const { return deviceBytes_; }
  hipStream_t stream() const { return stream_; }
  // waits for the engine's streams (its own and the eigensolver's side stream)
  int synchronize() {
    QTG_HIP_TRY(hipStreamSynchronize(stream_));
    if (side_) QTG_HIP_TRY(hipStreamSynchronize(side_));
    return QTG_OK;
  }
  void invalidate_graph();
  // Per-kernel timing with HIP events recorded on the launch stream around
  // every launch group: 0 sandwich, 1 heev, 2 apply, 3 finish, 4 convert, 5 combine.
  // 6 gm_grouped, 7 gc_grouped, 8 solve, 9 (unused: split-K sums run inside 6 and 7).
  static constexpr int kPhases = 10;
  int set_timing(bool on);
  int read_timing(double* ms, int64_t* launches, double* flops, double* bytes);

 private:
  template <typename T>
  int build_t(hipStream_t s);
  template <typename T>
  int apply_t(hipStream_t s, double ph_re, double ph_im);
  int plan();
  int plan_gemm();
  void dump_plan() const;
  int convert_psi(hipStream_t s);  // FP64 state -> FP32 copy (complex64 engines)
  int build_gemm(hipStream_t s);
  int apply_gemm(hipStream_t s, const double* psi64, double* dpsi64, const uint8_t* active, double ph_re,
                 double ph_im);
  // one launch of the grouped path; matrices always live in the FP64 pool
  enum Buf { B_PSI32 = 0, B_WORK32, B_POOL, B_PSI64, B_APPLY64, B_DPSI64, B_POOL64, B_WORK64, B_COUNT };
  // P64: FP64 operands and outputs; P32: FP32 tensors (matrices rounded on
  // load), FP32 mode-product outputs, FP64 contraction outputs; P32W: FP32
  // mode product writing FP64 (the stacked Apply sum)
  enum Prec { P64 = 0, P32 = 1, P32W = 2 };
  struct GLaunch {
    int kind;    // 0 mode products, 1 hole contractions, 2 solve
    int prec;    // Prec
    int64_t d0, nd, ntasks;
    int64_t wg0;       // first entry of the workgroup -> descriptor table
    int x0, x1, m, o;  // buffer ids
    bool scaled;       // multiply alphas by the launch scale (-i phase)
    double flops;
    double bytes;  // algorithmic HBM bytes (distinct inputs once, outputs)
    int waves;     // mode products: tasks per workgroup (0: kGemmWaves)
    bool lean;     // every descriptor takes the lean kernel form
    int64_t lds;   // fused hole layers (kind 3): dynamic LDS bytes per workgroup
  };
  bool gramInSolve_ = false;  // the solve launch forms S, A^H Y and the energy terms
  int run_launches(const std::vector<GLaunch>& L, void* const* bases, const uint8_t* active, double sre, double sim,
                   hipStream_t s) {
    return run_launches(L, 0, L.size(), bases, active, sre, sim, s);
  }
  // launches [i0, i1) of L
  int run_launches(const std::vector<GLaunch>& L, size_t i0, size_t i1, void* const* bases, const uint8_t* active,
                   double sre, double sim, hipStream_t s);
  int materialize();
  std::vector<std::pair<int, int>> restrict_in(int l, int a) const;
  std::vector<std::pair<int, int>> restrict_out(int l, int a) const;
  std::vector<std::vector<char>> inSub_;  // node x mode: leaf below the node
  void active_sets();
  int alloc(void** p, size_t bytes);
  void release();

  HTree tree_;
  std::string treeText_;
  HSop sop_;
  bool c64_ = false, planned_ = false, built_ = false, host_only_ = false;
  int precision_ = 0;  // 0 complex128, 1 complex64 bulk, 2 complex64 bulk with FP64 for large terms
  std::vector<SwItem> h_items_;
  std::vector<EigItem> h_eig_;
  std::vector<ApTerm> h_terms_;
  std::vector<ApChunk> h_chunks_;
  std::vector<FinNode> h_fin_;
  int nmax_ = 0, nspf_ = 1;
  std::vector<std::vector<char>> active_;
  std::vector<int64_t> leafOff_, rhoslot_, rhoinv_;
  struct Summand {
    cd coef;
    std::vector<int64_t> ids;  // matrix per slot (-1: identity)
  };
  std::vector<std::vector<Summand>> summ_;    // Apply summands per node
  std::vector<Summand> aggregate(const std::vector<Summand>& tl, const std::vector<int>& sdim, int64_t& off,
                                 double& combFlops);
  std::vector<CombOut> h_comb1_, h_comb2_;  // aggregated slots: partial sums, sums of partials
  std::vector<CombTerm> h_combt1_, h_combt2_;
  double combFlops12_ = 0.0;
  // energy gradient (grouped path)
  std::vector<int64_t> mslot_, gslot_;
  std::vector<CombOut> h_combG1_, h_combG2_, h_combG3_, h_combGE_;
  std::vector<CombTerm> h_combtG1_, h_combtG2_, h_combtG3_, h_combtGE_;
  std::vector<GLaunch> gGradPre_, gGradFin_;
  std::vector<std::vector<GLaunch>> gGradDepth_;
  double* d_grad64_ = nullptr;
  int backward_unscaled(hipStream_t s);
  int grad_ready();  // checks and allocations of the gradient path
  double* d_escal_ = nullptr;  // [E, 1 / <Psi|Psi>] of the last forward
  double* d_elog_ = nullptr;   // energies of a device descent
  int64_t elogCap_ = 0;
  // canonicalisation: QR items per depth (deepest first) and the parent
  // pushes of each depth, leg by leg
  std::vector<SmItem> h_canon_;
  SmItem* d_canon_ = nullptr;
  std::vector<std::pair<int64_t, int64_t>> canonRange_;
  std::vector<std::vector<GLaunch>> gCanon_;
  int64_t deviceBytes_ = 0;
  double gradFlops_ = 0.0;
  CombOut *d_combG1_ = nullptr, *d_combG2_ = nullptr, *d_combG3_ = nullptr, *d_combGE_ = nullptr;
  CombTerm *d_combtG1_ = nullptr, *d_combtG2_ = nullptr, *d_combtG3_ = nullptr, *d_combtGE_ = nullptr;
  std::vector<std::vector<int>> smode_, sdim_;  // slot legs and sizes per node
  double lflops_ = 0.0;                       // sweep flops of the LDS path
  // grouped path
  std::vector<GmDesc> h_gm_;
  std::vector<GcDesc> h_gc_;
  std::vector<HoleKet> h_hk_;   // fused hole layers: kets and their items
  std::vector<HoleItem> h_hi_;
  HoleKet* d_hk_ = nullptr;
  HoleItem* d_hi_ = nullptr;
  std::vector<int32_t> h_gmwg_, h_gcwg_;
  int32_t* d_gmwg_ = nullptr;
  int32_t* d_gcwg_ = nullptr;
  std::vector<SolveItem> h_solve_;
  std::vector<GLaunch> gBuild_, gHole_, gApply_;
  std::vector<std::pair<int64_t, int>> idSlots_;  // identity matrices in the pool (offset, n)
  std::vector<CombOut> h_comb3_;
  std::vector<CombTerm> h_combt3_;
  int64_t workSize_ = 0, work32Size_ = 0, energySlot_ = -1, nComb3_ = 0;
  void* d_work32_ = nullptr;
  // pool slots computed in FP64 by the split-precision form: those serving a
  // large-coefficient term, and the densities
  std::set<int64_t> highSlots_;
  GmDesc* d_gm_ = nullptr;
  GcDesc* d_gc_ = nullptr;
  SolveItem* d_solve_ = nullptr;
  CombOut* d_combOut3_ = nullptr;
  CombTerm* d_combTerm3_ = nullptr;
  void* d_work_ = nullptr;
  std::vector<std::vector<int64_t>> hslot_, holeslot_;
  std::vector<std::pair<int64_t, int64_t>> repRange_, holeRange_;
  int64_t poolSize_ = 0, pool64Size_ = 0, partialSize_ = 0;
  int64_t nSw_ = 0, nChunks_ = 0, nFin_ = 0, nEig_ = 0, nComb_ = 0, nComb2_ = 0;
  double flops_ = 0.0, refFlops_ = 0.0;

  DNode* d_nodes_ = nullptr;
  SwItem* d_sw_ = nullptr;
  EigItem* d_eig_ = nullptr;
  ApTerm* d_terms_ = nullptr;
  ApChunk* d_chunks_ = nullptr;
  FinNode* d_fin_ = nullptr;
  CombOut* d_combOut_ = nullptr;
  CombTerm* d_combTerm_ = nullptr;
  CombOut* d_combOut2_ = nullptr;
  CombTerm* d_combTerm2_ = nullptr;
  double* d_psi64_ = nullptr;
  double* d_dpsi64_ = nullptr;
  double* d_apply64_ = nullptr;
  float* d_psi32_ = nullptr;
  void* d_pool_ = nullptr;
  double* d_pool64_ = nullptr;
  void* d_partial_ = nullptr;
  double* d_energy_ = nullptr;
  std::vector<void*> device_buffers_;
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;  // density eigensolver, overlapping Apply
  hipEvent_t evFork_ = nullptr, evHeev_ = nullptr, evRho_ = nullptr;
  bool heevPending_ = false;

  hipGraph_t graph_h_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  double graph_ph_re_ = 0.0, graph_ph_im_ = 0.0;
  int derivative_eager(hipStream_t s, double ph_re, double ph_im);
  // timing state
  bool timing_ = false;
  int clockKHz_ = 0;  // wall-clock rate of the current device (kHz), read on first use
  struct Mark {
    hipEvent_t a, b;
    int phase;
  };
  std::vector<Mark> marks_;
  std::vector<hipEvent_t> evpool_;
  size_t evnext_ = 0;
  double tms_[kPhases] = {0}, tfl_[kPhases] = {0}, tby_[kPhases] = {0};
  int64_t tn_[kPhases] = {0};
  double phaseFlops_[kPhases] = {0};
  std::vector<double> layerFlops_;  // flops of each sandwich launch, in launch order
  std::vector<int> layerNmax_;      // largest node tensor of each sandwich launch
  int begin_mark(int phase, hipStream_t s);
  int end_mark(hipStream_t s);
  int collect_marks();
  // While timing is on: a fresh event pair for the next kernel's own start
  // and stop times (*kt stays null when timing is off).
  int kernel_mark(int phase, const KernelTiming** kt);
  KernelTiming ktiming_{};
};

}  // namespace qtg
