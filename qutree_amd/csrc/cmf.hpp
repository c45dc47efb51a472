// Device CMF integrator (see cmf.cpp).
#pragma once
#include <string>
#include <complex>
#include <utility>
#include <vector>

#include "cmf_ops.hpp"
#include "engine.hpp"
#include "gemm_ops.hpp"

namespace qtg {

struct CmfResult {
  std::vector<double> times, energies;
  std::vector<cd> autocorr;
  size_t steps = 0, rejected = 0, builds = 0, derivs = 0, ticks = 0;
  double dt_end = 0.0;  // step size on return (job.dt, CMFIntegrator.cpp:172)
};

struct EigenResult {
  std::vector<double> energies, propagator, converged;  // nstates per entry
  size_t steps = 0, rejected = 0, builds = 0, derivs = 0;
  double dt_end = 0.0;
};

struct NodeBS {
  enum State { NEED_START, START_MMID, IN_MMID, AWAIT_ERR, DONE };
  State st = NEED_START;
  double x = 0.0, xend = 0.0, h = 0.0, hh = 0.0;
  int i = 0, step = 0, nsub = 0;
  double xtab[11] = {0};
};

class Cmf {
 public:
  explicit Cmf(Engine& e) : E(e) {}
  ~Cmf();
  int init();
  // state in/out: P_ (device arena)
  cplx<double>* state() { return P_; }
  int64_t size() const { return total_; }
  hipStream_t stream() const { return s_; }
  int integrate(double time, double timeend, double dt, double out, double acc_cmf, double acc_bs,
                int64_t max_steps, CmfResult& res);
  // Eigenstates(IntegratorVariables&) (Eigenstates.cpp:60-103) from the state in P_
  int eigenstates(double dt, double out, int64_t iterations, double acc_cmf, double acc_bs, int64_t max_steps,
                  EigenResult& res);
  int nstates() const { return E.tree_.nodes.back().n; }
  double ph_re_ = 1.0, ph_im_ = 0.0;  // propagation phase (1 = real time)
  // Psi.write(job.ofname, job.append_) of CMFIntegrator::Integrate
  // (CMFIntegrator.cpp:46,153): the state is written as a TensorTree at the
  // start and at every output; empty name: off
  std::string outFile_;
  bool outAppend_ = true;
  // ticks replayed from HIP graphs (one per state buffer and phase); false
  // (default: measured faster on MI355X, where a graph launch per tick costs
  // more than the tick's six kernel launches) launches every tick eagerly
  bool tickGraph_ = false;
  // the op sequences a tick emits most run as fused ops (one pass over the
  // node each); false keeps one pass per op.  Results are the same.
  bool fuseOps_ = true;
  // a tick after an error tick is staged with both outcomes of each waiting
  // node's accept test and launched before the errors reach the host; false
  // waits for the errors first.  Results are the same.
  bool lookahead_ = true;
  // every node's Bulirsch-Stoer state machine runs on the device, inside the
  // launch that runs its vector programs; false runs them on the host
  // (above).  Results are the same.
  bool deviceBs_ = true;

 private:
  int eig_init();
  int apply_top(const cplx<double>* P, double ph_re, double ph_im, double* energy);
  int overlap_tree(const cplx<double>* bra, const cplx<double>* ket);
  int eigenstate(std::vector<double>& ev);
  int dalloc(void** p, size_t bytes);
  CmfBufs bufs(cplx<double>* P) const;
  int copy(cplx<double>* dst, const cplx<double>* src);
  int build(const cplx<double>* P);
  bool advance(int v, std::vector<VecOp>& pre, std::vector<VecOp>& post, double eps, int verdict);
  void finish_mmid(int v, std::vector<VecOp>& post);
  int cmfstep(cplx<double>* P, double t0, double t1, double eps);
  int cmfstep_dev(cplx<double>* P, double t0, double t1, double eps);
  struct TickLayout {
    size_t pre, post, ops;  // byte offsets in the staging; the active mask at 0
  };
  TickLayout tick_layout() const;
  // post: launch the post programs (false when the tick has none)
  int tick_launches(const CmfBufs& bb, bool post);
  int run_tick(cplx<double>* P, bool post);
  // a tick on this state buffer is launched eagerly (not from a graph)
  bool tick_eager(const cplx<double>* P) const;
  void drop_tick_graphs();
  hipGraphExec_t tickExec_[2] = {nullptr, nullptr};  // ticks on P_ and on P2_
  double tickPh_[2] = {0.0, 0.0};                    // the phase they were captured with
  int error(double* out);
  int orthogonal(cplx<double>* P);
  int write_state(const cplx<double>* P);
  int overlap_top(const cplx<double>* bra, const cplx<double>* ket, cplx<double>* out);
  int output(double time, CmfResult& res);

  Engine& E;
  hipStream_t s_ = nullptr;
  int NN_ = 0;
  int64_t total_ = 0;
  cplx<double>*P_ = nullptr, *P0_ = nullptr, *P2_ = nullptr, *PS_ = nullptr, *X_ = nullptr, *D_ = nullptr;
  cplx<double>* Y_[5] = {nullptr};
  cplx<double>* YT_[7] = {nullptr};
  double* err_ = nullptr;
  double* err_h_ = nullptr;
  int64_t* rhoslot_ = nullptr;
  void* host_stage_ = nullptr;  // two halves of stage_bytes_, used in turn by the ticks
  void* dev_stage_ = nullptr;
  size_t stage_bytes_ = 0;
  hipEvent_t evStage_[2] = {nullptr, nullptr};  // the upload from each half has been read
  hipEvent_t evErr_ = nullptr;                   // the node errors have reached err_h_
  // the device integrators, the ring of their per-tick records, the host
  // copy of one record and its event
  BsNode* d_bs_ = nullptr;
  uint8_t* d_rec_ = nullptr;
  uint8_t* h_rec_ = nullptr;
  hipEvent_t evRec_ = nullptr;
  int slot_ = 0;  // the record entry the next tick boundary writes
  cplx<double>* sw_ = nullptr;
  OrthoItem* d_oi_ = nullptr;
  SwApply* d_sa_ = nullptr;
  std::vector<std::pair<std::pair<int64_t, int64_t>, std::pair<int64_t, int64_t>>> orthoRange_;
  // orthogonal() on node tensors too large for LDS: the node kernel works in
  // global memory and the parents take the square roots by grouped mode
  // products (per height and leg position, chained through swWork_)
  bool bigOrtho_ = false;
  struct SwLaunch {
    int64_t d0, ntasks;
    bool toState;  // the last leg of each parent: writes the state
  };
  std::vector<std::vector<SwLaunch>> swLaunch_;  // per height
  GmDesc* d_swgm_ = nullptr;
  cplx<double>* swWork_ = nullptr;
  // the overlap tree on such trees: per height, mode-product levels (the
  // children's overlap matrices on the ket, into the workspace) and one
  // grouped contraction with the bra
  struct OvLaunch {
    int kind;  // 0 mode products, 1 contraction
    int64_t d0, ntasks;
  };
  std::vector<std::vector<OvLaunch>> ovLaunch_;
  GmDesc* d_ovgm_ = nullptr;
  GcDesc* d_ovgc_ = nullptr;
  int64_t swWorkSize_ = 0;
  cplx<double>* ovWork_ = nullptr;  // when the overlap tree needs more than swWork_
  cplx<double>* spool_ = nullptr;
  std::vector<int64_t> sslot_;
  SwItem* d_ov_ = nullptr;
  std::vector<std::pair<int64_t, int64_t>> ovRange_;
  std::vector<int> ovNmax_;
  double* energy_ = nullptr;
  // eigenstates: previous iterate, top matrix and its eigenvectors
  // (UT at eigm_ + n^2), eigenvalues, and the launch items
  cplx<double>* PL_ = nullptr;
  cplx<double>* eigm_ = nullptr;
  double* eigv_ = nullptr;
  SwItem* d_hx_ = nullptr;
  SwApply* d_rot_ = nullptr;
  OrthoItem* d_on_ = nullptr;
  std::vector<NodeBS> bs_;
  std::vector<double> dtbs_;
  bool failed_ = false;
  size_t nbuild_ = 0, nderiv_ = 0, nticks_ = 0;
  std::vector<void*> dev_;
};

}  // namespace qtg
