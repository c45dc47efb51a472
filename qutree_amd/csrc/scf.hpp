// Krylov-sweep SCF ground-state search on the device (SCF.cpp:36-348 of
// QuTree, see scf.cpp).
#pragma once
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "scf_ops.hpp"

namespace qtg {

class Scf {
 public:
  explicit Scf(Engine& e) : E(e) {}
  ~Scf();
  // plans the sweep for Krylov spaces of n_krylov vectors (capped per node
  // at the node's size); the engine's tree and Hamiltonian are used
  int init(int64_t n_krylov);
  // scf(): n_iter iterations, the first n_itp of them imaginary-time
  // updates with step beta; psi (engine layout) in and out; energies
  // (n_iter x sweep length, at most cap) receives the lowest Ritz value of
  // every microstep
  int run(double* psi_host, int64_t n_iter, int64_t n_itp, double beta, double* energies, int64_t cap);
  int64_t sweep_length() const { return (int64_t)sweep_.size(); }
  int64_t launches_per_iteration() const { return (int64_t)iter_.size(); }
  int64_t device_bytes() const { return bytes_; }
  // plan sizes: work and pool elements, GM and GC tasks
  int64_t work_elems() const { return workSize_; }
  int64_t pool_elems() const { return poolSize_; }
  int64_t gm_tasks_total() const { return (int64_t)h_gm_.size(); }
  int64_t gc_tasks_total() const { return (int64_t)h_gc_.size(); }
  bool graph_ = true;  // replay iterations from a captured HIP graph

 private:
  using Chain = std::vector<std::pair<int, int64_t>>;  // (leg, pool offset); leg -1: copy
  enum Kind { K_GM, K_GC, K_COMB, K_LBEGIN, K_LSTEP, K_HEEV, K_KCOMB, K_PERM, K_QR };
  struct Step {
    int kind = K_GM;
    int64_t d0 = 0, n = 0;  // descriptor range (GM, GC: tasks; COMB: outputs)
    int obuf = 1;           // GM output base: 0 psi, 1 work
    int waves = 0;
    int64_t a = 0, b = 0, c = 0, d = 0;  // offsets (see scf.cpp)
    int abuf = 1, cbuf = 1;
    int64_t N = 0, bef = 0, nn = 0, aft = 0;
    int j = 0, k = 0, dir = 0, slot = 0;
  };
  struct StackEntry {
    Chain chain;
    std::vector<CombTerm> block;
  };
  struct ApplyProg {
    int64_t base = 0, N = 0;       // stack (entry 0 = the input vector)
    std::vector<Step> comb, body;  // block assembly; chains + stacked product
  };

  int64_t palloc(int64_t n);
  int64_t walloc(int64_t n);
  int64_t leg_bef(int a, int k) const;
  int64_t leg_aft(int a, int k) const;
  GmDesc gm_desc(int a, int leg, int64_t mat, int xsel, int64_t x, int64_t out) const;
  GcDesc gc_desc(int a, int leg, int xsel, int64_t x, int ysel, int64_t y, int64_t out) const;
  void push_gm(std::vector<Step>& P, std::vector<GmDesc>& ds, int obuf);
  void push_gc(std::vector<Step>& P, std::vector<GcDesc>& ds);
  void push_comb(std::vector<Step>& P, const std::vector<std::pair<int64_t, std::vector<CombTerm>>>& outs, int n);
  std::vector<int64_t> emit_chains(std::vector<Step>& P, int a, int inbuf, int64_t in,
                                   const std::vector<std::pair<Chain, int64_t>>& reqs);
  // sum over entries of block_e x_last chain_e(input) into out (work); the
  // blocks are assembled into comb (dynamic) or once at init (identity)
  int emit_stack(std::vector<Step>& comb, std::vector<Step>& body, int a, int inbuf, int64_t in,
                 std::vector<StackEntry>& entries, bool entry0_is_input, int64_t stack_base, int64_t out);
  Chain r_chain(int l, int a, int skip) const;
  const ApplyProg& apply_prog(int a);
  const std::vector<Step>& up_prog(int a);
  const std::vector<Step>& down_prog(int a);
  int emit_microstep(int a, int next, int slot);
  int run_steps(const std::vector<Step>& P, hipStream_t s);
  // every access of every step inside its buffer, no two outputs of one
  // launch overlapping, no launch writing what it reads
  int validate() const;
  int validate_steps(const std::vector<Step>& P) const;
  using Key = std::vector<std::pair<int, int>>;
  Key restrict_key(int l, int a, bool inside) const;

  Engine& E;
  int64_t kry_ = 5;
  std::vector<int> sweep_;
  // model
  std::vector<std::vector<char>> under_;    // node x mode
  std::vector<std::vector<int>> cid_;       // term x mode: canonical leaf matrix (-1: none)
  std::vector<int64_t> leafOff_;            // canonical leaf matrix -> pool
  std::vector<std::vector<char>> act_;      // term x node: tail-cut sparse tree
  std::vector<std::vector<int>> nact_;      // node x term
  std::vector<std::map<Key, int64_t>> hmat_, hole_;
  std::vector<int64_t> hcorr_, hconcorr_;
  std::map<std::pair<int, int>, int64_t> ident_;  // (n, blocks) -> identity stack
  int64_t poolSize_ = 0, tmpGC_ = 0, rSlot_ = 0, T_ = 0, U_ = 0, S_ = 0, Smax_ = 0;
  // work layout
  int64_t Nmax_ = 0, kryOff_ = 0, haOff_ = 0, psiwOff_ = 0, permOff_ = 0, pushOff_ = 0, progBase_ = 0;
  int64_t progTop_ = 0, progMax_ = 0, workSize_ = 0;
  // programs
  std::map<int, ApplyProg> apply_;
  std::map<int, std::vector<Step>> up_, down_;
  std::vector<Step> init_, iter_;
  std::vector<GmDesc> h_gm_;
  std::vector<GcDesc> h_gc_;
  std::vector<CombOut> h_co_;
  std::vector<CombTerm> h_ct_;
  // device
  hipStream_t s_ = nullptr;
  double *psi_ = nullptr, *work_ = nullptr, *pool_ = nullptr, *ev_ = nullptr, *scal_ = nullptr, *elog_ = nullptr;
  ScfParams* par_ = nullptr;
  GmDesc* d_gm_ = nullptr;
  GcDesc* d_gc_ = nullptr;
  CombOut* d_co_ = nullptr;
  CombTerm* d_ct_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  int64_t bytes_ = 0;
  bool ready_ = false;
};

}  // namespace qtg
