// Krylov-sweep SCF ground-state search on the device (SCF.cpp:36-348 of
// QuTree, see scf.cpp).
#pragma once
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "tree_prog.hpp"

namespace qtg {

class Scf : public TreeProg {
 public:
  explicit Scf(Engine& e) : TreeProg(e) { name_ = "scf"; }
  ~Scf() override;
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
  bool graph_ = true;  // replay iterations from a captured HIP graph

 private:
  struct ApplyProg {
    int64_t base = 0, N = 0;       // stack (entry 0 = the input vector)
    std::vector<Step> comb, body;  // block assembly; chains + stacked product
  };
  using Key = std::vector<std::pair<int, int>>;
  Key restrict_key(int l, int a, bool inside) const;
  Chain r_chain(int l, int a, int skip) const;
  const ApplyProg& apply_prog(int a);
  const std::vector<Step>& up_prog(int a);
  const std::vector<Step>& down_prog(int a);
  int emit_microstep(int a, int next, int slot);
  int run_extra(const Step& st, hipStream_t s) override;
  int validate_extra(const Step& st, std::vector<Span>& rd, std::vector<Span>& wr) const override;

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
  int64_t tmpGC_ = 0, rSlot_ = 0, T_ = 0, U_ = 0;
  // work layout
  int64_t Nmax_ = 0, kryOff_ = 0, haOff_ = 0, psiwOff_ = 0, permOff_ = 0, pushOff_ = 0;
  // programs
  std::map<int, ApplyProg> apply_;
  std::map<int, std::vector<Step>> up_, down_;
  std::vector<Step> init_, iter_;
  // device (besides TreeProg's buffers)
  double *ev_ = nullptr, *scal_ = nullptr, *elog_ = nullptr;
  ScfParams* par_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool ready_ = false;
};

}  // namespace qtg
