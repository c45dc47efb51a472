// SOP application fitted into a tree of the state's shape, on the device
// (applyOperatorIteration, ApplyFirstOrder.cpp:43-147 of QuTree; see
// apply_fit.cpp).
#pragma once
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "tree_prog.hpp"

namespace qtg {

class ApplyFit : public TreeProg {
 public:
  explicit ApplyFit(Engine& e) : TreeProg(e) { name_ = "apply_fit"; }
  ~ApplyFit() override;
  int init();
  // psi (engine layout) in: the state; out: the fit of H|psi>.  The fit
  // starts from TensorTree(mt19937(seed), tree, false); at most max_iter
  // iterations; fids (cap) receives the fidelity after each; *iters their
  // number.
  int run(double* psi_host, uint64_t seed, int64_t max_iter, double* fids, int64_t cap, int64_t* iters);
  int64_t launches_first() const { return (int64_t)first_.size(); }
  int64_t launches_per_iteration() const { return (int64_t)iter_.size(); }
  bool graph_ = true;

 private:
  using Key = std::vector<std::pair<int, int>>;
  Key restrict_key(int l, int a, bool inside) const;
  Chain r_chain(int l, int a, int skip) const;
  int run_extra(const Step& st, hipStream_t s) override;
  int validate_extra(const Step& st, std::vector<Span>& rd, std::vector<Span>& wr) const override;

  std::vector<std::vector<char>> under_;  // node x mode
  std::vector<std::vector<int>> cid_;     // term x mode: canonical leaf matrix (-1: none)
  std::vector<int64_t> leafOff_;
  std::vector<char> inS_;                 // node in the operator's (tail-cut) sparse tree
  std::vector<std::map<Key, int64_t>> hmat_, hole_;
  bool single_ = false;  // one product: applied to the bottom tensors directly
  int64_t ketOff_ = 0;   // the state (ket) in work
  std::vector<Step> first_, iter_;
  double* fid_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool ready_ = false;
};

}  // namespace qtg
