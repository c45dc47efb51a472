// Fixed launch programs over one tree (tree_prog.cpp): the grouped
// mode-product / hole-contraction / combination launches of sweeps that
// visit nodes one after another (the SCF sweep, the fitted SOP
// application), planned once on the host and replayed.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "scf_ops.hpp"

namespace qtg {

class TreeProg {
 public:
  explicit TreeProg(Engine& e) : E(e) {}
  virtual ~TreeProg();
  int64_t device_bytes() const { return bytes_; }
  int64_t work_elems() const { return workSize_; }
  int64_t pool_elems() const { return poolSize_; }
  int64_t gm_tasks_total() const { return (int64_t)h_gm_.size(); }
  int64_t gc_tasks_total() const { return (int64_t)h_gc_.size(); }

 protected:
  using Chain = std::vector<std::pair<int, int64_t>>;  // (leg, pool offset); leg -1: copy
  enum Kind { K_GM, K_GC, K_COMB, K_LBEGIN, K_LSTEP, K_HEEV, K_KCOMB, K_PERM, K_QR, K_QRB, K_NORM };
  struct Step {
    int kind = K_GM;
    int64_t d0 = 0, n = 0;  // descriptor range (GM, GC: tasks; COMB: outputs; QRB: items)
    int obuf = 1;           // GM output base: 0 psi, 1 work
    int waves = 0;
    int64_t a = 0, b = 0, c = 0, d = 0;  // offsets (see the launching code)
    int abuf = 1, cbuf = 1;
    int64_t N = 0, bef = 0, nn = 0, aft = 0;
    int j = 0, k = 0, dir = 0, slot = 0;
  };
  struct StackEntry {
    Chain chain;
    std::vector<CombTerm> block;
  };
  // one node's chains: requests (chain, destination work offset or -1)
  // on the node tensor at `in` of base inbuf (0 psi, 1 work)
  struct ChainGroup {
    int a, inbuf;
    int64_t in;
    std::vector<std::pair<Chain, int64_t>> reqs;
  };
  struct Span {
    int buf;  // 0 psi, 1 work, 2 pool
    int64_t lo, hi;
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
  void push_comb_sized(std::vector<Step>& P, const std::vector<std::pair<int64_t, std::vector<CombTerm>>>& outs,
                       const std::vector<int>& n);
  // Mode-product chains of several nodes: depth d of every chain of every
  // group is one grouped launch; within a group, equal prefixes are shared.
  // Returns per group the output offset (work) of every request.
  std::vector<std::vector<int64_t>> emit_chains_multi(std::vector<Step>& P, const std::vector<ChainGroup>& groups);
  std::vector<int64_t> emit_chains(std::vector<Step>& P, int a, int inbuf, int64_t in,
                                   const std::vector<std::pair<Chain, int64_t>>& reqs);
  // The stacked product sum_e block_e x_last chain_e(input) into `out`
  // (work): emits the chains into body and returns the product's
  // descriptor in d; the blocks go to the identity stacks (all identity,
  // at most 4) or are assembled into comb at sblock (the shared block
  // region S_ when sblock < 0).  Returns QTG_ESTATE with no entries.
  int stack_desc(std::vector<Step>& comb, std::vector<Step>& body, int a, int inbuf, int64_t in,
                 std::vector<StackEntry>& entries, bool entry0_is_input, int64_t stack_base, int64_t out,
                 GmDesc& d, int64_t sblock = -1);
  int emit_stack(std::vector<Step>& comb, std::vector<Step>& body, int a, int inbuf, int64_t in,
                 std::vector<StackEntry>& entries, bool entry0_is_input, int64_t stack_base, int64_t out);
  int64_t qr_item(int64_t a, int64_t q, int64_t rows, int n);
  int run_steps(const std::vector<Step>& P, hipStream_t s);
  virtual int run_extra(const Step& st, hipStream_t s);
  int validate_steps(const std::vector<Step>& P) const;
  virtual int validate_extra(const Step& st, std::vector<Span>& rd, std::vector<Span>& wr) const;
  // allocates psi (the tree's arena), work, pool and the descriptor tables,
  // uploads the pool image (zeros, then the given complex entries) and the
  // descriptors
  int materialize(const std::vector<std::pair<int64_t, const std::vector<cd>*>>& poolInit);
  int dmalloc(void** p, size_t bytes);
  std::string name_ = "tree program";

  Engine& E;
  std::map<std::pair<int, int>, int64_t> ident_;  // (n, 4) -> four identity blocks
  int64_t poolSize_ = 0, S_ = 0, Smax_ = 0;
  int64_t progBase_ = 0, progTop_ = 0, progMax_ = 0, workSize_ = 0;
  std::vector<GmDesc> h_gm_;
  std::vector<GcDesc> h_gc_;
  std::vector<CombOut> h_co_;
  std::vector<CombTerm> h_ct_;
  std::vector<SmItem> h_qr_;
  hipStream_t s_ = nullptr;
  double *psi_ = nullptr, *work_ = nullptr, *pool_ = nullptr;
  GmDesc* d_gm_ = nullptr;
  GcDesc* d_gc_ = nullptr;
  CombOut* d_co_ = nullptr;
  CombTerm* d_ct_ = nullptr;
  SmItem* d_qr_ = nullptr;
  std::vector<void*> owned_;
  int64_t bytes_ = 0;
};

}  // namespace qtg
