// Device-resident engine for one tree and one SOP Hamiltonian: builds the
// work-item plans of HamiltonianRepresentation::build (density tree,
// represent, hole contraction, regularised rho inverse) and of
// LayerDerivative over all nodes, and launches them layer by layer.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace qtg {

namespace {

template <typename X>
int up(X** dst, const std::vector<X>& v) {
  *dst = nullptr;
  if (v.empty()) return QTG_OK;
  QTG_HIP_TRY(hipMalloc((void**)dst, sizeof(X) * v.size()));
  QTG_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
  return QTG_OK;
}

__global__ void to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// Holds the stream for about `ticks` periods of the constant-rate wall clock
// (bounded by an iteration cap, so it always ends).  With kernel timing on,
// an eager sweep starts behind it: the host enqueues the whole sweep while
// the device waits, and the events then time kernels, not host gaps.
__global__ void hold_kernel(uint64_t ticks) {
  const uint64_t t0 = wall_clock64();
  for (int i = 0; i < (1 << 20); ++i) {
    if (wall_clock64() - t0 >= ticks) break;
    __builtin_amdgcn_s_sleep(8);
  }
}

}  // namespace

Engine::~Engine() {
  invalidate_graph();
  release();
  for (auto e : evpool_) (void)hipEventDestroy(e);
  if (evFork_) (void)hipEventDestroy(evFork_);
  if (evHeev_) (void)hipEventDestroy(evHeev_);
  if (evRho_) (void)hipEventDestroy(evRho_);
  if (side_) (void)hipStreamDestroy(side_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::invalidate_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_h_) (void)hipGraphDestroy(graph_h_);
  graph_exec_ = nullptr;
  graph_h_ = nullptr;
}

void Engine::release() {
  for (void* p : device_buffers_)
    if (p) (void)hipFree(p);
  device_buffers_.clear();
  d_grad64_ = nullptr;
  d_escal_ = nullptr;
  d_elog_ = nullptr;
  elogCap_ = 0;
  deviceBytes_ = 0;
}

int Engine::alloc(void** p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
    set_error("engine: device allocation failed");
    return QTG_ENOMEM;
  }
  device_buffers_.push_back(*p);
  deviceBytes_ += (int64_t)std::max<size_t>(bytes, 16);
  return QTG_OK;
}

int Engine::create(const std::string& tree_text, int precision, bool host_only) {
  host_only_ = host_only;
  treeText_ = tree_text;
  try {
    tree_ = parse_tree(tree_text);
  } catch (std::exception& e) {
    set_error(std::string("tree: ") + e.what());
    return QTG_ESHAPE;
  }
  if (precision < 0 || precision > 2) {
    set_error("precision must be 0 (complex128), 1 (complex64 bulk) or 2 (complex64 bulk, FP64 large terms)");
    return QTG_EARG;
  }
  c64_ = precision != 0;
  precision_ = precision;
  nmax_ = 256;
  nspf_ = 1;
  for (auto& n : tree_.nodes) {
    nmax_ = std::max(nmax_, n.N);
    nspf_ = std::max(nspf_, n.n);
  }
  if (nspf_ > 64) {
    set_error("engine: more than 64 SPFs at a node is not supported");
    return QTG_ESHAPE;
  }
  return QTG_OK;
}

int Engine::materialize() {
  invalidate_graph();
  if (!stream_) QTG_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  if (!side_) QTG_HIP_TRY(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
  if (!evFork_) QTG_HIP_TRY(hipEventCreateWithFlags(&evFork_, hipEventDisableTiming));
  if (!evHeev_) QTG_HIP_TRY(hipEventCreateWithFlags(&evHeev_, hipEventDisableTiming));
  if (!evRho_) QTG_HIP_TRY(hipEventCreateWithFlags(&evRho_, hipEventDisableTiming));
  std::vector<DNode> dn;
  for (auto& n : tree_.nodes) {
    DNode d{};
    d.order = n.order();
    d.n = n.n;
    d.N = n.N;
    for (int i = 0; i < n.order(); ++i) d.dims[i] = n.dims[i];
    d.off = n.off;
    dn.push_back(d);
  }
  int rc = up(&d_nodes_, dn);
  if (rc) return rc;
  device_buffers_.push_back(d_nodes_);
  if ((rc = alloc((void**)&d_psi64_, sizeof(double) * 2 * tree_.total))) return rc;
  if ((rc = alloc((void**)&d_dpsi64_, sizeof(double) * 2 * tree_.total))) return rc;
  if ((rc = alloc((void**)&d_apply64_, sizeof(double) * 2 * tree_.total))) return rc;
  if (c64_ && (rc = alloc((void**)&d_psi32_, sizeof(float) * 2 * tree_.total))) return rc;
  if ((rc = alloc((void**)&d_energy_, sizeof(double) * 4))) return rc;
  return QTG_OK;
}

int Engine::set_hamiltonian(int64_t nterms, const double* coeff, const int64_t* term_off, const int64_t* fmode,
                            const int64_t* fkind, const int64_t* fpool, const double* pool) {
  std::string err;
  if (build_sop(tree_, nterms, coeff, term_off, fmode, fkind, fpool, pool, sop_, err)) {
    set_error("hamiltonian: " + err);
    return QTG_EARG;
  }
  apply_tuning();
  return plan();
}

// Active node sets: union of leaf-to-root paths of each term (SparseTree with tail).
void Engine::active_sets() {
  const int L = (int)sop_.terms.size();
  active_.assign(L, std::vector<char>(tree_.nodes.size(), 0));
  for (int l = 0; l < L; ++l)
    for (auto& lm : sop_.terms[l].leafmats) {
      int a = tree_.leafNode[lm.first];
      while (a >= 0) {
        active_[l][a] = 1;
        a = tree_.nodes[a].parent;
      }
    }
}

// Term restrictions: the factors of term l on the leaves inside (or outside)
// the subtree of node a.  An h-matrix depends only on the inside restriction,
// a hole matrix only on the outside one, so equal restrictions share a slot.
std::vector<std::pair<int, int>> Engine::restrict_in(int l, int a) const {
  std::vector<std::pair<int, int>> r;
  for (auto& lm : sop_.terms[l].leafmats)
    if (inSub_[a][lm.first]) r.push_back(lm);
  return r;
}

std::vector<std::pair<int, int>> Engine::restrict_out(int l, int a) const {
  std::vector<std::pair<int, int>> r;
  for (auto& lm : sop_.terms[l].leafmats)
    if (!inSub_[a][lm.first]) r.push_back(lm);
  return r;
}

int Engine::plan() {
  active_sets();
  const int L = (int)sop_.terms.size();
  const int NN = (int)tree_.nodes.size();
  const int F = (int)tree_.leafNode.size();
  inSub_.assign(NN, std::vector<char>(F, 0));
  for (auto& v : tree_.nodes) {
    if (v.bottom) inSub_[v.addr][v.leaf.mode] = 1;
    else
      for (int c : v.children)
        for (int m = 0; m < F; ++m)
          if (inSub_[c][m]) inSub_[v.addr][m] = 1;
  }
  for (auto& v : tree_.nodes)
    if ((v.bottom ? 1 : (int)v.children.size()) + (v.top() ? 0 : 1) > kMaxFactors) {
      set_error("engine: nodes with more than two children are not supported");
      return QTG_ESHAPE;
    }
  // FLOPs of the reference algorithm (one represent and one hole contraction
  // per (term, active non-top node), one density item per non-top node, Apply
  // per (term, active node), projector and rho^-1 per non-top node), the
  // count a sweep is measured against whatever this plan shares or merges.
  refFlops_ = 0.0;
  for (int l = 0; l < L; ++l)
    for (auto& v : tree_.nodes) {
      if (!active_[l][v.addr]) continue;
      double ch = 0.0;
      if (v.bottom) ch = 8.0 * v.N * v.dims[0];
      else
        for (int c : v.children)
          if (active_[l][c]) ch += 8.0 * v.N * tree_.nodes[c].n;
      if (v.top()) {
        refFlops_ += ch;
        continue;
      }
      refFlops_ += 2.0 * (ch + 8.0 * v.N * v.n);  // represent + apply
      const HNode& p = tree_.nodes[v.parent];
      for (int sb : p.children)
        if (sb != v.addr && active_[l][sb]) refFlops_ += 8.0 * p.N * tree_.nodes[sb].n;
      if (!p.top()) refFlops_ += 8.0 * p.N * p.n;
      refFlops_ += 8.0 * p.N * v.n;
    }
  for (auto& v : tree_.nodes) {
    if (v.top()) continue;
    const HNode& p = tree_.nodes[v.parent];
    refFlops_ += 8.0 * p.N * v.n + (p.top() ? 0.0 : 8.0 * p.N * p.n) + 4 * 8.0 * v.N * v.n;
  }
  auto nn = [&](int a) { return (int64_t)tree_.nodes[a].n * tree_.nodes[a].n; };
  // ---- pool layout (complex elements): leaf matrices, h slots, hole slots
  int64_t off = 0;
  leafOff_.clear();
  for (auto& m : sop_.leafMatrices) {
    leafOff_.push_back(off);
    off += (int64_t)m.size();
  }
  hslot_.assign(L, std::vector<int64_t>(NN, -1));
  holeslot_.assign(L, std::vector<int64_t>(NN, -1));
  rhoslot_.assign(NN, -1);
  // per node: (slot, representative term) of every distinct h / hole matrix
  std::vector<std::vector<std::pair<int64_t, int>>> hent(NN), qent(NN);
  typedef std::vector<std::pair<int, int>> Key;
  std::vector<std::map<Key, int64_t>> hkey(NN), qkey(NN);
  // large-coefficient terms (|c| >= 0.1 max |c|): their slots run in FP64 in
  // the split-precision form
  double cmax = 0.0;
  for (auto& tm : sop_.terms) cmax = std::max(cmax, std::abs(tm.coeff));
  std::vector<char> big(L, 0);
  for (int l = 0; l < L; ++l) big[l] = std::abs(sop_.terms[l].coeff) >= 0.1 * cmax;
  highSlots_.clear();
  for (int a = 0; a < NN; ++a) {
    const HNode& v = tree_.nodes[a];
    if (v.top()) continue;
    // density matrix = hole of the empty outside restriction
    rhoslot_[a] = off;
    highSlots_.insert(off);
    if (compress_) qkey[a][Key()] = off;
    qent[a].push_back({off, -1});
    off += nn(a);
    for (int l = 0; l < L; ++l) {
      if (!active_[l][a]) continue;
      if (compress_) {
        Key ki = restrict_in(l, a), ko = restrict_out(l, a);
        auto hi = hkey[a].find(ki);
        if (hi == hkey[a].end()) {
          hkey[a][ki] = off;
          hent[a].push_back({off, l});
          hslot_[l][a] = off;
          off += nn(a);
        } else {
          hslot_[l][a] = hi->second;
        }
        if (big[l]) highSlots_.insert(hslot_[l][a]);
        auto qi = qkey[a].find(ko);
        if (qi == qkey[a].end()) {
          qkey[a][ko] = off;
          qent[a].push_back({off, l});
          holeslot_[l][a] = off;
          off += nn(a);
        } else {
          holeslot_[l][a] = qi->second;
        }
        if (big[l]) highSlots_.insert(holeslot_[l][a]);
      } else {
        hslot_[l][a] = off;
        hent[a].push_back({off, l});
        off += nn(a);
        holeslot_[l][a] = off;
        qent[a].push_back({off, l});
        off += nn(a);
        if (big[l]) {
          highSlots_.insert(hslot_[l][a]);
          highSlots_.insert(holeslot_[l][a]);
        }
      }
    }
  }
  int64_t off64 = 0;
  rhoinv_.assign(NN, -1);
  for (int a = 0; a < NN; ++a) {
    rhoinv_[a] = off64;
    off64 += nn(a);
  }
  pool64Size_ = off64;

  auto leafmat = [&](int l, int mode) -> int64_t {
    for (auto& lm : sop_.terms[l].leafmats)
      if (lm.first == mode) return leafOff_[lm.second];
    return -1;
  };
  flops_ = 0.0;
  layerFlops_.clear();
  layerNmax_.clear();
  for (auto& x : phaseFlops_) x = 0.0;
  double mark = 0.0;
  // ---- represent: bottom-up by height
  std::vector<SwItem> items;
  repRange_.assign(tree_.maxHeight, {0, 0});
  for (int h = 0; h < tree_.maxHeight; ++h) {
    repRange_[h].first = (int64_t)items.size();
    for (auto& v : tree_.nodes) {
      if (v.height != h || v.top()) continue;
      for (auto& e : hent[v.addr]) {
        const int l = e.second;
        SwItem it{};
        it.node = v.addr;
        it.k = v.order() - 1;
        it.bra = it.ket = v.off;
        it.out = e.first;
        if (v.bottom) {
          it.fmode[it.nf] = 0;
          it.fmat[it.nf++] = leafmat(l, v.leaf.mode);
          flops_ += 8.0 * v.N * v.dims[0];
        } else {
          for (int c : v.children)
            if (active_[l][c]) {
              it.fmode[it.nf] = tree_.nodes[c].childIdx;
              it.fmat[it.nf++] = hslot_[l][c];
              flops_ += 8.0 * v.N * tree_.nodes[c].n;
            }
        }
        flops_ += 8.0 * v.N * v.n;
        items.push_back(it);
      }
    }
    repRange_[h].second = (int64_t)items.size();
    layerNmax_.push_back(256);
    for (int64_t i = repRange_[h].first; i < repRange_[h].second; ++i)
      layerNmax_.back() = std::max(layerNmax_.back(), tree_.nodes[items[i].node].N);
    layerFlops_.push_back(flops_ - mark);
    mark = flops_;
  }
  // ---- holes and density matrices: top-down by depth
  holeRange_.assign(tree_.maxDepth + 1, {0, 0});
  for (int dd = 1; dd <= tree_.maxDepth; ++dd) {
    holeRange_[dd].first = (int64_t)items.size();
    for (auto& v : tree_.nodes) {
      if (v.depth != dd) continue;
      const HNode& p = tree_.nodes[v.parent];
      for (auto& e : qent[v.addr]) {
        const int l = e.second;  // -1: density matrix
        SwItem it{};
        it.node = p.addr;
        it.k = v.childIdx;
        it.bra = it.ket = p.off;
        it.out = e.first;
        if (l >= 0)
          for (int sb : p.children)
            if (sb != v.addr && active_[l][sb]) {
              it.fmode[it.nf] = tree_.nodes[sb].childIdx;
              it.fmat[it.nf++] = hslot_[l][sb];
              flops_ += 8.0 * p.N * tree_.nodes[sb].n;
            }
        if (!p.top()) {
          it.fmode[it.nf] = p.order() - 1;
          it.fmat[it.nf++] = (l >= 0) ? holeslot_[l][p.addr] : rhoslot_[p.addr];
          flops_ += 8.0 * p.N * p.n;
        }
        flops_ += 8.0 * p.N * v.n;
        items.push_back(it);
      }
    }
    holeRange_[dd].second = (int64_t)items.size();
    layerNmax_.push_back(256);
    for (int64_t i = holeRange_[dd].first; i < holeRange_[dd].second; ++i)
      layerNmax_.back() = std::max(layerNmax_.back(), tree_.nodes[items[i].node].N);
    layerFlops_.push_back(flops_ - mark);
    mark = flops_;
  }
  nSw_ = (int64_t)items.size();
  phaseFlops_[0] = flops_;
  // ---- density eigen / inverse
  // density eigenproblems, largest first (one workgroup each)
  std::vector<EigItem> eig;
  std::vector<const HNode*> bySize;
  for (auto& v : tree_.nodes) bySize.push_back(&v);
  std::stable_sort(bySize.begin(), bySize.end(), [](const HNode* a, const HNode* b) { return a->n > b->n; });
  for (const HNode* vp : bySize) {
    const HNode& v = *vp;
    if (v.top()) continue;
    if (v.n > 64) {
      set_error("engine: more than 64 SPFs at a node is not supported");
      return QTG_ESHAPE;
    }
    EigItem e{};
    e.n = v.n;
    e.in = rhoslot_[v.addr];
    e.out = rhoinv_[v.addr];
    e.ev = -1;
    eig.push_back(e);
  }
  // ---- Apply at every node.  Each active term is a Kronecker product over
  // the node's slots (children or leaf operator, then the hole on the last
  // mode).  With compression, terms that agree on all slots but one are
  // merged into one summand whose free slot is the coefficient-weighted sum
  // of their matrices (greedy, largest group first).
  std::vector<ApTerm> terms;
  std::vector<ApChunk> chunks;
  std::vector<FinNode> fin;
  std::vector<CombOut>& combOut = h_comb1_;
  std::vector<CombOut>& combOut2 = h_comb2_;
  std::vector<CombTerm>& combTerm = h_combt1_;
  std::vector<CombTerm>& combTerm2 = h_combt2_;
  combOut.clear();
  combOut2.clear();
  combTerm.clear();
  combTerm2.clear();
  double combFlops = 0.0;
  int64_t part = 0;
  summ_.assign(NN, {});
  smode_.assign(NN, {});
  sdim_.assign(NN, {});
  for (auto& v : tree_.nodes) {
    FinNode fn{};
    fn.node = v.addr;
    fn.top = v.top() ? 1 : 0;
    fn.part0 = part;
    fn.rhoinv = rhoinv_[v.addr];
    // slots of node v
    std::vector<int> smode, sdim;
    if (v.bottom) {
      smode.push_back(0);
      sdim.push_back(v.dims[0]);
    } else {
      for (int c : v.children) {
        smode.push_back(tree_.nodes[c].childIdx);
        sdim.push_back(tree_.nodes[c].n);
      }
    }
    if (!v.top()) {
      smode.push_back(v.order() - 1);
      sdim.push_back(v.n);
    }
    const int S = (int)smode.size();
    std::vector<Summand> tl;
    for (int l = 0; l < L; ++l) {
      if (!active_[l][v.addr]) continue;
      Summand g;
      g.coef = sop_.terms[l].coeff;
      if (v.bottom) {
        g.ids.push_back(leafmat(l, v.leaf.mode));
      } else {
        for (int c : v.children) g.ids.push_back(active_[l][c] ? hslot_[l][c] : -1);
      }
      if (!v.top()) g.ids.push_back(holeslot_[l][v.addr]);
      tl.push_back(g);
    }
    std::vector<Summand> groups = compress_ ? aggregate(tl, sdim, off, combFlops) : tl;
    summ_[v.addr] = groups;
    smode_[v.addr] = smode;
    sdim_[v.addr] = sdim;
    std::vector<ApTerm> vt;
    for (auto& g : groups) {
      ApTerm t{};
      t.cre = g.coef.real();
      t.cim = g.coef.imag();
      for (int sl = 0; sl < S; ++sl) {
        if (g.ids[sl] < 0) continue;
        t.fmode[t.nf] = smode[sl];
        t.fmat[t.nf++] = g.ids[sl];
        flops_ += 8.0 * v.N * sdim[sl];
      }
      vt.push_back(t);
    }
    if (!v.top()) flops_ += 4 * 8.0 * v.N * v.n;  // projector and rho^-1
    const int per = compress_ ? chunkTermsCompressed_ : chunkTerms_;
    int nch = 0;
    for (size_t t0 = 0; t0 < vt.size() || (t0 == 0 && vt.empty()); t0 += per) {
      ApChunk c{};
      c.node = v.addr;
      c.t0 = (int32_t)(terms.size() + t0);
      c.t1 = (int32_t)(terms.size() + std::min(vt.size(), t0 + per));
      c.out = part;
      part += v.N;
      chunks.push_back(c);
      ++nch;
      if (vt.empty()) break;
    }
    terms.insert(terms.end(), vt.begin(), vt.end());
    fn.nchunk = nch;
    fin.push_back(fn);
  }
  poolSize_ = off;
  nChunks_ = (int64_t)chunks.size();
  nFin_ = (int64_t)fin.size();
  nComb_ = (int64_t)combOut.size();
  nComb2_ = (int64_t)combOut2.size();
  combFlops12_ = combFlops;
  {
    double proj = 0.0;
    for (auto& v : tree_.nodes)
      if (!v.top()) proj += 4 * 8.0 * v.N * v.n;
    phaseFlops_[2] = flops_ - phaseFlops_[0] - proj;
    phaseFlops_[3] = proj;
    phaseFlops_[5] = combFlops;
    flops_ += combFlops;
  }
  nEig_ = (int64_t)eig.size();
  partialSize_ = part;
  h_items_ = items;
  h_eig_ = eig;
  h_terms_ = terms;
  h_chunks_ = chunks;
  h_fin_ = fin;
  lflops_ = flops_;
  if (path_ == 0) {
    const size_t es = c64_ ? 8 : 16;
    if (precision_ == 2) {
      set_error("engine: the split-precision form runs on the grouped-contraction path");
      return QTG_EARG;
    }
    if (work_lds_bytes(nmax_, es) > kLdsBudget || finish_lds_bytes(nmax_, nspf_) > kLdsBudget || nspf_ > 64) {
      set_error("engine: node tensors of " + std::to_string(nmax_) + " elements / " + std::to_string(nspf_) +
                " SPFs exceed the LDS-resident kernels; use the grouped-contraction path");
      return QTG_ESHAPE;
    }
  } else {
    int grc = plan_gemm();
    if (grc) return grc;
  }
  if (host_only_) {
    planned_ = true;
    return QTG_OK;
  }
  release();
  int rc = materialize();
  if (rc) return rc;
  if ((rc = up(&d_sw_, items))) return rc;
  device_buffers_.push_back(d_sw_);
  if ((rc = up(&d_eig_, eig))) return rc;
  device_buffers_.push_back(d_eig_);
  if ((rc = up(&d_terms_, terms))) return rc;
  device_buffers_.push_back(d_terms_);
  if ((rc = up(&d_chunks_, chunks))) return rc;
  device_buffers_.push_back(d_chunks_);
  if ((rc = up(&d_fin_, fin))) return rc;
  device_buffers_.push_back(d_fin_);
  if ((rc = up(&d_combOut_, combOut))) return rc;
  device_buffers_.push_back(d_combOut_);
  if ((rc = up(&d_combTerm_, combTerm))) return rc;
  device_buffers_.push_back(d_combTerm_);
  if ((rc = up(&d_combOut2_, combOut2))) return rc;
  device_buffers_.push_back(d_combOut2_);
  if ((rc = up(&d_combTerm2_, combTerm2))) return rc;
  device_buffers_.push_back(d_combTerm2_);
  if (path_ == 1) {
    if ((rc = up(&d_gm_, h_gm_))) return rc;
    device_buffers_.push_back(d_gm_);
    if ((rc = up(&d_gc_, h_gc_))) return rc;
    device_buffers_.push_back(d_gc_);
    if ((rc = up(&d_hk_, h_hk_))) return rc;
    device_buffers_.push_back(d_hk_);
    if ((rc = up(&d_hi_, h_hi_))) return rc;
    device_buffers_.push_back(d_hi_);
    if ((rc = up(&d_solve_, h_solve_))) return rc;
    device_buffers_.push_back(d_solve_);
    if ((rc = up(&d_combOut3_, h_comb3_))) return rc;
    device_buffers_.push_back(d_combOut3_);
    if ((rc = up(&d_combTerm3_, h_combt3_))) return rc;
    device_buffers_.push_back(d_combTerm3_);
    for (auto pr : {std::make_pair(&d_combG1_, &h_combG1_), std::make_pair(&d_combG2_, &h_combG2_),
                    std::make_pair(&d_combG3_, &h_combG3_), std::make_pair(&d_combGE_, &h_combGE_)}) {
      if ((rc = up(pr.first, *pr.second))) return rc;
      device_buffers_.push_back(*pr.first);
    }
    for (auto pr : {std::make_pair(&d_combtG1_, &h_combtG1_), std::make_pair(&d_combtG2_, &h_combtG2_),
                    std::make_pair(&d_combtG3_, &h_combtG3_), std::make_pair(&d_combtGE_, &h_combtGE_)}) {
      if ((rc = up(pr.first, *pr.second))) return rc;
      device_buffers_.push_back(*pr.first);
    }
    if ((rc = up(&d_gmwg_, h_gmwg_))) return rc;
    device_buffers_.push_back(d_gmwg_);
    if ((rc = up(&d_gcwg_, h_gcwg_))) return rc;
    device_buffers_.push_back(d_gcwg_);
    if ((rc = up(&d_canon_, h_canon_))) return rc;
    device_buffers_.push_back(d_canon_);
  }
  // the grouped path keeps every matrix in FP64 (FP32 kernels round on load)
  const size_t es = (c64_ && path_ == 0) ? 8 : 16;
  if (path_ == 1 && (rc = alloc(&d_work_, 16 * std::max<int64_t>(workSize_, 1)))) return rc;
  if (path_ == 1 && c64_ && (rc = alloc(&d_work32_, 8 * std::max<int64_t>(work32Size_, 1)))) return rc;
  if ((rc = alloc(&d_pool_, es * poolSize_))) return rc;
  if ((rc = alloc((void**)&d_pool64_, 16 * pool64Size_))) return rc;
  if ((rc = alloc(&d_partial_, es * (path_ == 1 ? 1 : partialSize_)))) return rc;
  // leaf matrices (constant)
  std::vector<double> lm;
  for (auto& m : sop_.leafMatrices)
    for (auto& z : m) {
      lm.push_back(z.real());
      lm.push_back(z.imag());
    }
  auto put = [&](int64_t off, const std::vector<double>& v) -> int {
    if (v.empty()) return QTG_OK;
    if (es == 8) {
      std::vector<float> lf(v.begin(), v.end());
      QTG_HIP_TRY(hipMemcpy((char*)d_pool_ + 8 * off, lf.data(), sizeof(float) * lf.size(), hipMemcpyHostToDevice));
    } else {
      QTG_HIP_TRY(hipMemcpy((char*)d_pool_ + 16 * off, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    }
    return QTG_OK;
  };
  if ((rc = put(0, lm))) return rc;
  for (auto& id : idSlots_) {
    std::vector<double> e(2 * (size_t)id.second * id.second, 0.0);
    for (int i = 0; i < id.second; ++i) e[2 * ((size_t)i + (size_t)i * id.second)] = 1.0;
    if ((rc = put(id.first, e))) return rc;
  }
  planned_ = true;
  return QTG_OK;
}

// Greedy merge of Kronecker-product summands: terms that agree on every slot
// but one become one summand whose free slot is the coefficient-weighted sum
// of their matrices (largest group first).  The sums are combination work of
// the build; long sums are split into partial sums of <= 32 terms (first
// combine launch) that a second launch adds, so no workgroup walks a long
// serial chain of loads.
std::vector<Engine::Summand> Engine::aggregate(const std::vector<Summand>& tl, const std::vector<int>& sdim,
                                               int64_t& off, double& combFlops) {
  std::vector<Summand> groups;
  const int S = (int)sdim.size();
  std::vector<char> alive(tl.size(), 1);
  size_t nalive = tl.size();
  while (nalive > 1) {
    int best_s = -1;
    std::vector<int> best;
    for (int sl = 0; sl < S; ++sl) {
      std::map<std::vector<int64_t>, std::vector<int>> gm;
      for (size_t t = 0; t < tl.size(); ++t) {
        if (!alive[t]) continue;
        std::vector<int64_t> k = tl[t].ids;
        k.erase(k.begin() + sl);
        gm[k].push_back((int)t);
      }
      for (auto& kv : gm)
        if (kv.second.size() > std::max<size_t>(1, best.size())) {
          best = kv.second;
          best_s = sl;
        }
    }
    if (best_s < 0) break;
    const int64_t d = sdim[best_s];
    const size_t kCombChunk = 32;
    std::vector<int64_t> parts;
    for (size_t c0 = 0; c0 < best.size(); c0 += kCombChunk) {
      CombOut co{};
      co.out = off;
      co.n = (int32_t)d;
      co.t0 = (int32_t)h_combt1_.size();
      for (size_t q = c0; q < std::min(best.size(), c0 + kCombChunk); ++q) {
        const int t = best[q];
        CombTerm ct{};
        ct.cre = tl[t].coef.real();
        ct.cim = tl[t].coef.imag();
        ct.src = tl[t].ids[best_s];
        h_combt1_.push_back(ct);
        alive[t] = 0;
        --nalive;
      }
      co.t1 = (int32_t)h_combt1_.size();
      h_comb1_.push_back(co);
      for (int32_t q = co.t0; q < co.t1; ++q)
        if (h_combt1_[q].src >= 0 && highSlots_.count(h_combt1_[q].src)) highSlots_.insert(off);
      parts.push_back(off);
      off += d * d;
    }
    combFlops += 8.0 * d * d * best.size();
    int64_t result = parts[0];
    if (parts.size() > 1) {
      CombOut co{};
      co.out = off;
      co.n = (int32_t)d;
      co.t0 = (int32_t)h_combt2_.size();
      for (int64_t pidx : parts) {
        CombTerm ct{};
        ct.cre = 1.0;
        ct.cim = 0.0;
        ct.src = pidx;
        h_combt2_.push_back(ct);
      }
      co.t1 = (int32_t)h_combt2_.size();
      h_comb2_.push_back(co);
      combFlops += 8.0 * d * d * parts.size();
      for (int64_t pidx : parts)
        if (highSlots_.count(pidx)) highSlots_.insert(off);
      result = off;
      off += d * d;
    }
    Summand g;
    g.ids = tl[best[0]].ids;
    g.ids[best_s] = result;
    g.coef = cd(1.0, 0.0);
    groups.push_back(g);
  }
  for (size_t t = 0; t < tl.size(); ++t)
    if (alive[t]) groups.push_back(tl[t]);
  return groups;
}

int Engine::set_psi(const double* host, hipStream_t s) {
  QTG_HIP_TRY(hipMemcpyAsync(d_psi64_, host, sizeof(double) * 2 * tree_.total, hipMemcpyHostToDevice, s));
  QTG_HIP_TRY(hipStreamSynchronize(s));
  built_ = false;
  return QTG_OK;
}

int Engine::set_psi_device(const double* dev, hipStream_t s) {
  QTG_HIP_TRY(hipMemcpyAsync(d_psi64_, dev, sizeof(double) * 2 * tree_.total, hipMemcpyDeviceToDevice, s));
  built_ = false;
  return QTG_OK;
}

int Engine::set_timing(bool on) {
  timing_ = on;
  marks_.clear();
  evnext_ = 0;
  for (int i = 0; i < kPhases; ++i) tms_[i] = tfl_[i] = tby_[i] = 0.0, tn_[i] = 0;
  return QTG_OK;
}

int Engine::begin_mark(int phase, hipStream_t s) {
  if (!timing_) return QTG_OK;
  while (evpool_.size() < evnext_ + 2) {
    hipEvent_t e;
    QTG_HIP_TRY(hipEventCreate(&e));
    evpool_.push_back(e);
  }
  Mark m{evpool_[evnext_], evpool_[evnext_ + 1], phase};
  evnext_ += 2;
  QTG_HIP_TRY(hipEventRecord(m.a, s));
  marks_.push_back(m);
  return QTG_OK;
}

int Engine::kernel_mark(int phase, const KernelTiming** kt) {
  *kt = nullptr;
  if (!timing_) return QTG_OK;
  // fold finished marks first so the event pool stays bounded
  if (marks_.size() >= 512) {
    int rc = collect_marks();
    if (rc) return rc;
  }
  while (evpool_.size() < evnext_ + 2) {
    hipEvent_t e;
    QTG_HIP_TRY(hipEventCreate(&e));
    evpool_.push_back(e);
  }
  Mark m{evpool_[evnext_], evpool_[evnext_ + 1], phase};
  evnext_ += 2;
  marks_.push_back(m);
  ktiming_ = KernelTiming{m.a, m.b};
  *kt = &ktiming_;
  return QTG_OK;
}

int Engine::end_mark(hipStream_t s) {
  if (!timing_) return QTG_OK;
  QTG_HIP_TRY(hipEventRecord(marks_.back().b, s));
  // keep the event pool bounded: fold finished marks into the totals
  if (marks_.size() >= 512) return collect_marks();
  return QTG_OK;
}

int Engine::collect_marks() {
  for (auto& m : marks_) {
    QTG_HIP_TRY(hipEventSynchronize(m.b));
    float ms = 0.f;
    QTG_HIP_TRY(hipEventElapsedTime(&ms, m.a, m.b));
    tms_[m.phase] += ms;
    tn_[m.phase] += 1;
  }
  marks_.clear();
  evnext_ = 0;
  return QTG_OK;
}

int Engine::read_timing(double* ms, int64_t* launches, double* flops, double* bytes) {
  int rc = collect_marks();
  if (rc) return rc;
  for (int i = 0; i < kPhases; ++i) {
    if (ms) ms[i] = tms_[i];
    if (launches) launches[i] = tn_[i];
    if (flops) flops[i] = tfl_[i];
    if (bytes) bytes[i] = tby_[i];
  }
  return QTG_OK;
}

int launch_to_f32(const double* in, float* out, int64_t n, hipStream_t s) {
  if (n <= 0) return QTG_OK;
  hipLaunchKernelGGL(to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int Engine::convert_psi(hipStream_t s) {
  const int64_t n = 2 * tree_.total;
  int rc = begin_mark(4, s);
  if (rc) return rc;
  hipLaunchKernelGGL(to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_psi64_, d_psi32_, n);
  QTG_LAUNCH_CHECK();
  return end_mark(s);
}

template <typename T>
int Engine::build_t(hipStream_t s) {
  const void* psi = c64_ ? (const void*)d_psi32_ : (const void*)d_psi64_;
  int rc;
  if (c64_ && (rc = convert_psi(s))) return rc;
  size_t li = 0;
  auto sw = [&](const std::pair<int64_t, int64_t>& r) -> int {
    const double fl = li < layerFlops_.size() ? layerFlops_[li] : 0.0;
    const int lnmax = li < layerNmax_.size() ? layerNmax_[li] : nmax_;
    ++li;
    if (r.second == r.first) return QTG_OK;
    int e = begin_mark(0, s);
    if (e) return e;
    if ((e = launch_sandwich<T>(d_sw_ + r.first, r.second - r.first, d_nodes_, psi, psi, d_pool_, lnmax, s)))
      return e;
    if (timing_) tfl_[0] += fl;
    return end_mark(s);
  };
  for (auto& r : repRange_)
    if ((rc = sw(r))) return rc;
  for (size_t dd = 1; dd < holeRange_.size(); ++dd)
    if ((rc = sw(holeRange_[dd]))) return rc;
  if ((rc = begin_mark(1, s))) return rc;
  if ((rc = launch_heev<T>(d_eig_, nEig_, d_pool_, d_pool64_, eps_, nspf_, s))) return rc;
  if ((rc = end_mark(s))) return rc;
  // aggregated Apply matrices depend only on the h and hole matrices
  if (nComb_ > 0) {
    if ((rc = begin_mark(5, s))) return rc;
    if ((rc = launch_combine<T>(d_combOut_, nComb_, d_combTerm_, d_pool_, s))) return rc;
    if ((rc = launch_combine<T>(d_combOut2_, nComb2_, d_combTerm2_, d_pool_, s))) return rc;
    if ((rc = end_mark(s))) return rc;
    if (timing_) tfl_[5] += phaseFlops_[5];
  }
  built_ = true;
  return QTG_OK;
}

int Engine::build(hipStream_t s) {
  if (!planned_) {
    set_error("engine: no Hamiltonian");
    return QTG_ESTATE;
  }
  if (path_ == 1) return build_gemm(s);
  return c64_ ? build_t<float>(s) : build_t<double>(s);
}

template <typename T>
int Engine::apply_t(hipStream_t s, double ph_re, double ph_im) {
  const void* psi = c64_ ? (const void*)d_psi32_ : (const void*)d_psi64_;
  int rc = begin_mark(2, s);
  if (rc) return rc;
  if ((rc = launch_apply<T>(d_chunks_, nChunks_, d_terms_, d_nodes_, psi, d_pool_, d_partial_, nmax_, s))) return rc;
  if ((rc = end_mark(s))) return rc;
  if (timing_) tfl_[2] += phaseFlops_[2];
  if ((rc = begin_mark(3, s))) return rc;
  if ((rc = launch_finish<T>(d_fin_, nFin_, d_nodes_, d_psi64_, d_partial_, d_pool64_, d_dpsi64_, d_apply64_, ph_re,
                             ph_im, d_energy_, nmax_, nspf_, s)))
    return rc;
  if (timing_) tfl_[3] += phaseFlops_[3];
  return end_mark(s);
}

int Engine::layer_derivative_all(hipStream_t s, double ph_re, double ph_im) {
  if (!built_) {
    set_error("engine: build() first");
    return QTG_ESTATE;
  }
  if (path_ == 1) {
    return apply_gemm(s, d_psi64_, d_dpsi64_, nullptr, ph_re, ph_im);
  }
  return c64_ ? apply_t<float>(s, ph_re, ph_im) : apply_t<double>(s, ph_re, ph_im);
}

int Engine::derivative_eager(hipStream_t s, double ph_re, double ph_im) {
  if (timing_) {
    if (clockKHz_ == 0) {
      int dev = 0;
      QTG_HIP_TRY(hipGetDevice(&dev));
      QTG_HIP_TRY(hipDeviceGetAttribute(&clockKHz_, hipDeviceAttributeWallClockRate, dev));
    }
    hipLaunchKernelGGL(hold_kernel, dim3(1), dim3(64), 0, s, (uint64_t)clockKHz_ * 3);  // 3 ms
    QTG_LAUNCH_CHECK();
  }
  int rc = build(s);
  if (rc) return rc;
  return layer_derivative_all(s, ph_re, ph_im);
}

// One sweep = a fixed sequence of ~20 launches on fixed buffers, so it is
// captured once into a HIP graph and replayed (the phase is baked into the
// finish launch; a new phase re-captures).  Eager launches are used while
// kernel timing is on.
int Engine::derivative(hipStream_t s, double ph_re, double ph_im) {
  if (!planned_) {
    set_error("engine: no Hamiltonian");
    return QTG_ESTATE;
  }
  if (!graph_ || timing_ || s == nullptr) return derivative_eager(s, ph_re, ph_im);
  if (graph_exec_ && (ph_re != graph_ph_re_ || ph_im != graph_ph_im_)) invalidate_graph();
  if (!graph_exec_) {
    QTG_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = derivative_eager(s, ph_re, ph_im);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("graph capture: ") + hipGetErrorString(e));
      return QTG_EDEVICE;
    }
    graph_h_ = g;
    QTG_HIP_TRY(hipGraphInstantiate(&graph_exec_, graph_h_, nullptr, nullptr, 0));
    graph_ph_re_ = ph_re;
    graph_ph_im_ = ph_im;
  }
  QTG_HIP_TRY(hipGraphLaunch(graph_exec_, s));
  built_ = true;
  return QTG_OK;
}

int Engine::energy(hipStream_t s, double* out) {
  double e[4];
  const double* src = (path_ == 1) ? d_pool64_ + 2 * energySlot_ : d_energy_;
  QTG_HIP_TRY(hipMemcpyAsync(e, src, sizeof(e), hipMemcpyDeviceToHost, s));
  QTG_HIP_TRY(hipStreamSynchronize(s));
  // <A_top|Apply(A_top)> / <A_top|A_top>
  const cd num(e[0], e[1]), den(e[2], e[3]);
  *out = (num / den).real();
  return QTG_OK;
}

int Engine::get_matrix(int which, int term, int node, double* out, hipStream_t s) {
  if (node < 0 || node >= (int)tree_.nodes.size()) return QTG_EARG;
  const int n = tree_.nodes[node].n;
  int64_t off = -1;
  bool from64 = false;
  if (which == 0) off = rhoslot_[node];
  else if (which == 1) {
    off = rhoinv_[node];
    from64 = true;
  } else if (term >= 0 && term < (int)hslot_.size()) {
    off = (which == 2) ? hslot_[term][node] : holeslot_[term][node];
  }
  if (off < 0) return QTG_EARG;
  if (heevPending_) QTG_HIP_TRY(hipStreamWaitEvent(s, evHeev_, 0));  // rho^-1 of the side stream
  if (from64 || !c64_ || path_ == 1) {
    const void* base = from64 ? (const void*)d_pool64_ : d_pool_;
    QTG_HIP_TRY(hipMemcpyAsync(out, (const char*)base + 16 * off, 16 * (size_t)n * n, hipMemcpyDeviceToHost, s));
    QTG_HIP_TRY(hipStreamSynchronize(s));
  } else {
    std::vector<float> tmp(2 * (size_t)n * n);
    QTG_HIP_TRY(hipMemcpyAsync(tmp.data(), (const char*)d_pool_ + 8 * off, 8 * (size_t)n * n,
                               hipMemcpyDeviceToHost, s));
    QTG_HIP_TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < tmp.size(); ++i) out[i] = tmp[i];
  }
  return QTG_OK;
}

}  // namespace qtg
