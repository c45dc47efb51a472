// SOP application fitted into a tree of the state's shape, on the device.
//
// applyOperatorIteration (ApplyFirstOrder.cpp:43-147 of QuTree) fits
// |HPsi> ~ H|Psi> with the tree shape of Psi by alternating sweeps on the
// operator's tail-cut sparse tree S (SparseTree(sop, tree, false)):
//   - h-matrices <HPsi|h_l|Psi> of every term at every node of S below its
//     top (overlaps where the term has no mode below), bottom-up;
//   - hole matrices of every term, top-down (contraction with rho);
//   - a bottom-up sweep: each node becomes the thin Q of
//     sum_l hole_l x_last (neighbour matrices of term l) Psi (applyOperatorLocal),
//     after which its h-matrices are formed again with the new node
//     (representLayer); the top node becomes the normalised sum of the
//     terms' products (applyTop + gramSchmidt);
//   - the fidelity |<HPsi|H Psi>|^2 (fidelity, TreeApplyOperator.cpp:22-35),
//     and the stop when 1 - f < 1e-6 or f rose by less than 1e-10.
// Both reference files sum the bare products sop(i): the SOP's coefficients
// are not applied.  A single-product operator is applied to the bottom
// tensors directly (applyReference).
//
// Device form: the same grouped launches as the SCF sweep (tree_prog.cpp).
// Nodes of one height are independent in the sweep (a node needs only its
// children's new h-matrices), so each height is one set of launches: the
// chains of every node and term prefix (one grouped launch per depth), one
// stacked product per node against its hole blocks (all nodes in one
// launch), one batched QR, and the new h-matrices as hole contractions of
// the new node against the same chain products.  The h-matrices that start
// an iteration are those the previous sweep left (the same products of the
// same tensors), so only the first iteration forms them.  The fidelity is
// |U|^2 of the top node's sum U before its normalisation: with the sweep's
// h-matrices the fidelity's representation is the same, and
// <U / |U| | U> = |U|.  Restricted to operators whose sparse tree reaches
// the top node, with one state there.
#include "apply_fit.hpp"

#include <algorithm>
#include <cstring>
#include <string>

namespace qtg {

ApplyFit::~ApplyFit() {
  if (exec_) hipGraphExecDestroy(exec_);
}

ApplyFit::Key ApplyFit::restrict_key(int l, int a, bool inside) const {
  Key k;
  const auto& c = cid_[l];
  for (int m = 0; m < (int)c.size(); ++m)
    if (c[m] >= 0 && (under_[a][m] != 0) == inside) k.push_back({m, c[m]});
  return k;
}

// apply(mats, Phi, M, node) restricted to the children (all nodes of S
// below a carry h-matrices, overlaps where the term has no mode): the
// children's matrices in child order (skip: a child index left out), or the
// term's leaf matrix at a bottom node (none when it does not act there)
ApplyFit::Chain ApplyFit::r_chain(int l, int a, int skip) const {
  const HNode& nd = E.tree().nodes[a];
  Chain c;
  if (nd.bottom) {
    if (cid_[l][nd.leaf.mode] >= 0) c.push_back({0, leafOff_[cid_[l][nd.leaf.mode]]});
    return c;
  }
  for (int k = 0; k < (int)nd.children.size(); ++k) {
    const int ch = nd.children[k];
    if (k == skip || !inS_[ch]) continue;
    c.push_back({k, hmat_[ch].at(restrict_key(l, ch, true))});
  }
  return c;
}

int ApplyFit::init() {
  if (!E.planned()) {
    set_error("apply_fit: set an operator (Hamiltonian) first");
    return QTG_ESTATE;
  }
  if (E.complex64()) {
    set_error("apply_fit: needs a complex128 engine");
    return QTG_ESTATE;
  }
  const HTree& T = E.tree();
  const HSop& H = E.sop();
  const int nn = (int)T.nodes.size(), nm = (int)T.leafNode.size(), L = (int)H.terms.size();
  const int top = T.topAddr();
  if (L < 1) {
    set_error("apply_fit: the operator has no terms");
    return QTG_ESTATE;
  }
  under_.assign(nn, std::vector<char>(nm, 0));
  for (int a = 0; a < nn; ++a) {
    const HNode& nd = T.nodes[a];
    if (nd.bottom) under_[a][nd.leaf.mode] = 1;
    for (int c : nd.children)
      for (int m = 0; m < nm; ++m) under_[a][m] |= under_[c][m];
  }
  std::map<std::pair<int, std::string>, int> canon;
  std::vector<const std::vector<cd>*> cdata;
  cid_.assign(L, std::vector<int>(nm, -1));
  for (int l = 0; l < L; ++l)
    for (auto& lm : H.terms[l].leafmats) {
      const auto& M = H.leafMatrices[lm.second];
      auto key = std::make_pair(lm.first, std::string((const char*)M.data(), M.size() * sizeof(cd)));
      auto f = canon.find(key);
      int id;
      if (f == canon.end()) {
        id = (int)leafOff_.size();
        canon[key] = id;
        leafOff_.push_back(palloc((int64_t)M.size()));
        cdata.push_back(&M);
      } else {
        id = f->second;
      }
      cid_[l][lm.first] = id;
    }
  // SparseTree(sop, tree, false) (SparseTree.cpp:20-95): the union of all
  // terms' leaf-to-root paths, cut above its first node (from the top) with
  // more than one active child
  std::vector<char> on(nn, 0);
  for (int l = 0; l < L; ++l)
    for (int m = 0; m < nm; ++m) {
      if (cid_[l][m] < 0) continue;
      for (int a = T.leafNode[m];; a = T.nodes[a].parent) {
        on[a] = 1;
        if (T.nodes[a].top()) break;
      }
    }
  std::vector<int> snodes;
  for (int a = 0; a < nn; ++a)
    if (on[a]) snodes.push_back(a);
  inS_.assign(nn, 0);
  if (!snodes.empty()) {
    int m;
    for (m = (int)snodes.size() - 1; m > 0; --m) {
      int na = 0;
      for (int c : T.nodes[snodes[m]].children) na += on[c];
      if (na > 1) break;
    }
    ++m;
    for (int i = 0; i < m; ++i) inS_[snodes[i]] = 1;
  }
  int rc;
  ketOff_ = 0;
  progBase_ = progTop_ = progMax_ = T.total;
  if (L == 1) {
    // applyReference (MultiLeafOperator.cpp:73-93): every factor of the one
    // product applied to its bottom tensor
    single_ = true;
    std::vector<GmDesc> ds;
    for (int a = 0; a < nn; ++a) {
      const HNode& nd = T.nodes[a];
      if (nd.bottom && cid_[0][nd.leaf.mode] >= 0)
        ds.push_back(gm_desc(a, 0, leafOff_[cid_[0][nd.leaf.mode]], 1, ketOff_ + nd.off, nd.off));
    }
    push_gm(iter_, ds, 0);
  } else {
    if (!inS_[top] || T.nodes[top].n != 1) {
      set_error("apply_fit: supported for operators whose sparse tree reaches the top node, with one state there");
      return QTG_ESHAPE;
    }
    for (auto& nd : T.nodes)
      if (nd.order() > kMaxOrder) {
        set_error("apply_fit: nodes with more than 3 children are not supported");
        return QTG_ESHAPE;
      }
    // matrices: per node of S below the top, one h-matrix per restriction
    // below and one hole per restriction above
    hmat_.assign(nn, {});
    hole_.assign(nn, {});
    std::vector<std::map<Key, int>> rep(nn);              // a term of each restriction below
    std::vector<std::map<Key, std::vector<int>>> byIn(nn);  // the terms of each restriction below
    std::vector<std::map<Key, int>> repOut(nn);           // a term of each restriction above
    std::vector<std::vector<int64_t>> holeOf(nn);         // node x term: its hole
    for (int a = 0; a < nn; ++a) {
      if (!inS_[a]) continue;
      const HNode& nd = T.nodes[a];
      holeOf[a].assign(L, -1);
      for (int l = 0; l < L; ++l) {
        const Key hk = restrict_key(l, a, true);
        if (!rep[a].count(hk)) rep[a][hk] = l;
        byIn[a][hk].push_back(l);
        if (nd.top()) continue;
        const int64_t n2 = (int64_t)nd.n * nd.n;
        if (!hmat_[a].count(hk)) hmat_[a][hk] = palloc(n2);
        const Key ok = restrict_key(l, a, false);
        if (!hole_[a].count(ok)) {
          hole_[a][ok] = palloc(n2);
          repOut[a][ok] = l;
        }
        holeOf[a][l] = hole_[a][ok];
      }
      if (!nd.top() && (nd.N / nd.n) < nd.n) {
        set_error("apply_fit: node " + std::to_string(a) + " has fewer rows than SPFs");
        return QTG_ESHAPE;
      }
    }
    for (auto& nd : T.nodes)
      if (!ident_.count({nd.n, 4})) ident_[{nd.n, 4}] = palloc(4 * (int64_t)nd.n * nd.n);
    // block regions: one per node, one block per restriction below
    std::vector<int64_t> sblock(nn, -1);
    for (int a = 0; a < nn; ++a)
      if (inS_[a]) sblock[a] = palloc((int64_t)rep[a].size() * T.nodes[a].n * T.nodes[a].n);
    int maxH = 0, maxD = 0;
    for (int a = 0; a < nn; ++a)
      if (inS_[a]) {
        maxH = std::max(maxH, T.nodes[a].height);
        maxD = std::max(maxD, T.nodes[a].depth);
      }
    auto ket = [&](int a) { return ketOff_ + T.nodes[a].off; };
    // first iteration only: the h-matrices, bottom-up by height
    for (int h = 0; h <= maxH; ++h) {
      progTop_ = progBase_;
      std::vector<ChainGroup> groups;
      std::vector<std::pair<int, Key>> owner;
      std::vector<GcDesc> gcs;
      for (int a = 0; a < nn; ++a) {
        const HNode& nd = T.nodes[a];
        if (!inS_[a] || nd.top() || nd.height != h) continue;
        ChainGroup g{a, 1, ket(a), {}};
        for (auto& hk : hmat_[a]) {
          Chain c = r_chain(rep[a].at(hk.first), a, -1);
          if (c.empty()) gcs.push_back(gc_desc(a, nd.order() - 1, 0, nd.off, 1, ket(a), hk.second));
          else g.reqs.push_back({c, -1});
        }
        groups.push_back(g);
      }
      auto outs = emit_chains_multi(first_, groups);
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const int a = groups[gi].a;
        size_t ri = 0;
        for (auto& hk : hmat_[a]) {
          Chain c = r_chain(rep[a].at(hk.first), a, -1);
          if (c.empty()) continue;
          gcs.push_back(gc_desc(a, T.nodes[a].order() - 1, 0, T.nodes[a].off, 1, outs[gi][ri++], hk.second));
        }
      }
      push_gc(first_, gcs);
    }
    // holes, top-down by depth
    for (int d = 1; d <= maxD; ++d) {
      progTop_ = progBase_;
      std::vector<ChainGroup> groups;
      std::vector<int> xs;
      std::vector<std::vector<int64_t>> dest;
      std::vector<GcDesc> gcs;
      for (int x = 0; x < nn; ++x) {
        const HNode& nx = T.nodes[x];
        if (!inS_[x] || nx.top() || nx.depth != d) continue;
        const int p = nx.parent;
        const HNode& np = T.nodes[p];
        ChainGroup g{p, 1, ket(p), {}};
        std::vector<int64_t> ds;
        for (auto& ok : hole_[x]) {
          const int l = repOut[x].at(ok.first);  // a term with this restriction above x
          Chain c = r_chain(l, p, nx.childIdx);
          if (!np.top()) c.push_back({np.order() - 1, holeOf[p][l]});
          if (c.empty()) {
            gcs.push_back(gc_desc(p, nx.childIdx, 0, np.off, 1, ket(p), ok.second));
            continue;
          }
          g.reqs.push_back({c, -1});
          ds.push_back(ok.second);
        }
        groups.push_back(g);
        xs.push_back(x);
        dest.push_back(ds);
      }
      auto outs = emit_chains_multi(iter_, groups);
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const int p = groups[gi].a;
        for (size_t ri = 0; ri < outs[gi].size(); ++ri)
          gcs.push_back(gc_desc(p, T.nodes[xs[gi]].childIdx, 0, T.nodes[p].off, 1, outs[gi][ri], dest[gi][ri]));
      }
      push_gc(iter_, gcs);
    }
    // the blocks of every node: sum of the holes of the terms that share a
    // restriction below (applyOperatorLocal); the top: the number of such
    // terms times the 1 x 1 identity (applyTop)
    {
      std::vector<std::pair<int64_t, std::vector<CombTerm>>> outs;
      std::vector<int> ns;
      for (int a = 0; a < nn; ++a) {
        if (!inS_[a]) continue;
        const HNode& nd = T.nodes[a];
        int e = 0;
        for (auto& r : rep[a]) {
          std::vector<CombTerm> terms;
          for (int l : byIn[a].at(r.first)) terms.push_back({1.0, 0.0, nd.top() ? -1 : holeOf[a][l]});
          outs.push_back({sblock[a] + (int64_t)e * nd.n * nd.n, terms});
          ns.push_back(nd.n);
          ++e;
        }
      }
      push_comb_sized(iter_, outs, ns);
    }
    // the sweep, bottom-up by height
    for (int h = 0; h <= maxH; ++h) {
      progTop_ = progBase_;
      std::vector<ChainGroup> groups;
      std::vector<int64_t> sbase, tmp;
      std::vector<int> nodes;
      for (int a = 0; a < nn; ++a) {
        const HNode& nd = T.nodes[a];
        if (!inS_[a] || nd.height != h) continue;
        const int64_t sb = walloc((int64_t)rep[a].size() * nd.N);
        ChainGroup g{a, 1, ket(a), {}};
        int e = 0;
        for (auto& r : rep[a]) {
          Chain c = r_chain(r.second, a, -1);
          if (c.empty()) c.push_back({-1, 0});
          g.reqs.push_back({c, sb + (int64_t)e * nd.N});
          ++e;
        }
        groups.push_back(g);
        sbase.push_back(sb);
        tmp.push_back(walloc(nd.N));
        nodes.push_back(a);
      }
      emit_chains_multi(iter_, groups);
      std::vector<GmDesc> stk;
      for (size_t i = 0; i < nodes.size(); ++i) {
        const HNode& nd = T.nodes[nodes[i]];
        GmDesc d{};
        d.x = sbase[i];
        d.xsel = 1;
        d.m = sblock[nodes[i]];
        d.out = tmp[i];
        d.bef = (int32_t)(nd.N / nd.n);
        d.nin = (int32_t)(nd.n * rep[nodes[i]].size());
        d.nout = nd.n;
        d.aft = 1;
        d.op = GM_N;
        d.are = 1.0;
        d.node = -1;
        d.ksplit = 1;
        stk.push_back(d);
      }
      push_gm(iter_, stk, 1);
      if (h == maxH && nodes.size() == 1 && nodes[0] == top) {
        Step st;
        st.kind = K_NORM;
        st.a = tmp[0];
        st.c = T.nodes[top].off;
        st.N = T.nodes[top].N;
        iter_.push_back(st);
        continue;
      }
      Step q;
      q.kind = K_QRB;
      q.abuf = 1;
      q.cbuf = 0;
      q.d0 = (int64_t)h_qr_.size();
      for (size_t i = 0; i < nodes.size(); ++i) {
        const HNode& nd = T.nodes[nodes[i]];
        qr_item(tmp[i], nd.off, nd.N / nd.n, nd.n);
      }
      q.n = (int64_t)nodes.size();
      iter_.push_back(q);
      std::vector<GcDesc> gcs;
      for (size_t i = 0; i < nodes.size(); ++i) {
        const int a = nodes[i];
        const HNode& nd = T.nodes[a];
        int e = 0;
        for (auto& r : rep[a]) {
          gcs.push_back(gc_desc(a, nd.order() - 1, 0, nd.off, 1, sbase[i] + (int64_t)e * nd.N, hmat_[a].at(r.first)));
          ++e;
        }
      }
      push_gc(iter_, gcs);
    }
  }
  workSize_ = progMax_;
  if ((rc = validate_steps(first_)) || (rc = validate_steps(iter_))) return rc;
  if (E.host_only()) return QTG_OK;
  std::vector<std::pair<int64_t, const std::vector<cd>*>> init;
  for (size_t i = 0; i < cdata.size(); ++i) init.push_back({leafOff_[i], cdata[i]});
  if ((rc = materialize(init))) return rc;
  if ((rc = dmalloc((void**)&fid_, 16))) return rc;
  ready_ = true;
  return QTG_OK;
}

int ApplyFit::validate_extra(const Step& st, std::vector<Span>& rd, std::vector<Span>& wr) const {
  if (st.kind != K_NORM) {
    set_error("apply_fit plan check: unknown step");
    return QTG_ESTATE;
  }
  rd.push_back({1, st.a, st.a + st.N});
  wr.push_back({0, st.c, st.c + st.N});
  return QTG_OK;
}

int ApplyFit::run_extra(const Step& st, hipStream_t s) {
  if (st.kind != K_NORM) {
    set_error("apply_fit: unknown step");
    return QTG_ESTATE;
  }
  return launch_normalize(work_ + 2 * st.a, psi_ + 2 * st.c, st.N, fid_, s);
}

int ApplyFit::run(double* psi_host, uint64_t seed, int64_t max_iter, double* fids, int64_t cap, int64_t* iters) {
  if (!ready_) {
    set_error("apply_fit: not initialised");
    return QTG_ESTATE;
  }
  const HTree& T = E.tree();
  const size_t bytes = 16 * (size_t)T.total;
  // the state (ket) and the start: TensorTree(gen, tree, false), with the
  // nodes outside the operator's sparse tree taken from the state
  // (ApplyFirstOrder.cpp:56-62)
  std::vector<cd> start;
  fill_random(T, seed, false, start);
  const cd* P = reinterpret_cast<const cd*>(psi_host);
  for (auto& nd : T.nodes)
    if (single_ || !inS_[nd.addr]) std::copy(P + nd.off, P + nd.off + nd.N, start.begin() + nd.off);
  QTG_HIP_TRY(hipMemcpyAsync(work_ + 2 * ketOff_, psi_host, bytes, hipMemcpyHostToDevice, s_));
  QTG_HIP_TRY(hipMemcpyAsync(psi_, start.data(), bytes, hipMemcpyHostToDevice, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  int rc;
  int64_t n = 0;
  if (single_) {
    if ((rc = run_steps(iter_, s_))) return rc;
    if (cap > 0 && fids) fids[0] = 1.0;
    n = 1;
  } else {
    if ((rc = run_steps(first_, s_))) return rc;
    double f = 0.0, f_last = 1.0;
    for (int64_t it = 0; it < max_iter; ++it) {
      if (graph_) {
        if (!exec_) {
          hipGraph_t g = nullptr;
          QTG_HIP_TRY(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
          rc = run_steps(iter_, s_);
          hipError_t e = hipStreamEndCapture(s_, &g);
          if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
          }
          if (e != hipSuccess) {
            set_error(std::string("apply_fit: graph capture: ") + hipGetErrorString(e));
            return QTG_EDEVICE;
          }
          hipError_t ie = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
          hipGraphDestroy(g);
          if (ie != hipSuccess) {
            exec_ = nullptr;
            set_error(std::string("apply_fit: graph instantiate: ") + hipGetErrorString(ie));
            return QTG_EDEVICE;
          }
        }
        QTG_HIP_TRY(hipGraphLaunch(exec_, s_));
      } else if ((rc = run_steps(iter_, s_))) {
        return rc;
      }
      double v = 0.0;
      QTG_HIP_TRY(hipMemcpyAsync(&v, fid_, 8, hipMemcpyDeviceToHost, s_));
      QTG_HIP_TRY(hipStreamSynchronize(s_));
      f_last = f;
      f = v;
      if (fids && n < cap) fids[n] = f;
      ++n;
      if ((1.0 - f) < 1e-6 || (f - f_last) < 1e-10) break;
    }
  }
  QTG_HIP_TRY(hipMemcpyAsync(psi_host, psi_, bytes, hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  if (iters) *iters = n;
  return QTG_OK;
}

}  // namespace qtg
