// Krylov-sweep SCF ground-state search on the device.
//
// The sweep of scf() (SCF.cpp:270-348 of QuTree) visits the nodes in a
// depth-first walk (scf_sweep, SCF.cpp:14-34).  At each node it builds a
// Lanczos space of the node's effective Hamiltonian (solveKrylovSpace,
// SCF.cpp:134-196), replaces the node tensor by the lowest Krylov
// eigenvector or an imaginary-time step (SCF.cpp:198-235), moves the
// orthogonality centre to the next node with a QR decomposition and rebuilds
// the matrices that point toward it (buildSCF, HamiltonianRepresentation.h:
// 131-177).  The effective Hamiltonian is that of SCF.cpp:36-81 on the
// tail-cut sparse trees of HamiltonianRepresentation(H, tree, tree, false):
// every term active at more than one neighbour of the node applies its
// neighbour matrices, all other terms enter through the correction matrices
// hCorr (below a node) and hConCorr (above it) (HamiltonianRepresentation.cpp:
// 119-218).
//
// Device form.  Every node-level contraction is a grouped MFMA launch
// (gemm_ops.hip): the mode products of a chain of neighbour matrices, one
// launch per chain depth for all terms together, with equal prefixes
// computed once; the sum over terms as ONE stacked product along the node's
// last leg, sum_e B_e x_last T_e = [B_1 | B_2 | ...] x_last [T_1; T_2; ...],
// whose blocks B_e (coefficient-weighted hole or correction matrices, summed
// over the terms that share a chain T_e) are assembled once per visit; hole,
// h- and correction matrices as grouped hole contractions.  The Lanczos
// recurrence, the Krylov combination and the QR are single-workgroup
// kernels.  Matrices shared by terms with equal restrictions below (h) or
// above (hole) a node are stored once.  All shapes and offsets are fixed by
// the tree and the Hamiltonian, so one iteration is a fixed launch sequence,
// replayed from a captured HIP graph; the Lanczos scalars stay on the device.
#include "scf.hpp"

#include <algorithm>
#include <cstring>
#include <functional>
#include <string>

namespace qtg {

Scf::~Scf() {
  if (exec_) hipGraphExecDestroy(exec_);
}

Scf::Key Scf::restrict_key(int l, int a, bool inside) const {
  Key k;
  const auto& c = cid_[l];
  for (int m = 0; m < (int)c.size(); ++m)
    if (c[m] >= 0 && (under_[a][m] != 0) == inside) k.push_back({m, c[m]});
  return k;
}

// the neighbour matrices of term l below node a in child order (skip: a
// child index left out), or its leaf matrix at a bottom node
Scf::Chain Scf::r_chain(int l, int a, int skip) const {
  const HNode& nd = E.tree().nodes[a];
  Chain c;
  if (nd.bottom) {
    c.push_back({0, leafOff_[cid_[l][nd.leaf.mode]]});
    return c;
  }
  for (int k = 0; k < (int)nd.children.size(); ++k) {
    const int ch = nd.children[k];
    if (k == skip || !act_[l][ch]) continue;
    c.push_back({k, hmat_[ch].at(restrict_key(l, ch, true))});
  }
  return c;
}

// H_eff v at node a (SCF.cpp:36-81): the input vector sits at the stack
// base (entry 0), the result at haOff_.
const Scf::ApplyProg& Scf::apply_prog(int a) {
  auto it = apply_.find(a);
  if (it != apply_.end()) return it->second;
  const HSop& H = E.sop();
  const HNode& nd = E.tree().nodes[a];
  progTop_ = progBase_;
  std::map<Chain, std::vector<CombTerm>> groups;
  for (int l = 0; l < (int)H.terms.size(); ++l) {
    const bool on = nd.bottom ? (act_[l][a] != 0) : (nact_[a][l] > 1);
    if (!on) continue;
    const cd c = H.terms[l].coeff;
    const int64_t src = nd.top() ? -1 : hole_[a].at(restrict_key(l, a, false));
    groups[r_chain(l, a, -1)].push_back({c.real(), c.imag(), src});
  }
  if (!nd.bottom)
    for (int k = 0; k < (int)nd.children.size(); ++k)
      groups[Chain{{k, hcorr_[nd.children[k]]}}].push_back({1.0, 0.0, -1});
  std::vector<StackEntry> ents(1);
  if (!nd.top()) ents[0].block.push_back({1.0, 0.0, hconcorr_[a]});
  for (auto& g : groups) ents.push_back({g.first, g.second});
  ApplyProg P;
  P.N = nd.N;
  P.base = walloc((int64_t)ents.size() * nd.N);
  if (emit_stack(P.comb, P.body, a, 1, P.base, ents, true, P.base, haOff_)) P.N = -1;
  return apply_[a] = P;
}

// buildUp(node) (HamiltonianRepresentation.h:90-105): the h-matrices of
// every term at the node, then buildUpCorrection (.cpp:140-168).
const std::vector<Scf::Step>& Scf::up_prog(int a) {
  auto it = up_.find(a);
  if (it != up_.end()) return it->second;
  const HSop& H = E.sop();
  const HNode& nd = E.tree().nodes[a];
  const int last = nd.order() - 1;
  progTop_ = progBase_;
  std::vector<Step> P;
  std::map<Key, Chain> hk;
  for (int l = 0; l < (int)H.terms.size(); ++l)
    if (act_[l][a]) hk[restrict_key(l, a, true)] = r_chain(l, a, -1);
  std::vector<std::pair<Chain, int64_t>> reqs;
  for (auto& h : hk) reqs.push_back({h.second, -1});
  const int nch = (int)nd.children.size();
  int64_t sb = -1;
  if (!nd.bottom) {
    sb = walloc((int64_t)nch * nd.N);
    for (int k = 0; k < nch; ++k) reqs.push_back({Chain{{k, hcorr_[nd.children[k]]}}, sb + (int64_t)k * nd.N});
  }
  const std::vector<int64_t> outs = emit_chains(P, a, 0, nd.off, reqs);
  int64_t hA = -1;
  if (!nd.bottom) {
    hA = walloc(nd.N);
    GmDesc d = gm_desc(a, last, ident_.at({nd.n, 4}), 1, sb, hA);
    d.nin = nd.n * nch;  // the identity stack sums the children's corrections
    std::vector<GmDesc> ds{d};
    push_gm(P, ds, 1);
  }
  std::vector<GcDesc> gcs;
  size_t i = 0;
  for (auto& h : hk) gcs.push_back(gc_desc(a, last, 0, nd.off, 1, outs[i++], hmat_[a].at(h.first)));
  if (!nd.bottom) gcs.push_back(gc_desc(a, last, 0, nd.off, 1, hA, tmpGC_));
  push_gc(P, gcs);
  std::vector<CombTerm> terms;
  if (!nd.bottom) terms.push_back({1.0, 0.0, tmpGC_});
  for (int l = 0; l < (int)H.terms.size(); ++l) {
    if (nact_[nd.parent][l] != 1 || !act_[l][a]) continue;
    const cd c = H.terms[l].coeff;
    terms.push_back({c.real(), c.imag(), hmat_[a].at(restrict_key(l, a, true))});
  }
  push_comb(P, {{hcorr_[a], terms}}, nd.n);
  return up_[a] = P;
}

// buildDown(x) (HamiltonianRepresentation.h:107-129): the hole matrices of
// every term active at x, then buildDownCorrection (.cpp:170-207).
const std::vector<Scf::Step>& Scf::down_prog(int x) {
  auto it = down_.find(x);
  if (it != down_.end()) return it->second;
  const HSop& H = E.sop();
  const HNode& nx = E.tree().nodes[x];
  const int p = nx.parent;
  const HNode& np = E.tree().nodes[p];
  const int leg = nx.childIdx, plast = np.order() - 1;
  progTop_ = progBase_;
  std::vector<Step> comb, body;
  // holes
  std::map<Key, Chain> ok;
  for (int l = 0; l < (int)H.terms.size(); ++l) {
    if (!act_[l][x]) continue;
    const Key key = restrict_key(l, x, false);
    if (ok.count(key)) continue;
    Chain c;
    if (!key.empty()) {
      c = r_chain(l, p, leg);
      if (!np.top()) c.push_back({plast, hole_[p].at(restrict_key(l, p, false))});
    }
    ok[key] = c;
  }
  // the down correction: the stacked sum at p over the terms active at two
  // or more of p's neighbours but not at x, the other children's hCorr and
  // p's own hConCorr
  std::map<Chain, std::vector<CombTerm>> groups;
  for (int l = 0; l < (int)H.terms.size(); ++l) {
    if (nact_[p][l] < 2 || act_[l][x]) continue;
    const cd c = H.terms[l].coeff;
    const int64_t src = np.top() ? -1 : hole_[p].at(restrict_key(l, p, false));
    groups[r_chain(l, p, leg)].push_back({c.real(), c.imag(), src});
  }
  for (int k = 0; k < (int)np.children.size(); ++k)
    if (k != leg) groups[Chain{{k, hcorr_[np.children[k]]}}].push_back({1.0, 0.0, -1});
  if (!np.top()) groups[Chain{{-1, 0}}].push_back({1.0, 0.0, hconcorr_[p]});
  std::vector<StackEntry> ents;
  for (auto& g : groups) ents.push_back({g.first, g.second});
  const int64_t sb = walloc((int64_t)ents.size() * np.N);
  std::vector<std::pair<Chain, int64_t>> reqs;
  std::vector<Key> hkeys;
  for (auto& o : ok)
    if (!o.first.empty()) {
      reqs.push_back({o.second, -1});
      hkeys.push_back(o.first);
    }
  const std::vector<int64_t> houts = emit_chains(body, p, 0, np.off, reqs);
  int64_t hA = -1;
  if (!ents.empty()) {
    hA = walloc(np.N);
    // the stack's chains start from p's tensor in psi
    if (emit_stack(comb, body, p, 0, np.off, ents, false, sb, hA)) return down_[x] = {};
  }
  std::vector<GcDesc> gcs;
  if (ok.count(Key{})) gcs.push_back(gc_desc(p, leg, 0, np.off, 0, np.off, hole_[x].at(Key{})));
  for (size_t i = 0; i < hkeys.size(); ++i) gcs.push_back(gc_desc(p, leg, 0, np.off, 1, houts[i], hole_[x].at(hkeys[i])));
  if (hA >= 0) gcs.push_back(gc_desc(p, leg, 0, np.off, 1, hA, hconcorr_[x]));
  push_gc(body, gcs);
  if (hA < 0) push_comb(body, {{hconcorr_[x], {}}}, nx.n);
  std::vector<Step> P = comb;
  P.insert(P.end(), body.begin(), body.end());
  return down_[x] = P;
}

int Scf::emit_microstep(int a, int next, int slot) {
  const HTree& T = E.tree();
  const HNode& nd = T.nodes[a];
  const ApplyProg& A = apply_prog(a);
  if (A.N < 0) return QTG_ESTATE;
  const int kc = (int)std::min<int64_t>(kry_, nd.N);
  iter_.insert(iter_.end(), A.comb.begin(), A.comb.end());
  Step s;
  s.kind = K_LBEGIN;
  s.a = nd.off;
  s.b = kryOff_;
  s.c = A.base;
  s.N = nd.N;
  iter_.push_back(s);
  for (int j = 0; j < kc; ++j) {
    iter_.insert(iter_.end(), A.body.begin(), A.body.end());
    Step l;
    l.kind = K_LSTEP;
    l.j = j;
    l.k = kc;
    l.N = nd.N;
    l.a = haOff_;
    l.b = kryOff_;
    l.c = A.base;
    l.d = T_;
    iter_.push_back(l);
  }
  Step h;
  h.kind = K_HEEV;
  h.k = kc;
  h.a = T_;
  h.b = U_;
  iter_.push_back(h);
  Step c;
  c.kind = K_KCOMB;
  c.k = kc;
  c.N = nd.N;
  c.a = U_;
  c.b = kryOff_;
  c.c = psiwOff_;
  c.slot = slot;
  iter_.push_back(c);
  // QR toward the next node (adjacentIndex, SCF.cpp:237-256)
  const int last = nd.order() - 1;
  int out = last;
  if (next >= 0 && T.nodes[next].parent == a) out = T.nodes[next].childIdx;
  const int64_t bef = leg_bef(a, out), aft = leg_aft(a, out), n = nd.dims[out];
  if (bef * aft < n) {
    set_error("scf: node " + std::to_string(a) + " has fewer rows than columns toward leg " + std::to_string(out));
    return QTG_ESHAPE;
  }
  Step q;
  q.kind = K_QR;
  q.nn = n;
  q.bef = bef * aft;  // rows
  if (out == last) {
    q.a = psiwOff_;
    q.abuf = 1;
    q.c = nd.off;
    q.cbuf = 0;
    iter_.push_back(q);
  } else {
    Step p1;
    p1.kind = K_PERM;
    p1.a = psiwOff_;
    p1.abuf = 1;
    p1.c = permOff_;
    p1.cbuf = 1;
    p1.bef = bef;
    p1.nn = n;
    p1.aft = aft;
    p1.dir = 0;
    iter_.push_back(p1);
    q.a = q.c = permOff_;
    q.abuf = q.cbuf = 1;
    iter_.push_back(q);
    Step p2 = p1;
    p2.a = permOff_;
    p2.abuf = 1;
    p2.c = nd.off;
    p2.cbuf = 0;
    p2.dir = 1;
    iter_.push_back(p2);
  }
  if (next < 0) return QTG_OK;
  // r = contraction(Q, PsiW, out); Psi[next] = r x_in Psi[next]
  std::vector<GcDesc> gcs{gc_desc(a, out, 0, nd.off, 1, psiwOff_, rSlot_)};
  push_gc(iter_, gcs);
  const HNode& nn = T.nodes[next];
  const int in = (nd.parent == next) ? nd.childIdx : nn.order() - 1;
  std::vector<GmDesc> g1{gm_desc(next, in, rSlot_, 0, nn.off, pushOff_)};
  push_gm(iter_, g1, 1);
  std::vector<GmDesc> g2{gm_desc(next, -1, 0, 1, pushOff_, nn.off)};
  push_gm(iter_, g2, 0);
  const bool up = !nd.top() && nd.parent == next;
  const std::vector<Step>& R = up ? up_prog(a) : down_prog(next);
  if (R.empty()) return QTG_ESTATE;
  iter_.insert(iter_.end(), R.begin(), R.end());
  return QTG_OK;
}

int Scf::init(int64_t n_krylov) {
  if (!E.planned()) {
    set_error("scf: set a Hamiltonian first");
    return QTG_ESTATE;
  }
  if (E.complex64()) {
    set_error("scf: needs a complex128 engine");
    return QTG_ESTATE;
  }
  if (n_krylov < 1 || n_krylov > kMaxKrylov) {
    set_error("scf: Krylov size outside 1..64");
    return QTG_EARG;
  }
  kry_ = n_krylov;
  const HTree& T = E.tree();
  const HSop& H = E.sop();
  const int nn = (int)T.nodes.size(), nm = (int)T.leafNode.size(), L = (int)H.terms.size();
  for (auto& nd : T.nodes)
    if (nd.order() > kMaxOrder) {
      set_error("scf: nodes with more than 3 children are not supported");
      return QTG_ESHAPE;
    }
  // modes below each node (addresses put children first)
  under_.assign(nn, std::vector<char>(nm, 0));
  for (int a = 0; a < nn; ++a) {
    const HNode& nd = T.nodes[a];
    if (nd.bottom) under_[a][nd.leaf.mode] = 1;
    for (int c : nd.children)
      for (int m = 0; m < nm; ++m) under_[a][m] |= under_[c][m];
  }
  // leaf matrices, equal ones (same mode and entries) stored once
  std::map<std::pair<int, std::string>, int> canon;
  std::vector<const std::vector<cd>*> cdata;
  cid_.assign(L, std::vector<int>(nm, -1));
  for (int l = 0; l < L; ++l)
    for (auto& lm : H.terms[l].leafmats) {
      const auto& M = H.leafMatrices[lm.second];
      std::string bytes((const char*)M.data(), M.size() * sizeof(cd));
      auto key = std::make_pair(lm.first, bytes);
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
  // tail-cut sparse trees (SparseTree.cpp:33-95 with tail = false): the
  // union of the term's leaf-to-root paths, cut above its first node (from
  // the top) with more than one active child
  act_.assign(L, std::vector<char>(nn, 0));
  for (int l = 0; l < L; ++l) {
    std::vector<char> on(nn, 0);
    for (int m = 0; m < nm; ++m) {
      if (cid_[l][m] < 0) continue;
      for (int a = T.leafNode[m];; a = T.nodes[a].parent) {
        on[a] = 1;
        if (T.nodes[a].top()) break;
      }
    }
    std::vector<int> nodes;
    for (int a = 0; a < nn; ++a)
      if (on[a]) nodes.push_back(a);
    if (nodes.empty()) continue;
    int m;
    for (m = (int)nodes.size() - 1; m > 0; --m) {
      int na = 0;
      for (int c : T.nodes[nodes[m]].children) na += on[c];
      if (na > 1) break;
    }
    ++m;
    for (int i = 0; i < m; ++i) act_[l][nodes[i]] = 1;
  }
  // nActives (HamiltonianRepresentation.cpp:121-138)
  nact_.assign(nn, std::vector<int>(L, 0));
  for (int a = 0; a < nn; ++a) {
    const HNode& nd = T.nodes[a];
    for (int l = 0; l < L; ++l) {
      int c = 0;
      if (nd.bottom) c += cid_[l][nd.leaf.mode] >= 0;
      else
        for (int ch : nd.children) c += act_[l][ch];
      if (!nd.top() && act_[l][a]) ++c;
      nact_[a][l] = c;
    }
  }
  // matrices
  hmat_.assign(nn, {});
  hole_.assign(nn, {});
  hcorr_.assign(nn, -1);
  hconcorr_.assign(nn, -1);
  int nmax = 1;
  for (int a = 0; a < nn; ++a) {
    const HNode& nd = T.nodes[a];
    const int64_t n2 = (int64_t)nd.n * nd.n;
    nmax = std::max(nmax, nd.n);
    Nmax_ = std::max<int64_t>(Nmax_, nd.N);
    hcorr_[a] = palloc(n2);
    hconcorr_[a] = palloc(n2);
    if (nd.top()) continue;
    for (int l = 0; l < L; ++l) {
      if (!act_[l][a]) continue;
      Key hk = restrict_key(l, a, true), ok = restrict_key(l, a, false);
      if (!hmat_[a].count(hk)) hmat_[a][hk] = palloc(n2);
      if (!hole_[a].count(ok)) hole_[a][ok] = palloc(n2);
    }
  }
  tmpGC_ = palloc((int64_t)nmax * nmax);
  rSlot_ = palloc((int64_t)nmax * nmax);
  T_ = palloc(kry_ * kry_);
  U_ = palloc(kry_ * kry_);
  for (auto& nd : T.nodes)
    if (!ident_.count({nd.n, 4})) ident_[{nd.n, 4}] = palloc(4 * (int64_t)nd.n * nd.n);
  S_ = poolSize_;
  // work layout: Krylov vectors, H v, the updated tensor, QR and push
  // scratch, then the programs' stacks and temporaries
  kryOff_ = 0;
  haOff_ = kry_ * Nmax_;
  psiwOff_ = haOff_ + Nmax_;
  permOff_ = psiwOff_ + Nmax_;
  pushOff_ = permOff_ + Nmax_;
  progBase_ = pushOff_ + Nmax_;
  progMax_ = progBase_;
  // programs
  sweep_.clear();
  std::function<void(int)> walk = [&](int a) {
    sweep_.push_back(a);
    if (T.nodes[a].bottom) return;
    for (int c : T.nodes[a].children) {
      walk(c);
      sweep_.push_back(a);
    }
  };
  walk(T.topAddr());
  int rc;
  for (int a = 0; a < nn; ++a) {
    if (T.nodes[a].top()) continue;
    const auto& U = up_prog(a);
    init_.insert(init_.end(), U.begin(), U.end());
  }
  for (size_t l = 0; l < sweep_.size(); ++l)
    if ((rc = emit_microstep(sweep_[l], l + 1 < sweep_.size() ? sweep_[l + 1] : -1, (int)l))) return rc;
  workSize_ = progMax_;
  poolSize_ = S_ + Smax_;
  if ((rc = validate_steps(init_)) || (rc = validate_steps(iter_))) return rc;
  if (E.host_only()) return QTG_OK;  // plan and checks only
  std::vector<std::pair<int64_t, const std::vector<cd>*>> init;
  for (size_t i = 0; i < cdata.size(); ++i) init.push_back({leafOff_[i], cdata[i]});
  if ((rc = materialize(init))) return rc;
  if ((rc = dmalloc((void**)&ev_, 8 * (size_t)kry_)) || (rc = dmalloc((void**)&scal_, 16 * (size_t)kry_)) ||
      (rc = dmalloc((void**)&elog_, 8 * sweep_.size())) || (rc = dmalloc((void**)&par_, sizeof(ScfParams))))
    return rc;
  QTG_HIP_TRY(hipMemset(scal_, 0, 16 * (size_t)kry_));
  ready_ = true;
  return QTG_OK;
}

int Scf::validate_extra(const Step& st, std::vector<Span>& rd, std::vector<Span>& wr) const {
  auto fail = [&](const std::string& what) {
    set_error("scf plan check: " + what);
    return QTG_ESTATE;
  };
  switch (st.kind) {
    case K_LBEGIN:
      rd.push_back({0, st.a, st.a + st.N});
      wr.push_back({1, st.b, st.b + st.N});
      wr.push_back({1, st.c, st.c + st.N});
      return QTG_OK;
    case K_LSTEP:
      if (st.k < 1 || st.k > kry_ || st.j >= st.k) return fail("Krylov index");
      rd.push_back({1, st.a, st.a + st.N});
      wr.push_back({1, st.b, st.b + st.k * st.N});
      wr.push_back({1, st.c, st.c + st.N});
      wr.push_back({2, st.d, st.d + (int64_t)st.k * st.k});
      return QTG_OK;
    case K_HEEV:
      if (st.k < 1 || st.k > kry_) return fail("Krylov size");
      rd.push_back({2, st.a, st.a + (int64_t)st.k * st.k});
      wr.push_back({2, st.b, st.b + (int64_t)st.k * st.k});
      return QTG_OK;
    case K_KCOMB:
      if (st.slot < 0 || st.slot >= (int)sweep_.size()) return fail("energy slot");
      rd.push_back({2, st.a, st.a + (int64_t)st.k * st.k});
      rd.push_back({1, st.b, st.b + st.k * st.N});
      wr.push_back({1, st.c, st.c + st.N});
      return QTG_OK;
    default:
      return fail("unknown step");
  }
}

int Scf::run_extra(const Step& st, hipStream_t s) {
  auto Z = [](double* b, int64_t off) { return (void*)(b + 2 * off); };
  switch (st.kind) {
    case K_LBEGIN:
      return launch_lanczos_begin(Z(psi_, st.a), st.N, Z(work_, st.b), Z(work_, st.c), s);
    case K_LSTEP:
      return launch_lanczos_step(st.j, st.k, st.N, Z(work_, st.a), 1, Z(work_, st.b), Z(work_, st.c), scal_,
                                 Z(pool_, st.d), s);
    case K_HEEV: {
      SmItem u{};
      u.a = st.a;
      u.q = st.b;
      u.e = 0;
      u.m = st.k;
      u.n = st.k;
      const SmItem zero{};
      return launch_heev_batch(nullptr, u, zero, 1, st.k, pool_, pool_, ev_, nullptr, 0.0, s);
    }
    case K_KCOMB:
      return launch_krylov_combine(st.k, st.N, Z(pool_, st.a), ev_, Z(work_, st.b), Z(work_, st.c), par_, elog_,
                                   st.slot, s);
    default:
      set_error("scf: unknown step");
      return QTG_ESTATE;
  }
}

int Scf::run(double* psi_host, int64_t n_iter, int64_t n_itp, double beta, double* energies, int64_t cap) {
  if (!ready_) {
    set_error("scf: not initialised");
    return QTG_ESTATE;
  }
  const HTree& T = E.tree();
  const size_t bytes = 16 * (size_t)T.total;
  QTG_HIP_TRY(hipMemcpyAsync(psi_, psi_host, bytes, hipMemcpyHostToDevice, s_));
  int rc;
  if ((rc = run_steps(init_, s_))) return rc;
  const int64_t steps = (int64_t)sweep_.size();
  std::vector<double> elog(steps);
  for (int64_t it = 0; it < n_iter; ++it) {
    ScfParams p{};
    p.itp = it < n_itp ? 1 : 0;
    p.beta = beta;
    QTG_HIP_TRY(hipMemcpyAsync(par_, &p, sizeof(p), hipMemcpyHostToDevice, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
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
          set_error(std::string("scf: graph capture: ") + hipGetErrorString(e));
          return QTG_EDEVICE;
        }
        hipError_t ie = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) {
          exec_ = nullptr;
          set_error(std::string("scf: graph instantiate: ") + hipGetErrorString(ie));
          return QTG_EDEVICE;
        }
      }
      QTG_HIP_TRY(hipGraphLaunch(exec_, s_));
    } else if ((rc = run_steps(iter_, s_))) {
      return rc;
    }
    QTG_HIP_TRY(hipMemcpyAsync(elog.data(), elog_, 8 * steps, hipMemcpyDeviceToHost, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
    for (int64_t i = 0; i < steps; ++i)
      if (energies && it * steps + i < cap) energies[it * steps + i] = elog[i];
  }
  QTG_HIP_TRY(hipMemcpyAsync(psi_host, psi_, bytes, hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  return QTG_OK;
}

}  // namespace qtg
