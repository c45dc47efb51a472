// Fixed launch programs over one tree: planning helpers, the host-side
// check of every access, and the launcher.
//
// A program is a list of steps on three device buffers: psi (the tree's node
// tensors in the engine's layout), work (vectors and intermediates) and pool
// (n x n matrices).  Grouped mode products (GmDesc) read their tensor from
// psi or work and their matrix from the pool; grouped hole contractions
// (GcDesc) read two tensors from psi or work and write pool matrices; the
// combination kernel sums coefficient-weighted pool matrices.  Chains of
// mode products that share a prefix share its products, and depth d of all
// chains of a step is one grouped launch.  A sum over terms is one stacked
// product along the node's last leg:
//   sum_e B_e x_last T_e = [B_1 | B_2 | ...] x_last [T_1; T_2; ...],
// with the T_e laid out back to back in work and the blocks B_e back to back
// in the pool.
#include "tree_prog.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <tuple>

namespace qtg {

TreeProg::~TreeProg() {
  for (void* p : owned_)
    if (p) hipFree(p);
  if (s_) hipStreamDestroy(s_);
}

int TreeProg::dmalloc(void** p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
    set_error(name_ + ": device allocation failed");
    return QTG_ENOMEM;
  }
  owned_.push_back(*p);
  bytes_ += (int64_t)bytes;
  return QTG_OK;
}

int64_t TreeProg::palloc(int64_t n) {
  const int64_t o = poolSize_;
  poolSize_ += n;
  return o;
}

int64_t TreeProg::walloc(int64_t n) {
  const int64_t o = progTop_;
  progTop_ += n;
  progMax_ = std::max(progMax_, progTop_);
  return o;
}

int64_t TreeProg::leg_bef(int a, int k) const {
  int64_t b = 1;
  for (int i = 0; i < k; ++i) b *= E.tree().nodes[a].dims[i];
  return b;
}

int64_t TreeProg::leg_aft(int a, int k) const {
  const auto& d = E.tree().nodes[a].dims;
  int64_t f = 1;
  for (int i = k + 1; i < (int)d.size(); ++i) f *= d[i];
  return f;
}

GmDesc TreeProg::gm_desc(int a, int leg, int64_t mat, int xsel, int64_t x, int64_t out) const {
  const HNode& nd = E.tree().nodes[a];
  GmDesc d{};
  if (leg < 0) {  // copy of the whole tensor
    d.bef = (int32_t)(nd.N / nd.n);
    d.nin = d.nout = nd.n;
    d.aft = 1;
    d.op = GM_COPY;
    d.m = 0;
  } else {
    d.bef = (int32_t)leg_bef(a, leg);
    d.nin = d.nout = nd.dims[leg];
    d.aft = (int32_t)leg_aft(a, leg);
    d.op = GM_N;
    d.m = mat;
  }
  d.x = x;
  d.out = out;
  d.acc = 0;
  d.are = 1.0;
  d.aim = 0.0;
  d.node = -1;
  d.xsel = xsel;
  d.ksplit = 1;
  return d;
}

GcDesc TreeProg::gc_desc(int a, int leg, int xsel, int64_t x, int ysel, int64_t y, int64_t out) const {
  const HNode& nd = E.tree().nodes[a];
  GcDesc d{};
  d.x = x;
  d.y = y;
  d.out = out;
  d.bef = (int32_t)leg_bef(a, leg);
  d.nx = d.ny = nd.dims[leg];
  d.aft = (int32_t)leg_aft(a, leg);
  d.acc = 0;
  d.node = -1;
  d.ksplit = 1;
  d.sel = (xsel ? 1 : 0) | (ysel ? 2 : 0);
  return d;
}

void TreeProg::push_gm(std::vector<Step>& P, std::vector<GmDesc>& ds, int obuf) {
  if (ds.empty()) return;
  Step st;
  st.kind = K_GM;
  st.obuf = obuf;
  st.waves = kGemmWaves;
  for (auto& d : ds) {
    // long reductions (the stacked sums over terms) split along K into
    // chunks that sum inside their workgroup
    if (d.op != GM_COPY && d.nin > 2 * E.gmSplitK_) {
      d.ksplit = split_chunks(d.nin, kGemmWavesWide, E.gmSplitK_, kGemmWavesWide);
      if (d.ksplit > kGemmWaves) st.waves = kGemmWavesWide;
    }
  }
  const int64_t t = order_tasks(ds, gm_tasks);
  st.d0 = (int64_t)h_gm_.size();
  for (auto& d : ds) h_gm_.insert(h_gm_.end(), (size_t)gm_tasks(d), d);
  st.n = t;
  P.push_back(st);
}

void TreeProg::push_gc(std::vector<Step>& P, std::vector<GcDesc>& ds) {
  if (ds.empty()) return;
  // the split rule of the engine's grouped hole contractions
  int64_t tiles = 0;
  for (auto& d : ds) tiles += gc_tasks(d);
  const int64_t want = (E.splitWaves_ + tiles - 1) / tiles;
  for (auto& d : ds) {
    const int64_t K = (int64_t)d.bef * d.aft;
    d.ksplit = K > 2 * E.gcLongK_ ? split_chunks(K, kGemmWaves, 1) : split_chunks(K, want, E.splitMinK_);
  }
  const int64_t t = order_tasks(ds, gc_tasks);
  Step st;
  st.kind = K_GC;
  st.d0 = (int64_t)h_gc_.size();
  for (auto& d : ds) h_gc_.insert(h_gc_.end(), (size_t)gc_tasks(d), d);
  st.n = t;
  P.push_back(st);
}

void TreeProg::push_comb_sized(std::vector<Step>& P,
                               const std::vector<std::pair<int64_t, std::vector<CombTerm>>>& outs,
                               const std::vector<int>& n) {
  if (outs.empty()) return;
  Step st;
  st.kind = K_COMB;
  st.d0 = (int64_t)h_co_.size();
  for (size_t i = 0; i < outs.size(); ++i) {
    CombOut co{};
    co.out = outs[i].first;
    co.n = n[i];
    co.t0 = (int32_t)h_ct_.size();
    h_ct_.insert(h_ct_.end(), outs[i].second.begin(), outs[i].second.end());
    co.t1 = (int32_t)h_ct_.size();
    h_co_.push_back(co);
  }
  st.n = (int64_t)outs.size();
  P.push_back(st);
}

void TreeProg::push_comb(std::vector<Step>& P, const std::vector<std::pair<int64_t, std::vector<CombTerm>>>& outs,
                         int n) {
  push_comb_sized(P, outs, std::vector<int>(outs.size(), n));
}

std::vector<std::vector<int64_t>> TreeProg::emit_chains_multi(std::vector<Step>& P,
                                                               const std::vector<ChainGroup>& groups) {
  std::vector<std::map<Chain, int64_t>> loc(groups.size());
  size_t maxd = 0;
  for (auto& g : groups)
    for (auto& r : g.reqs) maxd = std::max(maxd, r.first.size());
  for (size_t d = 1; d <= maxd; ++d) {
    std::vector<GmDesc> ds;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const ChainGroup& g = groups[gi];
      const int64_t N = E.tree().nodes[g.a].N;
      for (int pass = 0; pass < 2; ++pass) {  // requests ending at this depth first: they own a destination
        for (auto& r : g.reqs) {
          const Chain& c = r.first;
          if (c.size() < d || (pass == 0) != (c.size() == d)) continue;
          Chain pre(c.begin(), c.begin() + d);
          if (loc[gi].count(pre)) continue;
          const int64_t out = (c.size() == d && r.second >= 0) ? r.second : walloc(N);
          loc[gi][pre] = out;
          int xsel = 1;
          int64_t x;
          if (d == 1) {
            xsel = g.inbuf;
            x = g.in;
          } else {
            x = loc[gi].at(Chain(c.begin(), c.begin() + (d - 1)));
          }
          ds.push_back(gm_desc(g.a, c[d - 1].first, c[d - 1].second, xsel, x, out));
        }
      }
    }
    push_gm(P, ds, 1);
  }
  std::vector<std::vector<int64_t>> outs(groups.size());
  for (size_t gi = 0; gi < groups.size(); ++gi)
    for (auto& r : groups[gi].reqs) outs[gi].push_back(r.first.empty() ? groups[gi].in : loc[gi].at(r.first));
  return outs;
}

std::vector<int64_t> TreeProg::emit_chains(std::vector<Step>& P, int a, int inbuf, int64_t in,
                                           const std::vector<std::pair<Chain, int64_t>>& reqs) {
  return emit_chains_multi(P, {ChainGroup{a, inbuf, in, reqs}})[0];
}

int TreeProg::stack_desc(std::vector<Step>& comb, std::vector<Step>& body, int a, int inbuf, int64_t in,
                         std::vector<StackEntry>& entries, bool entry0_is_input, int64_t stack_base, int64_t out,
                         GmDesc& d, int64_t sblock) {
  const HNode& nd = E.tree().nodes[a];
  const int64_t N = nd.N, n = nd.n;
  std::vector<std::pair<Chain, int64_t>> reqs;
  for (size_t e = 0; e < entries.size(); ++e) {
    if (e == 0 && entry0_is_input) continue;
    Chain c = entries[e].chain;
    if (c.empty()) c.push_back({-1, 0});  // the input itself: copy it into the stack
    reqs.push_back({c, stack_base + (int64_t)e * N});
  }
  emit_chains(body, a, inbuf, in, reqs);
  const size_t start = (entry0_is_input && !entries.empty() && entries[0].block.empty()) ? 1 : 0;
  const int64_t Eff = (int64_t)entries.size() - (int64_t)start;
  if (Eff <= 0) {
    set_error(name_ + ": a node with nothing to sum");
    return QTG_ESTATE;
  }
  bool ident = Eff <= 4;
  for (size_t e = start; e < entries.size() && ident; ++e) {
    const auto& b = entries[e].block;
    ident = b.size() == 1 && b[0].src < 0 && b[0].cre == 1.0 && b[0].cim == 0.0;
  }
  int64_t m;
  if (ident) {
    m = ident_.at({(int)n, 4});
  } else {
    m = sblock >= 0 ? sblock : S_;
    std::vector<std::pair<int64_t, std::vector<CombTerm>>> outs;
    for (size_t e = start; e < entries.size(); ++e) outs.push_back({m + (int64_t)(e - start) * n * n, entries[e].block});
    push_comb(comb, outs, (int)n);
    if (sblock < 0) Smax_ = std::max(Smax_, Eff * n * n);
  }
  d = GmDesc{};
  d.x = stack_base + (int64_t)start * N;
  d.xsel = 1;
  d.m = m;
  d.out = out;
  d.bef = (int32_t)(N / n);
  d.nin = (int32_t)(n * Eff);
  d.nout = (int32_t)n;
  d.aft = 1;
  d.op = GM_N;
  d.acc = 0;
  d.are = 1.0;
  d.aim = 0.0;
  d.node = -1;
  d.ksplit = 1;
  return QTG_OK;
}

int TreeProg::emit_stack(std::vector<Step>& comb, std::vector<Step>& body, int a, int inbuf, int64_t in,
                         std::vector<StackEntry>& entries, bool entry0_is_input, int64_t stack_base, int64_t out) {
  GmDesc d;
  int rc = stack_desc(comb, body, a, inbuf, in, entries, entry0_is_input, stack_base, out, d);
  if (rc) return rc;
  std::vector<GmDesc> ds{d};
  push_gm(body, ds, 1);
  return QTG_OK;
}

int64_t TreeProg::qr_item(int64_t a, int64_t q, int64_t rows, int n) {
  SmItem it{};
  it.a = a;
  it.q = q;
  it.e = -1;
  it.m = rows;
  it.n = n;
  h_qr_.push_back(it);
  return (int64_t)h_qr_.size() - 1;
}

namespace {
bool overlap(const TreeProg* /*unused*/, int ab, int64_t al, int64_t ah, int bb, int64_t bl, int64_t bh) {
  return ab == bb && al < bh && bl < ah;
}
}  // namespace

int TreeProg::validate_steps(const std::vector<Step>& P) const {
  const int64_t size[3] = {E.tree().total, workSize_, poolSize_};
  auto fail = [&](const std::string& what) {
    set_error(name_ + " plan check: " + what);
    return QTG_ESTATE;
  };
  auto in_range = [&](const Span& s) { return s.lo >= 0 && s.lo <= s.hi && s.hi <= size[s.buf]; };
  for (size_t si = 0; si < P.size(); ++si) {
    const Step& st = P[si];
    std::vector<Span> rd, wr;
    bool ordered = false;  // a single-workgroup kernel that orders its own accesses
    switch (st.kind) {
      case K_GM:
        for (int64_t t = 0; t < st.n; ++t) {
          const GmDesc& d = h_gm_.at(st.d0 + t);
          if (t > 0 && d.tile0 == h_gm_[st.d0 + t - 1].tile0) continue;  // the same descriptor
          const int64_t R = (int64_t)d.bef * d.aft;
          rd.push_back({d.xsel ? 1 : 0, d.x, d.x + R * d.nin});
          if (d.op != GM_COPY) rd.push_back({2, d.m, d.m + (int64_t)d.nin * d.nout});
          else if (d.nin != d.nout) return fail("copy with nin != nout");
          wr.push_back({st.obuf, d.out, d.out + R * d.nout});
        }
        break;
      case K_GC:
        for (int64_t t = 0; t < st.n; ++t) {
          const GcDesc& d = h_gc_.at(st.d0 + t);
          if (t > 0 && d.tile0 == h_gc_[st.d0 + t - 1].tile0) continue;
          const int64_t K = (int64_t)d.bef * d.aft;
          rd.push_back({(d.sel & 1) ? 1 : 0, d.x, d.x + K * d.nx});
          rd.push_back({(d.sel & 2) ? 1 : 0, d.y, d.y + K * d.ny});
          wr.push_back({2, d.out, d.out + (int64_t)d.nx * d.ny});
        }
        break;
      case K_COMB:
        for (int64_t i = 0; i < st.n; ++i) {
          const CombOut& co = h_co_.at(st.d0 + i);
          const int64_t n2 = (int64_t)co.n * co.n;
          wr.push_back({2, co.out, co.out + n2});
          for (int32_t t = co.t0; t < co.t1; ++t)
            if (h_ct_.at(t).src >= 0) rd.push_back({2, h_ct_[t].src, h_ct_[t].src + n2});
        }
        break;
      case K_PERM: {
        const int64_t N = st.bef * st.nn * st.aft;
        rd.push_back({st.abuf, st.a, st.a + N});
        wr.push_back({st.cbuf, st.c, st.c + N});
        break;
      }
      case K_QR:
        if (st.nn < 1 || st.nn > 64 || st.bef < st.nn) return fail("QR shape");
        rd.push_back({st.abuf, st.a, st.a + st.bef * st.nn});
        if (!(st.abuf == st.cbuf && st.a == st.c)) wr.push_back({st.cbuf, st.c, st.c + st.bef * st.nn});
        break;
      case K_QRB:
        for (int64_t i = 0; i < st.n; ++i) {
          const SmItem& it = h_qr_.at(st.d0 + i);
          if (it.n < 1 || it.n > 64 || it.m < it.n) return fail("QR shape");
          rd.push_back({st.abuf, it.a, it.a + it.m * it.n});
          if (!(st.abuf == st.cbuf && it.a == it.q)) wr.push_back({st.cbuf, it.q, it.q + it.m * it.n});
        }
        break;
      default: {
        int rc = validate_extra(st, rd, wr);
        if (rc) return rc;
        ordered = true;
      }
    }
    for (auto& r : rd)
      if (!in_range(r)) return fail("read outside its buffer in step " + std::to_string(si));
    for (auto& w : wr)
      if (!in_range(w)) return fail("write outside its buffer in step " + std::to_string(si));
    if (ordered) continue;
    // within one launch: outputs disjoint, and no output over an input
    std::vector<Span> w = wr;
    std::sort(w.begin(), w.end(), [](const Span& x, const Span& y) { return std::tie(x.buf, x.lo) < std::tie(y.buf, y.lo); });
    for (size_t i = 1; i < w.size(); ++i)
      if (overlap(this, w[i - 1].buf, w[i - 1].lo, w[i - 1].hi, w[i].buf, w[i].lo, w[i].hi))
        return fail("overlapping outputs in step " + std::to_string(si));
    // every read against the sorted, disjoint writes: the first write that
    // ends after the read starts is the only one that can overlap it
    for (auto& r : rd) {
      auto it = std::lower_bound(w.begin(), w.end(), r, [](const Span& x, const Span& y) {
        return std::tie(x.buf, x.hi) <= std::tie(y.buf, y.lo);
      });
      if (it != w.end() && overlap(this, it->buf, it->lo, it->hi, r.buf, r.lo, r.hi))
        return fail("a launch writes what it reads in step " + std::to_string(si));
    }
  }
  return QTG_OK;
}

int TreeProg::validate_extra(const Step&, std::vector<Span>&, std::vector<Span>&) const {
  set_error(name_ + " plan check: unknown step");
  return QTG_ESTATE;
}

int TreeProg::run_extra(const Step&, hipStream_t) {
  set_error(name_ + ": unknown step");
  return QTG_ESTATE;
}

int TreeProg::materialize(const std::vector<std::pair<int64_t, const std::vector<cd>*>>& poolInit) {
  int rc;
  QTG_HIP_TRY(hipStreamCreate(&s_));
  if ((rc = dmalloc((void**)&psi_, 16 * (size_t)E.tree().total)) ||
      (rc = dmalloc((void**)&work_, 16 * (size_t)workSize_)) || (rc = dmalloc((void**)&pool_, 16 * (size_t)poolSize_)) ||
      (rc = dmalloc((void**)&d_gm_, sizeof(GmDesc) * h_gm_.size())) ||
      (rc = dmalloc((void**)&d_gc_, sizeof(GcDesc) * h_gc_.size())) ||
      (rc = dmalloc((void**)&d_co_, sizeof(CombOut) * h_co_.size())) ||
      (rc = dmalloc((void**)&d_ct_, sizeof(CombTerm) * h_ct_.size())) ||
      (rc = dmalloc((void**)&d_qr_, sizeof(SmItem) * h_qr_.size())))
    return rc;
  std::vector<cd> host(poolSize_, cd(0.0, 0.0));
  for (auto& pi : poolInit) std::copy(pi.second->begin(), pi.second->end(), host.begin() + pi.first);
  for (auto& id : ident_) {
    const int n = id.first.first;
    for (int b = 0; b < id.first.second; ++b)
      for (int i = 0; i < n; ++i) host[id.second + (int64_t)b * n * n + i + (int64_t)i * n] = 1.0;
  }
  QTG_HIP_TRY(hipMemcpy(pool_, host.data(), 16 * (size_t)poolSize_, hipMemcpyHostToDevice));
  QTG_HIP_TRY(hipMemset(work_, 0, 16 * (size_t)workSize_));
  if (!h_gm_.empty()) QTG_HIP_TRY(hipMemcpy(d_gm_, h_gm_.data(), sizeof(GmDesc) * h_gm_.size(), hipMemcpyHostToDevice));
  if (!h_gc_.empty()) QTG_HIP_TRY(hipMemcpy(d_gc_, h_gc_.data(), sizeof(GcDesc) * h_gc_.size(), hipMemcpyHostToDevice));
  if (!h_co_.empty())
    QTG_HIP_TRY(hipMemcpy(d_co_, h_co_.data(), sizeof(CombOut) * h_co_.size(), hipMemcpyHostToDevice));
  if (!h_ct_.empty())
    QTG_HIP_TRY(hipMemcpy(d_ct_, h_ct_.data(), sizeof(CombTerm) * h_ct_.size(), hipMemcpyHostToDevice));
  if (!h_qr_.empty()) QTG_HIP_TRY(hipMemcpy(d_qr_, h_qr_.data(), sizeof(SmItem) * h_qr_.size(), hipMemcpyHostToDevice));
  return QTG_OK;
}

int TreeProg::run_steps(const std::vector<Step>& P, hipStream_t s) {
  auto Z = [](double* b, int64_t off) { return (void*)(b + 2 * off); };
  int rc = QTG_OK;
  const SmItem zero{};
  for (const Step& st : P) {
    switch (st.kind) {
      case K_GM:
        rc = launch_gm<double, double, double>(d_gm_ + st.d0, nullptr, st.n, st.n, psi_, work_, pool_,
                                               st.obuf ? (void*)work_ : (void*)psi_, nullptr, 1.0, 0.0, s, nullptr,
                                               st.waves ? st.waves : kGemmWaves);
        break;
      case K_GC:
        rc = launch_gc<double, double>(d_gc_ + st.d0, nullptr, st.n, st.n, psi_, work_, pool_, nullptr, s, nullptr);
        break;
      case K_COMB:
        rc = launch_combine<double>(d_co_ + st.d0, st.n, d_ct_, pool_, s);
        break;
      case K_PERM:
        rc = launch_permute(Z(st.abuf ? work_ : psi_, st.a), Z(st.cbuf ? work_ : psi_, st.c), st.bef, st.nn, st.aft,
                            st.dir, s);
        break;
      case K_QR: {
        SmItem u{};
        u.a = st.a;
        u.q = st.c;
        u.e = -1;
        u.m = st.bef;
        u.n = (int32_t)st.nn;
        rc = launch_qr_batch(nullptr, u, zero, 1, st.abuf ? (void*)work_ : (void*)psi_,
                             st.cbuf ? (void*)work_ : (void*)psi_, s);
        break;
      }
      case K_QRB:
        rc = launch_qr_batch(d_qr_ + st.d0, zero, zero, st.n, st.abuf ? (void*)work_ : (void*)psi_,
                             st.cbuf ? (void*)work_ : (void*)psi_, s);
        break;
      default:
        rc = run_extra(st, s);
    }
    if (rc) return rc;
  }
  return QTG_OK;
}

}  // namespace qtg
