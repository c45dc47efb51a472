// Grouped-contraction path of the engine: every node contraction of a sweep
// is a descriptor of the MFMA tile kernels (gemm_ops.hip), so node tensors of
// any size work and each launch spreads over many wavefronts.
//
// Build.  A sandwich item out = <Psi_p| F_1 x F_2 (x F_3) |Psi_p> over all
// legs but k (represent: k = last leg; holes and densities: k = the child's
// leg) is factored as
//     out(j,l) = sum_{legs != k} conj(Zb(.., j, ..)) Zk(.., l, ..),
//     Zb = F_b^H x_b Psi_p   (factor on the lowest leg other than k),
//     Zk = F_c x_c (F_d x_d Psi_p)   (the remaining factors),
// and equal Zb / Zk tensors of a layer are computed once.  With shared slots
// (equal term restrictions share one matrix) this turns the per-term work
// of TreeFunctions::represent / contraction into a few mode-product launches
// followed by one grouped hole-contraction launch per layer.
//
// Apply.  For node v with summands s (coefficient c_s, matrices on the
// node's legs), the last leg L is split off:
//     Apply(v) = sum_s c_s G_s x_L T_s,   T_s = (prod_{legs != L} F) Psi_v,
// and the T_s are stored one after another, so the sum over s is one mode
// product over the stacked index (s, p') with the stacked matrix
// [c_1 G_1 | c_2 G_2 | ...] (G_s the hole matrix, identity at the top).
//
// Finish (FP64).  S = A^H A, X = A^H Y (grouped hole contractions), X <-
// S^-1 X and W = rho^-1 X^T (batched solve), then
//     dpsi = f rho^-1 x_L Y - f W x_L A,   f = -i phase,
// which is -i phase rho^-1 (1 - P) Apply of HamiltonianRepresentation.cpp:53-72.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>

#include "engine.hpp"

namespace qtg {

namespace {


void view(const HNode& v, int k, int64_t& bef, int& nk, int64_t& aft) {
  bef = 1;
  aft = 1;
  for (int i = 0; i < k; ++i) bef *= v.dims[i];
  nk = v.dims[k];
  for (int i = k + 1; i < v.order(); ++i) aft *= v.dims[i];
}

GmDesc gm_desc(const HNode& v, int leg, int64_t x, int xsel, int64_t m, int64_t out, int op, int node) {
  GmDesc g{};
  int64_t bef, aft;
  int nk;
  view(v, leg, bef, nk, aft);
  g.x = x;
  g.m = m;
  g.out = out;
  g.bef = (int32_t)bef;
  g.nin = nk;
  g.nout = nk;
  g.aft = (int32_t)aft;
  g.op = op;
  g.acc = 0;
  g.are = 1.0;
  g.aim = 0.0;
  g.node = node;
  g.xsel = xsel;
  return g;
}

}  // namespace

int Engine::plan_gemm() {
  const int NN = (int)tree_.nodes.size();
  h_gm_.clear();
  h_gc_.clear();
  h_gmwg_.clear();
  h_gcwg_.clear();
  h_solve_.clear();
  gBuild_.clear();
  gHole_.clear();
  gApply_.clear();
  idSlots_.clear();
  h_comb3_.clear();
  h_combt3_.clear();
  int64_t poolOff = poolSize_;
  int64_t pool64Off = pool64Size_;
  workSize_ = 0;
  double flops = 0.0;
  for (auto& x : phaseFlops_) x = 0.0;

  auto push_gm = [&](std::vector<GLaunch>& L, std::vector<GmDesc>& descs, int prec, int x0, int x1, int m, int o,
                     bool scaled) {
    if (descs.empty()) return;
    GLaunch gl{};
    gl.kind = 0;
    gl.prec = prec;
    gl.d0 = (int64_t)h_gm_.size();
    gl.nd = (int64_t)descs.size();
    double fl = 0.0;
    gl.waves = kGemmWaves;
    // algorithmic bytes: every distinct input tensor and matrix once, every
    // output written once (and read once when accumulated)
    const double ein = prec == P64 ? 16.0 : 8.0, eout = prec == P32 ? 8.0 : 16.0;
    std::set<std::pair<int, int64_t>> seenX;
    std::set<int64_t> seenM;
    double by = 0.0;
    for (auto& d : descs) {
      if (seenX.insert({d.xsel, d.x}).second) by += ein * d.bef * (double)d.nin * d.aft;
      if (d.op != GM_COPY && seenM.insert(d.m).second) by += 16.0 * d.nin * (double)d.nout;
      by += eout * (d.acc ? 2.0 : 1.0) * d.bef * (double)d.nout * d.aft;
    }
    for (auto& d : descs) {
      // long contractions (stacked Apply summands) are split along K into
      // chunks that sum inside their workgroup
      if (d.op != GM_COPY && d.nin > 2 * gmSplitK_) {
        d.ksplit = split_chunks(d.nin, kGemmWavesWide, gmSplitK_, kGemmWavesWide);
        if (d.ksplit > kGemmWaves) gl.waves = kGemmWavesWide;
      }
      if (d.op != GM_COPY) fl += 8.0 * d.bef * d.aft * (double)d.nin * d.nout;
    }
    gl.lean = std::all_of(descs.begin(), descs.end(), [](const GmDesc& d) { return gm_lean_ok(d); });
    const int64_t t = order_tasks(descs, gm_tasks);
    // one descriptor per task (its copy keeps tile0), so a wavefront's
    // descriptor is one load at descs[task]
    for (auto& d : descs) h_gm_.insert(h_gm_.end(), (size_t)gm_tasks(d), d);
    gl.nd = t;
    gl.wg0 = -1;
    gl.ntasks = t;
    gl.x0 = x0;
    gl.x1 = x1;
    gl.m = m;
    gl.o = o;
    gl.scaled = scaled;
    gl.flops = fl;
    gl.bytes = by;
    flops += fl;
    L.push_back(gl);
  };
  auto push_gc = [&](std::vector<GLaunch>& L, std::vector<GcDesc>& descs, int prec, int x0, int x1, int o) {
    if (descs.empty()) return;
    GLaunch gl{};
    gl.kind = 1;
    gl.prec = prec;
    gl.d0 = (int64_t)h_gc_.size();
    gl.nd = (int64_t)descs.size();
    // descriptors sharing a bra tensor sit next to each other, so the
    // wavefronts of one workgroup mostly read the same X through L1
    std::stable_sort(descs.begin(), descs.end(), [](const GcDesc& a, const GcDesc& b) {
      return std::make_tuple(a.sel & 1, a.x, a.y) < std::make_tuple(b.sel & 1, b.x, b.y);
    });
    // launches with few tiles split their reductions along K (2 or 4
    // chunks of at least splitMinK_) toward about splitWaves_ wavefronts (a
    // 16 x 16 tile is one); beyond 2 gcLongK_ a reduction is always split
    int64_t tiles = 0;
    for (auto& d : descs) tiles += gc_tasks(d);
    const int64_t want = (splitWaves_ + tiles - 1) / tiles;
    const double ein = prec == P64 ? 16.0 : 8.0;
    std::set<std::pair<int, int64_t>> seen;
    double by = 0.0;
    for (auto& d : descs) {
      const double K = (double)d.bef * d.aft;
      if (seen.insert({d.sel & 1, d.x}).second) by += ein * K * d.nx;
      if (seen.insert({(d.sel >> 1) & 1, d.y}).second) by += ein * K * d.ny;
      by += 16.0 * (d.acc ? 2.0 : 1.0) * d.nx * (double)d.ny;
    }
    double fl = 0.0;
    for (auto& d : descs) {
      const int64_t K = (int64_t)d.bef * d.aft;
      d.ksplit = K > 2 * gcLongK_ ? split_chunks(K, kGemmWaves, 1) : split_chunks(K, want, splitMinK_);
      fl += 8.0 * d.bef * d.aft * (double)d.nx * d.ny;
    }
    gl.lean = std::all_of(descs.begin(), descs.end(), [](const GcDesc& d) { return gc_lean_ok(d); });
    const int64_t t = order_tasks(descs, gc_tasks);
    for (auto& d : descs) h_gc_.insert(h_gc_.end(), (size_t)gc_tasks(d), d);
    gl.nd = t;
    gl.wg0 = -1;
    gl.ntasks = t;
    gl.x0 = x0;
    gl.x1 = x1;
    gl.m = -1;
    gl.o = o;
    gl.scaled = false;
    gl.flops = fl;
    gl.bytes = by;
    flops += fl;
    L.push_back(gl);
  };

  // ---------------------------------------------------------------- build
  // items of a layer; rho: 1 keeps only density items, 0 drops them, -1 keeps all
  auto is_rho = [&](const SwItem& it) {
    const HNode& p = tree_.nodes[it.node];
    if (it.k >= (int)p.children.size() || p.bottom) return false;
    return it.out == rhoslot_[p.children[it.k]];
  };
  // Each item runs in FP64 (complex128 engines; in the split form the items
  // whose output serves a large-coefficient term, and the densities) or in
  // FP32 (the complex64 bulk); equal Z tensors are shared per precision.
  auto item_hi = [&](int64_t out) {
    return precision_ == 0 || (precision_ == 2 && highSlots_.count(out) > 0);
  };
  int64_t wbase = 0, wbase32 = 0, wmax = 0, wmax32 = 0;
  // descriptors of one layer: first factors, second factors, contractions
  // ([0] FP32, [1] FP64)
  struct LayerLists {
    std::vector<GmDesc> lv0[2], lv1[2];
    std::vector<GcDesc> con[2];
  };
  auto sandwich_layer = [&](const std::pair<int64_t, int64_t>& r, int rho, LayerLists& out) {
    if (r.second == r.first) return;
    struct Side {
      std::map<std::tuple<int, int, int64_t>, int64_t> zbra;  // (node, leg, mat) -> work offset
      std::map<std::vector<int64_t>, int64_t> zket1, zket2;   // (node, leg, mat[, leg, mat])
      std::vector<GmDesc> lv0, lv1;
      std::vector<GcDesc> con;
      int64_t w = 0;
    } sd[2];  // [0] FP32, [1] FP64
    sd[0].w = wbase32;
    sd[1].w = wbase;
    for (int64_t i = r.first; i < r.second; ++i) {
      const SwItem& it = h_items_[i];
      if (rho >= 0 && (int)is_rho(it) != rho) continue;
      Side& S = sd[item_hi(it.out) ? 1 : 0];
      const HNode& p = tree_.nodes[it.node];
      const int k = it.k;
      const int braleg = (k == 0) ? 1 : 0;
      int64_t xoff = p.off;
      int xsel = 0;
      std::vector<std::pair<int, int64_t>> ket;
      for (int f = 0; f < it.nf; ++f) {
        if (it.fmode[f] == braleg) {
          auto key = std::make_tuple((int)it.node, braleg, it.fmat[f]);
          auto z = S.zbra.find(key);
          if (z == S.zbra.end()) {
            S.lv0.push_back(gm_desc(p, braleg, p.off, 0, it.fmat[f], S.w, GM_H, it.node));
            z = S.zbra.emplace(key, S.w).first;
            S.w += p.N;
          }
          xoff = z->second;
          xsel = 1;
        } else {
          ket.push_back({it.fmode[f], it.fmat[f]});
        }
      }
      std::sort(ket.begin(), ket.end(), [](auto& a, auto& b) { return a.first > b.first; });
      int64_t yoff = p.off;
      int ysel = 0;
      if (!ket.empty()) {
        std::vector<int64_t> key1 = {it.node, ket[0].first, ket[0].second};
        auto z = S.zket1.find(key1);
        if (z == S.zket1.end()) {
          S.lv0.push_back(gm_desc(p, ket[0].first, p.off, 0, ket[0].second, S.w, GM_N, it.node));
          z = S.zket1.emplace(key1, S.w).first;
          S.w += p.N;
        }
        yoff = z->second;
        ysel = 1;
        if (ket.size() > 1) {
          std::vector<int64_t> key2 = key1;
          key2.push_back(ket[1].first);
          key2.push_back(ket[1].second);
          auto z2 = S.zket2.find(key2);
          if (z2 == S.zket2.end()) {
            S.lv1.push_back(gm_desc(p, ket[1].first, yoff, 1, ket[1].second, S.w, GM_N, it.node));
            z2 = S.zket2.emplace(key2, S.w).first;
            S.w += p.N;
          }
          yoff = z2->second;
        }
      }
      GcDesc g{};
      int64_t bef, aft;
      int nk;
      view(p, k, bef, nk, aft);
      g.x = xoff;
      g.y = yoff;
      g.out = it.out;
      g.bef = (int32_t)bef;
      g.nx = g.ny = nk;
      g.aft = (int32_t)aft;
      g.acc = 0;
      g.node = it.node;
      g.ksplit = 1;
      g.sel = xsel | (ysel << 1);
      S.con.push_back(g);
    }
    wmax32 = std::max(wmax32, sd[0].w);
    wmax = std::max(wmax, sd[1].w);
    for (int hp = 0; hp < 2; ++hp) {
      auto app = [](auto& dst, auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
      app(out.lv0[hp], sd[hp].lv0);
      app(out.lv1[hp], sd[hp].lv1);
      app(out.con[hp], sd[hp].con);
    }
  };
  auto push_layer = [&](LayerLists& x, std::vector<GLaunch>& L) {
    push_gm(L, x.lv0[1], P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
    push_gm(L, x.lv0[0], P32, B_PSI32, B_WORK32, B_POOL, B_WORK32, false);
    push_gm(L, x.lv1[1], P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
    push_gm(L, x.lv1[0], P32, B_PSI32, B_WORK32, B_POOL, B_WORK32, false);
    push_gc(L, x.con[1], P64, B_PSI64, B_WORK64, B_POOL);
    push_gc(L, x.con[0], P32, B_PSI32, B_WORK32, B_POOL);
  };
  // Represent (bottom-up) and the density tree (top-down) are independent
  // chains of the same depth: layer j of each runs in the same launches
  // (the density layer in its own workspace region), so the densities are
  // ready when represent ends.  The holes of the terms (top-down), which
  // use the densities as the parent factor of terms acting only below a
  // node, follow; the eigensolver of rho^-1 runs on a side stream beside them.
  std::vector<LayerLists> rep(repRange_.size()), hol(holeRange_.size()), rho(holeRange_.size());
  for (size_t i = 0; i < repRange_.size(); ++i) sandwich_layer(repRange_[i], -1, rep[i]);
  for (size_t dd = 1; dd < holeRange_.size(); ++dd) sandwich_layer(holeRange_[dd], 0, hol[dd]);
  wbase = wmax;
  wbase32 = wmax32;
  for (size_t dd = 1; dd < holeRange_.size(); ++dd) sandwich_layer(holeRange_[dd], 1, rho[dd]);
  workSize_ = std::max(workSize_, wmax);
  work32Size_ = std::max(work32Size_, wmax32);
  {
    const size_t nl = std::max(rep.size(), holeRange_.size() > 0 ? holeRange_.size() - 1 : 0);
    for (size_t j = 0; j < nl; ++j) {
      LayerLists m;
      for (LayerLists* src : {j < rep.size() ? &rep[j] : nullptr, j + 1 < rho.size() ? &rho[j + 1] : nullptr}) {
        if (!src) continue;
        for (int hp = 0; hp < 2; ++hp) {
          m.lv0[hp].insert(m.lv0[hp].end(), src->lv0[hp].begin(), src->lv0[hp].end());
          m.lv1[hp].insert(m.lv1[hp].end(), src->lv1[hp].begin(), src->lv1[hp].end());
          m.con[hp].insert(m.con[hp].end(), src->con[hp].begin(), src->con[hp].end());
        }
      }
      push_layer(m, gBuild_);
    }
    // Hole layers.  Fused form (complex128, every ket a single factor, every
    // output at most 16 x 16): one hole_fused launch per layer, each ket
    // workgroup building its Z in LDS and running the items that read it.
    // Measured on MI355X it pays only on layers of small parent tensors
    // (at most kFuseMaxN elements; larger kets serialise too much work in one
    // workgroup), so the other layers keep a mode-product and a
    // hole-contraction launch.  The bras of the fused layers read only the
    // state and the build's h-matrices: they ride in the mode-product launch
    // of the last two-launch layer before the first fused one (the first
    // layer is never fused).
    h_hk_.clear();
    h_hi_.clear();
    constexpr int64_t kFuseMaxN = 1200;
    const size_t NL = hol.size();
    std::vector<char> fl(NL, 0);
    std::vector<std::vector<HoleKet>> lk(NL);
    std::vector<std::vector<HoleItem>> li(NL);
    std::vector<int64_t> llds(NL, 0);
    std::vector<double> lfl(NL, 0.0), lby(NL, 0.0);
    std::vector<GmDesc> bras;
    std::map<std::tuple<int, int, int64_t>, int64_t> braOff;
    const int64_t bw0 = std::max(wmax, workSize_);  // bras after every other work tensor of the sweep
    int64_t bw = bw0;
    // the first layer keeps two launches: its mode products carry the bras
    for (size_t dd = 2; fuseHoles_ && precision_ == 0 && dd < holeRange_.size(); ++dd) {
      std::map<std::tuple<int, int, int64_t>, std::vector<HoleItem>> byKet;  // (node, leg, matrix)
      bool ok = true;
      std::vector<GmDesc> lb;  // this layer's new bras
      std::map<std::tuple<int, int, int64_t>, int64_t> lbo;
      int64_t lw = bw;
      double fl_ = 0.0, by_ = 0.0;
      for (int64_t i = holeRange_[dd].first; i < holeRange_[dd].second && ok; ++i) {
        const SwItem& it = h_items_[i];
        if (is_rho(it)) continue;
        const HNode& p = tree_.nodes[it.node];
        const int k = it.k, braleg = (k == 0) ? 1 : 0;
        int64_t bmat = -1, kmat = -1;
        int kleg = -1, nket = 0;
        for (int f = 0; f < it.nf; ++f) {
          if (it.fmode[f] == braleg) {
            bmat = it.fmat[f];
          } else {
            kleg = it.fmode[f];
            kmat = it.fmat[f];
            ++nket;
          }
        }
        int64_t bef, aft;
        int nk;
        view(p, k, bef, nk, aft);
        if (nket > 1 || nk > 16 || p.N > kFuseMaxN) {
          ok = false;
          break;
        }
        HoleItem hi{};
        if (bmat >= 0) {
          const auto key = std::make_tuple((int)it.node, braleg, bmat);
          auto f = braOff.find(key);
          if (f == braOff.end()) {
            f = lbo.find(key);
            if (f == lbo.end()) {
              lb.push_back(gm_desc(p, braleg, p.off, 0, bmat, lw, GM_H, it.node));
              f = lbo.emplace(key, lw).first;
              lw += p.N;
            }
          }
          hi.x = f->second;
          hi.sel = 1;
        } else {
          hi.x = p.off;
          hi.sel = 0;
        }
        hi.out = it.out;
        hi.bef = (int32_t)bef;
        hi.n = nk;
        hi.aft = (int32_t)aft;
        byKet[std::make_tuple((int)it.node, kmat >= 0 ? kleg : -1, kmat)].push_back(hi);
        fl_ += 8.0 * bef * aft * (double)nk * nk;
        by_ += 16.0 * ((double)bef * aft * nk + (double)nk * nk);
      }
      if (!ok || byKet.empty()) continue;
      fl[dd] = 1;
      bras.insert(bras.end(), lb.begin(), lb.end());
      braOff.insert(lbo.begin(), lbo.end());
      bw = lw;
      for (auto& kv : byKet) {
        const HNode& p = tree_.nodes[std::get<0>(kv.first)];
        HoleKet hk{};
        hk.x = p.off;
        hk.m = std::get<2>(kv.first);
        hk.node = std::get<0>(kv.first);
        if (hk.m >= 0) {
          int64_t bef, aft;
          int nl;
          view(p, std::get<1>(kv.first), bef, nl, aft);
          hk.bef = (int32_t)bef;
          hk.nleg = nl;
          hk.aft = (int32_t)aft;
          fl_ += 8.0 * bef * aft * (double)nl * nl;
          by_ += 16.0 * ((double)p.N + (double)nl * nl);
        } else {
          hk.bef = (int32_t)p.N;
          hk.nleg = 1;
          hk.aft = 1;
          by_ += 16.0 * p.N;
        }
        // at most 8 items (two rounds of four wavefronts) per workgroup: a ket
        // with more is built by several workgroups, each running a share
        const int32_t i0 = (int32_t)li[dd].size();
        li[dd].insert(li[dd].end(), kv.second.begin(), kv.second.end());
        const int32_t i1 = (int32_t)li[dd].size();
        for (int32_t a = i0; a < i1; a += 8) {
          hk.it0 = a;
          hk.it1 = std::min(i1, a + 8);
          lk[dd].push_back(hk);
        }
        llds[dd] = std::max(llds[dd], hole_fused_lds(p.N));
      }
      lfl[dd] = fl_;
      lby[dd] = by_;
    }
    workSize_ = std::max(workSize_, bw);
    size_t first = NL;
    for (size_t dd = 1; dd < NL; ++dd)
      if (fl[dd]) {
        first = dd;
        break;
      }
    size_t host = 0;  // the two-launch layer whose mode products carry the bras
    for (size_t dd = 1; dd < first && dd < NL; ++dd) host = dd;
    if (first < NL) hol[host].lv0[1].insert(hol[host].lv0[1].end(), bras.begin(), bras.end());
    for (size_t dd = 1; dd < NL; ++dd) {
      if (!fl[dd]) {
        push_layer(hol[dd], gHole_);
        continue;
      }
      const int32_t ib = (int32_t)h_hi_.size();
      for (HoleKet& hk : lk[dd]) {
        hk.it0 += ib;
        hk.it1 += ib;
      }
      GLaunch gl{};
      gl.kind = 3;
      gl.prec = P64;
      gl.d0 = (int64_t)h_hk_.size();
      gl.nd = gl.ntasks = (int64_t)lk[dd].size();
      gl.x0 = B_PSI64;
      gl.x1 = B_WORK64;
      gl.m = B_POOL;
      gl.o = B_POOL;
      gl.lds = llds[dd];
      gl.flops = lfl[dd];
      gl.bytes = lby[dd];
      flops += lfl[dd];
      h_hk_.insert(h_hk_.end(), lk[dd].begin(), lk[dd].end());
      h_hi_.insert(h_hi_.end(), li[dd].begin(), li[dd].end());
      gHole_.push_back(gl);
    }
  }

  // ---------------------------------------------------------------- apply
  std::map<int, int64_t> idmap;
  auto identity = [&](int n) {
    auto f = idmap.find(n);
    if (f != idmap.end()) return f->second;
    const int64_t o = poolOff;
    poolOff += (int64_t)n * n;
    idSlots_.push_back({o, n});
    idmap[n] = o;
    return o;
  };
  int maxLevels = 0;
  for (auto& v : tree_.nodes) maxLevels = std::max(maxLevels, v.order() - 1);
  // [precision][level]; [0] FP32, [1] FP64
  std::vector<std::vector<GmDesc>> levels[2] = {std::vector<std::vector<GmDesc>>(maxLevels),
                                                std::vector<std::vector<GmDesc>>(maxLevels)};
  std::vector<GmDesc> stacked[2];
  int64_t wa[2] = {0, 0};
  auto summand_hi = [&](const Summand& s) {
    if (precision_ == 0) return true;
    if (precision_ == 1) return false;
    for (int64_t id : s.ids)
      if (id >= 0 && highSlots_.count(id)) return true;
    return false;
  };
  for (auto& v : tree_.nodes) {
    const int L = v.order() - 1;
    const int nL = v.dims[L];
    const auto& smode = smode_[v.addr];
    std::vector<Summand> part[2];
    for (auto& s : summ_[v.addr]) part[summand_hi(s) ? 1 : 0].push_back(s);
    if (part[0].empty() && part[1].empty()) {
      // no active term: Apply(v) = 0
      GmDesc g = gm_desc(v, L, v.off, 0, 0, v.off, GM_COPY, v.addr);
      g.are = 0.0;
      stacked[1].push_back(g);
      continue;
    }
    for (int hp = 1; hp >= 0; --hp) {
      const auto& sm = part[hp];
      if (sm.empty()) continue;
      int64_t& w = wa[hp];
      const int64_t stack0 = w;
      w += (int64_t)sm.size() * v.N;
      const int64_t mstack = poolOff;
      poolOff += (int64_t)sm.size() * nL * nL;
      std::map<std::vector<int64_t>, int64_t> zmemo;  // chain key -> work offset
      for (size_t s = 0; s < sm.size(); ++s) {
        std::vector<int64_t> fac(L, -1);
        int64_t gL = -1;
        for (size_t sl = 0; sl < smode.size(); ++sl) {
          if (sm[s].ids[sl] < 0) continue;
          if (smode[sl] == L) gL = sm[s].ids[sl];
          else fac[smode[sl]] = sm[s].ids[sl];
        }
        // chain from the highest leg below L down to leg 0; the last level
        // writes into the stack slot and sits in the last launch
        int64_t cur = v.off;
        int cursel = 0;
        std::vector<int64_t> key = {v.addr};
        for (int leg = L - 1; leg >= 0; --leg) {
          const int lvl = maxLevels - 1 - leg;
          const bool last = (leg == 0);
          const int64_t slot = stack0 + (int64_t)s * v.N;
          if (fac[leg] < 0) {
            if (last) levels[hp][lvl].push_back(gm_desc(v, 0, cur, cursel, 0, slot, GM_COPY, v.addr));
            continue;
          }
          key.push_back(leg);
          key.push_back(fac[leg]);
          if (last) {
            levels[hp][lvl].push_back(gm_desc(v, leg, cur, cursel, fac[leg], slot, GM_N, v.addr));
          } else {
            auto z = zmemo.find(key);
            if (z == zmemo.end()) {
              const int64_t o = w;
              w += v.N;
              levels[hp][lvl].push_back(gm_desc(v, leg, cur, cursel, fac[leg], o, GM_N, v.addr));
              z = zmemo.emplace(key, o).first;
            }
            cur = z->second;
            cursel = 1;
          }
        }
        // stacked matrix block s = c_s G_s
        CombOut co{};
        co.out = mstack + (int64_t)s * nL * nL;
        co.n = nL;
        co.t0 = (int32_t)h_combt3_.size();
        CombTerm ct{};
        ct.cre = sm[s].coef.real();
        ct.cim = sm[s].coef.imag();
        ct.src = (gL >= 0) ? gL : identity(nL);
        h_combt3_.push_back(ct);
        co.t1 = (int32_t)h_combt3_.size();
        h_comb3_.push_back(co);
      }
      GmDesc g{};
      g.x = stack0;
      g.xsel = 1;
      g.m = mstack;
      g.out = v.off;
      g.bef = (int32_t)(v.N / nL);
      g.nin = (int32_t)(sm.size() * nL);
      g.nout = nL;
      g.aft = 1;
      g.op = GM_N;
      g.acc = (hp == 0 && !part[1].empty()) ? 1 : 0;  // the FP32 part adds to the FP64 part
      g.are = 1.0;
      g.aim = 0.0;
      g.node = v.addr;
      stacked[hp].push_back(g);
    }
  }
  workSize_ = std::max(workSize_, wa[1]);
  work32Size_ = std::max(work32Size_, wa[0]);
  const double fa0 = flops;
  for (int lv = 0; lv < maxLevels; ++lv) {
    push_gm(gApply_, levels[1][lv], P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
    push_gm(gApply_, levels[0][lv], P32, B_PSI32, B_WORK32, B_POOL, B_WORK32, false);
  }
  push_gm(gApply_, stacked[1], P64, B_PSI64, B_WORK64, B_POOL, B_APPLY64, false);
  push_gm(gApply_, stacked[0], P32W, B_PSI32, B_WORK32, B_POOL, B_APPLY64, false);
  phaseFlops_[2] = flops - fa0;

  // --------------------------------------------------------------- finish
  // With SPF counts up to 48 the solve launch also writes the projected
  // derivative of every node (and the top node's scaled Apply); larger nodes
  // use two more mode-product launches for it.
  const double ff0 = flops;
  const bool fusedFinish = nspf_ <= 48;
  // with SPF counts up to 16 the solve also forms S = A^H A, A^H Y and the
  // top node's energy terms, and the Gram launch is dropped
  gramInSolve_ = nspf_ <= 16;
  std::vector<GcDesc> fin;
  std::vector<GmDesc> g1, g2;
  energySlot_ = pool64Off;
  pool64Off += 2;
  for (auto& v : tree_.nodes) {
    const int L = v.order() - 1;
    const int n = v.dims[L];
    GcDesc g{};
    g.bef = (int32_t)(v.N / n);
    g.aft = 1;
    g.nx = g.ny = n;
    g.node = v.addr;
    g.ksplit = 1;
    g.x = g.y = v.off;
    if (v.top()) {
      // energy: <A|Y> and <A|A> of the top node (one state)
      g.nx = g.ny = 1;
      g.bef = (int32_t)v.N;
      g.out = energySlot_;
      g.sel = 2;
      fin.push_back(g);
      g.out = energySlot_ + 1;
      g.sel = 0;
      fin.push_back(g);
      if (fusedFinish) {
        SolveItem si{};
        si.s = si.r = si.w = -1;
        si.x = gramInSolve_ ? energySlot_ : -1;  // <A|Y>, <A|A>
        si.a = v.off;
        si.y = si.out = v.off;
        si.n = n;
        si.rows = (int32_t)(v.N / n);
        si.node = v.addr;
        h_solve_.push_back(si);
      } else {
        g1.push_back(gm_desc(v, L, v.off, 1, 0, v.off, GM_COPY, v.addr));
      }
      continue;
    }
    const int64_t so = pool64Off, xo = so + (int64_t)n * n, wo = xo + (int64_t)n * n;
    pool64Off = wo + (int64_t)n * n;
    g.out = so;
    g.sel = 0;  // S = A^H A
    fin.push_back(g);
    g.out = xo;
    g.sel = 2;  // X = A^H Y
    fin.push_back(g);
    SolveItem si{};
    si.s = so;
    si.x = xo;
    si.r = rhoinv_[v.addr];
    si.w = wo;
    si.n = n;
    si.rows = (int32_t)(v.N / n);
    si.node = v.addr;
    si.a = si.y = v.off;
    si.out = fusedFinish ? v.off : -1;
    h_solve_.push_back(si);
    if (fusedFinish) continue;
    g1.push_back(gm_desc(v, L, v.off, 1, rhoinv_[v.addr], v.off, GM_N, v.addr));
    GmDesc m2 = gm_desc(v, L, v.off, 0, wo, v.off, GM_N, v.addr);
    m2.are = -1.0;
    m2.acc = 1;
    g2.push_back(m2);
  }
  if (!gramInSolve_) push_gc(gApply_, fin, P64, B_PSI64, B_APPLY64, B_POOL64);
  {
    GLaunch gl{};
    gl.kind = 2;
    gl.prec = P64;
    gl.nd = (int64_t)h_solve_.size();
    gl.x0 = gl.x1 = gl.m = gl.o = B_POOL64;
    gl.scaled = fusedFinish;
    double fl = 0.0;
    for (auto& si : h_solve_) {
      if (si.s >= 0) fl += 8.0 * 2.0 * si.n * si.n * si.n;
      if (si.s >= 0 && si.out >= 0) fl += 8.0 * 2.0 * si.n * si.n * (double)si.rows;  // the projected derivative
      if (gramInSolve_) fl += 8.0 * 2.0 * (double)si.rows * si.n * (si.s >= 0 ? si.n : 1);  // S, A^H Y / energy
    }
    gl.flops = fl;
    flops += fl;
    gApply_.push_back(gl);
  }
  push_gm(gApply_, g1, P64, B_PSI64, B_APPLY64, B_POOL64, B_DPSI64, true);
  push_gm(gApply_, g2, P64, B_PSI64, B_APPLY64, B_POOL64, B_DPSI64, true);
  phaseFlops_[3] = flops - ff0;

  // Two combination launches per sweep: the aggregated slots, then the sums
  // over them together with the stacked Apply blocks.  A block whose matrix
  // is such a sum is expanded into the sum's terms (over aggregated slots),
  // so no output of the second launch feeds another.
  {
    std::map<int64_t, int32_t> sum2;  // output offset -> index in h_comb2_
    for (size_t i = 0; i < h_comb2_.size(); ++i) sum2[h_comb2_[i].out] = (int32_t)i;
    std::vector<CombOut> mo = h_comb2_;
    std::vector<CombTerm> mt = h_combt2_;
    for (const CombOut& co : h_comb3_) {
      CombOut m = co;
      m.t0 = (int32_t)mt.size();
      for (int32_t q = co.t0; q < co.t1; ++q) {
        const CombTerm& t = h_combt3_[q];
        auto e = (t.src >= 0) ? sum2.find(t.src) : sum2.end();
        if (e == sum2.end()) {
          mt.push_back(t);
          continue;
        }
        const CombOut& c2 = h_comb2_[e->second];
        for (int32_t r = c2.t0; r < c2.t1; ++r) {
          const CombTerm& u = h_combt2_[r];
          CombTerm x{};
          x.cre = t.cre * u.cre - t.cim * u.cim;
          x.cim = t.cre * u.cim + t.cim * u.cre;
          x.src = u.src;
          mt.push_back(x);
        }
      }
      m.t1 = (int32_t)mt.size();
      mo.push_back(m);
    }
    std::set<int64_t> outs;
    for (const CombOut& co : mo) outs.insert(co.out);
    for (const CombTerm& t : mt)
      if (t.src >= 0 && outs.count(t.src)) {
        set_error("plan: a combination reads another output of its launch");
        return QTG_ESTATE;
      }
    h_comb3_ = mo;
    h_combt3_ = mt;
  }
  // combine flops (aggregated slots; sums over them and stacked blocks)
  double comb = 0.0;
  for (auto& co : h_comb1_) comb += 8.0 * co.n * co.n * (double)(co.t1 - co.t0);
  for (auto& co : h_comb3_) comb += 8.0 * co.n * co.n * (double)(co.t1 - co.t0);
  phaseFlops_[5] = comb;
  flops += comb;
  nComb3_ = (int64_t)h_comb3_.size();
  flops_ = flops;

  // ------------------------------------------------------------- gradient
  // dE/dA_v* = (H_eff,v A_v - E rho_v x_L A_v) / <Psi|Psi> (orthonormal
  // lower nodes).  H_eff,v A_v = Apply(v) + M_v x_L A_v, where Apply(v) is
  // HamiltonianRepresentation::Apply (terms active at v) and M_v, the mean
  // field of the terms inactive at v, follows top-down:
  //   M_v = <Psi_p| [ sum_{t active at p, not at v} c_t (x_sib h_t) x hole_t(p)
  //                   + M_p x_par ] |Psi_p>  over all legs of p but v's.
  // The first part is an Apply-like stacked sum at p (computed for all nodes
  // at once), the second a mode product on the parent leg that accumulates
  // into it once M_p exists; then one grouped hole contraction per depth.
  h_combG1_.clear();
  h_combG2_.clear();
  h_combG3_.clear();
  h_combtG3_.clear();
  h_combGE_.clear();
  h_combtGE_.clear();
  gGradPre_.clear();
  gGradFin_.clear();
  gGradDepth_.assign(tree_.maxDepth + 1, {});
  mslot_.assign(NN, -1);
  for (auto& v : tree_.nodes)
    if (!v.top()) {
      mslot_[v.addr] = poolOff;
      poolOff += (int64_t)v.n * v.n;
    }
  {
    // the aggregation helper appends to the build's combination lists; keep
    // the gradient's apart (they run only when a gradient is asked for)
    std::vector<CombOut> c1 = h_comb1_, c2 = h_comb2_;
    std::vector<CombTerm> t1 = h_combt1_, t2 = h_combt2_;
    h_comb1_.clear();
    h_comb2_.clear();
    h_combt1_.clear();
    h_combt2_.clear();
    double gcomb = 0.0;
    std::vector<std::vector<GmDesc>> glevels(maxLevels);
    std::vector<GmDesc> gstacked;
    std::vector<std::vector<GmDesc>> gacc(tree_.maxDepth + 1);
    std::vector<std::vector<GcDesc>> gcon(tree_.maxDepth + 1);
    int64_t gw = 0;
    const int L_ = (int)sop_.terms.size();
    for (auto& v : tree_.nodes) {
      if (v.top()) continue;
      const HNode& p = tree_.nodes[v.parent];
      const int Lp = p.order() - 1;
      const int nL = p.dims[Lp];
      // summands at p of the terms active at p and inactive at v
      std::vector<Summand> tl;
      for (int l = 0; l < L_; ++l) {
        if (!active_[l][p.addr] || active_[l][v.addr]) continue;
        Summand g;
        g.coef = sop_.terms[l].coeff;
        for (int c : p.children) g.ids.push_back(active_[l][c] ? hslot_[l][c] : -1);
        if (!p.top()) g.ids.push_back(holeslot_[l][p.addr]);
        tl.push_back(g);
      }
      std::vector<Summand> sm = compress_ ? aggregate(tl, sdim_[p.addr], poolOff, gcomb) : tl;
      const int64_t yv = gw;  // Y_v (N_p elements)
      gw += p.N;
      if (sm.empty()) {
        GmDesc z = gm_desc(p, Lp, p.off, 0, 0, yv, GM_COPY, -1);
        z.are = 0.0;
        z.xsel = 0;
        gstacked.push_back(z);
      } else {
        const int64_t stack0 = gw;
        gw += (int64_t)sm.size() * p.N;
        const int64_t mstack = poolOff;
        poolOff += (int64_t)sm.size() * nL * nL;
        std::map<std::vector<int64_t>, int64_t> zmemo;
        for (size_t s = 0; s < sm.size(); ++s) {
          std::vector<int64_t> fac(Lp, -1);
          int64_t gL = -1;
          for (size_t sl = 0; sl < smode_[p.addr].size(); ++sl) {
            if (sm[s].ids[sl] < 0) continue;
            if (smode_[p.addr][sl] == Lp) gL = sm[s].ids[sl];
            else fac[smode_[p.addr][sl]] = sm[s].ids[sl];
          }
          int64_t cur = p.off;
          int cursel = 0;
          std::vector<int64_t> key = {p.addr};
          for (int leg = Lp - 1; leg >= 0; --leg) {
            const int lvl = maxLevels - 1 - leg;
            const bool last = (leg == 0);
            const int64_t slot = stack0 + (int64_t)s * p.N;
            if (fac[leg] < 0) {
              if (last) glevels[lvl].push_back(gm_desc(p, 0, cur, cursel, 0, slot, GM_COPY, -1));
              continue;
            }
            key.push_back(leg);
            key.push_back(fac[leg]);
            if (last) {
              glevels[lvl].push_back(gm_desc(p, leg, cur, cursel, fac[leg], slot, GM_N, -1));
            } else {
              auto z = zmemo.find(key);
              if (z == zmemo.end()) {
                const int64_t o = gw;
                gw += p.N;
                glevels[lvl].push_back(gm_desc(p, leg, cur, cursel, fac[leg], o, GM_N, -1));
                z = zmemo.emplace(key, o).first;
              }
              cur = z->second;
              cursel = 1;
            }
          }
          CombOut co{};
          co.out = mstack + (int64_t)s * nL * nL;
          co.n = nL;
          co.t0 = (int32_t)h_combtG3_.size();
          CombTerm ct{};
          ct.cre = sm[s].coef.real();
          ct.cim = sm[s].coef.imag();
          ct.src = (gL >= 0) ? gL : identity(nL);
          h_combtG3_.push_back(ct);
          co.t1 = (int32_t)h_combtG3_.size();
          h_combG3_.push_back(co);
        }
        GmDesc g{};
        g.x = stack0;
        g.xsel = 1;
        g.m = mstack;
        g.out = yv;
        g.bef = (int32_t)(p.N / nL);
        g.nin = (int32_t)(sm.size() * nL);
        g.nout = nL;
        g.aft = 1;
        g.op = GM_N;
        g.are = 1.0;
        g.node = -1;
        gstacked.push_back(g);
      }
      // Y_v += M_p x_par Psi_p, once M_p exists (depth of p)
      if (!p.top()) {
        GmDesc a = gm_desc(p, Lp, p.off, 0, mslot_[p.addr], yv, GM_N, -1);
        a.acc = 1;
        gacc[v.depth].push_back(a);
      }
      // M_v = <Psi_p| Y_v> over the legs of p but v's
      GcDesc c{};
      int64_t bef, aft;
      int nk;
      view(p, v.childIdx, bef, nk, aft);
      c.x = p.off;
      c.y = yv;
      c.out = mslot_[v.addr];
      c.bef = (int32_t)bef;
      c.nx = c.ny = nk;
      c.aft = (int32_t)aft;
      c.node = -1;
      c.ksplit = 1;
      c.sel = 2;  // X = Psi (first base), Y = workspace (second base)
      gcon[v.depth].push_back(c);
    }
    workSize_ = std::max(workSize_, gw);
    h_combG1_ = h_comb1_;
    h_combG2_ = h_comb2_;
    h_combtG1_ = h_combt1_;
    h_combtG2_ = h_combt2_;
    h_comb1_ = c1;
    h_comb2_ = c2;
    h_combt1_ = t1;
    h_combt2_ = t2;
    for (auto& lv : glevels) push_gm(gGradPre_, lv, P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
    push_gm(gGradPre_, gstacked, P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
    for (int d = 1; d <= tree_.maxDepth; ++d) {
      push_gm(gGradDepth_[d], gacc[d], P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
      push_gc(gGradDepth_[d], gcon[d], P64, B_PSI64, B_WORK64, B_POOL);
    }
    // G_v = M_v - E rho_v (E set at run time), then
    // grad = (Apply + G_v x_L A) / <A_top|A_top>; at the top (Apply - E A) / N
    std::vector<GmDesc> f1, f2;
    gslot_.clear();
    for (auto& v : tree_.nodes) {
      const int L = v.order() - 1;
      f1.push_back(gm_desc(v, L, v.off, 1, 0, v.off, GM_COPY, -1));
      // G_v = M_v - E rho_v; at the top (no parent leg, rho = 1) G = -E
      gslot_.push_back(poolOff);
      CombOut co{};
      co.out = poolOff;
      co.n = v.dims[L];
      co.t0 = (int32_t)h_combtGE_.size();
      CombTerm a{};
      a.cre = v.top() ? 0.0 : 1.0;
      a.src = v.top() ? -1 : mslot_[v.addr];
      CombTerm b{};
      b.cre = 0.0;  // -E, set at run time
      b.src = v.top() ? -1 : rhoslot_[v.addr];
      h_combtGE_.push_back(a);
      h_combtGE_.push_back(b);
      co.t1 = (int32_t)h_combtGE_.size();
      h_combGE_.push_back(co);
      GmDesc g = gm_desc(v, L, v.off, 0, poolOff, v.off, GM_N, -1);
      g.acc = 1;
      f2.push_back(g);
      poolOff += (int64_t)v.dims[L] * v.dims[L];
    }
    push_gm(gGradFin_, f1, P64, B_PSI64, B_APPLY64, B_POOL, B_DPSI64, true);
    push_gm(gGradFin_, f2, P64, B_PSI64, B_APPLY64, B_POOL, B_DPSI64, true);
    gradFlops_ = gcomb;
  }
  // ----------------------------------------------------- canonicalisation
  // The QR gauge on the device (qutree_amd.gauge.canonicalize_all's order
  // and arithmetic): per depth, deepest first, one batched Householder QR of
  // the non-top nodes of that depth (in place, R kept in the pool), then
  // their parents take R on each child leg, one mode-product launch per leg
  // position, chained through the workspace and ending in the state.
  h_canon_.clear();
  canonRange_.clear();
  gCanon_.clear();
  {
    std::vector<int64_t> rslot(NN, -1);
    for (auto& v : tree_.nodes)
      if (!v.top()) {
        rslot[v.addr] = poolOff;
        poolOff += (int64_t)v.n * v.n;
      }
    for (int d = tree_.maxDepth; d >= 1; --d) {
      const int64_t i0 = (int64_t)h_canon_.size();
      for (auto& v : tree_.nodes)
        if (!v.top() && v.depth == d) {
          SmItem it{};
          it.a = it.q = v.off;
          it.e = rslot[v.addr];
          it.m = v.N / v.n;
          it.n = v.n;
          h_canon_.push_back(it);
        }
      canonRange_.push_back({i0, (int64_t)h_canon_.size()});
      // per leg position: descriptors writing the workspace, and those
      // writing the state (the last leg of each parent)
      std::vector<std::vector<GmDesc>> toWork, toPsi;
      int64_t w = 0;
      for (auto& p : tree_.nodes) {
        if (p.bottom || p.depth != d - 1) continue;
        const int nc = (int)p.children.size();
        const int levels = std::max(nc, 2);
        if ((int)toWork.size() < levels) toWork.resize(levels), toPsi.resize(levels);
        const int64_t wa = w, wb = w + p.N;
        w += 2 * p.N;
        int64_t cur = p.off;
        int cursel = 0;
        for (int c = 0; c < levels; ++c) {
          const bool last = (c == levels - 1);
          const int64_t out = last ? p.off : ((c % 2 == 0) ? wa : wb);
          GmDesc g = (c < nc) ? gm_desc(p, c, cur, cursel, rslot[p.children[c]], out, GM_N, -1)
                              : gm_desc(p, p.order() - 1, cur, cursel, 0, out, GM_COPY, -1);
          (last ? toPsi : toWork)[c].push_back(g);
          cur = out;
          cursel = 1;
        }
      }
      workSize_ = std::max(workSize_, w);
      gCanon_.emplace_back();
      for (size_t c = 0; c < toWork.size(); ++c) {
        push_gm(gCanon_.back(), toWork[c], P64, B_PSI64, B_WORK64, B_POOL, B_WORK64, false);
        push_gm(gCanon_.back(), toPsi[c], P64, B_PSI64, B_WORK64, B_POOL, B_PSI64, false);
      }
    }
  }

  poolSize_ = poolOff;
  pool64Size_ = pool64Off;
  if (std::getenv("QTG_DUMP_PLAN")) dump_plan();
  return QTG_OK;
}

// QTG_DUMP_PLAN=1: the launch lists of the grouped path on stderr (one line
// per launch: list, kind, precision, tasks, FLOPs, algorithmic bytes, and
// the K range and output count of its descriptors).
void Engine::dump_plan() const {
  auto one = [&](const char* name, const std::vector<GLaunch>& L) {
    for (size_t i = 0; i < L.size(); ++i) {
      const GLaunch& g = L[i];
      int kmin = 1 << 30, kmax = 0;
      long outs = 0;
      if (g.kind == 0) {
        for (int64_t t = 0; t < g.ntasks; ++t) {
          const GmDesc& d = h_gm_[g.d0 + t];
          if (d.tile0 != t) continue;
          kmin = std::min(kmin, d.nin);
          kmax = std::max(kmax, d.nin);
          ++outs;
        }
      } else if (g.kind == 1) {
        std::set<std::pair<int, int64_t>> xs, ys;
        for (int64_t t = 0; t < g.ntasks; ++t) {
          const GcDesc& d = h_gc_[g.d0 + t];
          if (d.tile0 != t) continue;
          kmin = std::min(kmin, d.bef * d.aft);
          kmax = std::max(kmax, d.bef * d.aft);
          xs.insert({d.sel & 1, d.x});
          ys.insert({(d.sel >> 1) & 1, d.y});
          ++outs;
        }
        std::fprintf(stderr, "plan %-5s %2zu distinct bra %zu ket %zu\n", name, i, xs.size(), ys.size());
        std::map<std::tuple<int, int, int, int, int>, int> shapes;  // (nx, ny, bef, aft, ksplit) -> descriptors
        for (int64_t t = 0; t < g.ntasks; ++t) {
          const GcDesc& d = h_gc_[g.d0 + t];
          if (d.tile0 == t) ++shapes[std::make_tuple(d.nx, d.ny, d.bef, d.aft, d.ksplit)];
        }
        for (auto& kv : shapes)
          std::fprintf(stderr, "plan %-5s %2zu   nx %d ny %d bef %d aft %d ksplit %d: %d\n", name, i,
                       std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first),
                       std::get<4>(kv.first), kv.second);
      }
      std::fprintf(stderr, "plan %-5s %2zu kind %d prec %d tasks %7lld descs %6ld K %d..%d flops %.3e bytes %.3e\n",
                   name, i, g.kind, g.prec, (long long)g.ntasks, outs, kmin, kmax, g.flops, g.bytes);
    }
  };
  one("build", gBuild_);
  one("hole", gHole_);
  one("apply", gApply_);
  // the two combination launches: outputs, terms, longest term list
  auto comb = [](const char* name, const std::vector<CombOut>& co) {
    int64_t terms = 0, tmax = 0, elems = 0;
    for (const CombOut& c : co) {
      terms += c.t1 - c.t0;
      tmax = std::max<int64_t>(tmax, c.t1 - c.t0);
      elems += (int64_t)c.n * c.n;
    }
    std::fprintf(stderr, "plan %-8s outputs %6zu terms %8lld longest %6lld elements %9lld\n", name, co.size(),
                 (long long)terms, (long long)tmax, (long long)elems);
  };
  comb("combine1", h_comb1_);
  comb("combine3", h_comb3_);
}

int Engine::run_launches(const std::vector<GLaunch>& L, size_t i0, size_t i1, void* const* bases,
                         const uint8_t* active, double sre, double sim, hipStream_t s) {
  int rc;
  for (size_t li = i0; li < std::min(i1, L.size()); ++li) {
    const GLaunch& gl = L[li];
    if (gl.kind == 2 && heevPending_) {
      // rho^-1 from the side stream of the last build
      QTG_HIP_TRY(hipStreamWaitEvent(s, evHeev_, 0));
      heevPending_ = false;
    }
    const double a = gl.scaled ? sre : 1.0, b = gl.scaled ? sim : 0.0;
    const KernelTiming* kt = nullptr;  // set while kernel timing is on
    if (gl.kind == 0) {
      if ((rc = kernel_mark(6, &kt))) return rc;
      const GmDesc* d = d_gm_ + gl.d0;
      const int32_t* wg = gl.wg0 >= 0 ? d_gmwg_ + gl.wg0 : nullptr;
      void *x0 = bases[gl.x0], *x1 = bases[gl.x1], *m = bases[gl.m], *o = bases[gl.o];
      const int wv = gl.waves > 0 ? gl.waves : kGemmWaves;
      if (gl.prec == P64)
        rc = launch_gm<double, double, double>(d, wg, gl.nd, gl.ntasks, x0, x1, m, o, active, a, b, s, kt, wv,
                                               gl.lean);
      else if (gl.prec == P32)
        rc = launch_gm<float, double, float>(d, wg, gl.nd, gl.ntasks, x0, x1, m, o, active, a, b, s, kt, wv,
                                             gl.lean);
      else
        rc = launch_gm<float, double, double>(d, wg, gl.nd, gl.ntasks, x0, x1, m, o, active, a, b, s, kt, wv,
                                              gl.lean);
      if (rc) return rc;
      if (timing_) tfl_[6] += gl.flops, tby_[6] += gl.bytes;
    } else if (gl.kind == 1) {
      if ((rc = kernel_mark(7, &kt))) return rc;
      const GcDesc* d = d_gc_ + gl.d0;
      const int32_t* wg = gl.wg0 >= 0 ? d_gcwg_ + gl.wg0 : nullptr;
      void *x0 = bases[gl.x0], *x1 = bases[gl.x1], *o = bases[gl.o];
      if (gl.prec == P64)
        rc = launch_gc<double, double>(d, wg, gl.nd, gl.ntasks, x0, x1, o, active, s, kt, gl.lean);
      else
        rc = launch_gc<float, double>(d, wg, gl.nd, gl.ntasks, x0, x1, o, active, s, kt, gl.lean);
      if (rc) return rc;
      if (timing_) tfl_[7] += gl.flops, tby_[7] += gl.bytes;
    } else if (gl.kind == 3) {
      if ((rc = kernel_mark(9, &kt))) return rc;
      if ((rc = launch_hole_fused(d_hk_ + gl.d0, gl.nd, d_hi_, bases[gl.x0], bases[gl.x1], bases[gl.m], bases[gl.o],
                                  gl.lds, active, s, kt)))
        return rc;
      if (timing_) tfl_[9] += gl.flops, tby_[9] += gl.bytes;
    } else {
      if ((rc = kernel_mark(8, &kt))) return rc;
      if ((rc = launch_solve(d_solve_, gl.nd, (cplx<double>*)bases[B_POOL64], nspf_, active, s, kt,
                             (const cplx<double>*)bases[B_PSI64], (const cplx<double>*)bases[B_APPLY64],
                             (cplx<double>*)bases[B_DPSI64], a, b, gramInSolve_)))
        return rc;
      if (timing_) tfl_[8] += gl.flops;
    }
  }
  return QTG_OK;
}

int Engine::build_gemm(hipStream_t s) {
  int rc;
  if (c64_ && (rc = convert_psi(s))) return rc;
  void* bases[B_COUNT] = {d_psi32_, d_work32_, d_pool_, d_psi64_, d_apply64_, d_dpsi64_, d_pool64_, d_work_};
  // represent and the density tree, layer by layer in shared launches
  if ((rc = run_launches(gBuild_, bases, nullptr, 1.0, 0.0, s))) return rc;
  if (heevSide_) {
    // side stream: the eigensolver (it only feeds the solve of the finish,
    // which apply_gemm makes wait for it); engine stream: the holes, the
    // combinations and Apply
    QTG_HIP_TRY(hipEventRecord(evFork_, s));
    QTG_HIP_TRY(hipStreamWaitEvent(side_, evFork_, 0));
    if ((rc = begin_mark(1, side_))) return rc;
    if ((rc = launch_heev<double>(d_eig_, nEig_, d_pool_, d_pool64_, eps_, nspf_, side_))) return rc;
    if ((rc = end_mark(side_))) return rc;
    QTG_HIP_TRY(hipEventRecord(evHeev_, side_));
    heevPending_ = true;
  } else {
    // in line on the engine stream
    if ((rc = begin_mark(1, s))) return rc;
    if ((rc = launch_heev<double>(d_eig_, nEig_, d_pool_, d_pool64_, eps_, nspf_, s))) return rc;
    if ((rc = end_mark(s))) return rc;
  }
  if ((rc = run_launches(gHole_, bases, nullptr, 1.0, 0.0, s))) return rc;
  if ((rc = begin_mark(5, s))) return rc;
  // the aggregated slots, then the sums over them with the stacked blocks
  if ((rc = launch_combine<double>(d_combOut_, nComb_, d_combTerm_, d_pool_, s))) return rc;
  if ((rc = launch_combine<double>(d_combOut3_, nComb3_, d_combTerm3_, d_pool_, s))) return rc;
  if ((rc = end_mark(s))) return rc;
  if (timing_) tfl_[5] += phaseFlops_[5];
  built_ = true;
  return QTG_OK;
}

int Engine::apply_gemm(hipStream_t s, const double* psi64, double* dpsi64, const uint8_t* active, double ph_re,
                       double ph_im) {
  void* bases[B_COUNT] = {d_psi32_, d_work32_, d_pool_, (void*)psi64, d_apply64_, dpsi64, d_pool64_, d_work_};
  // f = -i * phase
  return run_launches(gApply_, bases, active, ph_im, -ph_re, s);
}

int Engine::apply_finish(const double* psi64, double* dpsi64, const uint8_t* active, double ph_re, double ph_im,
                         double* energy_dev, hipStream_t s) {
  if (!built_) {
    set_error("engine: build() first");
    return QTG_ESTATE;
  }
  int rc;
  if (path_ == 1) {
    if (c64_) {
      // the FP32 bulk reads a rounded copy of this state
      const int64_t n = 2 * tree_.total;
      if ((rc = launch_to_f32(psi64, d_psi32_, n, s))) return rc;
    }
    if ((rc = apply_gemm(s, psi64, dpsi64, active, ph_re, ph_im))) return rc;
    if (energy_dev)
      QTG_HIP_TRY(hipMemcpyAsync(energy_dev, d_pool64_ + 2 * energySlot_, 4 * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
    return QTG_OK;
  }
  if ((rc = launch_apply<double>(d_chunks_, nChunks_, d_terms_, d_nodes_, psi64, d_pool_, d_partial_, nmax_, s,
                                 active)))
    return rc;
  return launch_finish<double>(d_fin_, nFin_, d_nodes_, psi64, d_partial_, d_pool64_, dpsi64, nullptr, ph_re, ph_im,
                               energy_dev, nmax_, nspf_, s, active);
}

}  // namespace qtg
