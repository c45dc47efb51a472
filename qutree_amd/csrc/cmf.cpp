// Constant-mean-field integrator on the device (CMFIntegrator.cpp:25-199 and
// BS_integrator.h:31-256 of QuTree).
//
// Host: the CMF step-size control in FP64, exactly as the reference decides
// it.  Device: all tensor arithmetic and, by default, every node's
// Bulirsch-Stoer state machine (bs_tick_kernel).  The nodes' BS integrators
// run in lockstep "ticks": each tick every node that is still integrating
// takes exactly one derivative evaluation (LayerDerivative with the matrices
// of the last build), preceded and followed by its own vector ops (midpoint
// updates, rational extrapolation, error norm).  A tick is the masked Apply
// and one tick-boundary launch; the host reads a small per-tick record only
// where the ticks known to have a derivative run out.  The host state
// machine (advance, cmfstep) stays as an option: per-node programs staged in
// pinned memory and uploaded per tick.
#include "cmf.hpp"
#include "jacobi.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>

namespace qtg {

namespace {
constexpr int kMaxUse = kBsMaxUse, kMax = kBsMax;
constexpr double kShrink = kBsShrink, kGrow = kBsGrow;
const int kSeq[kMax] = {bs_seq(0), bs_seq(1), bs_seq(2), bs_seq(3), bs_seq(4), bs_seq(5),
                        bs_seq(6), bs_seq(7), bs_seq(8), bs_seq(9), bs_seq(10)};
constexpr int kOpsPerNode = 16;
constexpr int64_t kMaxTicks = int64_t(1) << 22;  // per interval of the device integrators
}  // namespace

Cmf::~Cmf() {
  drop_tick_graphs();
  for (void* p : dev_)
    if (p) (void)hipFree(p);
  if (host_stage_) (void)hipHostFree(host_stage_);
  if (err_h_) (void)hipHostFree(err_h_);
  if (h_rec_) (void)hipHostFree(h_rec_);
  if (evRec_) (void)hipEventDestroy(evRec_);
  for (hipEvent_t e : evStage_)
    if (e) (void)hipEventDestroy(e);
  if (evErr_) (void)hipEventDestroy(evErr_);
}

int Cmf::dalloc(void** p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
    set_error("cmf: device allocation failed");
    return QTG_ENOMEM;
  }
  dev_.push_back(*p);
  return QTG_OK;
}

int Cmf::init() {
  if (E.c64_ && E.path_ != 1) {
    set_error("cmf: complex64 engines integrate on the grouped-contraction path");
    return QTG_ESTATE;
  }
  if (!E.planned_) {
    set_error("cmf: engine has no Hamiltonian");
    return QTG_ESTATE;
  }
  const HTree& t = E.tree_;
  NN_ = (int)t.nodes.size();
  total_ = t.total;
  s_ = E.stream_;
  const size_t arena = sizeof(cplx<double>) * total_;
  int rc;
  for (cplx<double>** a : {&P_, &P0_, &P2_, &PS_})
    if ((rc = dalloc((void**)a, arena))) return rc;
  for (int k = 0; k < 5; ++k)
    if ((rc = dalloc((void**)&Y_[k], arena))) return rc;
  // A tick's derivative input is the midpoint state yvec[1] and its output
  // lands in yvec[2], where BS_integrator keeps the current derivative, so
  // the copies x <- y1 and y2 <- d vanish (every op that fills x fills y1
  // with the same values, and y2 is not read between a tick's Apply and the
  // post program that would copy d into it).
  X_ = Y_[1];
  D_ = Y_[2];
  for (int k = 0; k < 7; ++k)
    if ((rc = dalloc((void**)&YT_[k], arena))) return rc;
  if ((rc = dalloc((void**)&err_, sizeof(double) * NN_))) return rc;
  // rho slot per node (-1 at the top) for the node error
  std::vector<int64_t> rs(NN_, -1);
  for (auto& v : t.nodes)
    if (!v.top()) rs[v.addr] = E.rhoslot_[v.addr];
  if ((rc = dalloc((void**)&rhoslot_, sizeof(int64_t) * NN_))) return rc;
  QTG_HIP_TRY(hipMemcpy(rhoslot_, rs.data(), sizeof(int64_t) * NN_, hipMemcpyHostToDevice));
  // staging for per-tick programs and the active mask
  // (each half starts on a 256-byte boundary)
  stage_bytes_ = (2 * (sizeof(NodeProg) * NN_ + sizeof(VecOp) * NN_ * kOpsPerNode) + NN_ + 1024 + 255) & ~size_t(255);
  if (hipHostMalloc(&host_stage_, 2 * stage_bytes_) != hipSuccess) {
    set_error("cmf: pinned allocation failed");
    return QTG_ENOMEM;
  }
  if ((rc = dalloc(&dev_stage_, stage_bytes_))) return rc;
  QTG_HIP_TRY(hipHostMalloc((void**)&err_h_, sizeof(double) * NN_));
  for (hipEvent_t& e : evStage_) QTG_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  QTG_HIP_TRY(hipEventCreateWithFlags(&evErr_, hipEventDisableTiming));
  // the device integrators
  if ((rc = dalloc((void**)&d_bs_, sizeof(BsNode) * NN_))) return rc;
  if ((rc = dalloc((void**)&d_rec_, (size_t)kBsRing * 2 * NN_))) return rc;
  QTG_HIP_TRY(hipHostMalloc((void**)&h_rec_, 2 * (size_t)NN_));
  QTG_HIP_TRY(hipEventCreateWithFlags(&evRec_, hipEventDisableTiming));
  // orthogonal(): per height, the nodes to orthonormalise and the parents
  // receiving their square roots (children in node-address order)
  int64_t swoff = 0;
  std::vector<int64_t> swslot(NN_, -1);
  for (auto& v : t.nodes)
    if (!v.top()) {
      swslot[v.addr] = swoff;
      swoff += (int64_t)v.n * v.n;
    }
  if ((rc = dalloc((void**)&sw_, sizeof(cplx<double>) * std::max<int64_t>(swoff, 1)))) return rc;
  std::vector<OrthoItem> oi;
  std::vector<SwApply> sa;
  orthoRange_.clear();
  for (int h = 0; h < t.maxHeight; ++h) {
    std::pair<int64_t, int64_t> r1{(int64_t)oi.size(), 0}, r2{(int64_t)sa.size(), 0};
    for (auto& v : t.nodes)
      if (v.height == h && !v.top()) oi.push_back(OrthoItem{v.addr, 0, swslot[v.addr]});
    r1.second = (int64_t)oi.size();
    for (auto& p : t.nodes) {
      if (p.bottom) continue;
      SwApply it{};
      it.node = p.addr;
      std::vector<int> kids = p.children;
      std::sort(kids.begin(), kids.end());  // address order
      for (int c : kids)
        if (t.nodes[c].height == h) {
          it.fmode[it.nf] = t.nodes[c].childIdx;
          it.fmat[it.nf++] = swslot[c];
        }
      if (it.nf > 0) sa.push_back(it);
    }
    r2.second = (int64_t)sa.size();
    orthoRange_.push_back({r1, r2});
  }
  if ((rc = dalloc((void**)&d_oi_, sizeof(OrthoItem) * std::max<size_t>(oi.size(), 1)))) return rc;
  if ((rc = dalloc((void**)&d_sa_, sizeof(SwApply) * std::max<size_t>(sa.size(), 1)))) return rc;
  if (!oi.empty()) QTG_HIP_TRY(hipMemcpy(d_oi_, oi.data(), sizeof(OrthoItem) * oi.size(), hipMemcpyHostToDevice));
  if (!sa.empty()) QTG_HIP_TRY(hipMemcpy(d_sa_, sa.data(), sizeof(SwApply) * sa.size(), hipMemcpyHostToDevice));
  // Node tensors beyond the LDS budget of ortho_kernel: the global-memory
  // kernel, and the square roots pushed into the parents by grouped mode
  // products.  Per height, a parent takes the matrices of its children of
  // that height leg by leg (children in address order), chained through two
  // workspace regions and ending in the state; a parent with one such child
  // ends with a copy.
  // (QTG_ORTHO_GLOBAL=1 forces this path on any tree, which lets its parity
  // be checked where the LDS kernel also runs)
  const char* force = std::getenv("QTG_ORTHO_GLOBAL");
  bigOrtho_ = (force && force[0] == '1') ||
              (size_t)(E.nmax_ + 2 * jacobi_elems(E.nspf_) + kOrthoNsElems) * sizeof(cplx<double>) > kLdsBudget;
  swLaunch_.clear();
  if (bigOrtho_) {
    std::vector<GmDesc> all;
    int64_t wmax = 0;
    for (int h = 0; h < t.maxHeight; ++h) {
      std::vector<std::vector<GmDesc>> toWork, toState;
      int64_t w = 0;
      for (auto& p : t.nodes) {
        if (p.bottom) continue;
        std::vector<int> kids;
        for (int c : p.children)
          if (t.nodes[c].height == h) kids.push_back(c);
        if (kids.empty()) continue;
        std::sort(kids.begin(), kids.end());
        const int levels = std::max<int>((int)kids.size(), 2);
        if ((int)toWork.size() < levels) toWork.resize(levels), toState.resize(levels);
        const int64_t wa = w, wb = w + p.N;
        w += 2 * p.N;
        int64_t cur = p.off;
        int cursel = 0;
        for (int c = 0; c < levels; ++c) {
          const bool last = (c == levels - 1);
          const bool copy = c >= (int)kids.size();
          const int leg = copy ? p.order() - 1 : t.nodes[kids[c]].childIdx;
          GmDesc g{};
          int64_t bef = 1, aft = 1;
          for (int i = 0; i < leg; ++i) bef *= p.dims[i];
          for (int i = leg + 1; i < p.order(); ++i) aft *= p.dims[i];
          g.x = cur;
          g.xsel = cursel;
          g.m = copy ? 0 : swslot[kids[c]];
          g.out = last ? p.off : ((c % 2 == 0) ? wa : wb);
          g.bef = (int32_t)bef;
          g.nin = g.nout = p.dims[leg];
          g.aft = (int32_t)aft;
          g.op = copy ? GM_COPY : GM_N;
          g.are = 1.0;
          g.node = -1;
          g.ksplit = 1;
          (last ? toState : toWork)[c].push_back(g);
          cur = g.out;
          cursel = 1;
        }
      }
      wmax = std::max(wmax, w);
      swLaunch_.emplace_back();
      for (size_t c = 0; c < toWork.size(); ++c)
        for (int which = 0; which < 2; ++which) {
          auto& L = which ? toState[c] : toWork[c];
          if (L.empty()) continue;
          SwLaunch sl{(int64_t)all.size(), 0, which == 1};
          int64_t tasks = 0;
          for (auto& g : L) {
            g.tile0 = tasks;
            const int64_t nt = gm_tasks(g);
            tasks += nt;
            all.insert(all.end(), (size_t)nt, g);
          }
          sl.ntasks = tasks;
          swLaunch_.back().push_back(sl);
        }
    }
    if ((rc = dalloc((void**)&d_swgm_, sizeof(GmDesc) * std::max<size_t>(all.size(), 1)))) return rc;
    if (!all.empty()) QTG_HIP_TRY(hipMemcpy(d_swgm_, all.data(), sizeof(GmDesc) * all.size(), hipMemcpyHostToDevice));
    swWorkSize_ = std::max<int64_t>(wmax, 1);
    if ((rc = dalloc((void**)&swWork_, sizeof(cplx<double>) * swWorkSize_))) return rc;
  }
  // overlap tree (TreeFunctions::dotProduct, MatrixTreeFunctions_Implementation.h:16-39)
  int64_t soff = 0;
  sslot_.assign(NN_, -1);
  for (auto& v : t.nodes) {
    sslot_[v.addr] = soff;
    soff += (int64_t)v.n * v.n;
  }
  if ((rc = dalloc((void**)&spool_, sizeof(cplx<double>) * soff))) return rc;
  std::vector<SwItem> ov;
  ovRange_.clear();
  ovNmax_.clear();
  for (int h = 0; h <= t.maxHeight; ++h) {
    std::pair<int64_t, int64_t> r{(int64_t)ov.size(), 0};
    int nm = 256;
    for (auto& v : t.nodes) {
      if (v.height != h) continue;
      SwItem it{};
      it.node = v.addr;
      it.k = v.order() - 1;
      it.bra = it.ket = v.off;
      it.out = sslot_[v.addr];
      for (int c : v.children) {
        it.fmode[it.nf] = t.nodes[c].childIdx;
        it.fmat[it.nf++] = sslot_[c];
      }
      ov.push_back(it);
      nm = std::max(nm, v.N);
    }
    r.second = (int64_t)ov.size();
    ovRange_.push_back(r);
    ovNmax_.push_back(nm);
  }
  if ((rc = dalloc((void**)&d_ov_, sizeof(SwItem) * ov.size()))) return rc;
  QTG_HIP_TRY(hipMemcpy(d_ov_, ov.data(), sizeof(SwItem) * ov.size(), hipMemcpyHostToDevice));
  // Overlap tree for node tensors beyond the LDS-resident sandwich kernel:
  // per height, X_v = (S of the children on their legs) ket_v by mode
  // products chained through the workspace (a bottom node copies its ket),
  // then S_v = bra_v^H X_v over all legs but the last.
  ovLaunch_.clear();
  if (bigOrtho_) {
    std::vector<GmDesc> gall;
    std::vector<GcDesc> call;
    int64_t wmax = 0;
    for (int h = 0; h <= t.maxHeight; ++h) {
      std::vector<std::vector<GmDesc>> lev;
      std::vector<GcDesc> con;
      int64_t w = 0;
      for (auto& v : t.nodes) {
        if (v.height != h) continue;
        std::vector<int> kids = v.children;
        std::sort(kids.begin(), kids.end());
        const int levels = std::max<int>((int)kids.size(), 1);
        if ((int)lev.size() < levels) lev.resize(levels);
        const int64_t wa = w, wb = w + v.N;
        w += 2 * v.N;
        int64_t cur = v.off;
        int cursel = 0;
        for (int c = 0; c < levels; ++c) {
          const bool copy = kids.empty();
          const int leg = copy ? v.order() - 1 : t.nodes[kids[c]].childIdx;
          GmDesc g{};
          int64_t bef = 1, aft = 1;
          for (int i = 0; i < leg; ++i) bef *= v.dims[i];
          for (int i = leg + 1; i < v.order(); ++i) aft *= v.dims[i];
          g.x = cur;
          g.xsel = cursel;
          g.m = copy ? 0 : sslot_[kids[c]];
          g.out = (c % 2 == 0) ? wa : wb;
          g.bef = (int32_t)bef;
          g.nin = g.nout = v.dims[leg];
          g.aft = (int32_t)aft;
          g.op = copy ? GM_COPY : GM_N;
          g.are = 1.0;
          g.node = -1;
          g.ksplit = 1;
          lev[c].push_back(g);
          cur = g.out;
          cursel = 1;
        }
        GcDesc c{};
        c.x = v.off;   // bra (first base)
        c.y = cur;     // X (workspace, second base)
        c.out = sslot_[v.addr];
        c.bef = (int32_t)(v.N / v.n);
        c.aft = 1;
        c.nx = c.ny = v.n;
        c.node = -1;
        c.sel = 2;
        const int64_t K = v.N / v.n;
        c.ksplit = K > 1024 ? split_chunks(K, kGemmWaves, 1) : 1;
        con.push_back(c);
      }
      wmax = std::max(wmax, w);
      ovLaunch_.emplace_back();
      for (auto& L : lev) {
        if (L.empty()) continue;
        OvLaunch ol{0, (int64_t)gall.size(), 0};
        for (auto& g : L) {
          g.tile0 = ol.ntasks;
          const int64_t nt = gm_tasks(g);
          ol.ntasks += nt;
          gall.insert(gall.end(), (size_t)nt, g);
        }
        ovLaunch_.back().push_back(ol);
      }
      if (!con.empty()) {
        OvLaunch ol{1, (int64_t)call.size(), 0};
        ol.ntasks = order_tasks(con, gc_tasks);
        for (auto& c : con) call.insert(call.end(), (size_t)gc_tasks(c), c);
        ovLaunch_.back().push_back(ol);
      }
    }
    if ((rc = dalloc((void**)&d_ovgm_, sizeof(GmDesc) * std::max<size_t>(gall.size(), 1)))) return rc;
    if ((rc = dalloc((void**)&d_ovgc_, sizeof(GcDesc) * std::max<size_t>(call.size(), 1)))) return rc;
    if (!gall.empty()) QTG_HIP_TRY(hipMemcpy(d_ovgm_, gall.data(), sizeof(GmDesc) * gall.size(), hipMemcpyHostToDevice));
    if (!call.empty()) QTG_HIP_TRY(hipMemcpy(d_ovgc_, call.data(), sizeof(GcDesc) * call.size(), hipMemcpyHostToDevice));
    if (wmax > swWorkSize_) {
      if ((rc = dalloc((void**)&ovWork_, sizeof(cplx<double>) * wmax))) return rc;
    }
  }
  if ((rc = dalloc((void**)&energy_, sizeof(double) * 4))) return rc;
  return QTG_OK;
}

CmfBufs Cmf::bufs(cplx<double>* P) const {
  CmfBufs b{};
  b.b[B_P] = P;
  for (int k = 0; k < 5; ++k) b.b[B_Y0 + k] = Y_[k];
  for (int k = 0; k < 7; ++k) b.b[B_YT0 + k] = YT_[k];
  b.b[B_X] = X_;
  b.b[B_D] = D_;
  b.b[B_E1] = P_;
  b.b[B_E2] = P2_;
  return b;
}

int Cmf::copy(cplx<double>* dst, const cplx<double>* src) {
  QTG_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(cplx<double>) * total_, hipMemcpyDeviceToDevice, s_));
  return QTG_OK;
}

int Cmf::build(const cplx<double>* P) {
  int rc = copy((cplx<double>*)E.d_psi64_, P);
  if (rc) return rc;
  ++nbuild_;
  return E.build(s_);
}

// ---------------------------------------------------------------- BS ticks
static VecOp op0(int code, int a = 0, int b = 0, int c = 0, double h = 0.0) {
  VecOp o{};
  o.code = code;
  o.a = a;
  o.b = b;
  o.c = c;
  o.h = h;
  return o;
}

// Emits the node's ops up to and around its next derivative evaluation.
// Returns false when the node has finished its interval.
// verdict: the outcome of a pending accept test (1 accept, 0 reject) in
// place of the error read back from the device (-1).
bool Cmf::advance(int v, std::vector<VecOp>& pre, std::vector<VecOp>& post, double eps, int verdict) {
  NodeBS& b = bs_[v];
  double& dt = dtbs_[v];
  while (true) {
    switch (b.st) {
      case NodeBS::AWAIT_ERR: {
        if (verdict < 0 ? err_h_[v] < eps : verdict == 1) {
          b.x += b.h;
          if (b.i == kMaxUse - 1) dt = b.h * kShrink;
          else if (b.i == kMaxUse - 2) dt = b.h * kGrow;
          else dt = (b.h * kSeq[kMaxUse - 2]) / (1.0 * kSeq[b.i]);
          pre.push_back(op0(V_COPY, B_P, B_Y0));
          b.st = NodeBS::NEED_START;
        } else {
          ++b.i;
          if (b.i >= kMax) {
            b.h = b.h / 4.0;
            for (int k = 0; k < (kMax - kMaxUse) / 2.0; ++k) b.h /= 2.0;
            if (dt + b.h == dt) {
              failed_ = true;
              return false;
            }
            b.i = 0;
          }
          b.st = NodeBS::START_MMID;
        }
        continue;
      }
      case NodeBS::NEED_START: {
        // BS_integrator::Integrate loop condition
        if (!(b.xend - b.x > dt * 1e-6)) {
          b.st = NodeBS::DONE;
          return false;
        }
        if (b.x + dt > b.xend) dt = b.xend - b.x + 1e-10;
        // bsstep: clearmemory, y0 = y3 = y, y4 = f(y)
        std::fill(b.xtab, b.xtab + kMax, 0.0);
        pre.push_back(op0(V_COPY, B_Y0, B_P));
        pre.push_back(op0(V_COPY, B_Y0 + 3, B_P));
        for (int k = 1; k < 7; ++k) pre.push_back(op0(V_ZERO, B_YT0 + k));
        pre.push_back(op0(V_COPY, B_X, B_P));
        post.push_back(op0(V_COPY, B_Y0 + 4, B_D));
        b.h = dt;
        b.i = 0;
        b.st = NodeBS::START_MMID;
        return true;
      }
      case NodeBS::START_MMID: {
        b.step = kSeq[b.i];
        b.hh = b.h / (1.0 * b.step);
        pre.push_back(op0(V_COPY, B_Y0, B_Y0 + 3));
        pre.push_back(op0(V_AXPY, B_Y0 + 1, B_Y0 + 3, B_Y0 + 4, b.hh));
        pre.push_back(op0(V_COPY, B_X, B_Y0 + 1));
        post.push_back(op0(V_COPY, B_Y0 + 2, B_D));
        b.nsub = 1;
        b.st = NodeBS::IN_MMID;
        if (b.nsub == b.step) finish_mmid(v, post);
        return true;
      }
      case NodeBS::IN_MMID: {
        pre.push_back(op0(V_MMID_STEP, 0, 0, 0, b.hh));
        pre.push_back(op0(V_COPY, B_X, B_Y0 + 1));
        post.push_back(op0(V_COPY, B_Y0 + 2, B_D));
        ++b.nsub;
        if (b.nsub == b.step) finish_mmid(v, post);
        return true;
      }
      default:
        return false;
    }
  }
}

// The op sequences advance() emits most, each as one fused op (one pass
// over the node; the same arithmetic per element).  Other sequences stay.
static bool is_op(const VecOp& o, int code, int a = -1, int b = -1, int c = -1) {
  return o.code == code && (a < 0 || o.a == a) && (b < 0 || o.b == b) && (c < 0 || o.c == c);
}
// the bsstep start: y0 = y3 = P, ytab[1..6] = 0, x = P
static bool is_start(const std::vector<VecOp>& o, size_t i) {
  if (o.size() < i + 9) return false;
  if (!is_op(o[i], V_COPY, B_Y0, B_P) || !is_op(o[i + 1], V_COPY, B_Y0 + 3, B_P)) return false;
  for (int k = 1; k < 7; ++k)
    if (!is_op(o[i + 1 + k], V_ZERO, B_YT0 + k)) return false;
  return is_op(o[i + 8], V_COPY, B_X, B_P);
}
static void fuse_ops(std::vector<VecOp>& o) {
  std::vector<VecOp> r;
  r.reserve(o.size());
  size_t i = 0;
  while (i < o.size()) {
    if (is_op(o[i], V_COPY, B_P, B_Y0) && is_start(o, i + 1)) {  // accept, then the next start
      r.push_back(op0(V_START, B_Y0, 1));
      i += 10;
    } else if (is_start(o, i)) {
      r.push_back(op0(V_START, B_P, 0));
      i += 9;
    } else if (i + 3 <= o.size() && is_op(o[i], V_COPY, B_Y0, B_Y0 + 3) &&
               is_op(o[i + 1], V_AXPY, B_Y0 + 1, B_Y0 + 3, B_Y0 + 4) && is_op(o[i + 2], V_COPY, B_X, B_Y0 + 1)) {
      r.push_back(op0(V_MMID_BEGIN, 0, 0, 0, o[i + 1].h));
      i += 3;
    } else if (i + 2 <= o.size() && is_op(o[i], V_MMID_STEP) && is_op(o[i + 1], V_COPY, B_X, B_Y0 + 1)) {
      r.push_back(op0(V_MMID_STEP_X, 0, 0, 0, o[i].h));
      i += 2;
    } else if (i + 4 <= o.size() && is_op(o[i], V_COPY, B_Y0 + 2, B_D) && is_op(o[i + 1], V_MMID_FINAL) &&
               is_op(o[i + 2], V_RZEXTR) && is_op(o[i + 3], V_POST_EXTR)) {
      VecOp f = o[i + 2];  // i, mi and fx of the extrapolation
      f.code = V_FINAL_EXTR;
      f.h = o[i + 1].h;
      r.push_back(f);
      i += 4;
    } else {
      r.push_back(o[i]);
      ++i;
    }
  }
  o.swap(r);
}

void Cmf::finish_mmid(int v, std::vector<VecOp>& post) {
  NodeBS& b = bs_[v];
  post.push_back(op0(V_MMID_FINAL, 0, 0, 0, b.hh));
  // pow(h / n, 2) of the reference (BS_integrator.h:121) as a product: the
  // device integrators compute the same correctly rounded square
  const double hs = b.h / (1.0 * b.step);
  const double xest = hs * hs;
  b.xtab[b.i] = xest;
  VecOp r = op0(V_RZEXTR);
  r.i = b.i;
  r.mi = (b.i < kMaxUse) ? b.i + 1 : kMaxUse;
  for (int k = 1; k < r.mi; ++k) r.fx[k] = b.xtab[b.i - k] / xest;
  post.push_back(r);
  post.push_back(op0(V_POST_EXTR));
  post.push_back(op0(V_ERROR, B_Y0, B_Y0 + 1));
  b.st = NodeBS::AWAIT_ERR;
}

// Fixed layout of the device staging of a tick: the active mask, every
// node's pre and post program (empty ones have op0 == op1) and the ops they
// index.  The launches of a tick then read fixed addresses, so the tick can
// be replayed from a HIP graph; only the upload's length varies.
Cmf::TickLayout Cmf::tick_layout() const {
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  TickLayout L;
  L.pre = align((size_t)NN_);
  L.post = L.pre + sizeof(NodeProg) * NN_;
  L.ops = align(L.post + sizeof(NodeProg) * NN_);
  return L;
}

// The launches of one tick on the staged programs: pre programs, the masked
// Apply and finish, post programs.
int Cmf::tick_launches(const CmfBufs& bb, bool post) {
  const TickLayout L = tick_layout();
  unsigned char* ds = (unsigned char*)dev_stage_;
  const VecOp* ops = (const VecOp*)(ds + L.ops);
  const cplx<double>* pool = (const cplx<double>*)E.d_pool_;
  int rc = launch_programs((const NodeProg*)(ds + L.pre), NN_, ops, bb, E.d_nodes_, pool, rhoslot_, err_, s_, ds);
  if (rc) return rc;
  if ((rc = E.apply_finish((const double*)X_, (double*)D_, ds, ph_re_, ph_im_, nullptr, s_))) return rc;
  if (!post) return QTG_OK;
  return launch_programs((const NodeProg*)(ds + L.post), NN_, ops, bb, E.d_nodes_, pool, rhoslot_, err_, s_);
}

bool Cmf::tick_eager(const cplx<double>* P) const {
  return !tickGraph_ || E.timing_ || (P != P_ && P != P2_);
}

void Cmf::drop_tick_graphs() {
  for (hipGraphExec_t& g : tickExec_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// One tick on the engine stream: replayed from the graph of its state
// buffer (captured on first use, again after a phase change), or eagerly
// while graphs are off or the engine times its kernels.
int Cmf::run_tick(cplx<double>* P, bool post) {
  // rho^-1 of the last build comes from the engine's side stream: wait for
  // it here, outside any capture
  if (E.heevPending_) {
    QTG_HIP_TRY(hipStreamWaitEvent(s_, E.evHeev_, 0));
    E.heevPending_ = false;
  }
  const CmfBufs bb = bufs(P);
  if (tick_eager(P)) return tick_launches(bb, post);
  if (tickPh_[0] != ph_re_ || tickPh_[1] != ph_im_) {
    drop_tick_graphs();
    tickPh_[0] = ph_re_;
    tickPh_[1] = ph_im_;
  }
  const int w = (P == P2_) ? 1 : 0;
  if (!tickExec_[w]) {
    QTG_HIP_TRY(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    int rc = tick_launches(bb, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s_, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("cmf tick capture: ") + hipGetErrorString(e));
      return QTG_EDEVICE;
    }
    hipError_t ie = hipGraphInstantiate(&tickExec_[w], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      tickExec_[w] = nullptr;
      set_error(std::string("cmf tick graph: ") + hipGetErrorString(ie));
      return QTG_EDEVICE;
    }
  }
  QTG_HIP_TRY(hipGraphLaunch(tickExec_[w], s_));
  return QTG_OK;
}

// The same interval with every node's integrator on the device: a tick is
// the masked Apply and one bs_tick launch (the tick's post ops, the
// decisions, the next tick's pre ops and mask).  The host reads the
// per-tick record only where the ticks known to have a derivative run out,
// after it has launched the next tick (a tick without one does nothing).
int Cmf::cmfstep_dev(cplx<double>* P, double t0, double t1, double eps) {
  if (E.heevPending_) {
    QTG_HIP_TRY(hipStreamWaitEvent(s_, E.evHeev_, 0));
    E.heevPending_ = false;
  }
  const CmfBufs bb = bufs(P);
  const cplx<double>* pool = (const cplx<double>*)E.d_pool_;
  const size_t recB = 2 * (size_t)NN_;
  int slot = 0;  // the record of the next tick
  auto boundary = [&](int init) {
    slot = slot_;
    slot_ = (slot_ + 1) % kBsRing;
    return launch_bs_tick(d_bs_, NN_, bb, E.d_nodes_, pool, rhoslot_, err_, d_rec_ + slot * recB, init, t0, t1, eps,
                          s_);
  };
  int rc = boundary(1);
  if (rc) return rc;
  int64_t known = -1;  // ticks up to this one surely have a derivative
  for (int64_t t = 0;; ++t) {
    if (t > kMaxTicks) {
      set_error("cmf: an interval took more than 2^22 ticks");
      return QTG_ESTATE;
    }
    const bool check = t > known;
    const uint8_t* rec = d_rec_ + slot * recB;
    if (check) {
      // the tick's record (a copy on a side stream measured slower)
      QTG_HIP_TRY(hipMemcpyAsync(h_rec_, rec, recB, hipMemcpyDeviceToHost, s_));
      QTG_HIP_TRY(hipEventRecord(evRec_, s_));
    }
    if ((rc = E.apply_finish((const double*)X_, (double*)D_, rec, ph_re_, ph_im_, nullptr, s_))) return rc;
    if ((rc = boundary(0))) return rc;
    if (!check) continue;
    QTG_HIP_TRY(hipEventSynchronize(evRec_));
    int live = 0, lead = 0;
    for (int v = 0; v < NN_; ++v) {
      const uint8_t l = h_rec_[NN_ + v];
      if (l == kBsFailed) {
        set_error("BS integrator: step size underflow");
        return QTG_ESTATE;
      }
      live += h_rec_[v];
      lead = std::max(lead, (int)l);
    }
    if (live == 0) {
      // every node's interval ended with the tick before
      nticks_ += t;
      return QTG_OK;
    }
    known = t + lead - 1;
  }
}

int Cmf::cmfstep(cplx<double>* P, double t0, double t1, double eps) {
  if (deviceBs_) return cmfstep_dev(P, t0, t1, eps);
  for (int v = 0; v < NN_; ++v) {
    bs_[v].st = NodeBS::NEED_START;
    bs_[v].x = t0;
    bs_[v].xend = t1;
  }
  const CmfBufs bb = bufs(P);
  const TickLayout L = tick_layout();
  unsigned char* ds = (unsigned char*)dev_stage_;
  int half = 0;
  std::vector<NodeProg> ppre(NN_), ppost(NN_);
  std::vector<VecOp> ops, pre, post, bpre[2], bpost[2];
  std::vector<uint8_t> act(NN_);
  // Nodes whose accept test reads an error still on its way to the host.
  // The next tick is staged with both outcomes of each such node's test
  // (the program kernel picks one by the same comparison and writes the
  // node's Apply mask byte); the host commits the picked branch once the
  // error has arrived, while that tick runs.
  std::vector<uint8_t> pending(NN_, 0);
  struct Branch {
    NodeBS b;
    double dt;
    bool active, fail;
  };
  std::vector<Branch> brA(NN_), brR(NN_);
  std::vector<NodeBS> saveBs;
  std::vector<double> saveDt;
  bool errInFlight = false;
  // copies between arenas that are the same buffer (x = y1, d = y2)
  auto self_copy = [&](const VecOp& o) { return o.code == V_COPY && bb.b[o.a] == bb.b[o.b]; };
  auto tidy = [&](std::vector<VecOp>& o) {
    if (fuseOps_) fuse_ops(o);
    o.erase(std::remove_if(o.begin(), o.end(), self_copy), o.end());
  };
  auto has_error_op = [](const std::vector<VecOp>& o) {
    return std::any_of(o.begin(), o.end(), [](const VecOp& x) { return x.code == V_ERROR; });
  };
  auto prog = [&](int v, const std::vector<VecOp>& a, const std::vector<VecOp>& r, int32_t flags) {
    NodeProg g{v, (int)ops.size(), (int)(ops.size() + a.size()), flags, 0, 0, eps};
    ops.insert(ops.end(), a.begin(), a.end());
    g.alt0 = (int)ops.size();
    g.alt1 = (int)(ops.size() + r.size());
    ops.insert(ops.end(), r.begin(), r.end());
    return g;
  };
  auto wait_errors = [&]() -> int {
    QTG_HIP_TRY(hipEventSynchronize(evErr_));
    errInFlight = false;
    return QTG_OK;
  };
  while (true) {
    bool look = lookahead_ && errInFlight;
    if (look) {
      saveBs = bs_;
      saveDt = dtbs_;
    }
    int nact, npre, npost;
    bool need_err, branchAct;
    while (true) {
      ops.clear();
      nact = npre = npost = 0;
      need_err = branchAct = false;
      for (int v = 0; v < NN_; ++v) {
        act[v] = 0;
        ppre[v] = NodeProg{v, 0, 0, 0, 0, 0, eps};
        ppost[v] = NodeProg{v, 0, 0, 0, 0, 0, eps};
        if (bs_[v].st == NodeBS::DONE) continue;
        if (look && pending[v]) {
          const NodeBS b0 = bs_[v];
          const double dt0 = dtbs_[v];
          Branch br[2];
          for (int k = 0; k < 2; ++k) {  // 0 accept, 1 reject
            bpre[k].clear();
            bpost[k].clear();
            const bool d = advance(v, bpre[k], bpost[k], eps, k == 0 ? 1 : 0);
            br[k] = Branch{bs_[v], dtbs_[v], d, failed_};
            failed_ = false;
            bs_[v] = b0;
            dtbs_[v] = dt0;
            tidy(bpre[k]);
            tidy(bpost[k]);
            if (has_error_op(bpost[k])) look = false;  // a test right after a test: not staged ahead
          }
          if (!look) break;
          brA[v] = br[0];
          brR[v] = br[1];
          branchAct = branchAct || br[0].active;
          const int32_t fl = kSelect | (br[0].active ? kActAccept : 0) | (br[1].active ? kActReject : 0);
          ppre[v] = prog(v, bpre[0], bpre[1], fl | kMask);
          if (!bpost[0].empty() || !bpost[1].empty()) {
            ppost[v] = prog(v, bpost[0], bpost[1], fl);
            ++npost;
          }
          ++npre;
          continue;
        }
        pre.clear();
        post.clear();
        const bool d = advance(v, pre, post, eps, -1);
        if (failed_) {
          set_error("BS integrator: step size underflow");
          return QTG_ESTATE;
        }
        tidy(pre);
        tidy(post);
        if (!pre.empty()) {
          ppre[v] = NodeProg{v, (int)ops.size(), (int)(ops.size() + pre.size()), 0, 0, 0, eps};
          ops.insert(ops.end(), pre.begin(), pre.end());
          ++npre;
        }
        if (!d) continue;
        act[v] = 1;
        ++nact;
        ++nderiv_;
        npost += post.empty() ? 0 : 1;
        ppost[v] = NodeProg{v, (int)ops.size(), (int)(ops.size() + post.size()), 0, 0, 0, eps};
        if (has_error_op(post)) need_err = true;
        ops.insert(ops.end(), post.begin(), post.end());
      }
      // Staged ahead only when some node derivative is sure to run or
      // follows an accept; otherwise (every node's step ends in this tick)
      // the errors are read first and the tick is staged as it will run.
      if (look && (nact > 0 || branchAct)) break;
      if (!look && !errInFlight) break;
      bs_ = saveBs;
      dtbs_ = saveDt;
      nderiv_ -= nact;
      look = false;
      if (int rc = wait_errors()) return rc;
      std::fill(pending.begin(), pending.end(), 0);
    }
    if (!look && nact == 0 && npre == 0) break;
    const size_t oend = L.ops + sizeof(VecOp) * ops.size();
    if (oend > stage_bytes_) {
      set_error("cmf: staging overflow");
      return QTG_ESTATE;
    }
    // The host runs ahead of the device while no node waits for an error:
    // the two halves of the pinned staging take turns, and a half is
    // rewritten only after the device has read its last upload.  The device
    // region is reused in stream order.
    unsigned char* hs = (unsigned char*)host_stage_ + half * stage_bytes_;
    QTG_HIP_TRY(hipEventSynchronize(evStage_[half]));
    std::memcpy(hs, act.data(), NN_);
    std::memcpy(hs + L.pre, ppre.data(), sizeof(NodeProg) * NN_);
    std::memcpy(hs + L.post, ppost.data(), sizeof(NodeProg) * NN_);
    std::memcpy(hs + L.ops, ops.data(), sizeof(VecOp) * ops.size());
    QTG_HIP_TRY(hipMemcpyAsync(ds, hs, oend, hipMemcpyHostToDevice, s_));
    QTG_HIP_TRY(hipEventRecord(evStage_[half], s_));
    half ^= 1;
    int rc;
    if (!look && nact == 0) {
      // only trailing accept copies were left
      rc = launch_programs((const NodeProg*)(ds + L.pre), NN_, (const VecOp*)(ds + L.ops), bb, E.d_nodes_,
                           (const cplx<double>*)E.d_pool_, rhoslot_, err_, s_);
      if (rc) return rc;
      break;
    }
    ++nticks_;
    if ((rc = run_tick(P, npost > 0))) return rc;
    if (look) {
      // the errors of the previous tick: commit the branch each node took
      if ((rc = wait_errors())) return rc;
      for (int v = 0; v < NN_; ++v) {
        if (!pending[v]) continue;
        pending[v] = 0;
        const Branch& br = (err_h_[v] < eps) ? brA[v] : brR[v];
        if (br.fail) {
          set_error("BS integrator: step size underflow");
          return QTG_ESTATE;
        }
        bs_[v] = br.b;
        dtbs_[v] = br.dt;
        if (br.active) ++nderiv_;
      }
    }
    if (need_err) {
      // the next tick's decisions read the errors
      QTG_HIP_TRY(hipMemcpyAsync(err_h_, err_, sizeof(double) * NN_, hipMemcpyDeviceToHost, s_));
      if (lookahead_) {
        QTG_HIP_TRY(hipEventRecord(evErr_, s_));
        errInFlight = true;
        for (int v = 0; v < NN_; ++v)
          if (bs_[v].st == NodeBS::AWAIT_ERR) pending[v] = 1;
      } else {
        QTG_HIP_TRY(hipStreamSynchronize(s_));
      }
    }
  }
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  return QTG_OK;
}

// CMFIntegrator::Error: sqrt(sum_v LayerInterface::Error(Psi_v, Chi_v)^2)
int Cmf::error(double* out) {
  unsigned char* hs = (unsigned char*)host_stage_;
  unsigned char* ds = (unsigned char*)dev_stage_;
  QTG_HIP_TRY(hipEventSynchronize(evStage_[0]));
  const size_t progB = sizeof(NodeProg) * NN_;
  NodeProg* hp = (NodeProg*)hs;
  VecOp* ho = (VecOp*)(hs + progB);
  for (int v = 0; v < NN_; ++v) {
    hp[v] = NodeProg{v, v, v + 1, 0};
    ho[v] = op0(V_ERROR, B_E1, B_E2);
  }
  QTG_HIP_TRY(hipMemcpyAsync(ds, hs, progB + sizeof(VecOp) * NN_, hipMemcpyHostToDevice, s_));
  int rc = launch_programs((NodeProg*)ds, NN_, (VecOp*)(ds + progB), bufs(P_), E.d_nodes_,
                           (const cplx<double>*)E.d_pool_, rhoslot_, err_, s_);
  if (rc) return rc;
  QTG_HIP_TRY(hipMemcpyAsync(err_h_, err_, sizeof(double) * NN_, hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  double e = 0.0;
  for (int v = 0; v < NN_; ++v) e += err_h_[v] * err_h_[v];
  *out = std::sqrt(e);
  return QTG_OK;
}

int Cmf::orthogonal(cplx<double>* P) {
  int rc;
  if (bigOrtho_) {
    for (size_t h = 0; h < orthoRange_.size(); ++h) {
      const auto& r = orthoRange_[h];
      if ((rc = launch_ortho_global(d_oi_ + r.first.first, r.first.second - r.first.first, E.d_nodes_, P, sw_,
                                    E.nspf_, s_)))
        return rc;
      for (const SwLaunch& sl : swLaunch_[h])
        if ((rc = launch_gm<double, double, double>(d_swgm_ + sl.d0, nullptr, sl.ntasks, sl.ntasks, P, swWork_, sw_,
                                                    sl.toState ? (void*)P : (void*)swWork_, nullptr, 1.0, 0.0, s_)))
          return rc;
    }
    return QTG_OK;
  }
  for (auto& r : orthoRange_) {
    if ((rc = launch_ortho(d_oi_ + r.first.first, r.first.second - r.first.first, E.d_nodes_, P, sw_, E.nmax_,
                           E.nspf_, s_)))
      return rc;
    if ((rc = launch_apply_sw(d_sa_ + r.second.first, r.second.second - r.second.first, E.d_nodes_, P, sw_,
                              E.nmax_, s_)))
      return rc;
  }
  return QTG_OK;
}

int Cmf::overlap_top(const cplx<double>* bra, const cplx<double>* ket, cplx<double>* out) {
  if (bigOrtho_) {
    cplx<double>* work = ovWork_ ? ovWork_ : swWork_;
    for (auto& hl : ovLaunch_)
      for (const OvLaunch& ol : hl) {
        int rc = ol.kind == 0
                     ? launch_gm<double, double, double>(d_ovgm_ + ol.d0, nullptr, ol.ntasks, ol.ntasks, ket, work,
                                                         spool_, work, nullptr, 1.0, 0.0, s_)
                     : launch_gc<double, double>(d_ovgc_ + ol.d0, nullptr, ol.ntasks, ol.ntasks, bra, work, spool_,
                                                 nullptr, s_);
        if (rc) return rc;
      }
    QTG_HIP_TRY(hipMemcpyAsync(out, spool_ + sslot_[NN_ - 1], sizeof(cplx<double>), hipMemcpyDeviceToHost, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
    return QTG_OK;
  }
  for (size_t h = 0; h < ovRange_.size(); ++h) {
    auto& r = ovRange_[h];
    int rc = launch_sandwich<double>(d_ov_ + r.first, r.second - r.first, E.d_nodes_, bra, ket, spool_, ovNmax_[h],
                                     s_);
    if (rc) return rc;
  }
  QTG_HIP_TRY(hipMemcpyAsync(out, spool_ + sslot_[NN_ - 1], sizeof(cplx<double>), hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  return QTG_OK;
}

// CMFIntegrator::Output: Expectation with the current matrices / S_top, and
// the autocorrelation <Psi(0)|Psi(t)>.
// TensorTree<cd>::write(filename, append) (TensorTree_Implementation.h:
// 93-100): append opens for appending (created when missing); otherwise the
// file is opened in place, in | out, so a missing file is not created and
// an existing one is overwritten from its start without being truncated.
int Cmf::write_state(const cplx<double>* P) {
  if (outFile_.empty()) return QTG_OK;
  std::vector<double> host(2 * (size_t)total_);
  QTG_HIP_TRY(hipMemcpyAsync(host.data(), P, sizeof(double) * host.size(), hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  const int64_t n = qtg_tensor_tree_write(E.tree_text().c_str(), host.data(), nullptr, 0);
  if (n < 0) return (int)-n;
  std::vector<unsigned char> buf((size_t)n);
  if (qtg_tensor_tree_write(E.tree_text().c_str(), host.data(), buf.data(), n) != n) return QTG_EARG;
  std::FILE* f = std::fopen(outFile_.c_str(), outAppend_ ? "ab" : "r+b");
  if (!f) return QTG_OK;  // the reference's stream fails silently as well
  const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
  std::fclose(f);
  if (w != buf.size()) {
    set_error("cmf: writing " + outFile_ + " failed");
    return QTG_EARG;
  }
  return QTG_OK;
}

int Cmf::output(double time, CmfResult& res) {
  cplx<double> stop, ac;
  int rc = overlap_top(P_, P_, &stop);
  if (rc) return rc;
  if ((rc = overlap_top(PS_, P_, &ac))) return rc;
  // Apply at the top only
  std::vector<uint8_t> act(NN_, 0);
  act[NN_ - 1] = 1;
  unsigned char* ds = (unsigned char*)dev_stage_;
  const size_t off = 2 * (sizeof(NodeProg) * NN_ + sizeof(VecOp) * NN_ * kOpsPerNode);
  QTG_HIP_TRY(hipMemcpyAsync(ds + off, act.data(), NN_, hipMemcpyHostToDevice, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  if ((rc = E.apply_finish((const double*)P_, (double*)D_, ds + off, 1.0, 0.0, energy_, s_))) return rc;
  double e[4];
  QTG_HIP_TRY(hipMemcpyAsync(e, energy_, sizeof(e), hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  cd num(e[0], e[1]);
  const cd den(stop.re, stop.im);
  if (E.tree_.nodes[NN_ - 1].n > 1) {
    // several states: h(0,0) / S(0,0) of the first one, h(0,0) =
    // <Psi_0|H Psi_0> = i <Psi_0|D_0> with the top node's D = -i H Psi
    if ((rc = eig_init())) return rc;
    if ((rc = launch_sandwich<double>(d_hx_, 1, E.d_nodes_, P_, D_, eigm_, ovNmax_.back(), s_))) return rc;
    cplx<double> x;
    QTG_HIP_TRY(hipMemcpyAsync(&x, eigm_, sizeof(x), hipMemcpyDeviceToHost, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
    num = cd(-x.im, x.re);
  }
  res.times.push_back(time);
  res.energies.push_back((num / den).real());
  res.autocorr.push_back(cd(ac.re, ac.im));
  return QTG_OK;
}

int Cmf::integrate(double time, double timeend, double dt, double out, double acc_cmf, double acc_bs,
                   int64_t max_steps, CmfResult& res) {
  const double tconst = 0.25, max_increase = 2.25;
  bs_.assign(NN_, NodeBS());
  dtbs_.assign(NN_, dt / 2.0);
  failed_ = false;
  nbuild_ = nderiv_ = nticks_ = 0;
  if (deviceBs_) {
    // every node's integrator idle with step size dt / 2
    std::vector<BsNode> bn(NN_);
    for (BsNode& b : bn) {
      b = BsNode{};
      b.st = BS_DONE;
      b.dt = dt / 2.0;
    }
    QTG_HIP_TRY(hipMemcpyAsync(d_bs_, bn.data(), sizeof(BsNode) * NN_, hipMemcpyHostToDevice, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
  }
  double t_next = time + out;
  bool calc_mat = true;
  int rc;
  if ((rc = write_state(P_))) return rc;
  if ((rc = copy(PS_, P_))) return rc;
  if ((rc = build(P_))) return rc;
  calc_mat = false;
  if ((rc = output(time, res))) return rc;
  double err = 0.0;
  while (time + 1e-6 < timeend && res.steps < (size_t)max_steps) {
    const double dtmax = std::min(timeend - time, t_next - time);
    if (dt > dtmax) dt = dtmax;
    else if (2 * dt > dtmax) dt = std::min(dt, dtmax * 0.6);
    if ((rc = copy(P0_, P_))) return rc;
    if (calc_mat) {
      if ((rc = build(P_))) return rc;
      calc_mat = false;
    }
    const double ts = dt * tconst;
    if ((rc = cmfstep(P_, time, time + ts, acc_bs))) return rc;
    if ((rc = copy(P2_, P_))) return rc;
    if ((rc = cmfstep(P2_, time + ts, time + 2 * ts, acc_bs))) return rc;
    if ((rc = build(P2_))) return rc;
    if ((rc = cmfstep(P_, time + dt * tconst, time + dt * (1 - tconst), acc_bs))) return rc;
    if ((rc = copy(P2_, P0_))) return rc;
    if ((rc = cmfstep(P2_, time, time + dt, acc_bs))) return rc;
    if ((rc = build(P2_))) return rc;
    if ((rc = cmfstep(P_, time + dt * (1 - tconst), time + dt, acc_bs))) return rc;
    if ((rc = error(&err))) return rc;
    err *= 0.25;
    ++res.steps;
    if (err > acc_cmf) {
      if ((rc = copy(P_, P0_))) return rc;
      calc_mat = true;
      dt *= std::pow(acc_cmf / err, 1.0 / 3.0) * 0.8;
      ++res.rejected;
    } else {
      time += dt;
      if ((rc = orthogonal(P_))) return rc;
      if (time + 1e-10 >= t_next) {
        if ((rc = write_state(P_))) return rc;
        if ((rc = output(time, res))) return rc;
        t_next = time + out;
      }
      dt = dt * std::min(std::pow(acc_cmf / err, 1.0 / 3.0), max_increase);
    }
  }
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  if (deviceBs_) {
    // the derivative evaluations the device integrators counted
    std::vector<BsNode> bn(NN_);
    QTG_HIP_TRY(hipMemcpy(bn.data(), d_bs_, sizeof(BsNode) * NN_, hipMemcpyDeviceToHost));
    nderiv_ = 0;
    for (const BsNode& b : bn) nderiv_ += b.nderiv;
  }
  res.builds = nbuild_;
  res.derivs = nderiv_;
  res.ticks = nticks_;
  res.dt_end = dt;
  return QTG_OK;
}

// ------------------------------------------------------------- eigenstates
int Cmf::eig_init() {
  if (PL_) return QTG_OK;
  const HTree& t = E.tree_;
  const HNode& top = t.nodes[NN_ - 1];
  const int n = top.n;
  if (n > 64) {
    set_error("eigenstates: more than 64 states");
    return QTG_EARG;
  }
  int rc;
  if ((rc = dalloc((void**)&PL_, sizeof(cplx<double>) * total_))) return rc;
  // [H_top | UT | scratch n x n]
  if ((rc = dalloc((void**)&eigm_, sizeof(cplx<double>) * 3 * n * n))) return rc;
  if ((rc = dalloc((void**)&eigv_, sizeof(double) * n))) return rc;
  // <Psi_top| (H Psi)_top> on the state index
  SwItem hx{};
  hx.node = top.addr;
  hx.k = top.order() - 1;
  hx.nf = 0;
  hx.bra = hx.ket = top.off;
  hx.out = 0;
  // C(:,j) = sum_k UT(j,k) B(:,k) on the state index
  SwApply rot{};
  rot.node = top.addr;
  rot.nf = 1;
  rot.fmode[0] = top.order() - 1;
  rot.fmat[0] = (int64_t)n * n;
  // Gram-Schmidt of every node
  std::vector<OrthoItem> on(NN_);
  for (int a = 0; a < NN_; ++a) {
    on[a].node = a;
    on[a].sw = -1;
  }
  if ((rc = dalloc((void**)&d_hx_, sizeof(SwItem)))) return rc;
  if ((rc = dalloc((void**)&d_rot_, sizeof(SwApply)))) return rc;
  if ((rc = dalloc((void**)&d_on_, sizeof(OrthoItem) * NN_))) return rc;
  QTG_HIP_TRY(hipMemcpy(d_hx_, &hx, sizeof(hx), hipMemcpyHostToDevice));
  QTG_HIP_TRY(hipMemcpy(d_rot_, &rot, sizeof(rot), hipMemcpyHostToDevice));
  QTG_HIP_TRY(hipMemcpy(d_on_, on.data(), sizeof(OrthoItem) * NN_, hipMemcpyHostToDevice));
  return QTG_OK;
}

// Apply and the top node's derivative of P with the matrices of the last
// build, at the top node only: D_top = -i phase (H P)_top.
int Cmf::apply_top(const cplx<double>* P, double ph_re, double ph_im, double* energy) {
  std::vector<uint8_t> act(NN_, 0);
  act[NN_ - 1] = 1;
  unsigned char* ds = (unsigned char*)dev_stage_;
  const size_t off = 2 * (sizeof(NodeProg) * NN_ + sizeof(VecOp) * NN_ * kOpsPerNode);
  QTG_HIP_TRY(hipMemcpyAsync(ds + off, act.data(), NN_, hipMemcpyHostToDevice, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  return E.apply_finish((const double*)P, (double*)D_, ds + off, ph_re, ph_im, energy, s_);
}

// dotProduct tree (MatrixTreeFunctions_Implementation.h:16-39) of bra and
// ket; the top node's n x n matrix stays at spool_ + sslot_[top].
int Cmf::overlap_tree(const cplx<double>* bra, const cplx<double>* ket) {
  for (size_t h = 0; h < ovRange_.size(); ++h) {
    auto& r = ovRange_[h];
    int rc = launch_sandwich<double>(d_ov_ + r.first, r.second - r.first, E.d_nodes_, bra, ket, spool_, ovNmax_[h],
                                     s_);
    if (rc) return rc;
  }
  return QTG_OK;
}

// Eigenstate (Eigenstates.cpp:18-31): build, Expectation (the top node's
// <Psi_i|H|Psi_j>), its eigenvectors, Psi_top <- multStateArTB(U, Psi_top).
int Cmf::eigenstate(std::vector<double>& ev) {
  const HNode& top = E.tree_.nodes[NN_ - 1];
  const int n = top.n;
  int rc;
  if ((rc = copy((cplx<double>*)E.d_psi64_, P_))) return rc;
  if ((rc = E.build(s_))) return rc;
  // phase i: D = -i * i * H Psi = H Psi
  if ((rc = apply_top(P_, 0.0, 1.0, nullptr))) return rc;
  if ((rc = launch_sandwich<double>(d_hx_, 1, E.d_nodes_, P_, D_, eigm_, ovNmax_.back(), s_))) return rc;
  if ((rc = launch_eig_sorted(eigm_, n, eigm_ + (int64_t)n * n, eigv_, s_))) return rc;
  // the top node only: its LDS buffers are sized by its own tensor
  if ((rc = launch_apply_sw(d_rot_, 1, E.d_nodes_, P_, eigm_, E.tree_.nodes[NN_ - 1].N, s_))) return rc;
  ev.assign(n, 0.0);
  QTG_HIP_TRY(hipMemcpyAsync(ev.data(), eigv_, sizeof(double) * n, hipMemcpyDeviceToHost, s_));
  QTG_HIP_TRY(hipStreamSynchronize(s_));
  return QTG_OK;
}

int Cmf::eigenstates(double dt, double out, int64_t iterations, double acc_cmf, double acc_bs, int64_t max_steps,
                     EigenResult& res) {
  int rc;
  if ((rc = eig_init())) return rc;
  const int n = E.tree_.nodes[NN_ - 1].n;
  const cplx<double>* stop = spool_ + sslot_[NN_ - 1];
  ph_re_ = 0.0;
  ph_im_ = -1.0;
  std::vector<double> ev;
  if ((rc = eigenstate(ev))) return rc;
  res.energies.insert(res.energies.end(), ev.begin(), ev.end());
  std::vector<cd> ov((size_t)n * n);
  for (int64_t it = 0; it < iterations; ++it) {
    if ((rc = copy(PL_, P_))) return rc;
    CmfResult r;
    if ((rc = integrate(0.0, out - 1e-5, dt, out, acc_cmf, acc_bs, max_steps, r))) return rc;
    dt = r.dt_end;
    res.steps += r.steps;
    res.rejected += r.rejected;
    res.builds += r.builds;
    res.derivs += r.derivs;
    // propagatorEnergies (Eigenstates.cpp:8-16)
    if ((rc = overlap_tree(P_, P_))) return rc;
    if ((rc = launch_eig_sorted(stop, n, nullptr, eigv_, s_))) return rc;
    ev.assign(n, 0.0);
    QTG_HIP_TRY(hipMemcpyAsync(ev.data(), eigv_, sizeof(double) * n, hipMemcpyDeviceToHost, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
    for (double& e : ev) e = -std::log(e) / (2.0 * out);
    res.propagator.insert(res.propagator.end(), ev.begin(), ev.end());
    // orthonormal (TensorTree_Implementation.h:190-195)
    if ((rc = bigOrtho_ ? launch_ortho_global(d_on_, NN_, E.d_nodes_, P_, sw_, E.nspf_, s_)
                        : launch_ortho(d_on_, NN_, E.d_nodes_, P_, sw_, E.nmax_, E.nspf_, s_)))
      return rc;
    if ((rc = eigenstate(ev))) return rc;
    res.energies.insert(res.energies.end(), ev.begin(), ev.end());
    // overlap with the previous iterate: s(i) = sum_j |<Psi_i|Psi'_j>|^2
    if ((rc = overlap_tree(P_, PL_))) return rc;
    QTG_HIP_TRY(hipMemcpyAsync(ov.data(), stop, sizeof(cd) * n * n, hipMemcpyDeviceToHost, s_));
    QTG_HIP_TRY(hipStreamSynchronize(s_));
    for (int i = 0; i < n; ++i) {
      double sum = 0.0;
      for (int j = 0; j < n; ++j) sum += std::norm(ov[i + (size_t)j * n]);
      res.converged.push_back(sum);
    }
  }
  res.dt_end = dt;
  return QTG_OK;
}

}  // namespace qtg
