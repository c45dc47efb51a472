// Host model: tree text parser, HO-DVR leaves, SOP reduction, random states.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace qtg {

namespace {

struct PNode {
  int nstates = 0;
  bool leaf = false;
  int ldim = 0, ltype = 0, lmode = 0;
  std::vector<std::unique_ptr<PNode>> kids;
};

std::unique_ptr<PNode> read_node(std::istream& is) {
  auto p = std::make_unique<PNode>();
  long ns = 0, f = 0;
  if (!(is >> ns >> f) || ns <= 0 || f >= 0) throw std::runtime_error("tree text: expected '<nstates> -<nchildren>'");
  p->nstates = (int)ns;
  for (long i = 0; i < -f; ++i) {
    const auto mark = is.tellg();
    long a = 0, b = 0;
    if (!(is >> a >> b)) throw std::runtime_error("tree text: truncated");
    is.seekg(mark);
    if (b < 0) {
      p->kids.push_back(read_node(is));
    } else {
      auto l = std::make_unique<PNode>();
      long dim, type, mode;
      is >> dim >> type >> mode;
      if (dim <= 0 || mode < 0) throw std::runtime_error("tree text: bad leaf");
      l->leaf = true;
      l->ldim = (int)dim;
      l->ltype = (int)type;
      l->lmode = (int)mode;
      p->kids.push_back(std::move(l));
    }
  }
  return p;
}

// symmetric Jacobi eigensolver (real, small), ascending, first row >= 0
void sym_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (i == j ? dia : off) += A[i + j * n] * A[i + j * n];
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p + q * n];
        if (apq == 0.0) continue;
        const double th = (A[q + q * n] - A[p + p * n]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a + a * n] < A[b + b * n]; });
  std::vector<double> W((size_t)n * n);
  w.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    w[j] = A[idx[j] + idx[j] * n];
    const double sg = V[0 + idx[j] * n] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) W[i + j * n] = sg * V[i + idx[j] * n];
  }
  V = W;
}

// HO-DVR basis of dimension d: grid = eigenvalues of the HO position matrix,
// kinetic energy transformed from the HO basis (HO_Basis.cpp:3-112 of QuTree).
void init_ho(HLeaf& L) {
  const int d = L.dim;
  std::vector<double> X((size_t)d * d, 0.0);
  for (int k = 0; k + 1 < d; ++k) X[(k + 1) + k * d] = X[k + (k + 1) * d] = std::sqrt((k + 1) / (2.0 * L.omega));
  std::vector<double> w, U;
  sym_eig(X, d, w, U);
  L.x.resize(d);
  for (int i = 0; i < d; ++i) L.x[i] = w[i] + L.r0;
  L.trafo = U;
  std::vector<double> K((size_t)d * d, 0.0);  // <k|p^2/2|l> in the HO basis
  for (int k = 0; k < d; ++k) {
    K[k + k * d] = L.omega * (2 * k + 1) / 4.0;
    if (k + 2 < d) K[k + (k + 2) * d] = K[(k + 2) + k * d] = -L.omega * std::sqrt((k + 1.0) * (k + 2.0)) / 4.0;
  }
  L.kin.assign((size_t)d * d, 0.0);
  std::vector<double> KU((size_t)d * d, 0.0);
  for (int j = 0; j < d; ++j)
    for (int m = 0; m < d; ++m)
      for (int i = 0; i < d; ++i) KU[i + j * d] += K[i + m * d] * U[m + j * d];
  for (int j = 0; j < d; ++j)
    for (int m = 0; m < d; ++m)
      for (int i = 0; i < d; ++i) L.kin[i + j * d] += U[m + i * d] * KU[m + j * d];
  // momentum i sqrt(w/2) (a^+ - a), transformed the same way
  std::vector<cd> P((size_t)d * d, 0.0);
  for (int k = 0; k + 1 < d; ++k) {
    const double s = std::sqrt((k + 1) * L.omega / 2.0);
    P[(k + 1) + k * d] = cd(0.0, s);
    P[k + (k + 1) * d] = cd(0.0, -s);
  }
  L.p.assign((size_t)d * d, 0.0);
  std::vector<cd> PU((size_t)d * d, 0.0);
  for (int j = 0; j < d; ++j)
    for (int m = 0; m < d; ++m)
      for (int i = 0; i < d; ++i) PU[i + j * d] += P[i + m * d] * U[m + j * d];
  for (int j = 0; j < d; ++j)
    for (int m = 0; m < d; ++m)
      for (int i = 0; i < d; ++i) L.p[i + j * d] += U[m + i * d] * PU[m + j * d];
}

}  // namespace

HTree parse_tree(const std::string& text) {
  std::istringstream is(text);
  auto root = read_node(is);
  HTree t;
  // address nodes in post-order, last child first (Node::nextNode)
  std::function<int(PNode&, int, int)> visit = [&](PNode& p, int depth, int childIdx) -> int {
    const bool bottom = (p.kids.size() == 1 && p.kids[0]->leaf);
    std::vector<int> kids(p.kids.size(), -1);
    int h = 0;
    if (!bottom) {
      for (int k = (int)p.kids.size() - 1; k >= 0; --k) {
        if (p.kids[k]->leaf) throw std::runtime_error("tree text: a node mixes leaves and nodes");
        kids[k] = visit(*p.kids[k], depth + 1, k);
      }
    }
    HNode n;
    n.addr = (int)t.nodes.size();
    n.childIdx = childIdx;
    n.depth = depth;
    n.bottom = bottom;
    if (bottom) {
      n.leaf.dim = p.kids[0]->ldim;
      n.leaf.type = p.kids[0]->ltype;
      n.leaf.mode = p.kids[0]->lmode;
      n.dims = {n.leaf.dim};
    } else {
      n.children = kids;
      for (int c : kids) {
        n.dims.push_back(t.nodes[c].n);
        h = std::max(h, t.nodes[c].height + 1);
      }
    }
    n.height = h;
    n.dims.push_back(p.nstates);
    n.n = p.nstates;
    n.N = 1;
    for (int d : n.dims) n.N *= d;
    t.nodes.push_back(n);
    for (int c : kids)
      if (c >= 0) t.nodes[c].parent = n.addr;
    return n.addr;
  };
  visit(*root, 0, 0);
  t.nodes.back().parent = -1;
  int nleaves = 0;
  for (auto& n : t.nodes) {
    if (n.bottom) ++nleaves;
    if (n.order() > 4) throw std::runtime_error("tree text: node order above 4 is not supported");
  }
  t.leafNode.assign(nleaves, -1);
  for (auto& n : t.nodes)
    if (n.bottom) {
      if (n.leaf.mode >= nleaves || t.leafNode[n.leaf.mode] >= 0) throw std::runtime_error("tree text: bad leaf modes");
      t.leafNode[n.leaf.mode] = n.addr;
    }
  for (int m = 0; m < nleaves; ++m) {
    HLeaf& L = t.nodes[t.leafNode[m]].leaf;
    double a, b, c, e;
    if (is >> a >> b >> c >> e) {
      L.omega = a;
      L.r0 = b;
      L.wfr0 = c;
      L.wfomega = e;
    }
    if (L.type != 0) throw std::runtime_error("only HO-DVR leaves (type 0) are supported");
    init_ho(L);
  }
  int64_t off = 0;
  for (auto& n : t.nodes) {
    n.off = off;
    off += n.N;
    t.maxHeight = std::max(t.maxHeight, n.height);
    t.maxDepth = std::max(t.maxDepth, n.depth);
  }
  t.total = off;
  return t;
}

int build_sop(const HTree& tree, int64_t nterms, const double* coeff, const int64_t* term_off,
              const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool, HSop& out,
              std::string& err) {
  out = HSop();
  std::map<std::pair<int, std::vector<double>>, int> dedup;  // identical leaf matrices share an index
  for (int64_t l = 0; l < nterms; ++l) {
    HTerm T;
    T.coeff = cd(coeff[2 * l], coeff[2 * l + 1]);
    // merged matrix per mode in application order: M = F_last ... F_first
    std::map<int, std::vector<cd>> per_mode;
    for (int64_t f = term_off[l]; f < term_off[l + 1]; ++f) {
      const int mode = (int)fmode[f];
      if (mode < 0 || mode >= (int)tree.leafNode.size()) {
        err = "factor mode out of range";
        return 1;
      }
      const HLeaf& L = tree.nodes[tree.leafNode[mode]].leaf;
      const int d = L.dim;
      std::vector<cd> F((size_t)d * d, 0.0);
      switch (fkind[f]) {
        case 0:
          for (int i = 0; i < d; ++i) F[i + i * d] = L.x[i];
          break;
        case 1:
          for (int i = 0; i < d; ++i) F[i + i * d] = L.x[i] * L.x[i];
          break;
        case 2:
          for (int i = 0; i < d * d; ++i) F[i] = L.kin[i];
          break;
        case 3:
          F = L.p;
          break;
        case 4:
          for (int i = 0; i < d; ++i) F[i + i * d] = 1.0;
          break;
        case 5: {
          const double* m = pool + 2 * fpool[f];
          for (int i = 0; i < d * d; ++i) F[i] = cd(m[2 * i], m[2 * i + 1]);
          break;
        }
        case 6: {
          const double* a = pool + 2 * fpool[f];
          for (int i = 0; i < d * d; ++i) F[i] = a[0] * L.kin[i];
          for (int i = 0; i < d; ++i) {
            const double x = L.x[i];
            F[i + i * d] += a[1] * x + a[2] * x * x + a[3] * x * x * x * x;
          }
          break;
        }
        default:
          err = "unknown leaf operator kind";
          return 1;
      }
      auto it = per_mode.find(mode);
      if (it == per_mode.end()) {
        per_mode[mode] = F;
      } else {
        std::vector<cd> P((size_t)d * d, 0.0);  // F * previous
        const auto& G = it->second;
        for (int j = 0; j < d; ++j)
          for (int m = 0; m < d; ++m)
            for (int i = 0; i < d; ++i) P[i + j * d] += F[i + m * d] * G[m + j * d];
        it->second = P;
      }
    }
    for (auto& pm : per_mode) {
      std::vector<double> key;
      key.reserve(pm.second.size() * 2);
      for (auto& z : pm.second) {
        key.push_back(z.real());
        key.push_back(z.imag());
      }
      auto k = std::make_pair(pm.first, key);
      auto f = dedup.find(k);
      int idx;
      if (f == dedup.end()) {
        idx = (int)out.leafMatrices.size();
        out.leafMatrices.push_back(pm.second);
        out.leafMatrixMode.push_back(pm.first);
        dedup[k] = idx;
      } else {
        idx = f->second;
      }
      T.leafmats.emplace_back(pm.first, idx);
    }
    out.terms.push_back(T);
  }
  return 0;
}

void gram_schmidt(cd* A, int before, int n) {
  // iterated classical Gram-Schmidt per column (Tensor_Implementation.h:806-853 of QuTree)
  for (int k = 0; k < n; ++k) {
    double accum = 1.0;
    for (int iter = 0; iter < 15 && accum > 1e-12; ++iter) {
      accum = 0.0;
      for (int m = 0; m < k; ++m) {
        cd ov = 0.0;
        for (int i = 0; i < before; ++i) ov += std::conj(A[i + m * before]) * A[i + k * before];
        accum += std::abs(ov);
        for (int i = 0; i < before; ++i) A[i + k * before] -= ov * A[i + m * before];
      }
      double nrm = 0.0;
      for (int i = 0; i < before; ++i) nrm += std::norm(A[i + k * before]);
      if (nrm != 0.0) {
        nrm = std::sqrt(nrm);
        for (int i = 0; i < before; ++i) A[i + k * before] /= nrm;
      }
    }
  }
}

namespace {
// an mt19937 that counts its outputs, so a caller can continue the stream
struct CountingGen {
  typedef std::mt19937::result_type result_type;
  std::mt19937& g;
  int64_t n = 0;
  static constexpr result_type min() { return std::mt19937::min(); }
  static constexpr result_type max() { return std::mt19937::max(); }
  result_type operator()() {
    ++n;
    return g();
  }
};
}  // namespace

void fill_random(const HTree& tree, uint64_t seed, bool delta_lowest, std::vector<cd>& psi) {
  fill_random_stream(tree, seed, 0, delta_lowest, psi);
}

int64_t fill_random_stream(const HTree& tree, uint64_t seed, int64_t discard, bool delta_lowest, std::vector<cd>& psi) {
  std::mt19937 mt((std::mt19937::result_type)seed);
  mt.discard((unsigned long long)discard);
  CountingGen gen{mt};
  psi.assign((size_t)tree.total, 0.0);
  for (const HNode& n : tree.nodes) {
    cd* A = psi.data() + n.off;
    const int L = n.N / n.n;
    if (n.bottom) {
      const HLeaf& lf = n.leaf;
      const int d = lf.dim;
      for (int i = 0; i < d; ++i) {
        const double w = lf.trafo[0 + i * d] / std::exp(-0.5 * lf.omega * std::pow(lf.x[i] - lf.r0, 2));
        A[i] = w * std::exp(-0.5 * lf.wfomega * std::pow(lf.x[i] - lf.wfr0, 2));
      }
      for (int k = 1; k < n.n; ++k)
        for (int i = 0; i < d; ++i) A[i + k * d] = lf.x[i] * A[i + (k - 1) * d];
    } else {
      std::uniform_real_distribution<double> dist(-1.0, 1.0);
      for (int i = 0; i < n.N; ++i) A[i] = dist(gen);
      if (delta_lowest) {
        for (int i = 0; i < L; ++i) A[i] = 0.0;
        A[0] = 1.0;
      }
    }
    gram_schmidt(A, L, n.n);
  }
  return gen.n;
}

}  // namespace qtg
