// Host-side model of the engine: tree topology in QuTree's post-order
// addressing, HO-DVR leaf bases, and the sum-of-products Hamiltonian reduced
// to one merged complex matrix per (term, mode).
#pragma once
#include <complex>
#include <cstdint>
#include <map>
#include <random>
#include <string>
#include <vector>

namespace qtg {

using cd = std::complex<double>;

struct HLeaf {
  int dim = 0, type = 0, mode = 0;
  double omega = 1.0, r0 = 0.0, wfr0 = 0.0, wfomega = 1.0;
  std::vector<double> x;      // DVR grid
  std::vector<double> trafo;  // FBR -> DVR, column-major
  std::vector<double> kin;    // kinetic matrix in the DVR, column-major
  std::vector<cd> p;          // momentum in the DVR
};

struct HNode {
  int addr = -1, parent = -1, childIdx = 0, depth = 0, height = 0;
  std::vector<int> children;
  bool bottom = false;
  HLeaf leaf;
  std::vector<int> dims;  // tensor shape, last = number of SPFs
  int64_t off = 0;        // arena offset (complex elements)
  int N = 0, n = 0;
  bool top() const { return parent < 0; }
  int order() const { return (int)dims.size(); }
};

struct HTree {
  std::vector<HNode> nodes;  // by address: post-order, last child visited first
  std::vector<int> leafNode; // mode -> bottom node
  int64_t total = 0;         // arena size
  int maxHeight = 0, maxDepth = 0;
  int topAddr() const { return (int)nodes.size() - 1; }
};

// Parses the compact tree text of Tree(istream) (Node.cpp:101-165 of QuTree)
// plus one "omega r0 wfr0 wfomega" line per mode, and builds the HO-DVR bases.
HTree parse_tree(const std::string& text);

// Hamiltonian term: coefficient and one merged leaf matrix per acted-on mode.
struct HTerm {
  cd coeff;
  std::vector<std::pair<int, int>> leafmats;  // (mode, index into leaf-matrix list), modes ascending
};

struct HSop {
  std::vector<HTerm> terms;
  std::vector<std::vector<cd>> leafMatrices;  // d x d column-major
  std::vector<int> leafMatrixMode;
};

// factor kinds: 0 x, 1 x^2, 2 kin, 3 p, 4 identity, 5 explicit matrix, 6 polynomial
int build_sop(const HTree& tree, int64_t nterms, const double* coeff, const int64_t* term_off,
              const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool, HSop& out,
              std::string& err);

// TensorTree(rng, tree, delta_lowest): HO ground-state SPFs at the bottom,
// uniform[-1,1] real coefficients above, Gram-Schmidt per node.
void fill_random(const HTree& tree, uint64_t seed, bool delta_lowest, std::vector<cd>& psi);
// The same drawn from mt19937(seed) after `discard` outputs; returns the
// number of outputs it drew (TensorTree(mt19937&, tree, delta_lowest) on a
// generator that earlier states already drew from).
int64_t fill_random_stream(const HTree& tree, uint64_t seed, int64_t discard, bool delta_lowest,
                           std::vector<cd>& psi);

void gram_schmidt(cd* A, int before, int n);

}  // namespace qtg
