// Tree tensor network operator (TTNO) application on the device (see ttno.cpp).
#pragma once
#include <vector>

#include "engine.hpp"
#include "launch_set.hpp"

namespace qtg {

class Ttno {
 public:
  explicit Ttno(Engine& e) : E(e) {}
  ~Ttno();
  // bond[a] = bond dimension D_a of node a (1 at the top); B = operator
  // tensors in node-address order: bottom (dim^2, D_a) with the leaf
  // matrices as columns, upper (D_child0, ..., D_childk, D_a).
  int create(const int64_t* bond, const double* B);
  // out = TTNO application at every node to the ket's node tensors, with the
  // representation <bra| . |ket> (host arrays, engine layout); expect = <bra|H|ket>.
  int apply(const double* bra, const double* ket, double* out, double* expect, hipStream_t s);
  double flops() const { return ls_.flops; }

 private:
  Engine& E;
  std::vector<int64_t> D_, boff_, rep0_, con0_;
  int64_t poolSize_ = 0, workSize_ = 0, repTop_ = 0;
  std::vector<double> pool_;  // constant part: B tensors and their leg-permuted copies
  LaunchSet ls_;
  void *d_pool_ = nullptr, *d_work_ = nullptr, *d_bra_ = nullptr, *d_ket_ = nullptr, *d_out_ = nullptr;
};

}  // namespace qtg
