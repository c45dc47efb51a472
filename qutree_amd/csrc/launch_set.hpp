// An ordered list of grouped MFMA launches (mode products, hole
// contractions) over a set of buffers identified by small integers; used by
// the tree-operator (TTNO) path.  Descriptor offsets are relative to the
// buffers a run binds.
#pragma once
#include <vector>

#include "gemm_ops.hpp"

namespace qtg {

struct LaunchSet {
  struct Launch {
    int kind;  // 0 mode products, 1 hole contractions
    int64_t d0, nd, ntasks, wg0;
    int x0, x1, m, o;
    double flops;
  };
  std::vector<GmDesc> gm;
  std::vector<GcDesc> gc;
  std::vector<int32_t> gmwg, gcwg;
  std::vector<Launch> launches;
  double flops = 0.0;

  GmDesc* d_gm = nullptr;
  GcDesc* d_gc = nullptr;
  int32_t* d_gmwg = nullptr;
  int32_t* d_gcwg = nullptr;

  ~LaunchSet() { release(); }
  void release();
  // appends one launch (no-op for an empty list); x0 / x1 are the input
  // buffer ids selected by each descriptor, m the matrix buffer, o the output
  void push_gm(std::vector<GmDesc> descs, int x0, int x1, int m, int o);
  void push_gc(std::vector<GcDesc> descs, int x0, int x1, int o);
  int upload();
  // complex128 launches over bases[buffer id]
  int run(void* const* bases, hipStream_t s) const;
};

}  // namespace qtg
