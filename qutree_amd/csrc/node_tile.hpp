// Device helpers shared by the tree kernels: the (before, n_k, after) view of
// a node tensor and the mode-k product on LDS-resident node tensors.
#pragma once
#include "common.hpp"
#include "tree_ops.hpp"

namespace qtg {

constexpr int TB = 256;

// dst[e] = src[e] for e < N (N >= 1) by a workgroup of NT threads: eight
// loads per thread are in flight before their LDS stores, instead of one
// load round trip per element a thread copies.
template <int NT, typename T>
__device__ inline void stage_to_lds(cplx<T>* dst, const cplx<T>* __restrict__ src, int N) {
  for (int base = 0; base < N; base += 8 * NT) {
    cplx<T> v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[min(base + u * NT + (int)threadIdx.x, N - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = base + u * NT + (int)threadIdx.x;
      if (e < N) dst[e] = v[u];
    }
  }
}

__device__ inline void node_view(const DNode& nd, int k, int& bef, int& nk, int& aft) {
  bef = 1;
  aft = 1;
  for (int i = 0; i < k; ++i) bef *= nd.dims[i];
  nk = nd.dims[k];
  for (int i = k + 1; i < nd.order; ++i) aft *= nd.dims[i];
}

// dst = M x_f src on an LDS-resident node tensor (M is n_f x n_f).
template <typename T>
__device__ void mode_apply(cplx<T>* dst, const cplx<T>* src, const cplx<T>* __restrict__ M, const DNode& nd,
                           int f) {
  int bef, nf, aft;
  node_view(nd, f, bef, nf, aft);
  const int N = nd.N;
  for (int e = threadIdx.x; e < N; e += TB) {
    const int b = e % bef;
    const int q = e / bef;
    const int k = q % nf;
    const int a = q / nf;
    const cplx<T>* s = src + b + a * bef * nf;
    cplx<T> acc = czero<T>();
    for (int m = 0; m < nf; ++m) cfma(acc, M[k + m * nf], s[m * bef]);
    dst[e] = acc;
  }
}

}  // namespace qtg
