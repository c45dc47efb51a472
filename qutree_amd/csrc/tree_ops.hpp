// Work-item records of the grouped tree kernels (device-resident arrays built
// by the host engine) and their launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qtg {

constexpr int kMaxOrder = 4;  // node tensors of order <= 4 (<= 3 children)
constexpr int kMaxFactors = 3;

struct DNode {
  int32_t order, n, N, pad;
  int32_t dims[kMaxOrder];
  int64_t off;  // offset of the node tensor in the arena (complex elements)
};

// out = contraction_k(bra, F ket), F = prod of nf mode factors (pool matrices)
struct SwItem {
  int32_t node, k, nf, pad;
  int32_t fmode[kMaxFactors];
  int32_t pad2;
  int64_t fmat[kMaxFactors];
  int64_t bra, ket;  // tensor offsets (complex elements) in the bra / ket arenas
  int64_t out;       // output matrix offset in the pool
};

// one summand c * F A_v of Apply at a node; the last factor is applied while
// accumulating
struct ApTerm {
  int32_t nf, pad;
  int32_t fmode[kMaxFactors];
  int32_t pad2;
  int64_t fmat[kMaxFactors];
  double cre, cim;
};

struct ApChunk {
  int32_t node, t0, t1, pad;
  int64_t out;  // partial-sum slab offset (complex elements)
};

struct FinNode {
  int32_t node, nchunk, top, pad;
  int64_t part0;   // first partial slab of the node
  int64_t rhoinv;  // offset of rho^-1 in the FP64 pool
};

// out (n x n) = sum_t coef_t M_t, M_t = pool matrix at src (identity if src < 0)
struct CombOut {
  int64_t out;
  int32_t n, t0, t1, pad;
};
struct CombTerm {
  double cre, cim;
  int64_t src;
};

struct EigItem {
  int32_t n, pad;
  int64_t in;   // density matrix (bulk pool)
  int64_t out;  // regularised inverse (FP64 pool)
  int64_t ev;   // eigenvalues (FP64 pool, complex slots) or -1
};

template <typename T>
int launch_sandwich(const SwItem* items, int64_t count, const DNode* nodes, const void* bra, const void* ket,
                    void* pool, int nmax, hipStream_t s);
template <typename T>
int launch_apply(const ApChunk* chunks, int64_t nchunks, const ApTerm* terms, const DNode* nodes, const void* psi,
                 const void* pool, void* partial, int nmax, hipStream_t s, const uint8_t* active = nullptr);
template <typename T>
int launch_finish(const FinNode* fin, int64_t count, const DNode* nodes, const void* psi64, const void* partial,
                  const void* pool64, void* dpsi, void* apply_out, double ph_re, double ph_im, void* energy,
                  int nmax, int nspf, hipStream_t s, const uint8_t* active = nullptr);
// dynamic LDS of finish_kernel: Y and A (FP64, nmax each) and two nspf^2 matrices
inline size_t finish_lds_bytes(int nmax, int nspf) {
  return (2 * (size_t)nmax + 2 * (size_t)nspf * nspf) * 16;
}
// dynamic LDS of the sandwich and apply kernels: three node-sized buffers
inline size_t work_lds_bytes(int nmax, size_t elem) { return 3 * (size_t)nmax * elem; }
constexpr size_t kLdsBudget = 160 * 1024 - 8 * 1024;  // leave room for static shared arrays
template <typename T>
int launch_combine(const CombOut* outs, int64_t count, const CombTerm* terms, void* pool, hipStream_t s);
template <typename T>
int launch_heev(const EigItem* items, int64_t count, const void* pool, void* pool64, double eps, int nspf,
                hipStream_t s);

}  // namespace qtg
