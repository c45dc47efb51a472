// Batched FP64 small-matrix kernels (smallmat.hip).
#pragma once
#include <cstdint>

#include "common.hpp"

namespace qtg {

// One matrix of a batch: input at a, output at q (complex-element offsets
// from the launch's base pointers), eigenvalues at e (heev), shape m x n
// (square ops use n).
struct SmItem {
  int64_t a, q, e, m;
  int32_t n, pad;
};

// Every launcher takes either a device item array (items) or, with items
// null, a uniform batch: item b is u with a, q, e advanced by b times st's.
// heev (U, ev; either may be null) and/or the regularised inverse R (eps)
int launch_heev_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, int nmax,
                      const void* in, void* U, double* ev, void* R, double eps, hipStream_t s);
int launch_inverse_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, int nmax,
                         const void* in, void* out, hipStream_t s);
// thin Q of every (m x n) item, n <= 64, m >= n; in and out may alias.
// With R non-null, an item with e >= 0 also writes its n x n upper
// triangular R (A = Q R) at R + e.
int launch_qr_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, const void* in, void* out,
                    hipStream_t s, void* R = nullptr);

}  // namespace qtg
