// Wavefront Jacobi eigensolver for small Hermitian matrices (FP64, n <= 64)
// held in LDS, column-major with leading dimension jacobi_ld(n) (even order,
// see jacobi_dim).  All lanes of a
// 64-lane workgroup call it.  Tournament (round-robin) ordering: each step
// rotates up to 32 disjoint (p,q) pairs whose parameters come from the same
// matrix; one pass applies both sides at once, each lane owning 2 x 2 blocks
// (rows of pair i, columns of pair j), and a wavefront vote skips steps in
// which no pair rotates.
// Player at position j of step t: j == 0 -> 0, else 1 + (j - 1 + t) mod (m - 1).  A pair is skipped when
// |a_pq| <= eps_mach sqrt(|a_pp a_qq|) (the relative-accuracy criterion of
// Demmel and Veselic) or |a_pq| <= 1e-24 max_i |a_ii|; a rotated pair gets
// a_pq = a_qp = 0 exactly, so the off-diagonal carries no rounding residue of
// the rotation itself and the sweeps converge quadratically.  Sweeps stop
// when no pair rotates or after 30 sweeps.
// On return A is diagonal (eigenvalues on the diagonal, unsorted) and
// U holds the eigenvectors as columns: A_in = U diag U^H.
#pragma once
#include "common.hpp"

namespace qtg {

constexpr int EW = 64;

// Matrices are padded to an even order m (a zero row and column when n is
// odd) and stored with leading dimension m + 1; a caller's LDS holds
// jacobi_elems(nmax) complex elements per matrix.
__host__ __device__ constexpr int jacobi_dim(int n) { return (n + 1) & ~1; }
__host__ __device__ constexpr int jacobi_ld(int n) { return jacobi_dim(n) + 1; }
__host__ __device__ constexpr int jacobi_elems(int n) { return jacobi_dim(n) * jacobi_ld(n); }

struct JacobiScratch {
  double pc[32], ps[32];
  cplx<double> pph[32];
  int pp[32], pq[32];
};

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, EW);
  return v;
}

// Rotation of tournament pair i at step `step` (m players): its (p, q) with
// p < q and, when |a_pq| passes the criteria above, the parameters of
// G = [[c, s ph], [-s conj(ph), c]].  Returns whether the pair rotates.
__device__ inline bool jacobi_pair(const cplx<double>* A, int n, int ld, int m, int step, int i, double floor_abs,
                                   int& p, int& q, double& c, double& sn, cplx<double>& ph) {
  const int j0 = i, j1 = m - 1 - i;
  p = (j0 == 0) ? 0 : 1 + (j0 - 1 + step) % (m - 1);
  q = (j1 == 0) ? 0 : 1 + (j1 - 1 + step) % (m - 1);
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
  c = 1.0;
  sn = 0.0;
  ph = mk<double>(1.0, 0.0);
  if (q >= n) return false;
  const cplx<double> apq = A[p + q * ld];
  const double aa = sqrt(apq.re * apq.re + apq.im * apq.im);
  const double app = A[p + p * ld].re, aqq = A[q + q * ld].re;
  if (!(aa > 2.2e-16 * sqrt(fabs(app * aqq)) && aa > floor_abs && aa > 1e-300)) return false;
  const double theta = (aqq - app) / (2.0 * aa);
  const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
  c = 1.0 / sqrt(t * t + 1.0);
  sn = t * c;
  ph = mk<double>(apq.re / aa, apq.im / aa);
  return true;
}

// jacobi_pair with fewer FP64 divides and square roots: 1/|a_pq| and
// 1/sqrt(1 + t^2) are reciprocal square roots, the tests compare squares.
__device__ inline bool jacobi_pair_fast(const cplx<double>* A, int n, int ld, int m, int step, int i,
                                        double floor2, int& p, int& q, double& c, double& sn, cplx<double>& ph) {
  const int j0 = i, j1 = m - 1 - i;
  p = (j0 == 0) ? 0 : 1 + (j0 - 1 + step) % (m - 1);
  q = (j1 == 0) ? 0 : 1 + (j1 - 1 + step) % (m - 1);
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
  c = 1.0;
  sn = 0.0;
  ph = mk<double>(1.0, 0.0);
  if (q >= n) return false;
  const cplx<double> apq = A[p + q * ld];
  const double r2 = apq.re * apq.re + apq.im * apq.im;
  const double app = A[p + p * ld].re, aqq = A[q + q * ld].re;
  // |a_pq| > eps sqrt|a_pp a_qq|, |a_pq| > floor, |a_pq| > 1e-140
  if (!(r2 > 4.84e-32 * fabs(app * aqq) && r2 > floor2 && r2 > 1e-280)) return false;
  const double inv = rsqrt(r2);
  const double theta = 0.5 * (aqq - app) * inv;
  const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
  c = rsqrt(t * t + 1.0);
  sn = t * c;
  ph = mk<double>(apq.re * inv, apq.im * inv);
  return true;
}

// The 2 x 2 block (rows p, q of pair bi; columns pc, qc of pair bj) of
// A <- G_bi^H A G_bj and U <- U G_bj, with G = [[c, s ph], [-s conj(ph), c]]
// on (p, q); a rotated diagonal block gets exact zeros off the diagonal and a
// real diagonal.  Pairs with an index beyond n are padding.
__device__ inline void jacobi_block(cplx<double>* A, cplx<double>* U, int n, int ld, bool diag, int p, int q,
                                    double ci, double si, cplx<double> hi, int pc, int qc, double cj, double sj,
                                    cplx<double> hj) {
  if (p >= n) return;
  const bool qv = q < n, qcv = qc < n;
  // rows p, q x columns pc, qc
  cplx<double> a00 = A[p + pc * ld];
  cplx<double> a01 = qcv ? A[p + qc * ld] : czero<double>();
  cplx<double> a10 = qv ? A[q + pc * ld] : czero<double>();
  cplx<double> a11 = (qv && qcv) ? A[q + qc * ld] : czero<double>();
  cplx<double> u00 = U[p + pc * ld];
  cplx<double> u01 = qcv ? U[p + qc * ld] : czero<double>();
  cplx<double> u10 = qv ? U[q + pc * ld] : czero<double>();
  cplx<double> u11 = (qv && qcv) ? U[q + qc * ld] : czero<double>();
  // right: X <- X G_bj (columns pc, qc)
  auto right = [&](cplx<double>& x0, cplx<double>& x1) {
    const cplx<double> a = x0, b = x1;
    x0 = mk<double>(cj * a.re - sj * (hj.re * b.re + hj.im * b.im), cj * a.im - sj * (hj.re * b.im - hj.im * b.re));
    x1 = mk<double>(sj * (hj.re * a.re - hj.im * a.im) + cj * b.re, sj * (hj.re * a.im + hj.im * a.re) + cj * b.im);
  };
  // left: X <- G_bi^H X (rows p, q)
  auto left = [&](cplx<double>& x0, cplx<double>& x1) {
    const cplx<double> a = x0, b = x1;
    x0 = mk<double>(ci * a.re - si * (hi.re * b.re - hi.im * b.im), ci * a.im - si * (hi.re * b.im + hi.im * b.re));
    x1 = mk<double>(si * (hi.re * a.re + hi.im * a.im) + ci * b.re, si * (hi.re * a.im - hi.im * a.re) + ci * b.im);
  };
  if (sj != 0.0) {
    right(a00, a01);
    right(a10, a11);
    right(u00, u01);
    right(u10, u11);
  }
  if (si != 0.0) {
    left(a00, a10);
    left(a01, a11);
  }
  if (diag && si != 0.0) {
    a01 = czero<double>();
    a10 = czero<double>();
    a00.im = 0.0;
    a11.im = 0.0;
  }
  A[p + pc * ld] = a00;
  if (qcv) A[p + qc * ld] = a01;
  if (qv) A[q + pc * ld] = a10;
  if (qv && qcv) A[q + qc * ld] = a11;
  U[p + pc * ld] = u00;
  if (qcv) U[p + qc * ld] = u01;
  if (qv) U[q + pc * ld] = u10;
  if (qv && qcv) U[q + qc * ld] = u11;
}

// The fused step's block update: rows p, q and columns pc, qc all lie in the
// padded m x m matrices, so no index is guarded.
__device__ __forceinline__ void jacobi_block_full(cplx<double>* A, cplx<double>* U, int ld, bool diag, int p, int q,
                                                  double ci, double si, cplx<double> hi, int pc, int qc, double cj,
                                                  double sj, cplx<double> hj) {
  cplx<double> a00 = A[p + pc * ld], a01 = A[p + qc * ld], a10 = A[q + pc * ld], a11 = A[q + qc * ld];
  cplx<double> u00 = U[p + pc * ld], u01 = U[p + qc * ld], u10 = U[q + pc * ld], u11 = U[q + qc * ld];
  auto right = [&](cplx<double>& x0, cplx<double>& x1) {
    const cplx<double> a = x0, b = x1;
    x0 = mk<double>(cj * a.re - sj * (hj.re * b.re + hj.im * b.im), cj * a.im - sj * (hj.re * b.im - hj.im * b.re));
    x1 = mk<double>(sj * (hj.re * a.re - hj.im * a.im) + cj * b.re, sj * (hj.re * a.im + hj.im * a.re) + cj * b.im);
  };
  auto left = [&](cplx<double>& x0, cplx<double>& x1) {
    const cplx<double> a = x0, b = x1;
    x0 = mk<double>(ci * a.re - si * (hi.re * b.re - hi.im * b.im), ci * a.im - si * (hi.re * b.im + hi.im * b.re));
    x1 = mk<double>(si * (hi.re * a.re + hi.im * a.im) + ci * b.re, si * (hi.re * a.im - hi.im * a.re) + ci * b.im);
  };
  if (sj != 0.0) {
    right(a00, a01);
    right(a10, a11);
    right(u00, u01);
    right(u10, u11);
    U[p + pc * ld] = u00;
    U[p + qc * ld] = u01;
    U[q + pc * ld] = u10;
    U[q + qc * ld] = u11;
  }
  if (si != 0.0) {
    left(a00, a10);
    left(a01, a11);
    if (diag) {
      a01 = czero<double>();
      a10 = czero<double>();
      a00.im = 0.0;
      a11.im = 0.0;
    }
  }
  if (sj != 0.0 || si != 0.0) {
    A[p + pc * ld] = a00;
    A[p + qc * ld] = a01;
    A[q + pc * ld] = a10;
    A[q + qc * ld] = a11;
  }
}

// ld must be jacobi_ld(n); the padding row and column (odd n) are set here.
__device__ __forceinline__ void jacobi_eig(cplx<double>* A, cplx<double>* U, int n, int ld, JacobiScratch& js) {
  const int md = jacobi_dim(n);
  for (int o = threadIdx.x; o < md * md; o += EW) {
    const int i = o % md, j = o / md;
    U[i + j * ld] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
    if (i >= n || j >= n) A[i + j * ld] = czero<double>();
  }
  __syncthreads();
  // Hermitise (the input is Hermitian up to rounding)
  for (int o = threadIdx.x; o < n * n; o += EW) {
    const int i = o % n, j = o / n;
    if (i < j) {
      const cplx<double> a = A[i + j * ld], b = A[j + i * ld];
      const cplx<double> m = mk<double>(0.5 * (a.re + b.re), 0.5 * (a.im - b.im));
      A[i + j * ld] = m;
      A[j + i * ld] = cconj(m);
    } else if (i == j) {
      A[i + i * ld].im = 0.0;
    }
  }
  __syncthreads();
  if (n < 2) return;
  const int m = (n + 1) & ~1;  // even player count (a dummy when n is odd)
  const int np = m / 2;
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = fmax(scale, fabs(A[i + i * ld].re));
  const double floor_abs = 1e-24 * scale;
  if (np * np <= EW) {
    // one 2 x 2 block per lane: every lane derives the rotations of its row
    // and column pairs itself (the same inputs, so the same parameters in
    // every lane), all reads of a step precede its writes in the wavefront,
    // and a step needs one barrier
    const int bi = threadIdx.x % np, bj = threadIdx.x / np;
    const bool mine = threadIdx.x < np * np;
    const double floor2 = floor_abs * floor_abs;
    for (int sweep = 0; sweep < 30; ++sweep) {
      int rotated = 0;
      for (int step = 0; step < m - 1; ++step) {
        int p = 0, q = 0, pc = 0, qc = 0;
        double ci = 1.0, si = 0.0, cj = 1.0, sj = 0.0;
        cplx<double> hi = mk<double>(1.0, 0.0), hj = hi;
        int flag = 0;
        if (mine) {
          if (jacobi_pair_fast(A, n, ld, m, step, bi, floor2, p, q, ci, si, hi)) flag = 1;
          jacobi_pair_fast(A, n, ld, m, step, bj, floor2, pc, qc, cj, sj, hj);
        }
        // the wavefront's reads of this step stay ahead of its writes
        asm volatile("" ::: "memory");
        if (!__any(flag)) continue;  // uniform: no pair rotates in this step
        rotated = 1;
        if (mine) jacobi_block_full(A, U, ld, bi == bj, p, q, ci, si, hi, pc, qc, cj, sj, hj);
        __syncthreads();
      }
      if (!rotated) break;
    }
    return;
  }
  for (int sweep = 0; sweep < 30; ++sweep) {
    int rotated = 0;
    for (int step = 0; step < m - 1; ++step) {
      // phase 1: lane i < np computes the rotation of pair i
      int flag = 0;
      for (int i = threadIdx.x; i < np; i += EW) {
        int p, q;
        double c, sn;
        cplx<double> ph;
        if (jacobi_pair(A, n, ld, m, step, i, floor_abs, p, q, c, sn, ph)) flag = 1;
        js.pp[i] = p;
        js.pq[i] = q;
        js.pc[i] = c;
        js.ps[i] = sn;
        js.pph[i] = ph;
      }
      if (!__any(flag)) continue;  // uniform: no pair rotates in this step
      rotated = 1;
      __syncthreads();
      // phase 2: every 2x2 block (rows of pair bi, columns of pair bj) gets
      // A <- G_bi^H A G_bj, and U <- U G_bj on the same rows, where
      // G = [[c, s ph], [-s conj(ph), c]] on (p, q)
      for (int o = threadIdx.x; o < np * np; o += EW) {
        const int bi = o % np, bj = o / np;
        jacobi_block(A, U, n, ld, bi == bj, js.pp[bi], js.pq[bi], js.pc[bi], js.ps[bi], js.pph[bi], js.pp[bj],
                     js.pq[bj], js.pc[bj], js.ps[bj], js.pph[bj]);
      }
      __syncthreads();
    }
    if (!rotated) break;
  }
}

}  // namespace qtg
