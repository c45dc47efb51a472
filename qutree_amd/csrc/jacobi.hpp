// Wavefront Jacobi eigensolver for small Hermitian matrices (FP64, n <= 64)
// held in LDS, column-major with leading dimension ld.  All lanes of a
// 64-lane workgroup call it.  Tournament (round-robin) ordering: each step
// rotates up to 32 disjoint (p,q) pairs whose parameters come from the same
// matrix, columns first, then rows.  Player at position j of step t:
// j == 0 -> 0, else 1 + (j - 1 + t) mod (m - 1).  A pair is skipped when
// |a_pq| <= eps_mach sqrt(|a_pp a_qq|) (the relative-accuracy criterion of
// Demmel and Veselic); sweeps stop when no pair rotates or after 20 sweeps.
// On return A is diagonal (eigenvalues on the diagonal, unsorted) and
// U holds the eigenvectors as columns: A_in = U diag U^H.
#pragma once
#include "common.hpp"

namespace qtg {

constexpr int EW = 64;

struct JacobiScratch {
  double pc[32], ps[32];
  cplx<double> pph[32];
  int pp[32], pq[32];
};

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, EW);
  return v;
}

__device__ inline void jacobi_eig(cplx<double>* A, cplx<double>* U, int n, int ld, JacobiScratch& js) {
  for (int o = threadIdx.x; o < n * n; o += EW) {
    const int i = o % n, j = o / n;
    U[i + j * ld] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
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
  for (int sweep = 0; sweep < 20; ++sweep) {
    int rotated = 0;
    for (int step = 0; step < m - 1; ++step) {
      for (int i = threadIdx.x; i < np; i += EW) {
        const int j0 = i, j1 = m - 1 - i;
        int p = (j0 == 0) ? 0 : 1 + (j0 - 1 + step) % (m - 1);
        int q = (j1 == 0) ? 0 : 1 + (j1 - 1 + step) % (m - 1);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        js.pp[i] = p;
        js.pq[i] = q;
        js.ps[i] = 0.0;
        if (q < n) {
          const cplx<double> apq = A[p + q * ld];
          const double aa = sqrt(apq.re * apq.re + apq.im * apq.im);
          const double app = A[p + p * ld].re, aqq = A[q + q * ld].re;
          if (aa > 2.2e-16 * sqrt(fabs(app * aqq)) && aa > 1e-300) {
            const double theta = (aqq - app) / (2.0 * aa);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0);
            js.pc[i] = c;
            js.ps[i] = t * c;
            js.pph[i] = mk<double>(apq.re / aa, apq.im / aa);
          }
        }
      }
      __syncthreads();
      int any = 0;
      for (int i = 0; i < np; ++i) any |= (js.ps[i] != 0.0);
      if (any) {
        rotated = 1;
        // columns: A <- A G, U <- U G  (G = [[c, s ph], [-s conj(ph), c]] on (p,q))
        for (int o = threadIdx.x; o < np * n; o += EW) {
          const int i = o / n, r = o % n;
          const double s_ = js.ps[i];
          if (s_ == 0.0) continue;
          const int p = js.pp[i], q = js.pq[i];
          const double c = js.pc[i];
          const cplx<double> ph = js.pph[i];
          const cplx<double> arp = A[r + p * ld], arq = A[r + q * ld];
          A[r + p * ld] = mk<double>(c * arp.re - s_ * (ph.re * arq.re + ph.im * arq.im),
                                     c * arp.im - s_ * (ph.re * arq.im - ph.im * arq.re));
          A[r + q * ld] = mk<double>(s_ * (ph.re * arp.re - ph.im * arp.im) + c * arq.re,
                                     s_ * (ph.re * arp.im + ph.im * arp.re) + c * arq.im);
          const cplx<double> urp = U[r + p * ld], urq = U[r + q * ld];
          U[r + p * ld] = mk<double>(c * urp.re - s_ * (ph.re * urq.re + ph.im * urq.im),
                                     c * urp.im - s_ * (ph.re * urq.im - ph.im * urq.re));
          U[r + q * ld] = mk<double>(s_ * (ph.re * urp.re - ph.im * urp.im) + c * urq.re,
                                     s_ * (ph.re * urp.im + ph.im * urp.re) + c * urq.im);
        }
        __syncthreads();
        // rows: A <- G^H A  (new p = c*apk - s*ph*aqk ; new q = s*conj(ph)*apk + c*aqk)
        for (int o = threadIdx.x; o < np * n; o += EW) {
          const int i = o / n, k = o % n;
          const double s_ = js.ps[i];
          if (s_ == 0.0) continue;
          const int p = js.pp[i], q = js.pq[i];
          const double c = js.pc[i];
          const cplx<double> ph = js.pph[i];
          const cplx<double> apk = A[p + k * ld], aqk = A[q + k * ld];
          A[p + k * ld] = mk<double>(c * apk.re - s_ * (ph.re * aqk.re - ph.im * aqk.im),
                                     c * apk.im - s_ * (ph.re * aqk.im + ph.im * aqk.re));
          A[q + k * ld] = mk<double>(s_ * (ph.re * apk.re + ph.im * apk.im) + c * aqk.re,
                                     s_ * (ph.re * apk.im - ph.im * apk.re) + c * aqk.im);
        }
      }
      __syncthreads();
    }
    if (!rotated) break;
  }
}

}  // namespace qtg
