// Wavefront Jacobi eigensolver for small Hermitian matrices (FP64, n <= 64)
// held in LDS, column-major with leading dimension ld.  All lanes of a
// 64-lane workgroup call it.  Tournament (round-robin) ordering: each step
// rotates up to 32 disjoint (p,q) pairs whose parameters come from the same
// matrix, columns first, then rows; the lanes of one pair compute its
// rotation in registers and a wavefront vote replaces any shared flag.
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

struct JacobiScratch {};  // rotation parameters live in registers

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
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = fmax(scale, fabs(A[i + i * ld].re));
  const double floor_abs = 1e-24 * scale;
  // lane -> (pair i, row group g): the G lanes of a pair compute its
  // rotation redundantly in registers and share the row / column updates
  const int G = EW / np;
  const int pi = threadIdx.x % np, g = threadIdx.x / np;
  const bool lane_on = g < G;
  for (int sweep = 0; sweep < 30; ++sweep) {
    int rotated = 0;
    for (int step = 0; step < m - 1; ++step) {
      const int j0 = pi, j1 = m - 1 - pi;
      int p = (j0 == 0) ? 0 : 1 + (j0 - 1 + step) % (m - 1);
      int q = (j1 == 0) ? 0 : 1 + (j1 - 1 + step) % (m - 1);
      if (p > q) {
        const int t = p;
        p = q;
        q = t;
      }
      double c = 1.0, sn = 0.0;
      cplx<double> ph = mk<double>(1.0, 0.0);
      if (lane_on && q < n) {
        const cplx<double> apq = A[p + q * ld];
        const double aa = sqrt(apq.re * apq.re + apq.im * apq.im);
        const double app = A[p + p * ld].re, aqq = A[q + q * ld].re;
        if (aa > 2.2e-16 * sqrt(fabs(app * aqq)) && aa > floor_abs && aa > 1e-300) {
          const double theta = (aqq - app) / (2.0 * aa);
          const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(t * t + 1.0);
          sn = t * c;
          ph = mk<double>(apq.re / aa, apq.im / aa);
        }
      }
      // the parameter reads above precede every write below in this
      // wavefront's program order; one barrier keeps that explicit
      __syncthreads();
      const bool act = lane_on && sn != 0.0;
      if (__any(act)) {
        rotated = 1;
        // columns: A <- A G, U <- U G  (G = [[c, s ph], [-s conj(ph), c]] on (p,q))
        if (act)
          for (int r = g; r < n; r += G) {
            const cplx<double> arp = A[r + p * ld], arq = A[r + q * ld];
            A[r + p * ld] = mk<double>(c * arp.re - sn * (ph.re * arq.re + ph.im * arq.im),
                                       c * arp.im - sn * (ph.re * arq.im - ph.im * arq.re));
            A[r + q * ld] = mk<double>(sn * (ph.re * arp.re - ph.im * arp.im) + c * arq.re,
                                       sn * (ph.re * arp.im + ph.im * arp.re) + c * arq.im);
            const cplx<double> urp = U[r + p * ld], urq = U[r + q * ld];
            U[r + p * ld] = mk<double>(c * urp.re - sn * (ph.re * urq.re + ph.im * urq.im),
                                       c * urp.im - sn * (ph.re * urq.im - ph.im * urq.re));
            U[r + q * ld] = mk<double>(sn * (ph.re * urp.re - ph.im * urp.im) + c * urq.re,
                                       sn * (ph.re * urp.im + ph.im * urp.re) + c * urq.im);
          }
        __syncthreads();
        // rows: A <- G^H A  (new p = c*apk - s*ph*aqk ; new q = s*conj(ph)*apk + c*aqk)
        if (act)
          for (int k = g; k < n; k += G) {
            const cplx<double> apk = A[p + k * ld], aqk = A[q + k * ld];
            cplx<double> np_ = mk<double>(c * apk.re - sn * (ph.re * aqk.re - ph.im * aqk.im),
                                          c * apk.im - sn * (ph.re * aqk.im + ph.im * aqk.re));
            cplx<double> nq_ = mk<double>(sn * (ph.re * apk.re + ph.im * apk.im) + c * aqk.re,
                                          sn * (ph.re * apk.im - ph.im * apk.re) + c * aqk.im);
            // the rotated pair: exact zeros off the diagonal, real diagonal
            if (k == q) {
              np_ = czero<double>();
              nq_.im = 0.0;
            }
            if (k == p) {
              nq_ = czero<double>();
              np_.im = 0.0;
            }
            A[p + k * ld] = np_;
            A[q + k * ld] = nq_;
          }
        __syncthreads();
      }
    }
    if (!rotated) break;
  }
}

}  // namespace qtg
