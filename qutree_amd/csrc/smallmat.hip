// Batched FP64 small-matrix linear algebra of the reference path, exported
// through the C ABI (include/qutree_amd.h, "small matrices"):
//   heev      Hermitian eigendecomposition, ascending eigenvalues
//             (Matrix::cDiag, include/Core/Matrix_Implementation.h:284-308);
//   reginv    U diag(1 / (l + eps e^{-l/eps})) U^H (SpectralDecompositionTree
//             .cpp:30-48 with Vector_Implementation.h:266-280);
//   inverse   Gauss-Jordan with partial pivoting (Matrix::cInv,
//             Matrix_Implementation.h:311-328, an LU inverse in the reference);
//   qr        thin Q of a Householder QR (qr(Matrixcd),
//             Matrix_Implementation.h:795-806; the Householder convention of
//             LAPACK zgeqrf / zungqr, which Eigen's HouseholderQR shares).
// heev and reginv: one wavefront per matrix, the tournament Jacobi solver of
// jacobi.hpp in LDS.  inverse: one workgroup per matrix, [A | I] in LDS.
// qr: one workgroup per matrix working in place in the output (a node
// tensor's (rows x n) matrix does not fit LDS); wavefront 0 forms each
// reflector, every wavefront applies it to its own columns with lane-strided
// rows and shuffle reductions, so a column costs two barriers.
#include <algorithm>
#include <cstdint>
#include <string>

#include "common.hpp"
#include "jacobi.hpp"
#include "smallmat.hpp"

namespace qtg {

namespace {

typedef cplx<double> z;

// items[b], or for a uniform batch (items null) the template u with its
// offsets advanced by b times the strides in u.pad-free fields of st
__device__ inline SmItem item_of(const SmItem* __restrict__ items, const SmItem& u, const SmItem& st) {
  if (items) return items[blockIdx.x];
  SmItem it = u;
  const int64_t b = blockIdx.x;
  it.a += b * st.a;
  it.q += b * st.q;
  it.e += b * st.e;
  return it;
}

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(EW)
    heev_batch_kernel(const SmItem* __restrict__ items, SmItem u, SmItem st, const z* __restrict__ in, z* U,
                      double* ev, z* R, double eps, int nmax) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  z* A = reinterpret_cast<z*>(smem_raw);
  z* V = A + jacobi_elems(nmax);
  __shared__ JacobiScratch js;
  __shared__ int idx[64];
  __shared__ z ph[64];
  __shared__ double wgt[64];
  const SmItem it = item_of(items, u, st);
  const int n = it.n, ld = jacobi_ld(n);
  for (int o = threadIdx.x; o < n * n; o += EW) A[o % n + (o / n) * ld] = in[it.a + o];
  __syncthreads();
  jacobi_eig(A, V, n, ld, js);
  __syncthreads();
  if (threadIdx.x == 0) {
    // stable insertion sort: ascending, equal eigenvalues keep their order
    for (int j = 0; j < n; ++j) {
      int k = j;
      const double v = A[j + j * ld].re;
      while (k > 0 && A[idx[k - 1] + idx[k - 1] * ld].re > v) {
        idx[k] = idx[k - 1];
        --k;
      }
      idx[k] = j;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += EW) {
    // phase: the first component of every eigenvector real and >= 0
    const z u0 = V[0 + idx[j] * ld];
    const double a = sqrt(u0.re * u0.re + u0.im * u0.im);
    ph[j] = a > 1e-300 ? mk<double>(u0.re / a, -u0.im / a) : mk<double>(1.0, 0.0);
    const double l = A[idx[j] + idx[j] * ld].re;
    if (ev) ev[it.e + j] = l;
    wgt[j] = 1.0 / (l + eps * exp(-l / eps));
  }
  __syncthreads();
  if (U)
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int k = o % n, j = o / n;
      U[it.q + o] = cmul(V[k + idx[j] * ld], ph[j]);
    }
  if (R)
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int i = o % n, j = o / n;
      z acc = czero<double>();
      for (int k = 0; k < n; ++k) {
        const double w = wgt[k];
        const z u = V[i + idx[k] * ld], v = V[j + idx[k] * ld];
        acc.re += w * (u.re * v.re + u.im * v.im);
        acc.im += w * (u.im * v.re - u.re * v.im);
      }
      R[it.q + o] = acc;
    }
}

// Gauss-Jordan with partial pivoting on [A | I] in LDS: the pivot is the
// largest modulus in the column (lowest row on ties).
constexpr int kInvThreads = 256;
__global__ void __launch_bounds__(kInvThreads)
    inverse_batch_kernel(const SmItem* __restrict__ items, SmItem u, SmItem st, const z* __restrict__ in, z* out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const SmItem it = item_of(items, u, st);
  const int n = it.n;
  z* Sm = reinterpret_cast<z*>(smem_raw);
  z* Xm = Sm + n * n;
  __shared__ int piv;
  __shared__ z facs[64];
  for (int o = threadIdx.x; o < n * n; o += kInvThreads) {
    Sm[o] = in[it.a + o];
    Xm[o] = mk<double>((o % n) == (o / n) ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int col = 0; col < n; ++col) {
    if (threadIdx.x < 64) {
      double best = -1.0;
      int bi = col;
      for (int r = col + (int)threadIdx.x; r < n; r += 64) {
        const z v = Sm[r + col * n];
        const double a = v.re * v.re + v.im * v.im;
        if (a > best) {
          best = a;
          bi = r;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (threadIdx.x == 0) piv = bi;
    }
    __syncthreads();
    const int p = piv;
    if (p != col) {
      for (int k = threadIdx.x; k < 2 * n; k += kInvThreads) {
        z* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
        const z tmp = Mk[col];
        Mk[col] = Mk[p];
        Mk[p] = tmp;
      }
      __syncthreads();
    }
    const z dg = Sm[col + col * n];
    const double dd = dg.re * dg.re + dg.im * dg.im;
    const z dinv = mk<double>(dg.re / dd, -dg.im / dd);
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * n; k += kInvThreads) {
      z* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
      Mk[col] = cmul(Mk[col], dinv);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += kInvThreads) facs[r] = Sm[r + col * n];
    __syncthreads();
    for (int o = threadIdx.x; o < n * 2 * n; o += kInvThreads) {
      const int r = o % n, k = o / n;
      if (r == col) continue;
      z* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
      const z fr = facs[r], v = Mk[col];
      z w = Mk[r];
      w.re -= fr.re * v.re - fr.im * v.im;
      w.im -= fr.re * v.im + fr.im * v.re;
      Mk[r] = w;
    }
    __syncthreads();
  }
  for (int o = threadIdx.x; o < n * n; o += kInvThreads) out[it.q + o] = Xm[o];
}

constexpr int kQrThreads = 256;
constexpr int kQrWaves = kQrThreads / 64;

}  // namespace

// Householder QR of the m x n (m >= n, n <= 64) matrix at W (in place),
// then the thin Q in place.  Reflector j: v = [1; x / (alpha - beta)],
// beta = -sign(Re alpha) sqrt(|alpha|^2 + |x|^2) (minus for Re alpha = 0),
// tau = (beta - alpha) / beta; tau = 0 when x = 0 and alpha is real
// (zlarfg without its underflow rescaling).  Q = H_0 ... H_{n-1} [I; 0] by
// backward accumulation (zung2r).  Called by all kQrThreads threads.
namespace {
__device__ void qr_inplace(z* W, int64_t m, int n, z* tau, z* R) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = 0; j < n; ++j) {
    z* cj = W + (int64_t)j * m;
    if (w == 0) {
      double s = 0.0;
      for (int64_t r = j + 1 + lane; r < m; r += 64) s += cj[r].re * cj[r].re + cj[r].im * cj[r].im;
      s = wsum(s);
      const z alpha = cj[j];
      z t = czero<double>();
      if (!(s == 0.0 && alpha.im == 0.0)) {
        const double nr = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + s);
        const double beta = alpha.re >= 0.0 ? -nr : nr;
        t = mk<double>((beta - alpha.re) / beta, -alpha.im / beta);
        // 1 / (alpha - beta)
        const z dlt = mk<double>(alpha.re - beta, alpha.im);
        const double dd = dlt.re * dlt.re + dlt.im * dlt.im;
        const z sc = mk<double>(dlt.re / dd, -dlt.im / dd);
        for (int64_t r = j + 1 + lane; r < m; r += 64) cj[r] = cmul(cj[r], sc);
        if (lane == 0) cj[j] = mk<double>(beta, 0.0);
      }
      if (lane == 0) tau[j] = t;
    }
    __syncthreads();
    const z t = tau[j];
    if (t.re != 0.0 || t.im != 0.0) {
      // columns k > j: A(j:m, k) -= conj(tau) v (v^H A(j:m, k))
      for (int k = j + 1 + w; k < n; k += kQrWaves) {
        z* ck = W + (int64_t)k * m;
        double sr = 0.0, si = 0.0;
        for (int64_t r = j + lane; r < m; r += 64) {
          const z v = r == j ? mk<double>(1.0, 0.0) : cj[r];
          const z a = ck[r];
          sr += v.re * a.re + v.im * a.im;
          si += v.re * a.im - v.im * a.re;
        }
        sr = wsum(sr);
        si = wsum(si);
        // f = conj(tau) * s
        const z f = mk<double>(t.re * sr + t.im * si, t.re * si - t.im * sr);
        for (int64_t r = j + lane; r < m; r += 64) {
          const z v = r == j ? mk<double>(1.0, 0.0) : cj[r];
          z a = ck[r];
          a.re -= v.re * f.re - v.im * f.im;
          a.im -= v.re * f.im + v.im * f.re;
          ck[r] = a;
        }
      }
    }
    __syncthreads();
  }
  // R (n x n, upper triangular) when asked for: the diagonal and above of
  // the reduced matrix
  if (R) {
    for (int o = threadIdx.x; o < n * n; o += kQrThreads) {
      const int i = o % n, j = o / n;
      R[o] = i <= j ? W[i + (int64_t)j * m] : czero<double>();
    }
    __syncthreads();
  }
  // thin Q, backward
  for (int i = n - 1; i >= 0; --i) {
    z* ci = W + (int64_t)i * m;
    const z t = tau[i];
    if (i < n - 1) {
      // Q(i:m, k) -= tau v (v^H Q(i:m, k)) for k > i
      for (int k = i + 1 + w; k < n; k += kQrWaves) {
        z* ck = W + (int64_t)k * m;
        double sr = 0.0, si = 0.0;
        for (int64_t r = i + lane; r < m; r += 64) {
          const z v = r == i ? mk<double>(1.0, 0.0) : ci[r];
          const z a = ck[r];
          sr += v.re * a.re + v.im * a.im;
          si += v.re * a.im - v.im * a.re;
        }
        sr = wsum(sr);
        si = wsum(si);
        const z f = mk<double>(t.re * sr - t.im * si, t.re * si + t.im * sr);  // tau * s
        for (int64_t r = i + lane; r < m; r += 64) {
          const z v = r == i ? mk<double>(1.0, 0.0) : ci[r];
          z a = ck[r];
          a.re -= v.re * f.re - v.im * f.im;
          a.im -= v.re * f.im + v.im * f.re;
          ck[r] = a;
        }
      }
    }
    __syncthreads();
    // column i: -tau v below the diagonal, 1 - tau on it, zeros above
    const z mt = mk<double>(-t.re, -t.im);
    for (int64_t r = threadIdx.x; r < m; r += kQrThreads) {
      if (r < i) ci[r] = czero<double>();
      else if (r == i) ci[r] = mk<double>(1.0 - t.re, -t.im);
      else ci[r] = cmul(mt, ci[r]);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kQrThreads)
    qr_batch_kernel(const SmItem* __restrict__ items, SmItem u, SmItem st, const z* in, z* out, z* Rb) {
  __shared__ z tau[64];
  const SmItem it = item_of(items, u, st);
  const int64_t m = it.m;
  const int n = it.n;
  z* W = out + it.q;
  const z* A = in + it.a;
  if (A != W) {
    for (int64_t o = threadIdx.x; o < m * n; o += kQrThreads) W[o] = A[o];
    __syncthreads();
  }
  qr_inplace(W, m, n, tau, (Rb && it.e >= 0) ? Rb + it.e : nullptr);
}

}  // namespace

int launch_heev_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, int nmax,
                      const void* in, void* U, double* ev, void* R, double eps, hipStream_t s) {
  if (count <= 0) return QTG_OK;
  if (nmax < 1 || nmax > 64) {
    set_error("heev: order outside 1..64");
    return QTG_ESHAPE;
  }
  const size_t lds = 2 * (size_t)jacobi_elems(nmax) * sizeof(z);
  hipLaunchKernelGGL(heev_batch_kernel, dim3((unsigned)count), dim3(EW), lds, s, items, u, st, (const z*)in, (z*)U,
                     ev, (z*)R, eps, nmax);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_inverse_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, int nmax,
                         const void* in, void* out, hipStream_t s) {
  if (count <= 0) return QTG_OK;
  if (nmax < 1 || nmax > 64) {
    set_error("inverse: order outside 1..64");
    return QTG_ESHAPE;
  }
  const size_t lds = 2 * (size_t)nmax * nmax * sizeof(z);
  hipLaunchKernelGGL(inverse_batch_kernel, dim3((unsigned)count), dim3(kInvThreads), lds, s, items, u, st,
                     (const z*)in, (z*)out);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_qr_batch(const SmItem* items, const SmItem& u, const SmItem& st, int64_t count, const void* in, void* out,
                    hipStream_t s, void* R) {
  if (count <= 0) return QTG_OK;
  hipLaunchKernelGGL(qr_batch_kernel, dim3((unsigned)count), dim3(kQrThreads), 0, s, items, u, st, (const z*)in,
                     (z*)out, (z*)R);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

}  // namespace qtg

using namespace qtg;

namespace {
int sm_check(const void* A, const void* out, int64_t m, int64_t n, int64_t count, const char* what) {
  if (!A || !out || count < 0 || count > INT32_MAX) {
    set_error(std::string(what) + ": null pointer or bad count");
    return QTG_EARG;
  }
  if (n < 1 || n > 64 || m < n || m > ((int64_t)1 << 40) / n) {
    set_error(std::string(what) + ": need 1 <= n <= 64 and n <= m");
    return QTG_ESHAPE;
  }
  return QTG_OK;
}
SmItem uniform(int64_t m, int64_t n) {
  SmItem u{};
  u.m = m;
  u.n = (int32_t)n;
  return u;
}
SmItem strides(int64_t sa, int64_t sq, int64_t se) {
  SmItem st{};
  st.a = sa;
  st.q = sq;
  st.e = se;
  return st;
}
}  // namespace

extern "C" {

int qtg_heev_batch_z(const double* A, double* U, double* lambda, int64_t n, int64_t count, qtg_stream stream) {
  int rc = sm_check(A, U ? (const void*)U : (const void*)lambda, n, n, count, "heev");
  if (rc) return rc;
  return launch_heev_batch(nullptr, uniform(n, n), strides(n * n, n * n, n), count, (int)n, A, U, lambda, nullptr,
                           1.0, S(stream));
}

int qtg_reg_inverse_batch_z(const double* A, double* R, int64_t n, int64_t count, double eps, qtg_stream stream) {
  int rc = sm_check(A, R, n, n, count, "reg_inverse");
  if (rc) return rc;
  if (!(eps > 0.0)) {
    set_error("reg_inverse: eps must be positive");
    return QTG_EARG;
  }
  return launch_heev_batch(nullptr, uniform(n, n), strides(n * n, n * n, 0), count, (int)n, A, nullptr, nullptr, R,
                           eps, S(stream));
}

int qtg_inverse_batch_z(const double* A, double* Ainv, int64_t n, int64_t count, qtg_stream stream) {
  int rc = sm_check(A, Ainv, n, n, count, "inverse");
  if (rc) return rc;
  return launch_inverse_batch(nullptr, uniform(n, n), strides(n * n, n * n, 0), count, (int)n, A, Ainv,
                              S(stream));
}

int qtg_qr_batch_z(const double* A, double* Q, int64_t m, int64_t n, int64_t count, qtg_stream stream) {
  int rc = sm_check(A, Q, m, n, count, "qr");
  if (rc) return rc;
  SmItem u = uniform(m, n);
  u.e = -1;  // no R
  return launch_qr_batch(nullptr, u, strides(m * n, m * n, 0), count, A, Q, S(stream), nullptr);
}

}  // extern "C"
