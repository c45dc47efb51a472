// Vector kernels of the Krylov-sweep SCF (SCF.cpp:134-235): the Lanczos
// recurrence, the tridiagonal matrix, the Krylov combination of the updated
// node tensor, and the mode-to-matrix permutation of a QR toward an inner
// leg.  A node tensor of a tree has at most a few hundred thousand entries,
// so each kernel runs as one workgroup that sweeps the vector in a
// grid-stride loop and reduces in LDS; the contractions around them run in
// the grouped MFMA kernels.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "scf_ops.hpp"

namespace qtg {

namespace {

constexpr int kTB = 1024;
using z = cplx<double>;

// sum of v over the workgroup (every thread gets the result)
__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x < 64) {
    s = threadIdx.x < (kTB / 64) ? red[threadIdx.x] : 0.0;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (threadIdx.x == 0) red[kTB / 64] = s;
  }
  __syncthreads();
  return red[kTB / 64];
}

// v0 = psi / |psi| into the Krylov slot 0 and the apply input slot
__global__ void __launch_bounds__(kTB)
    lanczos_begin_kernel(const z* __restrict__ psi, int64_t N, z* __restrict__ kry0, z* __restrict__ in) {
  __shared__ double red[kTB / 64 + 1];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += kTB) s += psi[i].re * psi[i].re + psi[i].im * psi[i].im;
  const double nrm = sqrt(block_sum(s, red));
  for (int64_t i = threadIdx.x; i < N; i += kTB) {
    const z x = mk<double>(psi[i].re / nrm, psi[i].im / nrm);
    kry0[i] = x;
    in[i] = x;
  }
}

// Step j of the recurrence: alpha_j = Re <v_j|H v_j>, w = H v_j - alpha_j v_j
// - beta_{j-1} v_{j-1}; below the last step beta_j = |w| and v_{j+1} = w /
// beta_j, at the last step the k x k tridiagonal matrix is written to T.
// H v_j is the sum of `parts` partial products `ha`, N apart.
__global__ void __launch_bounds__(kTB)
    lanczos_step_kernel(int j, int k, int64_t N, const z* __restrict__ ha, int parts, z* kry, z* __restrict__ in,
                        double* __restrict__ scal, z* __restrict__ T) {
  __shared__ double red[kTB / 64 + 1];
  const z* v = kry + (int64_t)j * N;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += kTB) {
    double hr = 0.0, hi = 0.0;
    for (int p = 0; p < parts; ++p) {
      hr += ha[i + p * N].re;
      hi += ha[i + p * N].im;
    }
    s += v[i].re * hr + v[i].im * hi;
  }
  const double alpha = block_sum(s, red);
  const double bprev = j > 0 ? scal[k + j - 1] : 0.0;
  const z* vp = kry + (int64_t)(j > 0 ? j - 1 : 0) * N;
  z* w = kry + (int64_t)(j + 1 < k ? j + 1 : j) * N;  // the last step's w is not kept
  double s2 = 0.0;
  if (j + 1 < k) {
    for (int64_t i = threadIdx.x; i < N; i += kTB) {
      double hr = 0.0, hi = 0.0;
      for (int p = 0; p < parts; ++p) {
        hr += ha[i + p * N].re;
        hi += ha[i + p * N].im;
      }
      double wr = hr - alpha * v[i].re, wi = hi - alpha * v[i].im;
      if (j > 0) {
        wr -= bprev * vp[i].re;
        wi -= bprev * vp[i].im;
      }
      w[i] = mk<double>(wr, wi);
      s2 += wr * wr + wi * wi;
    }
  }
  if (j + 1 < k) {
    const double beta = sqrt(block_sum(s2, red));
    for (int64_t i = threadIdx.x; i < N; i += kTB) {
      const z x = mk<double>(w[i].re / beta, w[i].im / beta);
      w[i] = x;
      in[i] = x;
    }
    if (threadIdx.x == 0) {
      scal[j] = alpha;
      scal[k + j] = beta;
    }
  } else {
    __syncthreads();
    if (threadIdx.x == 0) scal[j] = alpha;
    __syncthreads();
    for (int e = threadIdx.x; e < k * k; e += kTB) {
      const int r = e % k, c = e / k;
      double x = 0.0;
      if (r == c) x = (r == j) ? alpha : scal[r];
      else if (r == c + 1) x = scal[k + c];
      else if (c == r + 1) x = scal[k + r];
      T[e] = mk<double>(x, 0.0);
    }
  }
}

// out = sum_l c_l v_l with c from the eigenvectors U (k x k, column-major,
// eigenvalues ev ascending): the lowest eigenvector (diagonalizeKrylov) or,
// when par->itp is set, the imaginary-time weights of imaginaryTimePropagation
// followed by a normalisation.  elog[slot] receives ev[0].
__global__ void __launch_bounds__(kTB)
    krylov_combine_kernel(int k, int64_t N, const z* __restrict__ U, const double* __restrict__ ev,
                          const z* __restrict__ kry, z* __restrict__ out, const ScfParams* __restrict__ par,
                          double* __restrict__ elog, int slot) {
  __shared__ double red[kTB / 64 + 1];
  __shared__ z c[kMaxKrylov];
  const int itp = par->itp;
  if (threadIdx.x < k) {
    const int l = threadIdx.x;
    if (!itp) {
      c[l] = U[l];
    } else {
      const double beta = par->beta;
      // U(l,0) conj(U(0,0)) + sum_j U(l,j) e^{-beta (E_j - E_0)} conj(U(0,j))
      z acc = cmul(U[l], cconj(U[0]));
      for (int jj = 0; jj < k; ++jj) {
        const double e = exp(-beta * (ev[jj] - ev[0]));
        const z t = cmul(U[l + (int64_t)k * jj], cconj(U[(int64_t)k * jj]));
        acc.re += e * t.re;
        acc.im += e * t.im;
      }
      c[l] = acc;
    }
  }
  if (threadIdx.x == 0) elog[slot] = ev[0];
  __syncthreads();
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += kTB) {
    double xr = 0.0, xi = 0.0;
    for (int l = 0; l < k; ++l) {
      const z vv = kry[i + (int64_t)l * N];
      xr += c[l].re * vv.re - c[l].im * vv.im;
      xi += c[l].re * vv.im + c[l].im * vv.re;
    }
    out[i] = mk<double>(xr, xi);
    s += xr * xr + xi * xi;
  }
  if (itp) {
    const double nrm = sqrt(block_sum(s, red));
    for (int64_t i = threadIdx.x; i < N; i += kTB) out[i] = mk<double>(out[i].re / nrm, out[i].im / nrm);
  }
}

// dir 0: M(b + f bef, a) = A(b, a, f) (toMatrix(A, mode)); dir 1: the inverse
__global__ void permute_kernel(const z* __restrict__ in, z* __restrict__ out, int64_t bef, int64_t n, int64_t aft,
                               int dir) {
  const int64_t N = bef * n * aft;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x) {
    // t indexes the tensor (b fastest, then a, then f)
    const int64_t b = t % bef, a = (t / bef) % n, f = t / (bef * n);
    const int64_t m = b + f * bef + a * (bef * aft);
    if (dir == 0) out[m] = in[t];
    else out[t] = in[m];
  }
}

// out = in / |in|, nrm2[0] = |in|^2 (gramSchmidt of one column,
// Tensor_Implementation.h:806-853)
__global__ void __launch_bounds__(kTB)
    normalize_kernel(const z* __restrict__ in, z* __restrict__ out, int64_t N, double* __restrict__ nrm2) {
  __shared__ double red[kTB / 64 + 1];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += kTB) s += in[i].re * in[i].re + in[i].im * in[i].im;
  const double t = block_sum(s, red);
  const double nrm = sqrt(t);
  for (int64_t i = threadIdx.x; i < N; i += kTB) out[i] = mk<double>(in[i].re / nrm, in[i].im / nrm);
  if (threadIdx.x == 0) nrm2[0] = t;
}

}  // namespace

int launch_normalize(const void* in, void* out, int64_t N, double* nrm2, hipStream_t s) {
  hipLaunchKernelGGL(normalize_kernel, dim3(1), dim3(kTB), 0, s, (const z*)in, (z*)out, N, nrm2);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_lanczos_begin(const void* psi, int64_t N, void* kry0, void* in, hipStream_t s) {
  hipLaunchKernelGGL(lanczos_begin_kernel, dim3(1), dim3(kTB), 0, s, (const z*)psi, N, (z*)kry0, (z*)in);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_lanczos_step(int j, int k, int64_t N, const void* ha, int parts, void* kry, void* in, double* scal,
                        void* T, hipStream_t s) {
  if (k < 1 || k > kMaxKrylov || j < 0 || j >= k) {
    set_error("lanczos: Krylov size outside 1..64");
    return QTG_EARG;
  }
  hipLaunchKernelGGL(lanczos_step_kernel, dim3(1), dim3(kTB), 0, s, j, k, N, (const z*)ha, parts, (z*)kry, (z*)in,
                     scal, (z*)T);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_krylov_combine(int k, int64_t N, const void* U, const double* ev, const void* kry, void* out,
                          const ScfParams* par, double* elog, int slot, hipStream_t s) {
  if (k < 1 || k > kMaxKrylov) {
    set_error("krylov combine: Krylov size outside 1..64");
    return QTG_EARG;
  }
  hipLaunchKernelGGL(krylov_combine_kernel, dim3(1), dim3(kTB), 0, s, k, N, (const z*)U, ev, (const z*)kry, (z*)out,
                     par, elog, slot);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_permute(const void* in, void* out, int64_t bef, int64_t n, int64_t aft, int dir, hipStream_t s) {
  const int64_t N = bef * n * aft;
  const int tb = 256;
  const int64_t blocks = std::min<int64_t>((N + tb - 1) / tb, 1024);
  hipLaunchKernelGGL(permute_kernel, dim3((unsigned)blocks), dim3(tb), 0, s, (const z*)in, (z*)out, bef, n, aft, dir);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

}  // namespace qtg
