// Small kernels of the device-side energy-gradient path (gradient.cpp): the
// energy scalars taken from the top node's <A|HA>, <A|A> without a host
// round trip, the gradient step and its scaling.
#include "grad_ops.hpp"

namespace qtg {

namespace {

// escal[0] = E = Re(<A|HA> / <A|A>), escal[1] = 1 / Re<A|A>; every thread
// computes them from the two pool entries and writes -E into the second
// term of every (M_v, -E rho_v) combination pair.
__global__ void energy_scalars_kernel(const cplx<double>* __restrict__ e, double* escal, CombTerm* terms,
                                      int64_t npairs) {
  const cplx<double> num = e[0], den = e[1];
  const double dd = den.re * den.re + den.im * den.im;
  const double E = (num.re * den.re + num.im * den.im) / dd;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    escal[0] = E;
    escal[1] = 1.0 / den.re;
  }
  if (i < npairs) terms[2 * i + 1].cre = -E;
}

// x[i] += a * escal[1] * y[i] over n doubles (a = -eta: a gradient step on
// the unscaled gradient), or x[i] *= escal[1] when y is null
__global__ void scaled_axpy_kernel(double* x, const double* __restrict__ y, int64_t n, double a,
                                   const double* __restrict__ escal) {
  const double s = escal[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = y ? x[i] + a * s * y[i] : x[i] * s;
}

__global__ void copy_scalar_kernel(const double* __restrict__ src, double* dst) {
  if (threadIdx.x == 0) *dst = *src;
}

// dst = src, n2 double2 pairs (16-byte loads and stores), grid-stride
__global__ void __launch_bounds__(256) copy_vec_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                       int64_t n2) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

}  // namespace

int launch_energy_scalars(const cplx<double>* e, double* escal, CombTerm* terms, int64_t npairs, hipStream_t s) {
  const unsigned blocks = (unsigned)std::max<int64_t>(1, (npairs + 255) / 256);
  hipLaunchKernelGGL(energy_scalars_kernel, dim3(blocks), dim3(256), 0, s, e, escal, terms, npairs);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_scaled_axpy(double* x, const double* y, int64_t n, double a, const double* escal, hipStream_t s) {
  if (n <= 0) return QTG_OK;
  const unsigned blocks = (unsigned)std::min<int64_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(scaled_axpy_kernel, dim3(blocks), dim3(256), 0, s, x, y, n, a, escal);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_copy_vec(const double* src, double* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return QTG_OK;
  if ((n & 1) || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15)) {
    set_error("copy: n must be even and both pointers 16-byte aligned");
    return QTG_EARG;
  }
  const int64_t n2 = n / 2;
  const int64_t blocks = std::min<int64_t>((n2 + 255) / 256, 8192);
  hipLaunchKernelGGL(copy_vec_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const double2*)src, (double2*)dst, n2);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_copy_scalar(const double* src, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(copy_scalar_kernel, dim3(1), dim3(64), 0, s, src, dst);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

}  // namespace qtg
