// Shared device/host helpers of the gfx950 engine: complex arithmetic on
// interleaved (re, im) pairs, status handling and launch checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/qutree_amd.h"

namespace qtg {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> mk(T r, T i) {
  cplx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__host__ __device__ inline cplx<T> czero() {
  return mk<T>(T(0), T(0));
}
// acc += a * b
template <typename T>
__device__ inline void cfma(cplx<T>& acc, const cplx<T>& a, const cplx<T>& b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
}
// acc += conj(a) * b
template <typename T>
__device__ inline void cfmac(cplx<T>& acc, const cplx<T>& a, const cplx<T>& b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(-a.im, b.re, acc.im);
}
template <typename T>
__host__ __device__ inline cplx<T> cmul(const cplx<T>& a, const cplx<T>& b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(const cplx<T>& a, const cplx<T>& b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ inline cplx<T> cconj(const cplx<T>& a) {
  return mk<T>(a.re, -a.im);
}

// per-thread last error of the C ABI
void set_error(const std::string& msg);
const char* get_error();

#define QTG_HIP_TRY(expr)                                                                  \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      ::qtg::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                 \
      return QTG_EDEVICE;                                                                  \
    }                                                                                      \
  } while (0)

#define QTG_LAUNCH_CHECK()                                                                 \
  do {                                                                                     \
    hipError_t _e = hipGetLastError();                                                     \
    if (_e != hipSuccess) {                                                                \
      ::qtg::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));            \
      return QTG_EDEVICE;                                                                  \
    }                                                                                      \
  } while (0)

inline hipStream_t S(qtg_stream s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace qtg
