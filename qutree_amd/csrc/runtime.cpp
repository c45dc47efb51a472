// Runtime part of the C ABI: device selection, memory, streams, timers and
// the per-thread error string.
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "grad_ops.hpp"

namespace qtg {
static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* get_error() { return g_error.c_str(); }
static thread_local hipEvent_t g_t0 = nullptr, g_t1 = nullptr;
}  // namespace qtg

using namespace qtg;

extern "C" {

const char* qtg_last_error(void) { return get_error(); }

int qtg_init(int device) {
  int n = 0;
  QTG_HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("qtg_init: no such device");
    return QTG_EDEVICE;
  }
  QTG_HIP_TRY(hipSetDevice(device));
  // host waits spin instead of yielding: the CMF integrator synchronises
  // once per lockstep tick (~100 us apart), where a yield costs tens of us
  (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
  (void)hipGetLastError();
  return QTG_OK;
}

int qtg_device_alloc(void** ptr, size_t bytes) {
  if (!ptr) return QTG_EARG;
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    set_error("hipMalloc failed");
    return QTG_ENOMEM;
  }
  return QTG_OK;
}

int qtg_device_free(void* ptr) {
  QTG_HIP_TRY(hipFree(ptr));
  return QTG_OK;
}

int qtg_copy_d(double* dst, const double* src, int64_t n, qtg_stream stream) {
  if (!dst || !src || n < 0) return QTG_EARG;
  return launch_copy_vec(src, dst, n, S(stream));
}

int qtg_memcpy_h2d(void* dst, const void* src, size_t bytes, qtg_stream stream) {
  QTG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
  QTG_HIP_TRY(hipStreamSynchronize(S(stream)));
  return QTG_OK;
}

int qtg_memcpy_d2h(void* dst, const void* src, size_t bytes, qtg_stream stream) {
  QTG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
  QTG_HIP_TRY(hipStreamSynchronize(S(stream)));
  return QTG_OK;
}

int qtg_memset(void* dst, int value, size_t bytes, qtg_stream stream) {
  QTG_HIP_TRY(hipMemsetAsync(dst, value, bytes, S(stream)));
  return QTG_OK;
}

int qtg_stream_create(qtg_stream* stream) {
  hipStream_t s;
  QTG_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return QTG_OK;
}

int qtg_stream_destroy(qtg_stream stream) {
  QTG_HIP_TRY(hipStreamDestroy(S(stream)));
  return QTG_OK;
}

int qtg_synchronize(qtg_stream stream) {
  // NULL: wait for all work on the device, including engine-owned streams
  if (stream) QTG_HIP_TRY(hipStreamSynchronize(S(stream)));
  else QTG_HIP_TRY(hipDeviceSynchronize());
  return QTG_OK;
}

int qtg_timer_start(qtg_stream stream) {
  if (!g_t0) {
    QTG_HIP_TRY(hipEventCreate(&g_t0));
    QTG_HIP_TRY(hipEventCreate(&g_t1));
  }
  QTG_HIP_TRY(hipEventRecord(g_t0, S(stream)));
  return QTG_OK;
}

int qtg_timer_stop(qtg_stream stream, float* ms) {
  if (!g_t0) {
    set_error("timer not started");
    return QTG_ESTATE;
  }
  QTG_HIP_TRY(hipEventRecord(g_t1, S(stream)));
  QTG_HIP_TRY(hipEventSynchronize(g_t1));
  QTG_HIP_TRY(hipEventElapsedTime(ms, g_t0, g_t1));
  return QTG_OK;
}

}  // extern "C"
