// Vector kernels of the Krylov-sweep SCF (scf_ops.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qtg {

constexpr int kMaxKrylov = 64;

// device-resident parameters of one SCF iteration
struct ScfParams {
  int32_t itp, pad;  // 1: imaginary-time update, 0: lowest Krylov eigenvector
  double beta;       // imaginary-time step
};

// Complex FP64 vectors of N entries; every pointer is device memory.
int launch_lanczos_begin(const void* psi, int64_t N, void* kry0, void* in, hipStream_t s);
// scal: alpha[0..k) then beta[0..k); T: k x k tridiagonal, written at j = k - 1
int launch_lanczos_step(int j, int k, int64_t N, const void* ha, int parts, void* kry, void* in, double* scal,
                        void* T, hipStream_t s);
int launch_krylov_combine(int k, int64_t N, const void* U, const double* ev, const void* kry, void* out,
                          const ScfParams* par, double* elog, int slot, hipStream_t s);
// out = in / |in|, nrm2[0] = |in|^2 (one workgroup)
int launch_normalize(const void* in, void* out, int64_t N, double* nrm2, hipStream_t s);
// dir 0: tensor (bef, n, aft) -> matrix (bef aft, n); dir 1: back
int launch_permute(const void* in, void* out, int64_t bef, int64_t n, int64_t aft, int dir, hipStream_t s);

}  // namespace qtg
