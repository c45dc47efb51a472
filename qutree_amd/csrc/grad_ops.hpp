// Kernels of the device-side energy-gradient path (grad_ops.hip).
#pragma once
#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "tree_ops.hpp"

namespace qtg {

// From e[0] = <A|HA>, e[1] = <A|A> of the top node: escal[0] = E,
// escal[1] = 1 / <A|A>, and terms[2 i + 1].cre = -E for i < npairs.
int launch_energy_scalars(const cplx<double>* e, double* escal, CombTerm* terms, int64_t npairs, hipStream_t s);
// x += a escal[1] y (n doubles); y null: x *= escal[1]
int launch_scaled_axpy(double* x, const double* y, int64_t n, double a, const double* escal, hipStream_t s);
int launch_copy_scalar(const double* src, double* dst, hipStream_t s);
// dst = src for n doubles (n even, 16-byte aligned pointers)
int launch_copy_vec(const double* src, double* dst, int64_t n, hipStream_t s);

}  // namespace qtg
