// Energy functional E(A) = <Psi|H|Psi> / <Psi|Psi> on the device as a forward
// / backward pair, and gradient descent with the QR gauge, with no host round
// trip inside a step.
//
// forward:  the sweep (Derivative at phase 1: density tree, rho^-1,
//           h-matrices, holes, Apply at every node, the top node's <A|HA> and
//           <A|A>), then E and 1/<A|A> computed on the device, with -E written
//           into the gradient's combination terms.
// backward: dE/dA_v* = (Apply(v) + M_v x_L A_v - E rho_v x_L A_v) / <A|A>
//           (engine_gemm.cpp, gradient section): the mean fields M_v of the
//           terms inactive at v top-down, then the assembly; scaled by
//           1/<A|A> on the device.
// No reference counterpart; rho^-1 (1 - P) of the gradient is the
// reference's imaginary-time LayerDerivative (HamiltonianRepresentation.cpp:
// 45-72 with phase -i, Eigenstates.cpp:75).
#include "engine.hpp"
#include "grad_ops.hpp"
#include "smallmat.hpp"

namespace qtg {

int Engine::grad_ready() {
  if (c64_ || path_ != 1) {
    set_error("engine: the energy gradient needs a complex128 engine on the grouped path");
    return QTG_ESTATE;
  }
  if (!planned_) {
    set_error("engine: no Hamiltonian");
    return QTG_ESTATE;
  }
  if (tree_.nodes.back().n != 1) {
    set_error("engine: the energy gradient takes one state at the top node");
    return QTG_ESHAPE;
  }
  int rc;
  if (!d_grad64_ && (rc = alloc((void**)&d_grad64_, sizeof(double) * 2 * tree_.total))) return rc;
  if (!d_escal_ && (rc = alloc((void**)&d_escal_, sizeof(double) * 2))) return rc;
  return QTG_OK;
}

int Engine::forward(hipStream_t s, double* energy_host) {
  int rc = grad_ready();
  if (rc) return rc;
  if ((rc = derivative(s, 1.0, 0.0))) return rc;
  const auto* e = reinterpret_cast<const cplx<double>*>(d_pool64_) + energySlot_;
  if ((rc = launch_energy_scalars(e, d_escal_, d_combtGE_, (int64_t)h_combtGE_.size() / 2, s))) return rc;
  if (energy_host) {
    QTG_HIP_TRY(hipMemcpyAsync(energy_host, d_escal_, sizeof(double), hipMemcpyDeviceToHost, s));
    QTG_HIP_TRY(hipStreamSynchronize(s));
  }
  return QTG_OK;
}

// dE/dA* without the 1/<A|A> factor, into d_grad64_
int Engine::backward_unscaled(hipStream_t s) {
  int rc;
  if (gradRecompute_ && (rc = forward(s, nullptr))) return rc;
  void* bases[B_COUNT] = {d_psi32_, d_work32_, d_pool_, d_psi64_, d_apply64_, d_grad64_, d_pool64_, d_work_};
  if ((rc = launch_combine<double>(d_combG1_, (int64_t)h_combG1_.size(), d_combtG1_, d_pool_, s))) return rc;
  if ((rc = launch_combine<double>(d_combG2_, (int64_t)h_combG2_.size(), d_combtG2_, d_pool_, s))) return rc;
  if ((rc = launch_combine<double>(d_combG3_, (int64_t)h_combG3_.size(), d_combtG3_, d_pool_, s))) return rc;
  if ((rc = run_launches(gGradPre_, bases, nullptr, 1.0, 0.0, s))) return rc;
  for (size_t d = 1; d < gGradDepth_.size(); ++d)
    if ((rc = run_launches(gGradDepth_[d], bases, nullptr, 1.0, 0.0, s))) return rc;
  if ((rc = launch_combine<double>(d_combGE_, (int64_t)h_combGE_.size(), d_combtGE_, d_pool_, s))) return rc;
  return run_launches(gGradFin_, bases, nullptr, 1.0, 0.0, s);
}

int Engine::backward(hipStream_t s, double* grad_host) {
  int rc = grad_ready();
  if (rc) return rc;
  if ((rc = backward_unscaled(s))) return rc;
  if ((rc = launch_scaled_axpy(d_grad64_, nullptr, 2 * tree_.total, 1.0, d_escal_, s))) return rc;
  if (grad_host) {
    QTG_HIP_TRY(hipMemcpyAsync(grad_host, d_grad64_, sizeof(double) * 2 * tree_.total, hipMemcpyDeviceToHost, s));
    QTG_HIP_TRY(hipStreamSynchronize(s));
  }
  return QTG_OK;
}

int Engine::gradient(hipStream_t s, double* grad_host, double* energy) {
  int rc = forward(s, energy);
  if (rc) return rc;
  const bool keep = gradRecompute_;
  gradRecompute_ = false;  // the forward just ran
  rc = backward(s, grad_host);
  gradRecompute_ = keep;
  return rc;
}

int Engine::canonicalize(hipStream_t s) {
  void* bases[B_COUNT] = {d_psi32_, d_work32_, d_pool_, d_psi64_, d_apply64_, d_dpsi64_, d_pool64_, d_work_};
  int rc;
  for (size_t d = 0; d < canonRange_.size(); ++d) {
    const auto& r = canonRange_[d];
    if (r.second > r.first &&
        (rc = launch_qr_batch(d_canon_ + r.first, SmItem{}, SmItem{}, r.second - r.first, d_psi64_, d_psi64_, s,
                              d_pool_)))
      return rc;
    if ((rc = run_launches(gCanon_[d], bases, nullptr, 1.0, 0.0, s))) return rc;
  }
  built_ = false;
  return QTG_OK;
}

int Engine::gradient_descent(hipStream_t s, int64_t steps, double eta, double* energies_host, double* psi_host) {
  if (steps < 0) {
    set_error("gradient descent: negative step count");
    return QTG_EARG;
  }
  int rc = grad_ready();
  if (rc) return rc;
  if (elogCap_ < steps + 1) {
    if ((rc = alloc((void**)&d_elog_, sizeof(double) * (steps + 1)))) return rc;
    elogCap_ = steps + 1;
  }
  const bool keep = gradRecompute_;
  gradRecompute_ = false;  // every backward follows its forward directly
  for (int64_t i = 0; i < steps && !rc; ++i) {
    if ((rc = forward(s, nullptr))) break;
    if ((rc = launch_copy_scalar(d_escal_, d_elog_ + i, s))) break;
    if ((rc = backward_unscaled(s))) break;
    // A <- A - eta dE/dA* = A - (eta / <A|A>) (unscaled gradient)
    if ((rc = launch_scaled_axpy(d_psi64_, d_grad64_, 2 * tree_.total, -eta, d_escal_, s))) break;
    rc = canonicalize(s);
  }
  gradRecompute_ = keep;
  if (rc) return rc;
  if ((rc = forward(s, nullptr))) return rc;
  if ((rc = launch_copy_scalar(d_escal_, d_elog_ + steps, s))) return rc;
  if (energies_host)
    QTG_HIP_TRY(hipMemcpyAsync(energies_host, d_elog_, sizeof(double) * (steps + 1), hipMemcpyDeviceToHost, s));
  if (psi_host)
    QTG_HIP_TRY(hipMemcpyAsync(psi_host, d_psi64_, sizeof(double) * 2 * tree_.total, hipMemcpyDeviceToHost, s));
  QTG_HIP_TRY(hipStreamSynchronize(s));
  return QTG_OK;
}

}  // namespace qtg
