// Device records and launchers of the CMF integrator: per-node vector
// programs of the Bulirsch-Stoer state machines and the orthogonal() sweep.
#pragma once
#include "common.hpp"
#include "tree_ops.hpp"

namespace qtg {

// arenas addressed by buffer id (each holds a whole tree in node-address layout)
enum CmfBuf : int32_t {
  B_P = 0,    // state being integrated in this CMF substep
  B_Y0 = 1,   // yvec[0..4] of BS_integrator
  B_YT0 = 6,  // ytab[0..6]
  B_X = 13,   // derivative input
  B_D = 14,   // derivative output
  B_E1 = 15,  // error operands
  B_E2 = 16,
  kCmfBufs = 17
};
struct CmfBufs {
  cplx<double>* b[kCmfBufs];
};

enum VecCode : int32_t {
  V_COPY = 0,        // a <- b
  V_AXPY = 1,        // a <- b + h c
  V_ZERO = 2,        // a <- 0
  V_MMID_STEP = 3,   // t = y0 + 2h y2; y0 <- y1; y1 <- t; y2 <- t
  V_MMID_FINAL = 4,  // y0 <- (y0 + y1 + h y2) / 2
  V_RZEXTR = 5,      // rational extrapolation step i (rzextr)
  V_POST_EXTR = 6,   // y0 <- y1; y1 <- y1 + y2
  V_ERROR = 7,       // err[node] = LayerInterface::Error(a, b)
  // fused forms of the op sequences a tick emits most (one pass over the
  // node instead of one per op; the same arithmetic per element)
  V_START = 8,        // v = buffer a; P <- v when b; y0, y3, x <- v; ytab[1..6] <- 0
  V_MMID_BEGIN = 9,   // y0 <- y3; t = y3 + h y4; y1 <- t; x <- t
  V_MMID_STEP_X = 10, // V_MMID_STEP, then x <- y1
  V_FINAL_EXTR = 11   // y2 <- d; V_MMID_FINAL; V_RZEXTR (i, mi, fx); V_POST_EXTR
};

struct VecOp {
  int32_t code, a, b, c;
  int32_t i, mi, pad0, pad1;
  double h;
  double fx[7];
};

// A node's program: ops [op0, op1).  With kSelect in flags the node waits
// for the error of its last extrapolation: ops [op0, op1) when err[node] <
// eps (accept), else [alt0, alt1) (reject), and a pre program with kMask
// writes the node's byte of the Apply mask (kActAccept / kActReject).
struct NodeProg {
  int32_t node, op0, op1, flags;
  int32_t alt0, alt1;
  double eps;
};
enum NodeProgFlags : int32_t { kSelect = 1, kActAccept = 2, kActReject = 4, kMask = 8 };

struct OrthoItem {
  int32_t node, pad;
  int64_t sw;  // offset of S^{1/2 T} (n x n) in the square-root pool
};

struct SwApply {
  int32_t node, nf;
  int32_t fmode[4];
  int64_t fmat[4];
};

// Bulirsch-Stoer constants of BS_integrator (BS_integrator.h:34-37: maxuse_,
// max_, shrink_, grow_, and its step sequence), shared by the host and the
// device integrators.
constexpr int kBsMaxUse = 7, kBsMax = 11;
constexpr double kBsShrink = 0.95, kBsGrow = 1.2;
__host__ __device__ inline int bs_seq(int i) {
  constexpr int s[kBsMax] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
  return s[i];
}
enum BsState : int32_t { BS_NEED_START = 0, BS_START_MMID, BS_IN_MMID, BS_AWAIT_ERR, BS_DONE };

// One node's Bulirsch-Stoer integrator on the device: the state machine of
// the host integrator, its step size, the post ops its last derivative
// evaluation left for the next tick boundary, a failure flag (step size
// underflow; the node then stays done) and its derivative evaluations.
struct BsNode {
  int32_t st, i, step, nsub;
  double x, xend, h, hh, dt;
  double xtab[kBsMax];
  int32_t npost, fail;
  unsigned long long nderiv;
  VecOp post[2];
};
constexpr int kBsHeadWords = 20;  // 8-byte words of BsNode before post
// The per-tick record of the device integrators, a ring of kBsRing ticks of
// 2 * nodes bytes: the tick's Apply mask (1: the node takes a derivative),
// then per node the ticks from that one on in which it surely takes one
// (kBsFailed: its step size underflowed).
constexpr int kBsRing = 4;
constexpr uint8_t kBsFailed = 255;

// One tick boundary of every node's integrator (one workgroup per node): the
// post ops of the tick that ran (none when init), the accept test on the
// node error they computed, the next tick's pre op and the node's bytes of
// record entry rec (2 * count bytes).  init starts every node's interval
// [t0, t1] (the step size carries over).
int launch_bs_tick(BsNode* st, int64_t count, const CmfBufs& bufs, const DNode* nodes, const cplx<double>* pool,
                   const int64_t* rhoslot, double* err, uint8_t* rec, int init, double t0, double t1, double eps,
                   hipStream_t s);

// mask: the Apply mask that programs with kSelect | kMask write (may be null
// when no program has kMask)
int launch_programs(const NodeProg* progs, int64_t count, const VecOp* ops, const CmfBufs& bufs,
                    const DNode* nodes, const cplx<double>* pool, const int64_t* rhoslot, double* err,
                    hipStream_t s, uint8_t* mask = nullptr);
// complex elements of ortho_kernel's static Newton-Schulz buffers (five
// 16 x 16 matrices), to be counted with its dynamic LDS against the budget
constexpr int kOrthoNsElems = 5 * 256;
int launch_ortho(const OrthoItem* items, int64_t count, const DNode* nodes, cplx<double>* P, cplx<double>* sw,
                 int nmax, int nspf, hipStream_t s);
// The same for node tensors of any size: Phi in place in global memory.
int launch_ortho_global(const OrthoItem* items, int64_t count, const DNode* nodes, cplx<double>* P,
                        cplx<double>* sw, int nspf, hipStream_t s);
// Hermitian eigenproblem of one n x n matrix (n <= 64): ascending
// eigenvalues into ev, the eigenvectors transposed into UT (may be null).
int launch_eig_sorted(const cplx<double>* A, int n, cplx<double>* UT, double* ev, hipStream_t s);
int launch_apply_sw(const SwApply* items, int64_t count, const DNode* nodes, cplx<double>* P,
                    const cplx<double>* sw, int nmax, hipStream_t s);

}  // namespace qtg
