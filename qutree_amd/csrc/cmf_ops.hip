// Kernels of the CMF integrator (CMFIntegrator.cpp / BS_integrator.h of
// QuTree): the vector arithmetic of every node's Bulirsch-Stoer state machine,
// the rho-weighted node error (LayerInterface::Error) and orthogonal()
// (TensorTree_Implementation.h:143-171).  One workgroup per node.
#include "cmf_ops.hpp"

#include <cstddef>
#include "jacobi.hpp"
#include "node_tile.hpp"

namespace qtg {

using cd2 = cplx<double>;

__device__ inline cd2 cdiv(const cd2& a, const cd2& b) {
  const double d = b.re * b.re + b.im * b.im;
  return mk<double>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

__device__ inline double block_sum_d(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = TB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// b.b[i] for a wave-uniform i.  Indexing the by-value kernel argument with a
// runtime index would copy the whole pointer table to scratch memory; a
// switch reads each entry from the argument segment directly.
__device__ inline cd2* bp(const CmfBufs& b, int i) {
  switch (i) {
#define QTG_BP(k) \
  case k:         \
    return b.b[k];
    QTG_BP(0) QTG_BP(1) QTG_BP(2) QTG_BP(3) QTG_BP(4) QTG_BP(5) QTG_BP(6) QTG_BP(7) QTG_BP(8)
    QTG_BP(9) QTG_BP(10) QTG_BP(11) QTG_BP(12) QTG_BP(13) QTG_BP(14) QTG_BP(15) QTG_BP(16)
#undef QTG_BP
    default:
      return nullptr;
  }
}
static_assert(kCmfBufs == 17, "bp() covers every buffer");

// Elementwise ops: Op::load(e) reads every operand of element e, and
// Op::store(e, v) writes its results.  The buffers of one op are different
// arenas, but the compiler cannot prove it, so a store would order every later
// load behind it: a pass loads up to P elements per thread (the tail
// predicated) before storing any, so a node of up to P * TB elements takes
// one memory round trip per op.  P shrinks with the operands per element
// (16 for one complex value, 8 for three) to stay in registers.
template <class Op>
__device__ __forceinline__ void elementwise(int N, const Op& op) {
  constexpr int P = sizeof(typename Op::Val) <= 16 ? 16 : sizeof(typename Op::Val) <= 32 ? 12 : 8;
  for (int e0 = threadIdx.x; e0 < N; e0 += P * TB) {
    typename Op::Val v[P];
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (e0 + u * TB < N) v[u] = op.load(e0 + u * TB);
#pragma unroll
    for (int u = 0; u < P; ++u)
      if (e0 + u * TB < N) op.store(e0 + u * TB, v[u]);
  }
}

struct CopyOp {  // a <- b
  cd2* A;
  const cd2* B;
  struct Val {
    cd2 b;
  };
  __device__ Val load(int e) const { return {B[e]}; }
  __device__ void store(int e, const Val& v) const { A[e] = v.b; }
};
struct AxpyOp {  // a <- b + h c
  cd2* A;
  const cd2 *B, *C;
  double h;
  struct Val {
    cd2 b, c;
  };
  __device__ Val load(int e) const { return {B[e], C[e]}; }
  __device__ void store(int e, const Val& v) const {
    A[e] = mk<double>(v.b.re + h * v.c.re, v.b.im + h * v.c.im);
  }
};
struct MmidStepOp {  // t = y0 + 2h y2; y0 <- y1; y1 <- t; y2 <- t
  cd2 *y0, *y1, *y2;
  double h;
  struct Val {
    cd2 a, b, c;
  };
  __device__ Val load(int e) const { return {y0[e], y1[e], y2[e]}; }
  __device__ void store(int e, const Val& v) const {
    const cd2 t = mk<double>(v.a.re + 2.0 * h * v.c.re, v.a.im + 2.0 * h * v.c.im);
    y0[e] = v.b;
    y1[e] = t;
    y2[e] = t;
  }
};
struct MmidFinalOp {  // y0 <- (y0 + y1 + h y2) / 2
  cd2* y0;
  const cd2 *y1, *y2;
  double h;
  struct Val {
    cd2 a, b, c;
  };
  __device__ Val load(int e) const { return {y0[e], y1[e], y2[e]}; }
  __device__ void store(int e, const Val& v) const {
    y0[e] = mk<double>((v.a.re + v.b.re + h * v.c.re) / 2.0, (v.a.im + v.b.im + h * v.c.im) / 2.0);
  }
};
struct StartOp {  // v = src; P <- v (when wp); y0, y3, x <- v; ytab[1..6] <- 0
  const cd2* src;
  cd2 *P, *y0, *y3, *x;
  cd2* yt[6];
  bool wp;
  struct Val {
    cd2 v;
  };
  __device__ Val load(int e) const { return {src[e]}; }
  __device__ void store(int e, const Val& v) const {
    if (wp) P[e] = v.v;
    y0[e] = v.v;
    y3[e] = v.v;
    x[e] = v.v;
#pragma unroll
    for (int k = 0; k < 6; ++k) yt[k][e] = czero<double>();
  }
};
struct MmidBeginOp {  // y0 <- y3; t = y3 + h y4; y1 <- t; x <- t
  cd2 *y0, *y1, *x;
  const cd2 *y3, *y4;
  double h;
  struct Val {
    cd2 b, c;
  };
  __device__ Val load(int e) const { return {y3[e], y4[e]}; }
  __device__ void store(int e, const Val& v) const {
    y0[e] = v.b;
    const cd2 t = mk<double>(v.b.re + h * v.c.re, v.b.im + h * v.c.im);
    y1[e] = t;
    x[e] = t;
  }
};
struct MmidStepXOp {  // t = y0 + 2h y2; y0 <- y1; y1 <- t; y2 <- t; x <- t
  cd2 *y0, *y1, *y2, *x;
  double h;
  struct Val {
    cd2 a, b, c;
  };
  __device__ Val load(int e) const { return {y0[e], y1[e], y2[e]}; }
  __device__ void store(int e, const Val& v) const {
    const cd2 t = mk<double>(v.a.re + 2.0 * h * v.c.re, v.a.im + 2.0 * h * v.c.im);
    y0[e] = v.b;
    y1[e] = t;
    y2[e] = t;
    x[e] = t;
  }
};
struct PostExtrOp {  // y0 <- y1; y1 <- y1 + y2
  cd2 *y0, *y1;
  const cd2* y2;
  struct Val {
    cd2 a, c;
  };
  __device__ Val load(int e) const { return {y1[e], y2[e]}; }
  __device__ void store(int e, const Val& v) const {
    y0[e] = v.a;
    y1[e] = mk<double>(v.a.re + v.c.re, v.a.im + v.c.im);
  }
};

// One vector op on one node (the whole workgroup; the caller synchronises
// the workgroup before the next op).
__device__ __forceinline__ void exec_op(const VecOp* opp, const CmfBufs& bufs, const DNode* __restrict__ nodes,
                                        const cd2* __restrict__ pool, const int64_t* __restrict__ rhoslot,
                                        double* err, int node, const DNode& nd, double* red, cd2* rhos) {
  const int N = nd.N;
  const int64_t off = nd.off;
  const VecOp op = *opp;
  switch (op.code) {
    case V_COPY:
      elementwise(N, CopyOp{bp(bufs, op.a) + off, bp(bufs, op.b) + off});
      break;
    case V_AXPY:
      elementwise(N, AxpyOp{bp(bufs, op.a) + off, bp(bufs, op.b) + off, bp(bufs, op.c) + off, op.h});
      break;
    case V_ZERO: {
      cd2* A = bp(bufs, op.a) + off;
      for (int e = threadIdx.x; e < N; e += TB) A[e] = czero<double>();
      break;
    }
    case V_MMID_STEP:
      elementwise(N, MmidStepOp{bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 1) + off, bp(bufs, B_Y0 + 2) + off, op.h});
      break;
    case V_MMID_FINAL:
      elementwise(N, MmidFinalOp{bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 1) + off, bp(bufs, B_Y0 + 2) + off, op.h});
      break;
    case V_RZEXTR: {
      // BS_integrator::rzextr with yest = y0, ysav = y1, dy = y2
      cd2* yest = bp(bufs, B_Y0) + off;
      cd2* ysav = bp(bufs, B_Y0 + 1) + off;
      cd2* dy = bp(bufs, B_Y0 + 2) + off;
      cd2* yt0 = bp(bufs, B_YT0) + off;
      constexpr int kTab = 7;  // ytab[0..6]
      const int mi = op.mi < kTab ? op.mi : kTab;
      double fx[kTab];
#pragma unroll
      for (int k2 = 1; k2 < kTab; ++k2) fx[k2] = opp->fx[k2];
      cd2* yt[kTab];
#pragma unroll
      for (int k2 = 0; k2 < kTab; ++k2) yt[k2] = bp(bufs, B_YT0 + k2) + off;
      if (op.i == 0) {
        for (int e = threadIdx.x; e < N; e += TB) {
          const cd2 v = yest[e];
          ysav[e] = v;
          dy[e] = v;
          yt0[e] = v;
        }
        break;
      }
      // two elements per pass: both tableau columns are loaded before any
      // store, and the two extrapolation chains are independent
      for (int e0 = threadIdx.x; e0 < N; e0 += 2 * TB) {
        const int e1 = e0 + TB;
        const bool has1 = e1 < N;
        const int ee[2] = {e0, has1 ? e1 : e0};
        cd2 col[2][kTab], yy[2], v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int k2 = 1; k2 < kTab; ++k2) col[h][k2] = yt[k2 < mi ? k2 : 0][ee[h]];
          yy[h] = yest[ee[h]];
          v[h] = yt0[ee[h]];
        }
        const cd2 yst[2] = {yy[0], yy[1]};
        cd2 ddy[2] = {czero<double>(), czero<double>()}, c[2] = {yy[0], yy[1]};
#pragma unroll
        for (int k2 = 1; k2 < kTab; ++k2) {
          if (k2 < mi) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const cd2 bi = mk<double>(fx[k2] * v[h].re, fx[k2] * v[h].im);
              cd2 b = mk<double>(bi.re - c[h].re, bi.im - c[h].im);
              if (b.re != 0.0 || b.im != 0.0) {
                b = cdiv(mk<double>(c[h].re - v[h].re, c[h].im - v[h].im), b);
                ddy[h] = cmul(c[h], b);
                c[h] = cmul(bi, b);
              } else {
                ddy[h] = v[h];
              }
              v[h] = col[h][k2];
              col[h][k2] = ddy[h];
              yy[h] = mk<double>(yy[h].re + ddy[h].re, yy[h].im + ddy[h].im);
            }
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (h == 0 || has1) {
            const int e = ee[h];
            yt0[e] = yst[h];
#pragma unroll
            for (int k2 = 1; k2 < kTab; ++k2)
              if (k2 < mi) yt[k2][e] = col[h][k2];
            dy[e] = ddy[h];
            ysav[e] = yy[h];
          }
        }
      }
      break;
    }
    case V_POST_EXTR:
      elementwise(N, PostExtrOp{bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 1) + off, bp(bufs, B_Y0 + 2) + off});
      break;
    case V_START: {
      StartOp so{bp(bufs, op.a) + off, bp(bufs, B_P) + off, bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 3) + off,
                  bp(bufs, B_X) + off, {}, op.b != 0};
#pragma unroll
      for (int k2 = 0; k2 < 6; ++k2) so.yt[k2] = bp(bufs, B_YT0 + 1 + k2) + off;
      elementwise(N, so);
      break;
    }
    case V_MMID_BEGIN:
      elementwise(N, MmidBeginOp{bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 1) + off, bp(bufs, B_X) + off,
                                 bp(bufs, B_Y0 + 3) + off, bp(bufs, B_Y0 + 4) + off, op.h});
      break;
    case V_MMID_STEP_X:
      elementwise(N, MmidStepXOp{bp(bufs, B_Y0) + off, bp(bufs, B_Y0 + 1) + off, bp(bufs, B_Y0 + 2) + off,
                                 bp(bufs, B_X) + off, op.h});
      break;
    case V_FINAL_EXTR: {
      // y2 <- d; y0 <- (y0 + y1 + h y2) / 2 (the midpoint estimate yest);
      // rzextr at step i; then y0 <- ysav, y1 <- ysav + dy
      cd2* y0 = bp(bufs, B_Y0) + off;
      cd2* y1 = bp(bufs, B_Y0 + 1) + off;
      cd2* y2 = bp(bufs, B_Y0 + 2) + off;
      const cd2* dd = bp(bufs, B_D) + off;
      constexpr int kTab = 7;  // ytab[0..6]
      const int mi = op.mi < kTab ? op.mi : kTab;
      const double h = op.h;
      double fx[kTab];
#pragma unroll
      for (int k2 = 1; k2 < kTab; ++k2) fx[k2] = opp->fx[k2];
      cd2* yt[kTab];
#pragma unroll
      for (int k2 = 0; k2 < kTab; ++k2) yt[k2] = bp(bufs, B_YT0 + k2) + off;
      for (int e = threadIdx.x; e < N; e += TB) {
        const cd2 d = dd[e], a = y0[e], b = y1[e];
        cd2 col[kTab], v = czero<double>();
        if (op.i > 0) {
#pragma unroll
          for (int k2 = 1; k2 < kTab; ++k2) col[k2] = yt[k2 < mi ? k2 : 0][e];
          v = yt[0][e];
        }
        const cd2 yest = mk<double>((a.re + b.re + h * d.re) / 2.0, (a.im + b.im + h * d.im) / 2.0);
        cd2 ysav = yest, dy = yest;
        if (op.i > 0) {
          cd2 ddy = czero<double>(), c = yest, yy = yest;
#pragma unroll
          for (int k2 = 1; k2 < kTab; ++k2) {
            if (k2 < mi) {
              const cd2 bi = mk<double>(fx[k2] * v.re, fx[k2] * v.im);
              cd2 bb = mk<double>(bi.re - c.re, bi.im - c.im);
              if (bb.re != 0.0 || bb.im != 0.0) {
                bb = cdiv(mk<double>(c.re - v.re, c.im - v.im), bb);
                ddy = cmul(c, bb);
                c = cmul(bi, bb);
              } else {
                ddy = v;
              }
              v = col[k2];
              col[k2] = ddy;
              yy = mk<double>(yy.re + ddy.re, yy.im + ddy.im);
            }
          }
#pragma unroll
          for (int k2 = 1; k2 < kTab; ++k2)
            if (k2 < mi) yt[k2][e] = col[k2];
          dy = ddy;
          ysav = yy;
        }
        yt[0][e] = yest;
        y2[e] = dy;
        y0[e] = ysav;
        y1[e] = mk<double>(ysav.re + dy.re, ysav.im + dy.im);
      }
      break;
    }
    case V_ERROR: {
      const cd2* Phi = bp(bufs, op.a) + off;
      const cd2* Chi = bp(bufs, op.b) + off;
      const int n = nodes[node].dims[nd.order - 1], L = N / n;
      const int64_t rs = rhoslot[node];
      double delta = 0.0, nrm = 0.0;
      if (rs < 0) {
        for (int e = threadIdx.x; e < N; e += TB) {
          const double dr = Phi[e].re - Chi[e].re, di = Phi[e].im - Chi[e].im;
          delta += dr * dr + di * di;
          nrm += Phi[e].re * Phi[e].re + Phi[e].im * Phi[e].im;
        }
        delta = block_sum_d(delta, red);
        nrm = block_sum_d(nrm, red);
      } else {
        const cd2* rho = pool + rs;
        // Delta = sum_{i,n,m} Re( conj(d(i,n)) rho(n,m) d(i,m) )
        if (n <= 16) {
          // a thread per row i: d(i, :) in registers, rho in LDS
          for (int o = threadIdx.x; o < n * n; o += TB) rhos[o] = rho[o];
          __syncthreads();
          for (int i = threadIdx.x; i < L; i += TB) {
            cd2 d[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) {
              const int mm = m < n ? m : 0;
              const cd2 p = Phi[i + mm * L], q = Chi[i + mm * L];
              d[m] = mk<double>(p.re - q.re, p.im - q.im);
            }
#pragma unroll
            for (int nn = 0; nn < 16; ++nn) {
              if (nn < n) {
                double tr = 0.0, ti = 0.0;
#pragma unroll
                for (int m = 0; m < 16; ++m) {
                  if (m < n) {
                    const cd2 r = rhos[nn + m * n];
                    tr += r.re * d[m].re - r.im * d[m].im;
                    ti += r.re * d[m].im + r.im * d[m].re;
                  }
                }
                delta += d[nn].re * tr + d[nn].im * ti;
              }
            }
          }
        } else {
          for (int e = threadIdx.x; e < N; e += TB) {
            const int i = e % L, nn = e / L;
            const double dr = Phi[e].re - Chi[e].re, di = Phi[e].im - Chi[e].im;
            double tr = 0.0, ti = 0.0;
            for (int m = 0; m < n; ++m) {
              const cd2 r = rho[nn + m * n];
              const double er = Phi[i + m * L].re - Chi[i + m * L].re;
              const double ei = Phi[i + m * L].im - Chi[i + m * L].im;
              tr += r.re * er - r.im * ei;
              ti += r.re * ei + r.im * er;
            }
            delta += dr * tr + di * ti;
          }
        }
        delta = block_sum_d(delta, red);
        double t = 0.0, tim = 0.0;
        for (int m = 0; m < n; ++m) {
          t += rho[m + m * n].re;
          tim += rho[m + m * n].im;
        }
        nrm = sqrt(t * t + tim * tim);
      }
      if (threadIdx.x == 0) err[node] = sqrt(fabs(delta / nrm));
      break;
    }
    default:
      break;
  }
}

__global__ void __launch_bounds__(TB)
    program_kernel(const NodeProg* __restrict__ progs, const VecOp* __restrict__ ops, CmfBufs bufs,
                   const DNode* __restrict__ nodes, const cd2* __restrict__ pool, const int64_t* __restrict__ rhoslot,
                   double* err, uint8_t* mask) {
  __shared__ double red[TB];
  __shared__ cd2 rhos[16 * 16];
  const NodeProg pg = progs[blockIdx.x];
  int op0 = pg.op0, op1 = pg.op1;
  bool accept = true;
  if (pg.flags & kSelect) {
    // the accept test of the Bulirsch-Stoer step on the error that the
    // previous tick wrote (the same comparison the host makes)
    accept = err[pg.node] < pg.eps;
    if (!accept) {
      op0 = pg.alt0;
      op1 = pg.alt1;
    }
  }
  if ((pg.flags & kMask) && threadIdx.x == 0)
    mask[pg.node] = (pg.flags & (accept ? kActAccept : kActReject)) ? 1 : 0;
  if (op0 >= op1) return;  // an empty program (the whole workgroup)
  const DNode nd = nodes[pg.node];
  for (int k = op0; k < op1; ++k) {
    exec_op(&ops[k], bufs, nodes, pool, rhoslot, err, pg.node, nd, red, rhos);
    __syncthreads();
  }
}

// The transposed square root SW(j,k) = S^(1/2)(k,j) of a near-identity
// S = Phi^H Phi (n <= 16, rho = n max|S - I| < 1/2: the node tensors after
// a CMF step are orthonormal up to the integrator's error) by the coupled
// Newton-Schulz iteration T = (3I - Z Y) / 2, Y <- Y T, Z <- T Z from
// Y = S, Z = I (Y -> S^(1/2), Z -> S^(-1/2), quadratically: the error is
// below 1e-17 after the steps chosen).  It is the principal square root the
// eigendecomposition gives, to rounding.  One wavefront; returns false (and
// writes nothing) when S is not near enough the identity.
__device__ bool ortho_sqrt_ns(const cd2* A, int ld, int n, cd2* sw) {
  __shared__ cd2 nsb[5][256];
  static_assert(sizeof(nsb) == kOrthoNsElems * sizeof(cd2), "Newton-Schulz buffers");
  if (n > 16) return false;
  const int nn = n * n;
  double e = 0.0;
  for (int o = threadIdx.x; o < nn; o += EW) {
    const int i = o % n, j = o / n;
    const cd2 v = A[i + j * ld];
    e = fmax(e, fabs((i == j ? 1.0 : 0.0) - v.re) + fabs(v.im));
  }
  for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_xor(e, off, EW));
  const double rho = n * e;
  if (!(rho < 0.5)) return false;
  int nsk = 1;  // one step beyond the estimate
  for (double r = rho; r > 1e-17 && nsk < 9; r *= r) ++nsk;
  cd2 *Y = nsb[0], *Z = nsb[1], *T = nsb[2], *Yn = nsb[3], *Zn = nsb[4];
  for (int o = threadIdx.x; o < nn; o += EW) {
    const int i = o % n, j = o / n;
    Y[o] = A[i + j * ld];
    Z[o] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int k = 0; k < nsk; ++k) {
    for (int o = threadIdx.x; o < nn; o += EW) {
      const int i = o % n, j = o / n;
      cd2 acc = czero<double>();
      for (int l = 0; l < n; ++l) cfma(acc, Z[i + l * n], Y[l + j * n]);
      T[o] = mk<double>(0.5 * ((i == j ? 3.0 : 0.0) - acc.re), -0.5 * acc.im);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < nn; o += EW) {
      const int i = o % n, j = o / n;
      cd2 y = czero<double>(), z = czero<double>();
      for (int l = 0; l < n; ++l) {
        cfma(y, Y[i + l * n], T[l + j * n]);
        cfma(z, T[i + l * n], Z[l + j * n]);
      }
      Yn[o] = y;
      Zn[o] = z;
    }
    __syncthreads();
    cd2* t = Y;
    Y = Yn;
    Yn = t;
    t = Z;
    Z = Zn;
    Zn = t;
  }
  for (int o = threadIdx.x; o < nn; o += EW) {
    const int j = o % n, k = o / n;
    sw[j + k * n] = Y[k + j * n];
  }
  return true;
}

// ------------------------------------------------------------- orthogonal()
// S = Phi^H Phi, SW(j,k) = sum_l U(k,l) sqrt(max(l_l,0)) conj(U(j,l)) (the
// transposed square root), then iterated Gram-Schmidt of Phi in place
// (Tensor_Implementation.h:806-853).  One wavefront per node.  Items with
// sw < 0 take the Gram-Schmidt only (orthonormal(),
// TensorTree_Implementation.h:190-195).
__global__ void __launch_bounds__(EW)
    ortho_kernel(const OrthoItem* __restrict__ items, const DNode* __restrict__ nodes, cd2* P, cd2* sw,
                 int nmax, int nspf) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cd2* Phi = reinterpret_cast<cd2*>(smem_raw);
  cd2* A = Phi + nmax;
  cd2* U = A + jacobi_elems(nspf);
  __shared__ JacobiScratch js;
  __shared__ double lam[64];
  const OrthoItem it = items[blockIdx.x];
  const DNode nd = nodes[it.node];
  const int N = nd.N, n = nodes[it.node].dims[nd.order - 1], L = N / n, ld = jacobi_ld(n);
  stage_to_lds<EW>(Phi, P + nd.off, N);
  __syncthreads();
  if (it.sw >= 0) {  // uniform per workgroup
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int j = o % n, k = o / n;
      cd2 acc = czero<double>();
      for (int i = 0; i < L; ++i) cfmac(acc, Phi[i + j * L], Phi[i + k * L]);
      A[j + k * ld] = acc;
    }
    __syncthreads();
    if (ortho_sqrt_ns(A, ld, n, sw + it.sw)) goto gram_schmidt;
    jacobi_eig(A, U, n, ld, js);
    for (int i = threadIdx.x; i < n; i += EW) lam[i] = sqrt(fmax(A[i + i * ld].re, 0.0));
    __syncthreads();
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int j = o % n, k = o / n;
      cd2 acc = czero<double>();
      for (int l = 0; l < n; ++l) {
        const cd2 ukl = U[k + l * ld], ujl = U[j + l * ld];
        // U(k,l) * s_l * conj(U(j,l))
        acc.re += lam[l] * (ukl.re * ujl.re + ukl.im * ujl.im);
        acc.im += lam[l] * (ukl.im * ujl.re - ukl.re * ujl.im);
      }
      sw[it.sw + j + k * n] = acc;
    }
  }
gram_schmidt:
  // Gram-Schmidt: lane owns rows i = lane + 64 t, so every reduction is a
  // wavefront shuffle sum and no other lane touches its elements.
  for (int k = 0; k < n; ++k) {
    double accum = 1.0;
    for (int iter = 0; iter < 15 && accum > 1e-12; ++iter) {
      accum = 0.0;
      for (int m = 0; m < k; ++m) {
        double ore = 0.0, oim = 0.0;
        for (int i = threadIdx.x; i < L; i += EW) {
          const cd2 a = Phi[i + m * L], b = Phi[i + k * L];
          ore += a.re * b.re + a.im * b.im;
          oim += a.re * b.im - a.im * b.re;
        }
        ore = wave_sum(ore);
        oim = wave_sum(oim);
        accum += sqrt(ore * ore + oim * oim);
        for (int i = threadIdx.x; i < L; i += EW) {
          const cd2 a = Phi[i + m * L];
          cd2& b = Phi[i + k * L];
          b.re -= ore * a.re - oim * a.im;
          b.im -= ore * a.im + oim * a.re;
        }
      }
      double nr = 0.0;
      for (int i = threadIdx.x; i < L; i += EW) nr += Phi[i + k * L].re * Phi[i + k * L].re + Phi[i + k * L].im * Phi[i + k * L].im;
      nr = wave_sum(nr);
      if (nr != 0.0) {
        const double s = sqrt(nr);
        for (int i = threadIdx.x; i < L; i += EW) {
          Phi[i + k * L].re /= s;
          Phi[i + k * L].im /= s;
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < N; e += EW) P[nd.off + e] = Phi[e];
}

// orthogonal() for node tensors too large for LDS: the algorithm of
// ortho_kernel with Phi read and written in place in global memory (only the
// n x n Jacobi buffers in LDS).  One wavefront per node; a lane owns rows
// i = lane + 64 t, so the Gram-Schmidt reductions are shuffle sums.
__global__ void __launch_bounds__(EW)
    ortho_global_kernel(const OrthoItem* __restrict__ items, const DNode* __restrict__ nodes, cd2* P, cd2* sw,
                        int nspf) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cd2* A = reinterpret_cast<cd2*>(smem_raw);
  cd2* U = A + jacobi_elems(nspf);
  __shared__ JacobiScratch js;
  __shared__ double lam[64];
  const OrthoItem it = items[blockIdx.x];
  const DNode nd = nodes[it.node];
  const int N = nd.N, n = nd.dims[nd.order - 1], L = N / n, ld = jacobi_ld(n);
  cd2* Phi = P + nd.off;
  if (it.sw >= 0) {
    // S(j,k) = sum_i conj(Phi(i,j)) Phi(i,k): lanes split the rows, every
    // entry is a shuffle sum (coalesced column reads)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        for (int i = threadIdx.x; i < L; i += EW) {
          const cd2 a = Phi[i + j * L], b = Phi[i + k * L];
          re += a.re * b.re + a.im * b.im;
          im += a.re * b.im - a.im * b.re;
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if (threadIdx.x == 0) A[j + k * ld] = mk<double>(re, im);
      }
    __syncthreads();
    jacobi_eig(A, U, n, ld, js);
    for (int i = threadIdx.x; i < n; i += EW) lam[i] = sqrt(fmax(A[i + i * ld].re, 0.0));
    __syncthreads();
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int j = o % n, k = o / n;
      cd2 acc = czero<double>();
      for (int l = 0; l < n; ++l) {
        const cd2 ukl = U[k + l * ld], ujl = U[j + l * ld];
        acc.re += lam[l] * (ukl.re * ujl.re + ukl.im * ujl.im);
        acc.im += lam[l] * (ukl.im * ujl.re - ukl.re * ujl.im);
      }
      sw[it.sw + j + k * n] = acc;
    }
  }
  for (int k = 0; k < n; ++k) {
    double accum = 1.0;
    for (int iter = 0; iter < 15 && accum > 1e-12; ++iter) {
      accum = 0.0;
      for (int m = 0; m < k; ++m) {
        double ore = 0.0, oim = 0.0;
        for (int i = threadIdx.x; i < L; i += EW) {
          const cd2 a = Phi[i + m * L], b = Phi[i + k * L];
          ore += a.re * b.re + a.im * b.im;
          oim += a.re * b.im - a.im * b.re;
        }
        ore = wave_sum(ore);
        oim = wave_sum(oim);
        accum += sqrt(ore * ore + oim * oim);
        for (int i = threadIdx.x; i < L; i += EW) {
          const cd2 a = Phi[i + m * L];
          cd2 b = Phi[i + k * L];
          b.re -= ore * a.re - oim * a.im;
          b.im -= ore * a.im + oim * a.re;
          Phi[i + k * L] = b;
        }
      }
      double nr = 0.0;
      for (int i = threadIdx.x; i < L; i += EW) nr += Phi[i + k * L].re * Phi[i + k * L].re + Phi[i + k * L].im * Phi[i + k * L].im;
      nr = wave_sum(nr);
      if (nr != 0.0) {
        const double s = sqrt(nr);
        for (int i = threadIdx.x; i < L; i += EW) {
          cd2 b = Phi[i + k * L];
          b.re /= s;
          b.im /= s;
          Phi[i + k * L] = b;
        }
      }
    }
  }
}

// ------------------------------------------------------ top eigenproblems
// One wavefront: the Hermitian n x n matrix A (column-major) is diagonalised
// (jacobi.hpp); eigenvalues ascending into ev (stable order), and the
// eigenvector columns V(:,j) with the phase that makes V(0,j) real and
// non-negative are written transposed, UT(j,k) = V(k,j), ready for a mode
// product on the state index (multStateArTB, Tensor_Implementation.h:763-792:
// C(:,j) = sum_k V(k,j) B(:,k)).  UT may be null.
__global__ void __launch_bounds__(EW)
    eig_sorted_kernel(const cd2* __restrict__ Ain, int n, cd2* UT, double* ev) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cd2* A = reinterpret_cast<cd2*>(smem_raw);
  cd2* U = A + jacobi_elems(n);
  __shared__ JacobiScratch js;
  __shared__ int idx[64];
  __shared__ cd2 ph[64];
  const int ld = jacobi_ld(n);
  for (int o = threadIdx.x; o < n * n; o += EW) A[o % n + (o / n) * ld] = Ain[o];
  __syncthreads();
  jacobi_eig(A, U, n, ld, js);
  if (threadIdx.x == 0) {
    // stable insertion sort of the eigenvalues
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
    const cd2 u0 = U[0 + idx[j] * ld];
    const double a = sqrt(u0.re * u0.re + u0.im * u0.im);
    cd2 p = mk<double>(1.0, 0.0);
    if (a > 1e-300) p = mk<double>(u0.re / a, -u0.im / a);
    // real(V(0,j)) < 0 flips the column
    if (u0.re * p.re - u0.im * p.im < 0.0) p = mk<double>(-p.re, -p.im);
    ph[j] = p;
    ev[j] = A[idx[j] + idx[j] * ld].re;
  }
  __syncthreads();
  if (UT)
    for (int o = threadIdx.x; o < n * n; o += EW) {
      const int k = o % n, j = o / n;  // V(k,j)
      UT[j + k * n] = cmul(U[k + idx[j] * ld], ph[j]);
    }
}

// parent <- SW_c x_{childIdx(c)} parent for its children, in node-address order
__global__ void __launch_bounds__(TB)
    apply_sw_kernel(const SwApply* __restrict__ items, const DNode* __restrict__ nodes, cd2* P, const cd2* sw,
                    int nmax) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cd2* b0 = reinterpret_cast<cd2*>(smem_raw);
  cd2* b1 = b0 + nmax;
  const SwApply it = items[blockIdx.x];
  const DNode nd = nodes[it.node];
  const int N = nd.N;
  stage_to_lds<TB>(b0, P + nd.off, N);
  __syncthreads();
  cd2* cur = b0;
  cd2* nxt = b1;
  for (int f = 0; f < it.nf; ++f) {
    mode_apply<double>(nxt, cur, sw + it.fmat[f], nd, it.fmode[f]);
    __syncthreads();
    cd2* t = cur;
    cur = nxt;
    nxt = t;
  }
  for (int e = threadIdx.x; e < N; e += TB) P[nd.off + e] = cur[e];
}

// ------------------------------------------------- device BS integrators
namespace {
__device__ inline void bs_op(VecOp* o, int code, int a = 0, int b = 0, int c = 0, double h = 0.0) {
  o->code = code;
  o->a = a;
  o->b = b;
  o->c = c;
  o->i = o->mi = o->pad0 = o->pad1 = 0;
  o->h = h;
#pragma unroll
  for (int k = 0; k < 7; ++k) o->fx[k] = 0.0;
}

// The scalar state of one node's integrator, in registers.
struct BsRegs {
  int st, i, step, nsub;
  double x, xend, h, hh, dt;
};

// The host integrator's advance() for one node (BS_integrator::Integrate,
// bsstep and mmid, BS_integrator.h:113-256), emitting the fused op forms
// the host's peephole makes of the same sequences, without copies between
// aliased buffers.  Returns whether the node takes a derivative in the
// tick; pre: the tick's one pre op (npre 0 or 1); post: the ops after the
// derivative (npost); xtab: the node's extrapolation abscissae (xin: a copy
// of them as the tick found them, which is what is read).  lead:
// ticks from this one on in which the node surely takes a derivative.
__device__ bool bs_advance(BsRegs& b, const double* xin, double* xtab, double e, double eps, const CmfBufs& bufs,
                           VecOp* pre, int& npre, VecOp* post, int& npost, bool& fail, int& lead) {
#pragma clang fp contract(off)
  npre = npost = 0;
  lead = 0;
  bool acceptCopy = false;  // P <- y0 before the next start
  while (true) {
    if (b.st == BS_AWAIT_ERR) {
      if (e < eps) {
        b.x += b.h;
        if (b.i == kBsMaxUse - 1) b.dt = b.h * kBsShrink;
        else if (b.i == kBsMaxUse - 2) b.dt = b.h * kBsGrow;
        else b.dt = (b.h * bs_seq(kBsMaxUse - 2)) / (1.0 * bs_seq(b.i));
        acceptCopy = true;
        b.st = BS_NEED_START;
      } else {
        ++b.i;
        if (b.i >= kBsMax) {
          b.h = b.h / 4.0;
          for (int k = 0; k < (kBsMax - kBsMaxUse) / 2.0; ++k) b.h /= 2.0;
          if (b.dt + b.h == b.dt) {
            fail = true;
            b.st = BS_DONE;
            return false;
          }
          b.i = 0;
        }
        b.st = BS_START_MMID;
      }
      continue;
    }
    if (b.st == BS_NEED_START) {
      if (!(b.xend - b.x > b.dt * 1e-6)) {
        b.st = BS_DONE;
        if (acceptCopy) {
          bs_op(pre, V_COPY, B_P, B_Y0);
          npre = 1;
        }
        return false;
      }
      if (b.x + b.dt > b.xend) b.dt = b.xend - b.x + 1e-10;
      for (int k = 0; k < kBsMax; ++k) xtab[k] = 0.0;
      // y0 = y3 = x = y (after P <- y0 on an accept), ytab[1..6] = 0
      bs_op(pre, V_START, acceptCopy ? B_Y0 : B_P, acceptCopy ? 1 : 0);
      npre = 1;
      if (bufs.b[B_Y0 + 4] != bufs.b[B_D]) bs_op(post + npost++, V_COPY, B_Y0 + 4, B_D);
      b.h = b.dt;
      b.i = 0;
      b.st = BS_START_MMID;
      lead = 1 + bs_seq(0);
      return true;
    }
    if (b.st != BS_START_MMID && b.st != BS_IN_MMID) return false;
    if (b.st == BS_START_MMID) {
      b.step = bs_seq(b.i);
      b.hh = b.h / (1.0 * b.step);
      bs_op(pre, V_MMID_BEGIN, 0, 0, 0, b.hh);
      b.nsub = 1;
      b.st = BS_IN_MMID;
    } else {
      bs_op(pre, V_MMID_STEP_X, 0, 0, 0, b.hh);
      ++b.nsub;
    }
    npre = 1;
    if (b.nsub == b.step) {
      // the midpoint's final step, rzextr and the error (the copy of the
      // derivative into y2 is part of the fused op)
      const double hs = b.h / (1.0 * b.step);
      const double xest = hs * hs;
      xtab[b.i] = xest;
      VecOp* f = post + npost++;
      bs_op(f, V_FINAL_EXTR, 0, 0, 0, b.hh);
      f->i = b.i;
      f->mi = (b.i < kBsMaxUse) ? b.i + 1 : kBsMaxUse;
      for (int k = 1; k < f->mi; ++k) f->fx[k] = xin[b.i - k] / xest;
      bs_op(post + npost++, V_ERROR, B_Y0, B_Y0 + 1);
      b.st = BS_AWAIT_ERR;
      lead = 1;
    } else {
      if (bufs.b[B_Y0 + 2] != bufs.b[B_D]) bs_op(post + npost++, V_COPY, B_Y0 + 2, B_D);
      lead = 1 + b.step - b.nsub;
    }
    return true;
  }
}
}  // namespace

__global__ void __launch_bounds__(TB)
    bs_tick_kernel(BsNode* __restrict__ st, CmfBufs bufs, const DNode* __restrict__ nodes,
                   const cd2* __restrict__ pool, const int64_t* __restrict__ rhoslot, double* err, uint8_t* rec,
                   int init, double t0, double t1, double eps) {
  __shared__ double red[TB];
  __shared__ cd2 rhos[16 * 16];
  __shared__ VecOp pre;
  __shared__ int npre;
  // the node's state up to its post ops, read once for the workgroup (one
  // round trip, before thread 0 can rewrite it)
  __shared__ double head[kBsHeadWords];
  const int v = blockIdx.x;
  BsNode* b = st + v;
  const DNode nd = nodes[v];
  if (threadIdx.x < kBsHeadWords) head[threadIdx.x] = reinterpret_cast<const double*>(b)[threadIdx.x];
  __syncthreads();
  const BsNode& hd = *reinterpret_cast<const BsNode*>(head);
  const int np = init ? 0 : hd.npost;
  // the post ops of the tick that ran, then (k == np) the decisions and the
  // next tick's pre op: one call site of the op code
  for (int k = 0; k <= np; ++k) {
    const VecOp* opp = &pre;
    if (k < np) {
      opp = &b->post[k];
    } else {
      if (threadIdx.x == 0) {
        BsRegs r{hd.st, hd.i, hd.step, hd.nsub, hd.x, hd.xend, hd.h, hd.hh, hd.dt};
        if (init) {
          r.st = BS_NEED_START;
          r.x = t0;
          r.xend = t1;
        }
        // the node error the post ops just wrote (this thread wrote it)
        const double e = (r.st == BS_AWAIT_ERR) ? err[v] : 0.0;
        int nq = 0, lead = 0, n1 = 0;
        bool fail = false;
        const bool active = bs_advance(r, hd.xtab, b->xtab, e, eps, bufs, &pre, n1, b->post, nq, fail, lead);
        b->st = r.st;
        b->i = r.i;
        b->step = r.step;
        b->nsub = r.nsub;
        b->x = r.x;
        b->xend = r.xend;
        b->h = r.h;
        b->hh = r.hh;
        b->dt = r.dt;
        b->npost = nq;
        const bool failed = fail || hd.fail;
        if (fail) b->fail = 1;
        if (active) b->nderiv = hd.nderiv + 1;
        npre = n1;
        rec[v] = active ? 1 : 0;
        rec[gridDim.x + v] = failed ? kBsFailed : (uint8_t)(lead < 254 ? lead : 254);
      }
      __syncthreads();
      if (!npre) break;
    }
    exec_op(opp, bufs, nodes, pool, rhoslot, err, v, nd, red, rhos);
    __syncthreads();
  }
}

int launch_bs_tick(BsNode* st, int64_t count, const CmfBufs& bufs, const DNode* nodes, const cplx<double>* pool,
                   const int64_t* rhoslot, double* err, uint8_t* rec, int init, double t0, double t1, double eps,
                   hipStream_t s) {
  static_assert(offsetof(BsNode, post) == 8 * kBsHeadWords, "BsNode head");
  if (count == 0) return QTG_OK;
  hipLaunchKernelGGL(bs_tick_kernel, dim3((unsigned)count), dim3(TB), 0, s, st, bufs, nodes, pool, rhoslot, err, rec,
                     init, t0, t1, eps);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_programs(const NodeProg* progs, int64_t count, const VecOp* ops, const CmfBufs& bufs,
                    const DNode* nodes, const cplx<double>* pool, const int64_t* rhoslot, double* err,
                    hipStream_t s, uint8_t* mask) {
  if (count == 0) return QTG_OK;
  hipLaunchKernelGGL(program_kernel, dim3((unsigned)count), dim3(TB), 0, s, progs, ops, bufs, nodes, pool, rhoslot,
                     err, mask);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_ortho(const OrthoItem* items, int64_t count, const DNode* nodes, cplx<double>* P, cplx<double>* sw,
                 int nmax, int nspf, hipStream_t s) {
  if (count == 0) return QTG_OK;
  const size_t lds = ((size_t)nmax + 2 * (size_t)jacobi_elems(nspf)) * sizeof(cd2);
  hipLaunchKernelGGL(ortho_kernel, dim3((unsigned)count), dim3(EW), lds, s, items, nodes, P, sw, nmax, nspf);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_ortho_global(const OrthoItem* items, int64_t count, const DNode* nodes, cplx<double>* P,
                        cplx<double>* sw, int nspf, hipStream_t s) {
  if (count == 0) return QTG_OK;
  const size_t lds = 2 * (size_t)jacobi_elems(nspf) * sizeof(cd2);
  hipLaunchKernelGGL(ortho_global_kernel, dim3((unsigned)count), dim3(EW), lds, s, items, nodes, P, sw, nspf);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_eig_sorted(const cplx<double>* A, int n, cplx<double>* UT, double* ev, hipStream_t s) {
  if (n < 1 || n > 64) {
    set_error("eig_sorted: order outside 1..64");
    return QTG_EARG;
  }
  const size_t lds = 2 * (size_t)jacobi_elems(n) * sizeof(cd2);
  hipLaunchKernelGGL(eig_sorted_kernel, dim3(1), dim3(EW), lds, s, A, n, UT, ev);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_apply_sw(const SwApply* items, int64_t count, const DNode* nodes, cplx<double>* P,
                    const cplx<double>* sw, int nmax, hipStream_t s) {
  if (count == 0) return QTG_OK;
  const size_t lds = 2 * (size_t)nmax * sizeof(cd2);
  hipLaunchKernelGGL(apply_sw_kernel, dim3((unsigned)count), dim3(TB), lds, s, items, nodes, P, sw, nmax);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

}  // namespace qtg
