// Grouped tree kernels of the gfx950 engine.  One launch processes every
// (term, node) work item of one tree layer; items have uneven shapes and are
// described by small device records.
//
//  sandwich_kernel   out(j,l) = sum_{all indices but k} conj(Bra(..j..)) [F Ket](..l..)
//                    with F a product of <= 3 mode factors on modes != k.  This
//                    one form covers representLayer (k = last, F = children's
//                    h-matrices or the leaf operator), contractionLayer (k =
//                    child index, F = siblings' h-matrices and the parent's
//                    hole), the density tree (F = rho of the parent) and the
//                    overlap tree (F = children's overlaps).
//  apply_kernel      partial(v, chunk) = sum_{terms in chunk} c_t F_t A_v: the
//                    node-local SOP application of Apply(); factors are the
//                    children's h-matrices (or the leaf operator) and the
//                    node's hole matrix on the last mode.
//  finish_kernel     FP64: sums the partials of a node, applies the (1-P)
//                    projector with the overlap inverse, rho^-1 and -i*phase
//                    (LayerDerivative), and forms <A_top|Apply(A_top)> for the
//                    energy.
//  heev_reginv_kernel FP64 batched parallel-order (tournament) Jacobi
//                    eigensolver of the density matrices and the regularised
//                    inverse U diag(1/(l + eps e^{-l/eps})) U^H.
//
// Node tensors live in one arena (post-order by node address); matrices live
// in one pool; all offsets are in complex elements.  Work buffers are staged in
// LDS: the node tensor, and up to two intermediate products.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "jacobi.hpp"
#include "node_tile.hpp"
#include "tree_ops.hpp"

namespace qtg {

// out(j,l) = sum_r conj(X(r,j)) Y(r,l) over all indices but k.  scratch needs TB entries.
template <typename T>
__device__ void contract_k(cplx<T>* __restrict__ out, const cplx<T>* X, const cplx<T>* Y, const DNode& nd, int k,
                           cplx<T>* scratch, bool accumulate) {
  int bef, nk, aft;
  node_view(nd, k, bef, nk, aft);
  const int outs = nk * nk;
  const int R = bef * aft;
  if (outs >= TB) {
    for (int o = threadIdx.x; o < outs; o += TB) {
      const int j = o % nk, l = o / nk;
      cplx<T> acc = czero<T>();
      const int jo = j * bef, lo = l * bef, abn = bef * nk;
      for (int a = 0; a < aft; ++a)
        for (int b = 0; b < bef; ++b) {
          const int base = b + a * abn;
          cfmac(acc, X[base + jo], Y[base + lo]);
        }
      if (accumulate) acc = cadd(acc, out[o]);
      out[o] = acc;
    }
    return;
  }
  const int G = TB / outs;
  const int o = threadIdx.x % outs, g = threadIdx.x / outs;
  const int j = o % nk, l = o / nk;
  cplx<T> acc = czero<T>();
  if (g < G && g < R) {
    // rows r = g, g+G, ...; (b, a) = (r mod bef, r / bef) advanced incrementally
    int b = g % bef, a = g / bef;
    const int db = G % bef, da = G / bef;
    const int jo = j * bef, lo = l * bef, abn = bef * nk;
    for (int r = g; r < R; r += G) {
      const int base = b + a * abn;
      cfmac(acc, X[base + jo], Y[base + lo]);
      b += db;
      a += da;
      if (b >= bef) {
        b -= bef;
        ++a;
      }
    }
  }
  scratch[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < outs) {
    cplx<T> s = scratch[threadIdx.x];
    for (int gg = 1; gg < G; ++gg) s = cadd(s, scratch[gg * outs + threadIdx.x]);
    if (accumulate) s = cadd(s, out[threadIdx.x]);
    out[threadIdx.x] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(TB)
    sandwich_kernel(const SwItem* __restrict__ items, const DNode* __restrict__ nodes, const cplx<T>* bra_base,
                    const cplx<T>* ket_base, cplx<T>* pool, int nmax) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* buf0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf1 = buf0 + nmax;  // nmax >= TB (scratch of contract_k)
  cplx<T>* buf2 = buf1 + nmax;
  const SwItem it = items[blockIdx.x];
  const DNode nd = nodes[it.node];
  const int N = nd.N;
  // ket into buf0; the bra stays in global memory unless it is the ket
  stage_to_lds<TB>(buf0, ket_base + it.ket, N);
  const bool same = (bra_base == ket_base) && (it.bra == it.ket);
  __syncthreads();
  cplx<T>* cur = buf0;
  cplx<T>* bufs[2] = {buf1, buf2};
  int nb = 0;
  for (int f = 0; f < it.nf; ++f) {
    cplx<T>* dst = bufs[nb];
    mode_apply<T>(dst, cur, pool + it.fmat[f], nd, it.fmode[f]);
    __syncthreads();
    cur = dst;
    nb ^= 1;
  }
  // bra: buf0 still holds the ket tensor when bra == ket; otherwise stage the
  // bra into a buffer that is not the product buffer.
  const cplx<T>* bra = buf0;
  cplx<T>* scratch = (cur == buf1) ? buf2 : buf1;
  if (!same) {
    cplx<T>* b = (cur == buf0) ? buf1 : buf0;
    scratch = (cur == buf0) ? buf2 : ((cur == buf1) ? buf2 : buf1);
    __syncthreads();
    stage_to_lds<TB>(b, bra_base + it.bra, N);
    __syncthreads();
    bra = b;
  }
  contract_k<T>(pool + it.out, bra, cur, nd, it.k, scratch, false);
}

// ---------------------------------------------------------------------- apply
template <typename T, int EPT>
__global__ void __launch_bounds__(TB)
    apply_kernel(const ApChunk* __restrict__ chunks, const ApTerm* __restrict__ terms,
                 const DNode* __restrict__ nodes, const cplx<T>* psi, const cplx<T>* __restrict__ pool,
                 cplx<T>* partial, int nmax, const uint8_t* __restrict__ active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* buf0 = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf1 = buf0 + nmax;
  cplx<T>* buf2 = buf1 + nmax;
  const ApChunk ch = chunks[blockIdx.x];
  if (active && !active[ch.node]) return;  // uniform per workgroup
  const DNode nd = nodes[ch.node];
  const int N = nd.N;
  stage_to_lds<TB>(buf0, psi + nd.off, N);
  cplx<T> acc[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) acc[i] = czero<T>();
  __syncthreads();
  for (int t = ch.t0; t < ch.t1; ++t) {
    const ApTerm tm = terms[t];
    const cplx<T> c = mk<T>((T)tm.cre, (T)tm.cim);
    const cplx<T>* cur = buf0;
    int nb = 0;
    cplx<T>* bufs[2] = {buf1, buf2};
    for (int f = 0; f + 1 < tm.nf; ++f) {
      mode_apply<T>(bufs[nb], cur, pool + tm.fmat[f], nd, tm.fmode[f]);
      __syncthreads();
      cur = bufs[nb];
      nb ^= 1;
    }
    if (tm.nf == 0) {
#pragma unroll
      for (int i = 0; i < EPT; ++i) {
        const int e = threadIdx.x + i * TB;
        if (e < N) cfma(acc[i], c, cur[e]);
      }
    } else {
      const cplx<T>* M = pool + tm.fmat[tm.nf - 1];
      int bef, nf, aft;
      node_view(nd, tm.fmode[tm.nf - 1], bef, nf, aft);
#pragma unroll
      for (int i = 0; i < EPT; ++i) {
        const int e = threadIdx.x + i * TB;
        if (e < N) {
          const int b = e % bef, q = e / bef, k = q % nf, a = q / nf;
          const cplx<T>* s = cur + b + a * bef * nf;
          cplx<T> v = czero<T>();
          for (int m = 0; m < nf; ++m) cfma(v, M[k + m * nf], s[m * bef]);
          cfma(acc[i], c, v);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = threadIdx.x + i * TB;
    if (e < N) partial[ch.out + e] = acc[i];
  }
}

// --------------------------------------------------------------------- finish
// Block-wide FP64 complex sum over TB threads.
__device__ inline cplx<double> block_sum(cplx<double> v, cplx<double>* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = TB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  cplx<double> r = red[0];
  __syncthreads();
  return r;
}

template <typename T>
__global__ void __launch_bounds__(TB)
    finish_kernel(const FinNode* __restrict__ fin, const DNode* __restrict__ nodes,
                  const cplx<double>* __restrict__ psi64, const cplx<T>* __restrict__ partial,
                  const cplx<double>* __restrict__ pool64, cplx<double>* dpsi, cplx<double>* apply_out,
                  double ph_re, double ph_im, cplx<double>* energy, int nmax, const uint8_t* __restrict__ active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<double>* sY = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* sA = sY + nmax;
  cplx<double>* sm = sA + nmax;  // two n x n matrices (n = SPFs of the node)
  __shared__ cplx<double> red[TB];
  const FinNode fn = fin[blockIdx.x];
  if (active && !active[fn.node]) return;  // uniform per workgroup
  const DNode nd = nodes[fn.node];
  const int N = nd.N, n = nd.dims[nd.order - 1], L = N / n;
  for (int e = threadIdx.x; e < N; e += TB) {
    cplx<double> s = czero<double>();
    for (int c = 0; c < fn.nchunk; ++c) {
      const cplx<T> p = partial[fn.part0 + (int64_t)c * N + e];
      s.re += (double)p.re;
      s.im += (double)p.im;
    }
    sY[e] = s;
    sA[e] = psi64[nd.off + e];
  }
  __syncthreads();
  if (apply_out)
    for (int e = threadIdx.x; e < N; e += TB) apply_out[nd.off + e] = sY[e];
  // -i * phase
  const cplx<double> f = mk<double>(ph_im, -ph_re);
  if (fn.top) {
    cplx<double> ev = czero<double>(), nv = czero<double>();
    for (int e = threadIdx.x; e < N; e += TB) {
      cfmac(ev, sA[e], sY[e]);
      cfmac(nv, sA[e], sA[e]);
    }
    ev = block_sum(ev, red);
    nv = block_sum(nv, red);
    if (threadIdx.x == 0 && energy) {
      energy[0] = ev;
      energy[1] = nv;
    }
    for (int e = threadIdx.x; e < N; e += TB) dpsi[nd.off + e] = cmul(f, sY[e]);
    return;
  }
  cplx<double>* S = sm;          // n x n overlap A^H A
  cplx<double>* X = sm + n * n;  // A^H Y, then solution
  // overlaps
  for (int o = threadIdx.x; o < 2 * n * n; o += TB) {
    const int which = o / (n * n), oo = o % (n * n);
    const int j = oo % n, l = oo / n;
    const cplx<double>* R = which ? sY : sA;
    cplx<double> acc = czero<double>();
    for (int i = 0; i < L; ++i) cfmac(acc, sA[i + j * L], R[i + l * L]);
    (which ? X : S)[oo] = acc;
  }
  __syncthreads();
  // Gauss-Jordan with partial pivoting on [S | X] (column-major, n x n each)
  __shared__ int piv;
  __shared__ cplx<double> facs[64];
  for (int col = 0; col < n; ++col) {
    if (threadIdx.x == 0) {
      int p = col;
      double best = -1.0;
      for (int r = col; r < n; ++r) {
        const cplx<double> v = S[r + col * n];
        const double a = v.re * v.re + v.im * v.im;
        if (a > best) {
          best = a;
          p = r;
        }
      }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != col) {
      for (int k = threadIdx.x; k < 2 * n; k += TB) {
        cplx<double>* Mk = (k < n) ? S + k * n : X + (k - n) * n;
        const cplx<double> t = Mk[col];
        Mk[col] = Mk[p];
        Mk[p] = t;
      }
      __syncthreads();
    }
    const cplx<double> d = S[col + col * n];
    const double dd = d.re * d.re + d.im * d.im;
    const cplx<double> dinv = mk<double>(d.re / dd, -d.im / dd);
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * n; k += TB) {
      cplx<double>* Mk = (k < n) ? S + k * n : X + (k - n) * n;
      Mk[col] = cmul(Mk[col], dinv);
    }
    __syncthreads();
    // eliminate: every row r != col: row_r -= S(r,col) * row_col.  Read the
    // factors first so that the column col itself is updated last.
    for (int r = threadIdx.x; r < n; r += TB) facs[r] = S[r + col * n];
    __syncthreads();
    for (int o = threadIdx.x; o < n * 2 * n; o += TB) {
      const int r = o % n, k = o / n;
      if (r == col) continue;
      cplx<double>* Mk = (k < n) ? S + k * n : X + (k - n) * n;
      const cplx<double> fr = facs[r];
      const cplx<double> v = Mk[col];
      cplx<double> w = Mk[r];
      w.re -= fr.re * v.re - fr.im * v.im;
      w.im -= fr.re * v.im + fr.im * v.re;
      Mk[r] = w;
    }
    __syncthreads();
  }
  // Y <- Y - A X (each lane updates only its own elements; A is read-only)
  for (int e = threadIdx.x; e < N; e += TB) {
    const int i = e % L, k = e / L;
    cplx<double> acc = sY[e];
    for (int m = 0; m < n; ++m) {
      const cplx<double> a = sA[i + m * L], x = X[m + k * n];
      acc.re -= a.re * x.re - a.im * x.im;
      acc.im -= a.re * x.im + a.im * x.re;
    }
    sY[e] = acc;
  }
  __syncthreads();
  // dpsi(i,k) = f * sum_l rhoinv(k,l) Y(i,l)
  const cplx<double>* Ri = pool64 + fn.rhoinv;
  for (int e = threadIdx.x; e < N; e += TB) {
    const int i = e % L, k = e / L;
    cplx<double> acc = czero<double>();
    for (int l = 0; l < n; ++l) cfma(acc, Ri[k + l * n], sY[i + l * L]);
    dpsi[nd.off + e] = cmul(f, acc);
  }
}

// ------------------------------------------- density eigen / regularised inverse
// One wavefront per density matrix.  R = U diag(1/(l + eps e^{-l/eps})) U^H
// (SpectralDecompositionTree.cpp:30-48).  When rho is comfortably positive
// definite the regulariser is below rounding and R is rho^-1 itself, which
// n Gauss-Jordan steps give (fast path, no eigenvalues); otherwise the
// tournament Jacobi solver (jacobi.hpp) diagonalises rho.
//
// Fast path: in-place Gauss-Jordan on the diagonal pivots of the hermitised
// rho, ping-ponging between the two LDS buffers.  It is taken only when
// every pivot is positive and finite and 1/||X||_F >= 50 eps for the result
// X: 1/||X||_F bounds the smallest eigenvalue l_min from below, so
// eps e^{-l/eps} <= e^{-50} l < 2e-22 l for every eigenvalue, far below the
// rounding of either algorithm, and R = X to rounding.  It also asks for a
// condition number below 1e6 (||X||_F tr(rho) <= 1e6, tr2 = tr(rho)^2), so
// that rho^-1 carries at most ~1e-10 relative rounding either way; worse
// conditioned densities (an SPF almost unoccupied) take the eigensolver.
__device__ inline bool gj_inverse_certified(cplx<double>*& S, cplx<double>*& D, int n, int ld, double eps,
                                            double tr2) {
  const int nn = n * n;
  for (int k = 0; k < n; ++k) {
    const cplx<double> p = S[k + k * ld];  // the same element in every lane: a uniform test
    if (!(p.re > 0.0) || !(p.re < 1e300) || !(fabs(p.im) < 1e300)) return false;
    const double dd = p.re * p.re + p.im * p.im;
    const cplx<double> pinv = mk<double>(p.re / dd, -p.im / dd);
    // four elements per lane and pass: every LDS read of a pass is issued
    // before its writes (they go to the other buffer)
    for (int base = 0; base < nn; base += 4 * EW) {
      cplx<double> m[4], sik[4], skj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int o = min(base + u * EW + (int)threadIdx.x, nn - 1);
        const int i = o % n, j = o / n;
        m[u] = S[i + j * ld];
        sik[u] = S[i + k * ld];
        skj[u] = S[k + j * ld];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int o = base + u * EW + (int)threadIdx.x;
        if (o >= nn) continue;
        const int i = o % n, j = o / n;
        cplx<double> r;
        if (i == k && j == k) {
          r = pinv;
        } else if (i == k) {
          r = cmul(m[u], pinv);
        } else if (j == k) {
          const cplx<double> t = cmul(m[u], pinv);
          r = mk<double>(-t.re, -t.im);
        } else {
          const cplx<double> t = cmul(cmul(sik[u], skj[u]), pinv);
          r = mk<double>(m[u].re - t.re, m[u].im - t.im);
        }
        D[i + j * ld] = r;
      }
    }
    __syncthreads();
    cplx<double>* t = S;
    S = D;
    D = t;
  }
  double f = 0.0;
  for (int o = threadIdx.x; o < nn; o += EW) {
    const cplx<double> x = S[o % n + (o / n) * ld];
    f += x.re * x.re + x.im * x.im;
  }
  f = wave_sum(f);
  return f < 1e300 && f * (50.0 * eps) * (50.0 * eps) <= 1.0 && f * tr2 <= 1e12;
}

template <typename T>
__global__ void __launch_bounds__(EW)
    heev_reginv_kernel(const EigItem* __restrict__ items, const cplx<T>* __restrict__ pool,
                       cplx<double>* pool64, double eps, int nmaxspf) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<double>* A = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* U = A + jacobi_elems(nmaxspf);
  __shared__ JacobiScratch js;
  __shared__ double lam[64], wgt[64];
  const EigItem it = items[blockIdx.x];
  const int n = it.n;
  const int ld = jacobi_ld(n);
  if (it.ev < 0) {
    // the hermitised rho, eight element pairs per lane loaded before any
    // LDS write, so the global loads of a pass are in flight together
    const int nn = n * n;
    for (int base = 0; base < nn; base += 8 * EW) {
      cplx<T> a[8], b[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int o = min(base + u * EW + (int)threadIdx.x, nn - 1);
        const int i = o % n, j = o / n;
        a[u] = pool[it.in + i + j * n];
        b[u] = pool[it.in + j + i * n];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int o = base + u * EW + (int)threadIdx.x;
        if (o < nn)
          A[o % n + (o / n) * ld] =
              mk<double>(0.5 * ((double)a[u].re + (double)b[u].re), 0.5 * ((double)a[u].im - (double)b[u].im));
      }
    }
    __syncthreads();
    double tr = 0.0;
    for (int i = threadIdx.x; i < n; i += EW) tr += A[i + i * ld].re;
    tr = wave_sum(tr);
    cplx<double>* S = A;
    cplx<double>* D = U;
    if (gj_inverse_certified(S, D, n, ld, eps, tr * tr)) {
      for (int o = threadIdx.x; o < n * n; o += EW) pool64[it.out + o] = S[o % n + (o / n) * ld];
      return;
    }
    __syncthreads();
  }
  for (int o = threadIdx.x; o < n * n; o += EW) {
    const int i = o % n, j = o / n;
    const cplx<T> v = pool[it.in + o];
    A[i + j * ld] = mk<double>((double)v.re, (double)v.im);
  }
  __syncthreads();
  jacobi_eig(A, U, n, ld, js);
  for (int i = threadIdx.x; i < n; i += EW) {
    const double l = A[i + i * ld].re;
    lam[i] = l;
    wgt[i] = 1.0 / (l + eps * exp(-l / eps));
  }
  __syncthreads();
  if (it.ev >= 0)
    for (int i = threadIdx.x; i < n; i += EW) pool64[it.ev + i] = mk<double>(lam[i], 0.0);
  for (int o = threadIdx.x; o < n * n; o += EW) {
    const int i = o % n, j = o / n;
    cplx<double> acc = czero<double>();
    for (int k = 0; k < n; ++k) {
      const double w = wgt[k];
      const cplx<double> u = U[i + k * ld], v = U[j + k * ld];
      acc.re += w * (u.re * v.re + u.im * v.im);
      acc.im += w * (u.im * v.re - u.re * v.im);
    }
    pool64[it.out + o] = acc;
  }
}

// ------------------------------------------------------------ combinations
// Sum of term t into partial sum u: re/im[u] += c * (pool[src + e] or the
// identity element id when src < 0).
template <typename T>
__device__ inline void comb_term(const CombTerm& ct, const cplx<T>& m, double id, double& re, double& im) {
  const double mr = ct.src >= 0 ? (double)m.re : id, mi = ct.src >= 0 ? (double)m.im : 0.0;
  re += ct.cre * mr - ct.cim * mi;
  im += ct.cre * mi + ct.cim * mr;
}

// Element e of one output over its term list tl[0..nt): term 4q+u of a full
// group of four goes to partial sum u, the last nt % 4 terms to sum 0.
// Eight matrix loads are issued before their sums.  Terms without a source
// load element e of the output itself (in bounds, unused).
template <typename T>
__device__ inline cplx<T> comb_element(const CombTerm* tl, int nt, const cplx<T>* pool, int64_t out, int e,
                                       double id) {
  double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
  int t = 0;
  for (; t + 8 <= nt; t += 8) {
    cplx<T> m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t src = tl[t + u].src;
      m[u] = pool[(src >= 0 ? src : out) + e];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) comb_term(tl[t + u], m[u], id, re[u & 3], im[u & 3]);
  }
  for (; t + 4 <= nt; t += 4) {
    cplx<T> m[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t src = tl[t + u].src;
      m[u] = pool[(src >= 0 ? src : out) + e];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) comb_term(tl[t + u], m[u], id, re[u], im[u]);
  }
  for (; t < nt; ++t) {
    const int64_t src = tl[t].src;
    comb_term(tl[t], pool[(src >= 0 ? src : out) + e], id, re[0], im[0]);
  }
  return mk<T>((T)((re[0] + re[1]) + (re[2] + re[3])), (T)((im[0] + im[1]) + (im[2] + im[3])));
}

// One workgroup per output matrix.  Term lists of up to TB entries are
// staged in LDS first (one coalesced load), so the matrix loads do not wait
// on a term load each; longer lists are read from global memory.
template <typename T>
__global__ void __launch_bounds__(TB)
    combine_kernel(const CombOut* __restrict__ outs, const CombTerm* __restrict__ terms, cplx<T>* pool) {
  __shared__ CombTerm st[TB];
  const CombOut co = outs[blockIdx.x];
  const int n = co.n, nt = co.t1 - co.t0;
  if (nt <= TB) {
    if ((int)threadIdx.x < nt) st[threadIdx.x] = terms[co.t0 + threadIdx.x];
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += TB)
      pool[co.out + e] = comb_element<T>(st, nt, pool, co.out, e, (e % n == e / n) ? 1.0 : 0.0);
  } else {
    for (int e = threadIdx.x; e < n * n; e += TB)
      pool[co.out + e] = comb_element<T>(terms + co.t0, nt, pool, co.out, e, (e % n == e / n) ? 1.0 : 0.0);
  }
}

// ------------------------------------------------------------------- launches
template <typename T>
int launch_combine(const CombOut* outs, int64_t count, const CombTerm* terms, void* pool, hipStream_t s) {
  if (count == 0) return QTG_OK;
  hipLaunchKernelGGL(combine_kernel<T>, dim3((unsigned)count), dim3(TB), 0, s, outs, terms, (cplx<T>*)pool);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}
template int launch_combine<double>(const CombOut*, int64_t, const CombTerm*, void*, hipStream_t);
template int launch_combine<float>(const CombOut*, int64_t, const CombTerm*, void*, hipStream_t);

template <typename T>
int launch_sandwich(const SwItem* items, int64_t count, const DNode* nodes, const void* bra, const void* ket,
                    void* pool, int nmax, hipStream_t s) {
  if (count == 0) return QTG_OK;
  const size_t lds = 3 * (size_t)nmax * sizeof(cplx<T>);
  hipLaunchKernelGGL(sandwich_kernel<T>, dim3((unsigned)count), dim3(TB), lds, s, items, nodes,
                     (const cplx<T>*)bra, (const cplx<T>*)ket, (cplx<T>*)pool, nmax);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_apply(const ApChunk* chunks, int64_t nchunks, const ApTerm* terms, const DNode* nodes, const void* psi,
                 const void* pool, void* partial, int nmax, hipStream_t s, const uint8_t* active) {
  if (nchunks == 0) return QTG_OK;
  const size_t lds = 3 * (size_t)nmax * sizeof(cplx<T>);
  const int ept = (nmax + TB - 1) / TB;
#define QTG_APPLY(E)                                                                                     \
  hipLaunchKernelGGL((apply_kernel<T, E>), dim3((unsigned)nchunks), dim3(TB), lds, s, chunks, terms, nodes, \
                     (const cplx<T>*)psi, (const cplx<T>*)pool, (cplx<T>*)partial, nmax, active)
  if (ept <= 1) QTG_APPLY(1);
  else if (ept <= 2) QTG_APPLY(2);
  else if (ept <= 4) QTG_APPLY(4);
  else if (ept <= 8) QTG_APPLY(8);
  else if (ept <= 16) QTG_APPLY(16);
  else {
    set_error("apply: node tensor larger than 4096 elements is not supported by the fused kernel");
    return QTG_ESHAPE;
  }
#undef QTG_APPLY
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_finish(const FinNode* fin, int64_t count, const DNode* nodes, const void* psi64, const void* partial,
                  const void* pool64, void* dpsi, void* apply_out, double ph_re, double ph_im, void* energy,
                  int nmax, int nspf, hipStream_t s, const uint8_t* active) {
  if (count == 0) return QTG_OK;
  const size_t lds = finish_lds_bytes(nmax, nspf);
  hipLaunchKernelGGL(finish_kernel<T>, dim3((unsigned)count), dim3(TB), lds, s, fin, nodes,
                     (const cplx<double>*)psi64, (const cplx<T>*)partial, (const cplx<double>*)pool64,
                     (cplx<double>*)dpsi, (cplx<double>*)apply_out, ph_re, ph_im, (cplx<double>*)energy, nmax,
                     active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_heev(const EigItem* items, int64_t count, const void* pool, void* pool64, double eps, int nspf,
                hipStream_t s) {
  if (count == 0) return QTG_OK;
  const size_t lds = 2 * (size_t)jacobi_elems(nspf) * sizeof(cplx<double>);
  hipLaunchKernelGGL(heev_reginv_kernel<T>, dim3((unsigned)count), dim3(EW), lds, s, items, (const cplx<T>*)pool,
                     (cplx<double>*)pool64, eps, nspf);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template int launch_sandwich<double>(const SwItem*, int64_t, const DNode*, const void*, const void*, void*, int,
                                     hipStream_t);
template int launch_sandwich<float>(const SwItem*, int64_t, const DNode*, const void*, const void*, void*, int,
                                    hipStream_t);
template int launch_apply<double>(const ApChunk*, int64_t, const ApTerm*, const DNode*, const void*, const void*,
                                  void*, int, hipStream_t, const uint8_t*);
template int launch_apply<float>(const ApChunk*, int64_t, const ApTerm*, const DNode*, const void*, const void*,
                                 void*, int, hipStream_t, const uint8_t*);
template int launch_finish<double>(const FinNode*, int64_t, const DNode*, const void*, const void*, const void*,
                                   void*, void*, double, double, void*, int, int, hipStream_t, const uint8_t*);
template int launch_finish<float>(const FinNode*, int64_t, const DNode*, const void*, const void*, const void*,
                                  void*, void*, double, double, void*, int, int, hipStream_t, const uint8_t*);
template int launch_heev<double>(const EigItem*, int64_t, const void*, void*, double, int, hipStream_t);
template int launch_heev<float>(const EigItem*, int64_t, const void*, void*, double, int, hipStream_t);

}  // namespace qtg
