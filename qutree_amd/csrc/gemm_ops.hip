// Grouped complex contractions of node tensors on gfx950 matrix cores.
//
// Every tile is a 16 x 16 block of a complex product accumulated with the
// real 16x16x4 MFMA instructions (v_mfma_f64_16x16x4f64 for complex128,
// v_mfma_f32_16x16x4f32 for complex64): one complex k-step of four k values
// is four real products,
//   Cre += Are Bre - Aim Bim,   Cim += Are Bim + Aim Bre.
// Operand layout of both instructions (checked on the device): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; the 16 x 16 accumulator sits in
// four registers per lane, row (l / 16) + 4 r (f64) or 4 (l / 16) + r (f32),
// column l % 16.
//
// Operands are read straight from global memory (node tensors and matrices
// of one launch are a few MB and stay in L2 / MALL); each lane predicates its
// loads, so every wavefront runs the same number of k-steps and MFMA issue
// stays uniform.  Two accumulator pairs alternate between k-steps so that
// consecutive MFMAs do not depend on each other.
#include <hip/hip_ext.h>

#include <algorithm>

#include "gemm_ops.hpp"

namespace qtg {

namespace {

// Launches k; with kt set, through hipExtLaunchKernelGGL so that kt's events
// carry the kernel's own start and stop times (no marker packets around it).
template <typename K, typename... A>
void launch_timed(K k, dim3 grid, dim3 block, unsigned lds, hipStream_t s, const KernelTiming* kt, A... a) {
  if (kt) hipExtLaunchKernelGGL(k, grid, block, lds, s, kt->start, kt->stop, 0, a...);
  else hipLaunchKernelGGL(k, grid, block, lds, s, a...);
}


template <typename T>
struct MF;
template <>
struct MF<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct MF<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// (cr, ci) += a * b on one complex k-step (a is already conjugated if needed)
template <typename T>
__device__ inline void cmma(typename MF<T>::v4& cr, typename MF<T>::v4& ci, const cplx<T>& a, const cplx<T>& b) {
  cr = MF<T>::mma(a.re, b.re, cr);
  cr = MF<T>::mma(-a.im, b.im, cr);
  ci = MF<T>::mma(a.re, b.im, ci);
  ci = MF<T>::mma(a.im, b.re, ci);
}

template <typename D>
__device__ inline int find_desc(const D* __restrict__ descs, int ndesc, int64_t task) {
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= task) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

constexpr int kWaves = kGemmWaves;  // wavefronts (tasks) per workgroup

// Descriptor of a task: wgdesc[blockIdx.x] is the descriptor of the
// workgroup's first task (precomputed on the host), and a task lies at most
// kWaves - 1 descriptors further.
template <typename D>
__device__ inline int task_desc(const D* __restrict__ descs, int ndesc, const int32_t* __restrict__ wgdesc,
                                int64_t task) {
  int di = wgdesc[blockIdx.x];
  while (di + 1 < ndesc && descs[di + 1].tile0 <= task) ++di;
  return di;
}

// ------------------------------------------------------------- mode product
// Output tile: 16 rows of the flattened (b, a) index x 16 output columns i.
template <typename T, typename TM, typename TO>
__device__ inline void gm_task(const GmDesc& d, int64_t t, const cplx<T>* __restrict__ X,
                               const cplx<TM>* __restrict__ M, cplx<TO>* O, cplx<TO>* part, double sre,
                               double sim) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63;
  const int split = d.ksplit > 1 ? d.ksplit : 1;
  const int ks = (int)(t % split);
  t /= split;
  const int CT = (d.nout + 15) >> 4;
  const int64_t rt = t / CT;
  const int ct = (int)(t - rt * CT);
  const int64_t bef = d.bef, R = bef * d.aft;
  const int nin = d.nin, nout = d.nout;
  const int li = lane & 15, kk = lane >> 4;
  const T are = (T)(d.are * sre - d.aim * sim), aim = (T)(d.are * sim + d.aim * sre);

  if (d.op == GM_COPY) {
    // out(b,i,a) = alpha X(b,i,a): 16 rows x 16 columns, four per lane
    for (int q = 0; q < 4; ++q) {
      const int64_t r = rt * 16 + li;
      const int col = ct * 16 + kk * 4 + q;
      if (r < R && col < nout) {
        const int64_t b = r % bef, a = r / bef;
        const cplx<T> x = X[d.x + b + (int64_t)col * bef + a * bef * nin];
        const int64_t o = d.out + b + (int64_t)col * bef + a * bef * nout;
        cplx<TO> v = mk<TO>((TO)(are * x.re - aim * x.im), (TO)(are * x.im + aim * x.re));
        if (d.acc) v = cadd(v, O[o]);
        O[o] = v;
      }
    }
    return;
  }

  // A operand: row r = rt*16 + li of the flattened (b, a), column k
  const int64_t r = rt * 16 + li;
  const bool rv = r < R;
  const int64_t rb = rv ? r % bef : 0, ra = rv ? r / bef : 0;
  const cplx<T>* xa = X + d.x + rb + ra * bef * nin;
  // B operand: op(M)(j, k) for column j = ct*16 + li
  const int j = ct * 16 + li;
  const bool cv = j < nout;
  const cplx<TM>* mb;
  int64_t mks;
  if (d.op == GM_N) {
    mb = M + d.m + j;  // M(j, k) at j + k nout
    mks = nout;
  } else {
    mb = M + d.m + (int64_t)j * nin;  // M(k, j) at k + j nin
    mks = 1;
  }
  const bool conjm = d.op == GM_H;
  // K range of this split (a multiple of 4 long, except the last)
  const int kchunk = ((nin + split - 1) / split + 3) & ~3;
  const int kbeg = ks * kchunk, kend = min(nin, kbeg + kchunk);
  v4 cr0 = {0, 0, 0, 0}, ci0 = {0, 0, 0, 0}, cr1 = {0, 0, 0, 0}, ci1 = {0, 0, 0, 0};
  const cplx<T> z = czero<T>();
  // four k-steps per iteration: all loads are issued before the MFMAs
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    cplx<T> av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kx = k0 + 4 * u + kk;
      const bool kv = kx < kend;
      av[u] = (rv && kv) ? xa[(int64_t)kx * bef] : z;
      if (cv && kv) {
        const cplx<TM> mv = mb[(int64_t)kx * mks];
        bv[u] = mk<T>((T)mv.re, (T)mv.im);
      } else {
        bv[u] = z;
      }
      if (conjm) bv[u].im = -bv[u].im;
    }
    cmma<T>(cr0, ci0, av[0], bv[0]);
    if (k0 + 4 < kend) cmma<T>(cr1, ci1, av[1], bv[1]);
    if (k0 + 8 < kend) cmma<T>(cr0, ci0, av[2], bv[2]);
    if (k0 + 12 < kend) cmma<T>(cr1, ci1, av[3], bv[3]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t rr = rt * 16 + MF<T>::drow(lane, q);
    if (rr < R && cv) {
      const T vr = cr0[q] + cr1[q], vi = ci0[q] + ci1[q];
      cplx<TO> v = mk<TO>((TO)(are * vr - aim * vi), (TO)(are * vi + aim * vr));
      if (split > 1) {
        part[d.part + (int64_t)ks * R * nout + rr + (int64_t)j * R] = v;
      } else {
        const int64_t b = rr % bef, a = rr / bef;
        const int64_t o = d.out + b + (int64_t)j * bef + a * bef * nout;
        if (d.acc) v = cadd(v, O[o]);
        O[o] = v;
      }
    }
  }
}

// Sums the partial tiles of split descriptors in chunk order.  The split
// descriptors carry, in tile0, the first 256-element block of their output:
// each workgroup owns one block of one descriptor.
template <typename TO>
__global__ void __launch_bounds__(256)
    gm_reduce_kernel(const GmDesc* __restrict__ descs, int ndesc, cplx<TO>* O, const cplx<TO>* __restrict__ part,
                     const uint8_t* __restrict__ active) {
  const GmDesc d = descs[find_desc(descs, ndesc, blockIdx.x)];
  if (active && d.node >= 0 && !active[d.node]) return;
  const int64_t bef = d.bef, R = bef * d.aft, n = R * d.nout;
  const int64_t e = (blockIdx.x - d.tile0) * 256 + threadIdx.x;
  if (e < n) {
    cplx<TO> s = part[d.part + e];
    for (int c = 1; c < d.ksplit; ++c) s = cadd(s, part[d.part + c * n + e]);
    const int64_t r = e % R, jj = e / R;
    const int64_t o = d.out + r % bef + jj * bef + (r / bef) * bef * d.nout;
    if (d.acc) s = cadd(s, O[o]);
    O[o] = s;
  }
}

template <typename T, typename TM, typename TO>
__global__ void __launch_bounds__(64 * kWaves)
    gm_grouped(const GmDesc* __restrict__ descs, const int32_t* __restrict__ wgdesc, int ndesc, int64_t ntasks,
               const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1, const cplx<TM>* __restrict__ M,
               cplx<TO>* O, cplx<TO>* part, const uint8_t* __restrict__ active, double sre, double sim) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (task >= ntasks) return;  // uniform per wavefront
  const GmDesc d = descs[task_desc(descs, ndesc, wgdesc, task)];
  if (active && d.node >= 0 && !active[d.node]) return;
  gm_task<T, TM, TO>(d, task - d.tile0, d.xsel ? X1 : X0, M, O, part, sre, sim);
}

template <typename T>
__global__ void __launch_bounds__(64 * kWaves)
    gm_single(GmDesc d, int64_t ntasks, const cplx<T>* __restrict__ X, const cplx<T>* __restrict__ M, cplx<T>* O) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (task >= ntasks) return;
  gm_task<T, T, T>(d, task, X, M, O, O, 1.0, 0.0);  // no K split here: the workspace is never written
}

// --------------------------------------------------------- hole contraction
// Output tile: rows j (bra index), columns l (ket index); K = flattened (b, a).
template <typename T, typename TO>
__device__ inline void gc_task(const GcDesc& d, int64_t t, const cplx<T>* __restrict__ X,
                               const cplx<T>* __restrict__ Y, cplx<TO>* O, cplx<TO>* work) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63;
  const int split = d.ksplit > 1 ? d.ksplit : 1;
  const int ks = (int)(t % split);
  t /= split;
  const int TJ = (d.nx + 15) >> 4;
  const int jt = (int)(t % TJ), lt = (int)(t / TJ);
  const int64_t bef = d.bef, K = bef * d.aft;
  const int nx = d.nx, ny = d.ny;
  // K range of this chunk (a multiple of 4 long, except the last)
  const int64_t chunk = ((K + split - 1) / split + 3) & ~(int64_t)3;
  const int64_t kbeg = ks * chunk, kend = std::min(K, kbeg + chunk);
  const int li = lane & 15, kk = lane >> 4;
  const int jx = jt * 16 + li, ly = lt * 16 + li;
  const bool xv = jx < nx, yv = ly < ny;
  const cplx<T>* xp = X + d.x + (int64_t)jx * bef;
  const cplx<T>* yp = Y + d.y + (int64_t)ly * bef;
  const int64_t xs = bef * nx, ys = bef * ny;  // stride of a
  int64_t kc = kbeg + kk;
  int64_t b = kc % bef, a = kc / bef;
  v4 cr0 = {0, 0, 0, 0}, ci0 = {0, 0, 0, 0}, cr1 = {0, 0, 0, 0}, ci1 = {0, 0, 0, 0};
  const cplx<T> z = czero<T>();
  // four k-steps per iteration; the loads of the next iteration are issued
  // before the MFMAs of the current one (double-buffered registers)
  auto load4 = [&](cplx<T>* xr, cplx<T>* yr) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool kv = kc < kend;
      xr[u] = (xv && kv) ? xp[b + a * xs] : z;
      yr[u] = (yv && kv) ? yp[b + a * ys] : z;
      kc += 4;
      b += 4;
      if (b >= bef) {
        const int64_t q = b / bef;
        a += q;
        b -= q * bef;
      }
    }
  };
  cplx<T> xc[4], yc[4];
  if (kbeg < kend) load4(xc, yc);
  for (int64_t k0 = kbeg; k0 < kend; k0 += 16) {
    cplx<T> xn[4], yn[4];
    const bool more = k0 + 16 < kend;
    if (more) load4(xn, yn);
#pragma unroll
    for (int u = 0; u < 4; ++u) xc[u].im = -xc[u].im;
    cmma<T>(cr0, ci0, xc[0], yc[0]);
    if (k0 + 4 < kend) cmma<T>(cr1, ci1, xc[1], yc[1]);
    if (k0 + 8 < kend) cmma<T>(cr0, ci0, xc[2], yc[2]);
    if (k0 + 12 < kend) cmma<T>(cr1, ci1, xc[3], yc[3]);
    if (more) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xc[u] = xn[u];
        yc[u] = yn[u];
      }
    }
  }
  const int col = ly;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = jt * 16 + MF<T>::drow(lane, q);
    if (row < nx && yv) {
      cplx<TO> v = mk<TO>((TO)(cr0[q] + cr1[q]), (TO)(ci0[q] + ci1[q]));
      const int64_t e = row + (int64_t)col * nx;
      if (split > 1) {
        work[d.part + (int64_t)ks * nx * ny + e] = v;
      } else {
        if (d.acc) v = cadd(v, O[d.out + e]);
        O[d.out + e] = v;
      }
    }
  }
}

template <typename T, typename TO>
__global__ void __launch_bounds__(64 * kWaves)
    gc_grouped(const GcDesc* __restrict__ descs, const int32_t* __restrict__ wgdesc, int ndesc, int64_t ntasks,
               const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1, cplx<TO>* O, cplx<TO>* work,
               const uint8_t* __restrict__ active) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (task >= ntasks) return;
  const GcDesc d = descs[task_desc(descs, ndesc, wgdesc, task)];
  if (active && d.node >= 0 && !active[d.node]) return;
  gc_task<T, TO>(d, task - d.tile0, (d.sel & 1) ? X1 : X0, (d.sel & 2) ? X1 : X0, O, work);
}

template <typename T>
__global__ void __launch_bounds__(64 * kWaves)
    gc_single(GcDesc d, int64_t ntasks, const cplx<T>* __restrict__ X, const cplx<T>* __restrict__ Y, cplx<T>* O) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (task >= ntasks) return;
  gc_task<T, T>(d, task, X, Y, O, O);  // no K split here: the workspace is never written
}

// as gm_reduce_kernel, for hole contractions
template <typename T>
__global__ void __launch_bounds__(256)
    gc_reduce_kernel(const GcDesc* __restrict__ descs, int ndesc, cplx<T>* O, const cplx<T>* __restrict__ work,
                     const uint8_t* __restrict__ active) {
  const GcDesc d = descs[find_desc(descs, ndesc, blockIdx.x)];
  if (active && d.node >= 0 && !active[d.node]) return;
  const int64_t n = (int64_t)d.nx * d.ny;
  const int64_t e = (blockIdx.x - d.tile0) * 256 + threadIdx.x;
  if (e < n) {
    cplx<T> s = work[d.part + e];
    for (int c = 1; c < d.ksplit; ++c) s = cadd(s, work[d.part + c * n + e]);
    if (d.acc) s = cadd(s, O[d.out + e]);
    O[d.out + e] = s;
  }
}

// ------------------------------------------------------------------- solve
// One workgroup per system: Gauss-Jordan with partial pivoting on [S | X] in
// LDS; the pivot (largest modulus, lowest row on ties) comes from an argmax
// over the first wavefront.
__global__ void __launch_bounds__(256)
    solve_kernel(const SolveItem* __restrict__ items, cplx<double>* pool, const uint8_t* __restrict__ active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const SolveItem it = items[blockIdx.x];
  if (active && it.node >= 0 && !active[it.node]) return;
  const int n = it.n;
  cplx<double>* Sm = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* Xm = Sm + n * n;
  __shared__ int piv;
  __shared__ cplx<double> facs[64];
  for (int o = threadIdx.x; o < n * n; o += 256) {
    Sm[o] = pool[it.s + o];
    Xm[o] = pool[it.x + o];
  }
  __syncthreads();
  for (int col = 0; col < n; ++col) {
    if (threadIdx.x < 64) {
      double best = -1.0;
      int bi = col;
      for (int r = col + (int)threadIdx.x; r < n; r += 64) {
        const cplx<double> v = Sm[r + col * n];
        const double a = v.re * v.re + v.im * v.im;
        if (a > best) {
          best = a;
          bi = r;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (threadIdx.x == 0) piv = bi;
    }
    __syncthreads();
    const int p = piv;
    if (p != col) {
      for (int k = threadIdx.x; k < 2 * n; k += 256) {
        cplx<double>* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
        const cplx<double> tmp = Mk[col];
        Mk[col] = Mk[p];
        Mk[p] = tmp;
      }
      __syncthreads();
    }
    const cplx<double> dg = Sm[col + col * n];
    const double dd = dg.re * dg.re + dg.im * dg.im;
    const cplx<double> dinv = mk<double>(dg.re / dd, -dg.im / dd);
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * n; k += 256) {
      cplx<double>* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
      Mk[col] = cmul(Mk[col], dinv);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += 256) facs[r] = Sm[r + col * n];
    __syncthreads();
    for (int o = threadIdx.x; o < n * 2 * n; o += 256) {
      const int r = o % n, k = o / n;
      if (r == col) continue;
      cplx<double>* Mk = (k < n) ? Sm + k * n : Xm + (k - n) * n;
      const cplx<double> fr = facs[r];
      const cplx<double> v = Mk[col];
      cplx<double> w = Mk[r];
      w.re -= fr.re * v.re - fr.im * v.im;
      w.im -= fr.re * v.im + fr.im * v.re;
      Mk[r] = w;
    }
    __syncthreads();
  }
  for (int o = threadIdx.x; o < n * n; o += 256) pool[it.x + o] = Xm[o];
  if (it.r < 0) return;
  // W(k,m) = sum_l R(k,l) X(m,l)
  const cplx<double>* Rm = pool + it.r;
  for (int o = threadIdx.x; o < n * n; o += 256) {
    const int k = o % n, m = o / n;
    cplx<double> acc = czero<double>();
    for (int l = 0; l < n; ++l) cfma(acc, Rm[k + l * n], Xm[m + l * n]);
    pool[it.w + o] = acc;
  }
}

}  // namespace

template <typename T, typename TM, typename TO>
int launch_gm(const GmDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, const void* mbase, void* obase, void* part, const uint8_t* active, double sre,
              double sim, hipStream_t s, const KernelTiming* kt) {
  if (ntasks <= 0 || ndesc <= 0) return QTG_OK;
  const unsigned blocks = (unsigned)((ntasks + kWaves - 1) / kWaves);
  launch_timed(gm_grouped<T, TM, TO>, dim3(blocks), dim3(64 * kWaves), 0, s, kt, descs, wgdesc, (int)ndesc, ntasks,
               (const cplx<T>*)x0, (const cplx<T>*)x1, (const cplx<TM>*)mbase, (cplx<TO>*)obase, (cplx<TO>*)part,
               active, sre, sim);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T, typename TO>
int launch_gc(const GcDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, void* obase, void* work, const uint8_t* active, hipStream_t s,
              const KernelTiming* kt) {
  if (ntasks <= 0 || ndesc <= 0) return QTG_OK;
  const unsigned blocks = (unsigned)((ntasks + kWaves - 1) / kWaves);
  launch_timed(gc_grouped<T, TO>, dim3(blocks), dim3(64 * kWaves), 0, s, kt, descs, wgdesc, (int)ndesc, ntasks,
               (const cplx<T>*)x0, (const cplx<T>*)x1, (cplx<TO>*)obase, (cplx<TO>*)work, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_gm_single(const GmDesc& d, const void* x, const void* m, void* o, hipStream_t s) {
  const int64_t nt = gm_tasks(d);
  if (nt <= 0) return QTG_OK;
  hipLaunchKernelGGL(gm_single<T>, dim3((unsigned)((nt + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0, s, d, nt,
                     (const cplx<T>*)x, (const cplx<T>*)m, (cplx<T>*)o);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_gc_single(const GcDesc& d, const void* x, const void* y, void* o, hipStream_t s) {
  if (d.ksplit > 1) {
    set_error("single hole contraction: no K split without a workspace");
    return QTG_EARG;
  }
  const int64_t nt = gc_tasks(d);
  if (nt <= 0) return QTG_OK;
  hipLaunchKernelGGL(gc_single<T>, dim3((unsigned)((nt + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0, s, d, nt,
                     (const cplx<T>*)x, (const cplx<T>*)y, (cplx<T>*)o);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename TO>
int launch_gm_reduce(const GmDesc* descs, int64_t ndesc, int64_t nblocks, void* obase, const void* part,
                     const uint8_t* active, hipStream_t s, const KernelTiming* kt) {
  if (ndesc <= 0 || nblocks <= 0) return QTG_OK;
  launch_timed(gm_reduce_kernel<TO>, dim3((unsigned)nblocks), dim3(256), 0, s, kt, descs, (int)ndesc,
               (cplx<TO>*)obase, (const cplx<TO>*)part, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_gc_reduce(const GcDesc* descs, int64_t ndesc, int64_t nblocks, void* obase, const void* work,
                     const uint8_t* active, hipStream_t s, const KernelTiming* kt) {
  if (ndesc <= 0 || nblocks <= 0) return QTG_OK;
  launch_timed(gc_reduce_kernel<T>, dim3((unsigned)nblocks), dim3(256), 0, s, kt, descs, (int)ndesc,
               (cplx<T>*)obase, (const cplx<T>*)work, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_solve(const SolveItem* items, int64_t n, cplx<double>* pool, int nmax, const uint8_t* active,
                 hipStream_t s, const KernelTiming* kt) {
  if (n <= 0) return QTG_OK;
  const size_t lds = 2 * sizeof(cplx<double>) * (size_t)nmax * nmax;
  launch_timed(solve_kernel, dim3((unsigned)n), dim3(256), (unsigned)lds, s, kt, items, pool, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template int launch_gm<double, double, double>(const GmDesc*, const int32_t*, int64_t, int64_t, const void*,
                                               const void*, const void*, void*, void*, const uint8_t*, double,
                                               double, hipStream_t, const KernelTiming*);
template int launch_gm<float, float, float>(const GmDesc*, const int32_t*, int64_t, int64_t, const void*,
                                            const void*, const void*, void*, void*, const uint8_t*, double, double,
                                            hipStream_t, const KernelTiming*);
template int launch_gm<float, double, float>(const GmDesc*, const int32_t*, int64_t, int64_t, const void*,
                                             const void*, const void*, void*, void*, const uint8_t*, double,
                                             double, hipStream_t, const KernelTiming*);
template int launch_gm<float, double, double>(const GmDesc*, const int32_t*, int64_t, int64_t, const void*,
                                              const void*, const void*, void*, void*, const uint8_t*, double,
                                              double, hipStream_t, const KernelTiming*);
template int launch_gc<double, double>(const GcDesc*, const int32_t*, int64_t, int64_t, const void*, const void*,
                                       void*, void*, const uint8_t*, hipStream_t, const KernelTiming*);
template int launch_gc<float, float>(const GcDesc*, const int32_t*, int64_t, int64_t, const void*, const void*,
                                     void*, void*, const uint8_t*, hipStream_t, const KernelTiming*);
template int launch_gc<float, double>(const GcDesc*, const int32_t*, int64_t, int64_t, const void*, const void*,
                                      void*, void*, const uint8_t*, hipStream_t, const KernelTiming*);
template int launch_gm_single<double>(const GmDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gm_single<float>(const GmDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gc_single<double>(const GcDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gc_single<float>(const GcDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gm_reduce<double>(const GmDesc*, int64_t, int64_t, void*, const void*, const uint8_t*,
                                      hipStream_t, const KernelTiming*);
template int launch_gm_reduce<float>(const GmDesc*, int64_t, int64_t, void*, const void*, const uint8_t*,
                                     hipStream_t, const KernelTiming*);
template int launch_gc_reduce<double>(const GcDesc*, int64_t, int64_t, void*, const void*, const uint8_t*,
                                      hipStream_t, const KernelTiming*);
template int launch_gc_reduce<float>(const GcDesc*, int64_t, int64_t, void*, const void*, const uint8_t*,
                                     hipStream_t, const KernelTiming*);

}  // namespace qtg
