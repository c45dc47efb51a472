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
#include <type_traits>

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

constexpr int kWaves = kGemmWaves;  // wavefronts (tasks) per workgroup

// Wavefront index in the workgroup, as a wave-uniform (scalar) value, so that
// descriptor loads and address arithmetic stay on the scalar unit.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// base[off] when ok, else zero.  The load itself is unconditional, from
// base[0] when !ok (base is the start of an operand, always in bounds): a
// guarded load would let the compiler select between the operand and a stack
// zero and fall back to generic (flat) loads.
template <typename T>
__device__ inline cplx<T> ldz(const cplx<T>* __restrict__ base, int64_t off, bool ok) {
  const cplx<T> v = base[ok ? off : 0];
  return ok ? v : czero<T>();
}

// Descriptor of a task: one load from the per-task table built on the host.
__device__ inline int task_desc(const int32_t* __restrict__ wgdesc, int64_t task) { return wgdesc[task]; }

// Split-K inside a workgroup: the ksplit (2 or 4) chunks of a split tile are
// consecutive tasks starting at a multiple of ksplit, so they are wavefronts
// of one workgroup.  Each leaves its (alpha-scaled, output-precision) partial
// tile in LDS and the chunk-0 wavefront sums them in chunk order after the
// workgroup barrier: deterministic, no atomics, no second launch.  Every
// wavefront of a workgroup reaches the barrier, including idle ones.
template <typename TO>
__device__ inline void reduce_chunks(cplx<TO>* red, bool mine, int split, int ks, cplx<TO> (&v)[4]) {
  const int wave = wave_id(), lane = threadIdx.x & 63;
  if (mine && split > 1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[wave * 256 + q * 64 + lane] = v[q];
  }
  __syncthreads();
  if (mine && split > 1 && ks == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cplx<TO> x = red[wave * 256 + q * 64 + lane];
      for (int c = 1; c < split; ++c) x = cadd(x, red[(wave + c) * 256 + q * 64 + lane]);
      v[q] = x;
    }
  }
}

// ------------------------------------------------------------- mode product
// Output tile: 16 rows of the flattened (b, a) index x 16 output columns i.
// Tile index and K chunk of task t of descriptor d.
struct GmTile {
  int64_t rt;
  int ct, ks;
};
// A descriptor's tasks number fewer than 2^31 (a launch's grid), so the tile
// arithmetic runs in 32 bits; ksplit is a power of two.
__device__ inline GmTile gm_tile(const GmDesc& d, int64_t t64) {
  const unsigned split = d.ksplit > 1 ? (unsigned)d.ksplit : 1u;
  unsigned t = (unsigned)t64;
  GmTile g;
  g.ks = (int)(t & (split - 1));
  t >>= __builtin_ctz(split);
  const unsigned CT = (unsigned)(d.nout + 15) >> 4;
  const unsigned rt = t / CT;
  g.rt = rt;
  g.ct = (int)(t - rt * CT);
  return g;
}

// (b, a) = (r % bef, r / bef) of a flattened (b, a) row: 32-bit division
// when the rows of the view fit in 31 bits (every node tensor), else 64-bit.
// R is uniform across the wavefront, so is the branch.
__device__ inline void row_split(int64_t r, int64_t bef, int64_t R, int64_t& b, int64_t& a) {
  if (R < ((int64_t)1 << 31)) {
    const unsigned ur = (unsigned)r, ub = (unsigned)bef, q = ur / ub;
    a = q;
    b = ur - q * ub;
  } else {
    a = r / bef;
    b = r - a * bef;
  }
}

// out(b,i,a) = alpha X(b,i,a) for one 16 x 16 tile, four columns per lane
template <typename T, typename TO>
__device__ inline void gm_copy(const GmDesc& d, const GmTile& g, const cplx<T>* __restrict__ X, cplx<TO>* O,
                               double sre, double sim) {
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const int64_t bef = d.bef, R = bef * d.aft;
  const T are = (T)(d.are * sre - d.aim * sim), aim = (T)(d.are * sim + d.aim * sre);
  for (int q = 0; q < 4; ++q) {
    const int64_t r = g.rt * 16 + li;
    const int col = g.ct * 16 + kk * 4 + q;
    if (r < R && col < d.nout) {
      int64_t b, a;
      row_split(r, bef, R, b, a);
      const cplx<T> x = X[d.x + b + (int64_t)col * bef + a * bef * d.nin];
      const int64_t o = d.out + b + (int64_t)col * bef + a * bef * d.nout;
      cplx<TO> v = mk<TO>((TO)(are * x.re - aim * x.im), (TO)(are * x.im + aim * x.re));
      if (d.acc) v = cadd(v, O[o]);
      O[o] = v;
    }
  }
}

// Output rows are contiguous in memory in runs of bef elements.  With
// bef >= 4 the tile is accumulated transposed (the operands swap roles in
// the MFMA), so that a lane holds one row and the 16 lanes of a store write
// 16 consecutive rows; with bef < 4 a lane holds one column (consecutive
// columns are bef apart).
__device__ inline bool gm_rows_on_lanes(const GmDesc& d) { return d.bef >= 4; }

// This chunk's share of the product tile, alpha-scaled, in the output
// precision: v[q] is row rt*16 + lane % 16, column ct*16 + drow(lane, q)
// when gm_rows_on_lanes, else row rt*16 + drow(lane, q), column
// ct*16 + lane % 16.
template <typename T, typename TM, typename TO>
__device__ inline void gm_product(const GmDesc& d, const GmTile& g, const cplx<T>* __restrict__ X,
                                  const cplx<TM>* __restrict__ M, double sre, double sim, cplx<TO> (&v)[4]) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63;
  const int split = d.ksplit > 1 ? d.ksplit : 1;
  const int64_t bef = d.bef, R = bef * d.aft;
  const int nin = d.nin, nout = d.nout;
  const int li = lane & 15, kk = lane >> 4;
  const T are = (T)(d.are * sre - d.aim * sim), aim = (T)(d.are * sim + d.aim * sre);
  // A operand: row r = rt*16 + li of the flattened (b, a), column k
  const int64_t r = g.rt * 16 + li;
  const bool rv = r < R;
  int64_t rb, ra;
  row_split(rv ? r : 0, bef, R, rb, ra);
  const int64_t xoff = rb + ra * bef * nin;
  const cplx<T>* xa = X + d.x;
  // B operand: op(M)(j, k) for column j = ct*16 + li
  const int j = g.ct * 16 + li;
  const bool cv = j < nout;
  const cplx<TM>* mb = M + d.m;
  // M(j, k) at j + k nout (N), M(k, j) at k + j nin (T, H)
  const int64_t moff = !cv ? 0 : d.op == GM_N ? j : (int64_t)j * nin;
  const int64_t mks = d.op == GM_N ? nout : 1;
  const bool conjm = d.op == GM_H;
  // K range of this chunk (a multiple of 4 long, except the last)
  const int kchunk = ((nin + split - 1) / split + 3) & ~3;
  const int kbeg = g.ks * kchunk, kend = min(nin, kbeg + kchunk);
  v4 cr0 = {0, 0, 0, 0}, ci0 = {0, 0, 0, 0}, cr1 = {0, 0, 0, 0}, ci1 = {0, 0, 0, 0};
  // A lane's X value is element (row li, k kk) of X, its M value element
  // (k kk, column li) of op(M)^T: passed as the MFMA's (A, B) they give the
  // tile, as (B, A) its transpose.
  const bool tr = gm_rows_on_lanes(d);
  // four k-steps per iteration: all loads are issued before the MFMAs
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    cplx<T> av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kx = k0 + 4 * u + kk;
      const bool kv = kx < kend;
      const cplx<T> xv = ldz(xa, xoff + (int64_t)kx * bef, rv && kv);
      const cplx<TM> mv = ldz(mb, moff + (int64_t)kx * mks, cv && kv);
      const cplx<T> m = mk<T>((T)mv.re, conjm ? -(T)mv.im : (T)mv.im);
      av[u] = tr ? m : xv;
      bv[u] = tr ? xv : m;
    }
    cmma<T>(cr0, ci0, av[0], bv[0]);
    if (k0 + 4 < kend) cmma<T>(cr1, ci1, av[1], bv[1]);
    if (k0 + 8 < kend) cmma<T>(cr0, ci0, av[2], bv[2]);
    if (k0 + 12 < kend) cmma<T>(cr1, ci1, av[3], bv[3]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T vr = cr0[q] + cr1[q], vi = ci0[q] + ci1[q];
    v[q] = mk<TO>((TO)(are * vr - aim * vi), (TO)(are * vi + aim * vr));
  }
}

template <typename T, typename TO>
__device__ inline void gm_store(const GmDesc& d, const GmTile& g, const cplx<TO> (&v)[4], cplx<TO>* O) {
  const int lane = threadIdx.x & 63;
  const int64_t bef = d.bef, R = bef * d.aft;
  if (gm_rows_on_lanes(d)) {
    // one row per lane, its four columns drow(lane, q)
    const int64_t r = g.rt * 16 + (lane & 15);
    if (r >= R) return;
    int64_t b, a;
    row_split(r, bef, R, b, a);
    const int64_t o0 = d.out + b + a * bef * d.nout;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int jq = g.ct * 16 + MF<T>::drow(lane, q);
      if (jq < d.nout) {
        const int64_t o = o0 + (int64_t)jq * bef;
        cplx<TO> x = v[q];
        if (d.acc) x = cadd(x, O[o]);
        O[o] = x;
      }
    }
    return;
  }
  const int j = g.ct * 16 + (lane & 15);
  if (j >= d.nout) return;
  // the lane's four rows are rr0 + q * st: (b, a) of the first by division,
  // the others by a carry step
  const int st = MF<T>::drow(0, 1);
  int64_t rr = g.rt * 16 + MF<T>::drow(lane, 0);
  int64_t b, a;
  row_split(rr < R ? rr : 0, bef, R, b, a);
  const int qs = st / (int)min(bef, (int64_t)64), rs = st % (int)min(bef, (int64_t)64);
  const int64_t ostr = bef * d.nout;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (rr < R) {
      const int64_t o = d.out + b + (int64_t)j * bef + a * ostr;
      cplx<TO> x = v[q];
      if (d.acc) x = cadd(x, O[o]);
      O[o] = x;
    }
    rr += st;
    b += rs;
    a += qs;
    if (b >= bef) {
      b -= bef;
      ++a;
    }
  }
}

// ------------------------------------------------- mode product, lean form
// A descriptor whose three tensors each span fewer than 2^27 elements (every
// node tensor of the tree sweeps) takes this form: 32-bit element offsets
// from wave-uniform bases (global loads and stores with a scalar base and a
// 32-bit lane offset), rows and columns beyond the tile's range clamped into
// it instead of predicated (their products land in tile rows and columns
// that are never stored), and only the K tail zeroed.  The tile is
// accumulated transposed when bef >= 4 (gm_rows_on_lanes).

// base[idx] for idx < 2^27: a 32-bit byte offset from the base
template <typename V>
__device__ inline V ld32(const V* base, uint32_t idx) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + (idx * (uint32_t)sizeof(V)));
}
template <typename V>
__device__ inline V* at32(V* base, uint32_t idx) {
  return reinterpret_cast<V*>(reinterpret_cast<char*>(base) + (idx * (uint32_t)sizeof(V)));
}

// Lean tile coordinates: no division when the output has one column tile
__device__ inline GmTile gm_tile_lean(const GmDesc& d, uint32_t t) {
  GmTile g;
  const int lg = d.ksplit > 1 ? __builtin_ctz((unsigned)d.ksplit) : 0;
  g.ks = (int)(t & ((1u << lg) - 1u));
  t >>= lg;
  const unsigned CT = (unsigned)(d.nout + 15) >> 4;
  if (CT == 1) {
    g.rt = t;
    g.ct = 0;
  } else {
    const unsigned rt = t / CT;
    g.rt = rt;
    g.ct = (int)(t - rt * CT);
  }
  return g;
}

// The K loop of the lean form.  TR: the tile is accumulated transposed (X
// as the MFMA's B operand); CONJ: op(M) = M^H.  One accumulator pair: the
// MFMAs of a k-step alternate between the real and imaginary part, and
// consecutive MFMAs into the same accumulator chain without a stall.
template <typename T, typename TM, bool TR, bool CONJ>
__device__ inline void gm_lean_loop(const cplx<T>* __restrict__ xa, uint32_t xrow, uint32_t bef,
                                    const cplx<TM>* __restrict__ mb, uint32_t mrow, uint32_t mks, int kbeg, int kend,
                                    int kk, typename MF<T>::v4& cr, typename MF<T>::v4& ci) {
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    cplx<T> xv[4], mv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kx = k0 + 4 * u + kk;
      const bool kv = kx < kend;
      const uint32_t kc = (uint32_t)(kv ? kx : kend - 1);
      const cplx<T> x = ld32(xa, xrow + kc * bef);
      const cplx<TM> m = ld32(mb, mrow + kc * mks);
      xv[u] = mk<T>(kv ? x.re : (T)0, kv ? x.im : (T)0);
      mv[u] = mk<T>((T)m.re, CONJ ? -(T)m.im : (T)m.im);
    }
    // every operand is taken here, so the compiler keeps the eight loads
    // together ahead of the MFMAs instead of sinking them into the k-step
    // guards below (one memory round trip per iteration, not four)
#pragma unroll
    for (int u = 0; u < 4; ++u) asm volatile("" ::"v"(xv[u].re), "v"(xv[u].im), "v"(mv[u].re), "v"(mv[u].im));
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u > 0 && k0 + 4 * u >= kend) break;
      const cplx<T>& a = TR ? mv[u] : xv[u];
      const cplx<T>& b = TR ? xv[u] : mv[u];
      cr = MF<T>::mma(a.re, b.re, cr);
      ci = MF<T>::mma(a.re, b.im, ci);
      cr = MF<T>::mma(-a.im, b.im, cr);
      ci = MF<T>::mma(a.im, b.re, ci);
    }
  }
}

template <typename T, typename TM, typename TO>
__device__ inline void gm_lean(const GmDesc& d, const GmTile& g, const cplx<T>* __restrict__ X,
                               const cplx<TM>* __restrict__ M, double sre, double sim, cplx<TO> (&v)[4]) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const uint32_t bef = (uint32_t)d.bef, R = bef * (uint32_t)d.aft;
  const int nin = d.nin, nout = d.nout;
  const int lg = d.ksplit > 1 ? __builtin_ctz((unsigned)d.ksplit) : 0;
  // K range of this chunk (a multiple of 4 long, except the last)
  const int kchunk = (((nin + (1 << lg) - 1) >> lg) + 3) & ~3;
  const int kbeg = g.ks * kchunk, kend = min(nin, kbeg + kchunk);
  // the lane's row (clamped) and its (b, a)
  const uint32_t r = (uint32_t)g.rt * 16u + (uint32_t)li;
  const uint32_t rc = min(r, R - 1u);
  const uint32_t a = rc / bef, b = rc - a * bef;
  const uint32_t xrow = b + a * bef * (uint32_t)nin;
  // the lane's column of op(M) (clamped): M(j, k) at j + k nout (N), M(k, j) at k + j nin (T, H)
  const int j = min(g.ct * 16 + li, nout - 1);
  const uint32_t mrow = d.op == GM_N ? (uint32_t)j : (uint32_t)j * (uint32_t)nin;
  const uint32_t mks = d.op == GM_N ? (uint32_t)nout : 1u;
  const cplx<T>* xa = X + d.x;
  const cplx<TM>* mb = M + d.m;
  v4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  if (gm_rows_on_lanes(d)) {
    if (d.op == GM_H) gm_lean_loop<T, TM, true, true>(xa, xrow, bef, mb, mrow, mks, kbeg, kend, kk, cr, ci);
    else gm_lean_loop<T, TM, true, false>(xa, xrow, bef, mb, mrow, mks, kbeg, kend, kk, cr, ci);
  } else {
    if (d.op == GM_H) gm_lean_loop<T, TM, false, true>(xa, xrow, bef, mb, mrow, mks, kbeg, kend, kk, cr, ci);
    else gm_lean_loop<T, TM, false, false>(xa, xrow, bef, mb, mrow, mks, kbeg, kend, kk, cr, ci);
  }
  const double fre = d.are * sre - d.aim * sim, fim = d.are * sim + d.aim * sre;
  if (fre == 1.0 && fim == 0.0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = mk<TO>((TO)cr[q], (TO)ci[q]);
  } else {
    const T are = (T)fre, aim = (T)fim;
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = mk<TO>((TO)(are * cr[q] - aim * ci[q]), (TO)(are * ci[q] + aim * cr[q]));
  }
}

// out(b,i,a) = alpha X(b,i,a) for one 16 x 16 tile in the lean form: one
// row per lane (16 consecutive rows per column), four columns per lane
template <typename T, typename TO>
__device__ inline void gm_lean_copy(const GmDesc& d, const GmTile& g, const cplx<T>* __restrict__ X, cplx<TO>* O,
                                    double sre, double sim) {
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const uint32_t bef = (uint32_t)d.bef, R = bef * (uint32_t)d.aft, nout = (uint32_t)d.nout;
  const uint32_t r = (uint32_t)g.rt * 16u + (uint32_t)li;
  if (r >= R) return;
  const uint32_t a = r / bef, b = r - a * bef;
  const uint32_t xo = b + a * bef * (uint32_t)d.nin, oo = b + a * bef * nout;
  const cplx<T>* xa = X + d.x;
  cplx<TO>* ob = O + d.out;
  const T are = (T)(d.are * sre - d.aim * sim), aim = (T)(d.are * sim + d.aim * sre);
  cplx<T> x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t col = min((uint32_t)(g.ct * 16 + kk * 4 + q), nout - 1u);
    x[q] = ld32(xa, xo + col * bef);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t col = (uint32_t)(g.ct * 16 + kk * 4 + q);
    if (col < nout) {
      cplx<TO>* o = at32(ob, oo + col * bef);
      cplx<TO> v = mk<TO>((TO)(are * x[q].re - aim * x[q].im), (TO)(are * x[q].im + aim * x[q].re));
      if (d.acc) v = cadd(v, *o);
      *o = v;
    }
  }
}

template <typename T, typename TO>
__device__ inline void gm_lean_store(const GmDesc& d, const GmTile& g, const cplx<TO> (&v)[4], cplx<TO>* O) {
  const int lane = threadIdx.x & 63;
  const uint32_t bef = (uint32_t)d.bef, R = bef * (uint32_t)d.aft, nout = (uint32_t)d.nout;
  cplx<TO>* ob = O + d.out;
  if (gm_rows_on_lanes(d)) {
    // one row per lane, its four columns drow(lane, q)
    const uint32_t r = (uint32_t)g.rt * 16u + (uint32_t)(lane & 15);
    if (r >= R) return;
    const uint32_t a = r / bef, b = r - a * bef;
    const uint32_t o0 = b + a * bef * nout;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t jq = (uint32_t)(g.ct * 16 + MF<T>::drow(lane, q));
      if (jq < nout) {
        cplx<TO>* o = at32(ob, o0 + jq * bef);
        cplx<TO> x = v[q];
        if (d.acc) x = cadd(x, *o);
        *o = x;
      }
    }
    return;
  }
  // bef < 4: one column per lane, its four rows drow(lane, q)
  const uint32_t j = (uint32_t)(g.ct * 16 + (lane & 15));
  if (j >= nout) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t rr = (uint32_t)g.rt * 16u + (uint32_t)MF<T>::drow(lane, q);
    if (rr < R) {
      const uint32_t a = rr / bef, b = rr - a * bef;
      cplx<TO>* o = at32(ob, b + j * bef + a * bef * nout);
      cplx<TO> x = v[q];
      if (d.acc) x = cadd(x, *o);
      *o = x;
    }
  }
}

// LEAN: every descriptor takes the lean form (checked on the host)
template <typename T, typename TM, typename TO, int W, bool LEAN>
__global__ void __launch_bounds__(64 * W)
    gm_grouped(const GmDesc* __restrict__ descs, const int32_t* __restrict__ wgdesc, int ndesc, int64_t ntasks,
               const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1, const cplx<TM>* __restrict__ M,
               cplx<TO>* O, const uint8_t* __restrict__ active, double sre, double sim) {
  __shared__ cplx<TO> red[W * 256];
  const int64_t task = (int64_t)blockIdx.x * W + wave_id();
  bool mine = task < ntasks;  // uniform per wavefront
  GmDesc d{};
  if (mine) {
    d = descs[wgdesc ? task_desc(wgdesc, task) : task];
    mine = !(active && d.node >= 0 && !active[d.node]);
  }
  const cplx<T>* X = d.xsel ? X1 : X0;
  GmTile g{0, 0, 0};
  const int split = mine && d.ksplit > 1 ? d.ksplit : 1;
  cplx<TO> v[4];
  const bool lean = mine && (LEAN || gm_lean_ok(d));
  if (lean) {
    g = gm_tile_lean(d, (uint32_t)(task - d.tile0));
    if (d.op == GM_COPY) {
      gm_lean_copy<T, TO>(d, g, X, O, sre, sim);
      mine = false;
    } else {
      gm_lean<T, TM, TO>(d, g, X, M, sre, sim, v);
    }
  } else if (!LEAN && mine) {
    g = gm_tile(d, task - d.tile0);
    if (d.op == GM_COPY) {
      gm_copy<T, TO>(d, g, X, O, sre, sim);
      mine = false;
    } else {
      gm_product<T, TM, TO>(d, g, X, M, sre, sim, v);
    }
  }
  reduce_chunks<TO>(red, mine, split, g.ks, v);
  if (mine && g.ks == 0) {
    if (LEAN || lean) gm_lean_store<T, TO>(d, g, v, O);
    else gm_store<T, TO>(d, g, v, O);
  }
}

// one descriptor, no K split
template <typename T>
__global__ void __launch_bounds__(64 * kWaves)
    gm_single(GmDesc d, int64_t ntasks, const cplx<T>* __restrict__ X, const cplx<T>* __restrict__ M, cplx<T>* O) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + wave_id();
  if (task >= ntasks) return;
  const GmTile g = gm_tile(d, task);
  if (d.op == GM_COPY) {
    gm_copy<T, T>(d, g, X, O, 1.0, 0.0);
    return;
  }
  cplx<T> v[4];
  gm_product<T, T, T>(d, g, X, M, 1.0, 0.0, v);
  gm_store<T, T>(d, g, v, O);
}

// ------------------------------------------------------ real mode product
// out(b,i,a) (+)= sum_l op(M)(i,l) X(b,l,a) on real FP64 data (the rmatvec /
// trmatvec kernels of LA_lib.f): the tile of gm_product with one real MFMA
// per k-step.  op is GM_N or GM_T.
__global__ void __launch_bounds__(64 * kWaves)
    gmr_single(GmDesc d, int64_t ntasks, const double* __restrict__ X, const double* __restrict__ M, double* O) {
  typedef MF<double>::v4 v4;
  const int64_t task = (int64_t)blockIdx.x * kWaves + wave_id();
  if (task >= ntasks) return;
  const GmTile g = gm_tile(d, task);
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const int64_t bef = d.bef, R = bef * d.aft;
  const int nin = d.nin, nout = d.nout;
  const int64_t r = g.rt * 16 + li;
  const bool rv = r < R;
  const int64_t xoff = rv ? r % bef + (r / bef) * bef * nin : 0;
  const int j = g.ct * 16 + li;
  const bool cv = j < nout;
  const int64_t moff = !cv ? 0 : d.op == GM_N ? j : (int64_t)j * nin;
  const int64_t mks = d.op == GM_N ? nout : 1;
  v4 c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
  for (int k0 = 0; k0 < nin; k0 += 8) {
    const int ka = k0 + kk, kb = k0 + 4 + kk;
    const bool va = ka < nin, vb = kb < nin;
    // unconditional in-bounds loads (element 0 when the lane is idle)
    const double xa = X[(rv && va) ? xoff + (int64_t)ka * bef : 0];
    const double ma = M[(cv && va) ? moff + (int64_t)ka * mks : 0];
    const double xb = X[(rv && vb) ? xoff + (int64_t)kb * bef : 0];
    const double mb = M[(cv && vb) ? moff + (int64_t)kb * mks : 0];
    c0 = MF<double>::mma((rv && va) ? xa : 0.0, (cv && va) ? ma : 0.0, c0);
    if (k0 + 4 < nin) c1 = MF<double>::mma((rv && vb) ? xb : 0.0, (cv && vb) ? mb : 0.0, c1);
  }
  if (!cv) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t rr = g.rt * 16 + MF<double>::drow(lane, q);
    if (rr < R) {
      const int64_t b = rr % bef, a = rr / bef;
      const int64_t o = d.out + b + (int64_t)j * bef + a * bef * nout;
      const double v = c0[q] + c1[q];
      O[o] = d.acc ? O[o] + v : v;
    }
  }
}

// --------------------------------------------------------- hole contraction
// Output tile: rows j (bra index), columns l (ket index); K = flattened (b, a).
struct GcTile {
  int jt, lt, ks;
};
__device__ inline GcTile gc_tile(const GcDesc& d, int64_t t) {
  const int split = d.ksplit > 1 ? d.ksplit : 1;
  GcTile g;
  g.ks = (int)(t % split);
  t /= split;
  const int TJ = (d.nx + 15) >> 4;
  g.jt = (int)(t % TJ);
  g.lt = (int)(t / TJ);
  return g;
}

// This chunk's share of the tile in the output precision: v[q] is row
// jt*16 + drow(lane, q), column lt*16 + lane % 16.
template <typename T, typename TO>
__device__ inline void gc_product(const GcDesc& d, const GcTile& g, const cplx<T>* __restrict__ X,
                                  const cplx<T>* __restrict__ Y, cplx<TO> (&v)[4]) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63;
  const int split = d.ksplit > 1 ? d.ksplit : 1;
  // K = bef * aft < 2^31 for every node tensor, so the k walk runs in 32 bits
  const int bef = d.bef, K = d.bef * d.aft;
  const int nx = d.nx, ny = d.ny;
  // K range of this chunk (a multiple of 4 long, except the last)
  const int chunk = ((K + split - 1) / split + 3) & ~3;
  const int kbeg = g.ks * chunk, kend = min(K, kbeg + chunk);
  const int li = lane & 15, kk = lane >> 4;
  const int jx = g.jt * 16 + li, ly = g.lt * 16 + li;
  const bool xv = jx < nx, yv = ly < ny;
  const cplx<T>* xp = X + d.x;
  const cplx<T>* yp = Y + d.y;
  const int64_t xrow = (int64_t)jx * bef, yrow = (int64_t)ly * bef;
  const int xs = bef * nx, ys = bef * ny;  // stride of a
  // k = b + a bef advances by 4 per load: b by 4 % bef, a by 4 / bef, plus a carry
  const int q4 = 4 / bef, r4 = 4 % bef;
  int kc = kbeg + kk;
  int b = kc % bef, a = kc / bef;
  v4 cr0 = {0, 0, 0, 0}, ci0 = {0, 0, 0, 0}, cr1 = {0, 0, 0, 0}, ci1 = {0, 0, 0, 0};
  // four k-steps per iteration; the loads of the next iteration are issued
  // before the MFMAs of the current one (double-buffered registers)
  auto load4 = [&](cplx<T>* xr, cplx<T>* yr) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool kv = kc < kend;
      xr[u] = ldz(xp, xrow + b + (int64_t)a * xs, xv && kv);
      yr[u] = ldz(yp, yrow + b + (int64_t)a * ys, yv && kv);
      kc += 4;
      b += r4;
      a += q4;
      const bool carry = b >= bef;
      b = carry ? b - bef : b;
      a = carry ? a + 1 : a;
    }
  };
  cplx<T> xc[4], yc[4];
  if (kbeg < kend) load4(xc, yc);
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
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
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = mk<TO>((TO)(cr0[q] + cr1[q]), (TO)(ci0[q] + ci1[q]));
}

// ---------------------------------------------- hole contraction, lean form
// For X and Y tensors of fewer than 2^27 elements and a power-of-two K
// split: 32-bit element offsets from wave-uniform bases, the lane's bra row
// and ket column clamped into range (their products land in tile rows and
// columns that are never stored), and only the K tail zeroed.  The running
// position k = b + a bef of a lane advances by four per k-step as two
// 32-bit offsets (X and Y) with one carry.

__device__ inline GcTile gc_tile_lean(const GcDesc& d, uint32_t t) {
  const int lg = d.ksplit > 1 ? __builtin_ctz((unsigned)d.ksplit) : 0;
  GcTile g;
  g.ks = (int)(t & ((1u << lg) - 1u));
  t >>= lg;
  const unsigned TJ = (unsigned)(d.nx + 15) >> 4;
  if (TJ == 1) {
    g.jt = 0;
    g.lt = (int)t;
  } else {
    g.lt = (int)(t / TJ);
    g.jt = (int)(t - (unsigned)g.lt * TJ);
  }
  return g;
}

template <typename T, typename TO>
__device__ inline void gc_lean(const GcDesc& d, const GcTile& g, const cplx<T>* __restrict__ X,
                               const cplx<T>* __restrict__ Y, cplx<TO> (&v)[4]) {
  typedef typename MF<T>::v4 v4;
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const int lg = d.ksplit > 1 ? __builtin_ctz((unsigned)d.ksplit) : 0;
  const uint32_t bef = (uint32_t)d.bef, aft = (uint32_t)d.aft;
  const int K = d.bef * d.aft;
  // K range of this chunk (a multiple of 4 long, except the last)
  const int chunk = (((K + (1 << lg) - 1) >> lg) + 3) & ~3;
  const int kbeg = g.ks * chunk, kend = min(K, kbeg + chunk);
  const uint32_t jx = (uint32_t)min(g.jt * 16 + li, d.nx - 1), ly = (uint32_t)min(g.lt * 16 + li, d.ny - 1);
  const uint32_t xs = bef * (uint32_t)d.nx, ys = bef * (uint32_t)d.ny;  // stride of a
  const uint32_t xrow = jx * bef, yrow = ly * bef;
  // the last element of the lane's row: reads past the K range stay inside
  const uint32_t xlast = xrow + (bef - 1u) + (aft - 1u) * xs, ylast = yrow + (bef - 1u) + (aft - 1u) * ys;
  const cplx<T>* xp = X + d.x;
  const cplx<T>* yp = Y + d.y;
  int kc = kbeg + kk;
  uint32_t b = (uint32_t)kc % bef;
  const uint32_t a0 = (uint32_t)kc / bef;
  uint32_t px = b + a0 * xs, py = b + a0 * ys;
  const uint32_t q4 = 4u / bef, r4 = 4u % bef;
  const uint32_t dx = r4 + q4 * xs, dy = r4 + q4 * ys;
  v4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  // four k-steps per load: conj(X) zeroed past the chunk's end
  auto load4 = [&](cplx<T>* xr, cplx<T>* yr) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool kv = kc < kend;
      const cplx<T> x = ld32(xp, min(xrow + px, xlast));
      yr[u] = ld32(yp, min(yrow + py, ylast));
      xr[u] = mk<T>(kv ? x.re : (T)0, kv ? -x.im : (T)0);
      kc += 4;
      b += r4;
      px += dx;
      py += dy;
      if (b >= bef) {
        b -= bef;
        px += xs - bef;
        py += ys - bef;
      }
    }
  };
  cplx<T> xc[4], yc[4];
  if (kbeg < kend) load4(xc, yc);
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    cplx<T> xn[4], yn[4];
    const bool more = k0 + 16 < kend;
    if (more) load4(xn, yn);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u > 0 && k0 + 4 * u >= kend) break;
      cr = MF<T>::mma(xc[u].re, yc[u].re, cr);
      ci = MF<T>::mma(xc[u].re, yc[u].im, ci);
      cr = MF<T>::mma(-xc[u].im, yc[u].im, cr);
      ci = MF<T>::mma(xc[u].im, yc[u].re, ci);
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xc[u] = xn[u];
        yc[u] = yn[u];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = mk<TO>((TO)cr[q], (TO)ci[q]);
}

// Bra and ket of at most 8 columns (the bottom layers' 8 x 8 holes): lanes
// 8..15 of every row group take the second half of the chunk's K range, so
// the 16 x 16 MFMA tile holds the two halves' 8 x 8 products on its diagonal
// blocks, summed across lanes at the end: half the k-steps (MFMAs and loads)
// of gc_lean.  complex128 accumulator layout: lane l holds rows (l/16) + 4q,
// column l % 16, so rows 8..15 of column c + 8 sit in v[q + 2] of lane l + 8.
template <typename TO>
__device__ inline void gc_lean_packed(const GcDesc& d, const GcTile& g, const cplx<double>* __restrict__ X,
                                      const cplx<double>* __restrict__ Y, cplx<TO> (&v)[4]) {
  typedef MF<double>::v4 v4;
  const int lane = threadIdx.x & 63, li = lane & 15, kk = lane >> 4;
  const int half = li >> 3, jr = li & 7;
  const int lg = d.ksplit > 1 ? __builtin_ctz((unsigned)d.ksplit) : 0;
  const uint32_t bef = (uint32_t)d.bef, aft = (uint32_t)d.aft;
  const int K = d.bef * d.aft;
  const int chunk = (((K + (1 << lg) - 1) >> lg) + 3) & ~3;
  const int kbeg = g.ks * chunk, kend = min(K, kbeg + chunk);
  // the two halves: [kbeg, kbeg + h0) and [kbeg + h0, kend), h0 a multiple of 4
  const int h0 = ((((kend - kbeg) + 1) >> 1) + 3) & ~3;
  const int klim = half ? kend : min(kend, kbeg + h0);
  const uint32_t jx = (uint32_t)min(jr, d.nx - 1), ly = (uint32_t)min(jr, d.ny - 1);
  const uint32_t xs = bef * (uint32_t)d.nx, ys = bef * (uint32_t)d.ny;
  const uint32_t xrow = jx * bef, yrow = ly * bef;
  const uint32_t xlast = xrow + (bef - 1u) + (aft - 1u) * xs, ylast = yrow + (bef - 1u) + (aft - 1u) * ys;
  const cplx<double>* xp = X + d.x;
  const cplx<double>* yp = Y + d.y;
  int kc = kbeg + (half ? h0 : 0) + kk;
  uint32_t b = (uint32_t)kc % bef;
  const uint32_t a0 = (uint32_t)kc / bef;
  uint32_t px = b + a0 * xs, py = b + a0 * ys;
  const uint32_t q4 = 4u / bef, r4 = 4u % bef;
  const uint32_t dx = r4 + q4 * xs, dy = r4 + q4 * ys;
  v4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  auto load4 = [&](cplx<double>* xr, cplx<double>* yr) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool kv = kc < klim;
      const cplx<double> x = ld32(xp, min(xrow + px, xlast));
      yr[u] = ld32(yp, min(yrow + py, ylast));
      xr[u] = mk<double>(kv ? x.re : 0.0, kv ? -x.im : 0.0);
      kc += 4;
      b += r4;
      px += dx;
      py += dy;
      if (b >= bef) {
        b -= bef;
        px += xs - bef;
        py += ys - bef;
      }
    }
  };
  cplx<double> xc[4], yc[4];
  if (h0 > 0) load4(xc, yc);
  for (int s0 = 0; s0 < h0; s0 += 16) {
    cplx<double> xn[4], yn[4];
    const bool more = s0 + 16 < h0;
    if (more) load4(xn, yn);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u > 0 && s0 + 4 * u >= h0) break;
      cr = MF<double>::mma(xc[u].re, yc[u].re, cr);
      ci = MF<double>::mma(xc[u].re, yc[u].im, ci);
      cr = MF<double>::mma(-xc[u].im, yc[u].im, cr);
      ci = MF<double>::mma(xc[u].im, yc[u].re, ci);
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xc[u] = xn[u];
        yc[u] = yn[u];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double r2 = __shfl_down(cr[q + 2], 8, 16), i2 = __shfl_down(ci[q + 2], 8, 16);
    v[q] = mk<TO>((TO)(cr[q] + r2), (TO)(ci[q] + i2));
  }
  v[2] = czero<TO>();
  v[3] = czero<TO>();
}

template <typename T, typename TO>
__device__ inline void gc_store(const GcDesc& d, const GcTile& g, const cplx<TO> (&v)[4], cplx<TO>* O) {
  const int lane = threadIdx.x & 63;
  const int col = g.lt * 16 + (lane & 15);
  if (col >= d.ny) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = g.jt * 16 + MF<T>::drow(lane, q);
    if (row < d.nx) {
      const int64_t e = d.out + row + (int64_t)col * d.nx;
      cplx<TO> x = v[q];
      if (d.acc) x = cadd(x, O[e]);
      O[e] = x;
    }
  }
}

template <typename T, typename TO, bool LEAN>
__global__ void __launch_bounds__(64 * kWaves)
    gc_grouped(const GcDesc* __restrict__ descs, const int32_t* __restrict__ wgdesc, int ndesc, int64_t ntasks,
               const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1, cplx<TO>* O,
               const uint8_t* __restrict__ active) {
  __shared__ cplx<TO> red[kWaves * 256];
  const int64_t task = (int64_t)blockIdx.x * kWaves + wave_id();
  bool mine = task < ntasks;  // uniform per wavefront
  GcDesc d{};
  if (mine) {
    d = descs[wgdesc ? task_desc(wgdesc, task) : task];
    mine = !(active && d.node >= 0 && !active[d.node]);
  }
  GcTile g{0, 0, 0};
  const int split = mine && d.ksplit > 1 ? d.ksplit : 1;
  cplx<TO> v[4];
  if (mine) {
    const cplx<T>* xb = (d.sel & 1) ? X1 : X0;
    const cplx<T>* yb = (d.sel & 2) ? X1 : X0;
    if (LEAN || gc_lean_ok(d)) {
      g = gc_tile_lean(d, (uint32_t)(task - d.tile0));
      if constexpr (std::is_same<T, double>::value) {
        if (d.nx <= 8 && d.ny <= 8) gc_lean_packed<TO>(d, g, xb, yb, v);
        else gc_lean<T, TO>(d, g, xb, yb, v);
      } else {
        gc_lean<T, TO>(d, g, xb, yb, v);
      }
    } else if (!LEAN) {
      g = gc_tile(d, task - d.tile0);
      gc_product<T, TO>(d, g, xb, yb, v);
    }
  }
  reduce_chunks<TO>(red, mine, split, g.ks, v);
  if (mine && g.ks == 0) gc_store<T, TO>(d, g, v, O);
}

// one descriptor, no K split
template <typename T>
__global__ void __launch_bounds__(64 * kWaves)
    gc_single(GcDesc d, int64_t ntasks, const cplx<T>* __restrict__ X, const cplx<T>* __restrict__ Y, cplx<T>* O) {
  const int64_t task = (int64_t)blockIdx.x * kWaves + wave_id();
  if (task >= ntasks) return;
  const GcTile g = gc_tile(d, task);
  cplx<T> v[4];
  gc_product<T, T>(d, g, X, Y, v);
  gc_store<T, T>(d, g, v, O);
}

// ------------------------------------------------------------ fused holes
// One workgroup per ket: the four wavefronts build Z = M x_leg Psi_p in LDS
// tile by tile (the lean mode product), then run the ket's hole
// contractions with the bra from global memory and Z from LDS (the lean
// hole contraction, both K halves in one tile for 8 x 8 outputs).  Two
// items split K in two, one item in four (chunks summed in LDS as in the
// grouped kernels); three or more take one wavefront each in turn.
__global__ void __launch_bounds__(256)
    hole_fused_kernel(const HoleKet* __restrict__ kets, const HoleItem* __restrict__ items,
                      const cplx<double>* __restrict__ psi, const cplx<double>* __restrict__ work,
                      const cplx<double>* __restrict__ pool, cplx<double>* opool,
                      const uint8_t* __restrict__ active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<double>* red = reinterpret_cast<cplx<double>*>(smem_raw);  // 4 x 256 partial tiles
  cplx<double>* Z = red + 4 * 256;
  const HoleKet kt = kets[blockIdx.x];
  if (active && kt.node >= 0 && !active[kt.node]) return;  // the whole workgroup
  const int wave = wave_id();
  const int N = kt.bef * kt.nleg * kt.aft;
  if (kt.m < 0) {
    for (int base = 0; base < N; base += 8 * 256) {
      cplx<double> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = psi[kt.x + min(base + u * 256 + (int)threadIdx.x, N - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + u * 256 + (int)threadIdx.x;
        if (e < N) Z[e] = v[u];
      }
    }
  } else {
    GmDesc d{};
    d.x = kt.x;
    d.m = kt.m;
    d.out = 0;
    d.bef = kt.bef;
    d.nin = d.nout = kt.nleg;
    d.aft = kt.aft;
    d.op = GM_N;
    d.are = 1.0;
    d.ksplit = 1;
    const int tiles = ((kt.bef * kt.aft + 15) >> 4) * ((kt.nleg + 15) >> 4);
    for (int t = wave; t < tiles; t += 4) {
      const GmTile g = gm_tile_lean(d, (uint32_t)t);
      cplx<double> v[4];
      gm_lean<double, double, double>(d, g, psi, pool, 1.0, 0.0, v);
      gm_lean_store<double, double>(d, g, v, Z);
    }
  }
  __syncthreads();
  const int nit = kt.it1 - kt.it0;
  if (nit >= 3) {
    for (int i = wave; i < nit; i += 4) {
      const HoleItem it = items[kt.it0 + i];
      GcDesc gd{};
      gd.x = it.x;
      gd.out = it.out;
      gd.bef = it.bef;
      gd.nx = gd.ny = it.n;
      gd.aft = it.aft;
      gd.ksplit = 1;
      const GcTile g{0, 0, 0};
      cplx<double> v[4];
      const cplx<double>* X = (it.sel & 1) ? work : psi;
      if (it.n <= 8) gc_lean_packed<double>(gd, g, X, Z, v);
      else gc_lean<double, double>(gd, g, X, Z, v);
      gc_store<double, double>(gd, g, v, opool);
    }
    return;
  }
  const int S = nit == 2 ? 2 : 4;
  const int ii = wave / S, ks = wave % S;
  const bool mine = ii < nit;
  GcDesc gd{};
  const GcTile g{0, 0, ks};
  cplx<double> v[4];
  if (mine) {
    const HoleItem it = items[kt.it0 + ii];
    gd.x = it.x;
    gd.out = it.out;
    gd.bef = it.bef;
    gd.nx = gd.ny = it.n;
    gd.aft = it.aft;
    gd.ksplit = S;
    const cplx<double>* X = (it.sel & 1) ? work : psi;
    if (it.n <= 8) gc_lean_packed<double>(gd, g, X, Z, v);
    else gc_lean<double, double>(gd, g, X, Z, v);
  }
  reduce_chunks<double>(red, mine, S, ks, v);
  if (mine && ks == 0) gc_store<double, double>(gd, g, v, opool);
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

// Systems with n <= 48: ping-pong Gauss-Jordan.  Each pair of column steps
// reads the previous [S | X] buffer and writes the next one, so it costs one
// barrier: every thread forms its elements of the scaled and eliminated
// matrix directly from the old values (near-identity Gram matrices formed in
// the launch take Newton-Schulz steps instead, see below).  S = A^H A is Hermitian positive definite
// (the node's columns are independent), so its diagonal pivots are positive
// and elimination needs no row exchanges; for the near-orthonormal A of a
// sweep or a CMF substep S is diagonally dominant, and the partial pivoting
// of solve_kernel picks the same diagonal pivots, so the arithmetic per
// element is solve_kernel's.
constexpr int kSolveThreads = 512;  // 8 wavefronts: the derivative tiles of a node run side by side

__global__ void __launch_bounds__(kSolveThreads)
    solve_pp_kernel(const SolveItem* __restrict__ items, cplx<double>* pool, const uint8_t* __restrict__ active,
                    const cplx<double>* __restrict__ abase, const cplx<double>* __restrict__ ybase,
                    cplx<double>* obase, double fre, double fim, int nmax, int gram) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const SolveItem it = items[blockIdx.x];
  if (active && it.node >= 0 && !active[it.node]) return;
  const int n = it.n, nn = n * n;
  const cplx<double> f = mk<double>(fre, fim);
  // the buffers are addressed from the LDS base (an array of pointers would
  // turn every access into a generic one); the Gram partials follow the two
  // [S | X] buffers
  cplx<double>* lds = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* part = lds + 4 * nmax * nmax;
  const int lane = threadIdx.x & 63;
  if (it.s < 0) {
    // the top node: D = f Y; with gram also <A|Y> and <A|A>
    cplx<double> ay = czero<double>(), aa = czero<double>();
    for (int e = threadIdx.x; e < it.rows * n; e += kSolveThreads) {
      const cplx<double> y = ybase[it.y + e];
      obase[it.out + e] = cmul(f, y);
      if (gram && it.x >= 0) {
        const cplx<double> a = abase[it.a + e];
        cfmac(ay, a, y);  // conj(a) y
        cfmac(aa, a, a);
      }
    }
    if (gram && it.x >= 0) {
      part[threadIdx.x] = ay;
      part[kSolveThreads + threadIdx.x] = aa;
      __syncthreads();
      for (int h = kSolveThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
          part[threadIdx.x] = cadd(part[threadIdx.x], part[threadIdx.x + h]);
          part[kSolveThreads + threadIdx.x] = cadd(part[kSolveThreads + threadIdx.x], part[kSolveThreads + threadIdx.x + h]);
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        pool[it.x] = part[0];
        pool[it.x + 1] = part[kSolveThreads];
      }
    }
    return;
  }
  // R is needed only after the solve: its load overlaps the Gram products
  const bool rpf = it.r >= 0 && nn <= kSolveThreads;
  cplx<double> rpre = czero<double>();
  if (rpf && (int)threadIdx.x < nn) rpre = pool[it.r + threadIdx.x];
  if (gram) {
    // S = A^H A (wavefronts 0-3) and X = A^H Y (4-7), each one 16 x 16 tile
    // summed over four K chunks of the rows, partials added in chunk order
    const int w = wave_id(), prod = w >> 2;
    GcDesc gd{};
    gd.x = it.a;
    gd.y = prod ? it.y : it.a;
    gd.bef = it.rows;
    gd.aft = 1;
    gd.nx = gd.ny = n;
    gd.ksplit = 4;
    const GcTile gt{0, 0, w & 3};
    cplx<double> v[4];
    gc_product<double, double>(gd, gt, abase, prod ? ybase : abase, v);
#pragma unroll
    for (int q = 0; q < 4; ++q) part[w * 256 + q * 64 + lane] = v[q];
    __syncthreads();
    for (int o = threadIdx.x; o < 512; o += kSolveThreads) {
      const int pr = o >> 8, idx = o & 255, q = idx >> 6, l = idx & 63;
      const int row = MF<double>::drow(l, q), col = l & 15;
      if (row < n && col < n) {
        cplx<double> acc = part[(pr * 4) * 256 + idx];
        for (int c = 1; c < 4; ++c) acc = cadd(acc, part[(pr * 4 + c) * 256 + idx]);
        lds[pr * nn + row + col * n] = acc;
      }
    }
  } else {
    for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
      lds[o] = pool[it.s + o];
      lds[nn + o] = pool[it.x + o];
    }
  }
  __syncthreads();
  int cur = 0;
  // Near-orthonormal A (every sweep, and the CMF substeps): S = I - E with
  // rho = n max|E_ij| < 1/2 bounding the spectral radius of E, so the
  // Newton-Schulz iteration Z <- Z (2I - S Z) from Z = I converges to S^-1
  // quadratically (error rho^(2^k) after k steps); k is chosen for 1e-17 and
  // X <- Z X.  Otherwise the elimination below runs.  (Gram mode, n <= 16.)
  int nsk = -1;
  if (gram) {
    double e = 0.0;
    for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
      const int i = o % n, j = o / n;
      const cplx<double> v = lds[o];
      e = fmax(e, fabs((i == j ? 1.0 : 0.0) - v.re) + fabs(v.im));
    }
    // the maximum over the wavefront, then over the 8 wavefronts (the
    // partials are read before the barrier that precedes any write of part)
    for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_xor(e, off, 64));
    double* red = reinterpret_cast<double*>(part);
    if (lane == 0) red[wave_id()] = e;
    __syncthreads();
    double em = red[0];
#pragma unroll
    for (int w = 1; w < kSolveThreads / 64; ++w) em = fmax(em, red[w]);
    const double rho = n * em;
    if (rho < 0.5) {
      nsk = 0;
      // rho^(2^k) <= 1e-17
      for (double r = rho; r > 1e-17 && nsk < 8; r *= r) ++nsk;
    }
  }
  if (nsk >= 0) {
    const cplx<double>* S = lds;
    const cplx<double>* X0 = lds + nn;
    // Z alternates between the second [S | X] buffer and the partials; T = S Z
    cplx<double>* Z = lds + 2 * nn;
    cplx<double>* T = lds + 3 * nn;
    cplx<double>* Zn = part;
    for (int o = threadIdx.x; o < nn; o += kSolveThreads) Z[o] = mk<double>(o % n == o / n ? 1.0 : 0.0, 0.0);
    __syncthreads();
    for (int k = 0; k < nsk; ++k) {
      for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
        const int i = o % n, j = o / n;
        cplx<double> acc = czero<double>();
        for (int l = 0; l < n; ++l) cfma(acc, S[i + l * n], Z[l + j * n]);
        T[o] = acc;
      }
      __syncthreads();
      for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
        const int i = o % n, j = o / n;
        cplx<double> acc = czero<double>();
        for (int l = 0; l < n; ++l) cfma(acc, Z[i + l * n], T[l + j * n]);
        const cplx<double> z = Z[o];
        Zn[o] = mk<double>(2.0 * z.re - acc.re, 2.0 * z.im - acc.im);
      }
      __syncthreads();
      cplx<double>* t = Z;
      Z = Zn;
      Zn = t;
    }
    // X <- Z X into the second buffer's X half
    for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
      const int i = o % n, j = o / n;
      cplx<double> acc = czero<double>();
      for (int l = 0; l < n; ++l) cfma(acc, Z[i + l * n], X0[l + j * n]);
      T[o] = acc;
    }
    __syncthreads();
    cur = 1;
  }
  // Two columns per barrier: each thread derives the values of the first
  // column step it needs (its row, the second pivot row and column) with the
  // single step's arithmetic, then applies the second step to them.
  for (int c = 0; nsk < 0 && c < n; c += 2) {
    const cplx<double>* M = lds + cur * 2 * nn;
    cplx<double>* N = lds + (cur ^ 1) * 2 * nn;
    const cplx<double>* Mc = M + c * n;  // pivot column
    const cplx<double> dg = Mc[c];
    const double dd = dg.re * dg.re + dg.im * dg.im;
    const cplx<double> dinv = mk<double>(dg.re / dd, -dg.im / dd);
    // [S | X](rr, col) after column step c
    auto step1 = [&](int rr, int col) {
      const cplx<double>* Mk = M + col * n;
      const cplx<double> v = cmul(Mk[c], dinv);  // pivot row after scaling
      if (rr == c) return v;
      const cplx<double> fr = Mc[rr];
      cplx<double> w = Mk[rr];
      w.re -= fr.re * v.re - fr.im * v.im;
      w.im -= fr.re * v.im + fr.im * v.re;
      return w;
    };
    if (c + 1 == n) {
      for (int o = threadIdx.x; o < 2 * nn; o += kSolveThreads) N[o] = step1(o % n, o / n);
    } else {
      const int c1 = c + 1;
      const cplx<double> dg1 = step1(c1, c1);
      const double dd1 = dg1.re * dg1.re + dg1.im * dg1.im;
      const cplx<double> dinv1 = mk<double>(dg1.re / dd1, -dg1.im / dd1);
      for (int o = threadIdx.x; o < 2 * nn; o += kSolveThreads) {
        const int r = o % n, k = o / n;
        const cplx<double> v = cmul(step1(c1, k), dinv1);
        cplx<double> w;
        if (r == c1) {
          w = v;
        } else {
          const cplx<double> fr = step1(r, c1);
          w = step1(r, k);
          w.re -= fr.re * v.re - fr.im * v.im;
          w.im -= fr.re * v.im + fr.im * v.re;
        }
        N[o] = w;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  const cplx<double>* X = lds + cur * 2 * nn + nn;
  for (int o = threadIdx.x; o < nn; o += kSolveThreads) pool[it.x + o] = X[o];
  if (it.r < 0) return;
  // W(k, m) = sum_l R(k, l) X(m, l); R staged in the free buffer, W kept
  // beside it for the derivative tiles
  cplx<double>* Rm = lds + (cur ^ 1) * 2 * nn;
  cplx<double>* Wm = Rm + nn;
  if (rpf) {
    if ((int)threadIdx.x < nn) Rm[threadIdx.x] = rpre;
  } else {
    for (int o = threadIdx.x; o < nn; o += kSolveThreads) Rm[o] = pool[it.r + o];
  }
  __syncthreads();
  for (int o = threadIdx.x; o < nn; o += kSolveThreads) {
    const int k = o % n, m = o / n;
    cplx<double> acc = czero<double>();
    for (int l = 0; l < n; ++l) cfma(acc, Rm[k + l * n], X[m + l * n]);
    pool[it.w + o] = acc;
    Wm[o] = acc;
  }
  if (it.out < 0) return;
  __syncthreads();
  // D = f (R x_last Y) - f (W x_last A) on the node's (rows x n) tiles, one
  // wavefront per 16 x 16 tile: the two products of a tile load their Y and A
  // operands together, R and W come from LDS, and each product keeps the
  // accumulators and alpha scaling of the grouped mode product.
  typedef MF<double>::v4 v4;
  const int li = lane & 15, kk = lane >> 4;
  const int rows = it.rows, CT = (n + 15) >> 4;
  const int64_t tiles = (int64_t)((rows + 15) >> 4) * CT;
  const cplx<double>* ya = ybase + it.y;
  const cplx<double>* aa = abase + it.a;
  for (int64_t t = wave_id(); t < tiles; t += kSolveThreads / 64) {
    const int rt = (int)(t / CT), ct = (int)(t - (int64_t)rt * CT);
    const int r = rt * 16 + li, j = ct * 16 + li;
    const bool rv = r < rows, cv = j < n;
    v4 yr0 = {0, 0, 0, 0}, yi0 = yr0, yr1 = yr0, yi1 = yr0, ar0 = yr0, ai0 = yr0, ar1 = yr0, ai1 = yr0;
    for (int k0 = 0; k0 < n; k0 += 16) {
      cplx<double> ay[4], av[4], by[4], bw[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kx = k0 + 4 * u + kk;
        const bool kv = kx < n;
        ay[u] = ldz(ya, r + (int64_t)kx * rows, rv && kv);
        av[u] = ldz(aa, r + (int64_t)kx * rows, rv && kv);
        const int mo = (cv && kv) ? j + kx * n : 0;
        const cplx<double> rm = Rm[mo], wm = Wm[mo];
        by[u] = (cv && kv) ? rm : czero<double>();
        bw[u] = (cv && kv) ? wm : czero<double>();
      }
      cmma<double>(yr0, yi0, ay[0], by[0]);
      cmma<double>(ar0, ai0, av[0], bw[0]);
      if (k0 + 4 < n) {
        cmma<double>(yr1, yi1, ay[1], by[1]);
        cmma<double>(ar1, ai1, av[1], bw[1]);
      }
      if (k0 + 8 < n) {
        cmma<double>(yr0, yi0, ay[2], by[2]);
        cmma<double>(ar0, ai0, av[2], bw[2]);
      }
      if (k0 + 12 < n) {
        cmma<double>(yr1, yi1, ay[3], by[3]);
        cmma<double>(ar1, ai1, av[3], bw[3]);
      }
    }
    if (!cv) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = rt * 16 + MF<double>::drow(lane, q);
      if (rr >= rows) continue;
      const double vr = yr0[q] + yr1[q], vi = yi0[q] + yi1[q];
      const double wr = ar0[q] + ar1[q], wi = ai0[q] + ai1[q];
      // f Y R^T and (-f) A W^T, each scaled as the mode product scales it
      const cplx<double> py = mk<double>(fre * vr - fim * vi, fre * vi + fim * vr);
      const cplx<double> pa = mk<double>(-fre * wr + fim * wi, -fre * wi - fim * wr);
      obase[it.out + rr + (int64_t)j * rows] = cadd(py, pa);
    }
  }
}

}  // namespace

template <typename T, typename TM, typename TO>
int launch_gm(const GmDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, const void* mbase, void* obase, const uint8_t* active, double sre, double sim,
              hipStream_t s, const KernelTiming* kt, int waves, bool lean) {
  if (ntasks <= 0 || ndesc <= 0) return QTG_OK;
#define QTG_GM(W, L)                                                                                             \
  launch_timed(gm_grouped<T, TM, TO, W, L>, dim3((unsigned)((ntasks + W - 1) / W)), dim3(64 * W), 0, s, kt, descs, \
               wgdesc, (int)ndesc, ntasks, (const cplx<T>*)x0, (const cplx<T>*)x1, (const cplx<TM>*)mbase,       \
               (cplx<TO>*)obase, active, sre, sim)
  if (waves == kGemmWavesWide) {
    if (lean) QTG_GM(kGemmWavesWide, true);
    else QTG_GM(kGemmWavesWide, false);
  } else {
    if (lean) QTG_GM(kWaves, true);
    else QTG_GM(kWaves, false);
  }
#undef QTG_GM
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T, typename TO>
int launch_gc(const GcDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, void* obase, const uint8_t* active, hipStream_t s, const KernelTiming* kt, bool lean) {
  if (ntasks <= 0 || ndesc <= 0) return QTG_OK;
  const unsigned blocks = (unsigned)((ntasks + kWaves - 1) / kWaves);
  if (lean)
    launch_timed(gc_grouped<T, TO, true>, dim3(blocks), dim3(64 * kWaves), 0, s, kt, descs, wgdesc, (int)ndesc,
                 ntasks, (const cplx<T>*)x0, (const cplx<T>*)x1, (cplx<TO>*)obase, active);
  else
    launch_timed(gc_grouped<T, TO, false>, dim3(blocks), dim3(64 * kWaves), 0, s, kt, descs, wgdesc, (int)ndesc,
                 ntasks, (const cplx<T>*)x0, (const cplx<T>*)x1, (cplx<TO>*)obase, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
int launch_gm_single(const GmDesc& d, const void* x, const void* m, void* o, hipStream_t s) {
  if (d.ksplit > 1) {
    set_error("single mode product: no K split");
    return QTG_EARG;
  }
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
    set_error("single hole contraction: no K split");
    return QTG_EARG;
  }
  const int64_t nt = gc_tasks(d);
  if (nt <= 0) return QTG_OK;
  hipLaunchKernelGGL(gc_single<T>, dim3((unsigned)((nt + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0, s, d, nt,
                     (const cplx<T>*)x, (const cplx<T>*)y, (cplx<T>*)o);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_gm_real(const GmDesc& d, const double* x, const double* m, double* o, hipStream_t s) {
  if (d.ksplit > 1 || (d.op != GM_N && d.op != GM_T)) {
    set_error("real mode product: op M or M^T, no K split");
    return QTG_EARG;
  }
  const int64_t nt = gm_tasks(d);
  if (nt <= 0) return QTG_OK;
  hipLaunchKernelGGL(gmr_single, dim3((unsigned)((nt + kWaves - 1) / kWaves)), dim3(64 * kWaves), 0, s, d, nt, x, m,
                     o);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_solve(const SolveItem* items, int64_t n, cplx<double>* pool, int nmax, const uint8_t* active,
                 hipStream_t s, const KernelTiming* kt, const cplx<double>* abase, const cplx<double>* ybase,
                 cplx<double>* obase, double fre, double fim, bool gram) {
  if (n <= 0) return QTG_OK;
  if (gram && nmax > 16) {
    set_error("solve: the Gram matrices are formed in the solve only for n <= 16");
    return QTG_EARG;
  }
  if (nmax <= 48) {
    // two [S | X] buffers, then (gram) the 8 partial 16 x 16 tiles, which
    // also hold the top node's 2 x kSolveThreads partial sums
    static_assert(2 * kSolveThreads <= 8 * 256, "gram partials");
    const size_t lds = sizeof(cplx<double>) * (4 * (size_t)nmax * nmax + (gram ? 8 * 256 : 0));
    launch_timed(solve_pp_kernel, dim3((unsigned)n), dim3(kSolveThreads), (unsigned)lds, s, kt, items, pool, active,
                 abase, ybase, obase, fre, fim, nmax, gram ? 1 : 0);
    QTG_LAUNCH_CHECK();
    return QTG_OK;
  }
  const size_t lds = 2 * sizeof(cplx<double>) * (size_t)nmax * nmax;
  launch_timed(solve_kernel, dim3((unsigned)n), dim3(256), (unsigned)lds, s, kt, items, pool, active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

int launch_hole_fused(const HoleKet* kets, int64_t nkets, const HoleItem* items, const void* psi, const void* work,
                      const void* pool, void* opool, int64_t lds_bytes, const uint8_t* active, hipStream_t s,
                      const KernelTiming* kt) {
  if (nkets <= 0) return QTG_OK;
  launch_timed(hole_fused_kernel, dim3((unsigned)nkets), dim3(256), (unsigned)lds_bytes, s, kt, kets, items,
               (const cplx<double>*)psi, (const cplx<double>*)work, (const cplx<double>*)pool, (cplx<double>*)opool,
               active);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

#define QTG_GM(T, TM, TO)                                                                                      \
  template int launch_gm<T, TM, TO>(const GmDesc*, const int32_t*, int64_t, int64_t, const void*, const void*,  \
                                    const void*, void*, const uint8_t*, double, double, hipStream_t,            \
                                    const KernelTiming*, int, bool);
QTG_GM(double, double, double)
QTG_GM(float, float, float)
QTG_GM(float, double, float)
QTG_GM(float, double, double)
#undef QTG_GM
#define QTG_GC(T, TO)                                                                                           \
  template int launch_gc<T, TO>(const GcDesc*, const int32_t*, int64_t, int64_t, const void*, const void*, void*, \
                                const uint8_t*, hipStream_t, const KernelTiming*, bool);
QTG_GC(double, double)
QTG_GC(float, float)
QTG_GC(float, double)
#undef QTG_GC
template int launch_gm_single<double>(const GmDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gm_single<float>(const GmDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gc_single<double>(const GcDesc&, const void*, const void*, void*, hipStream_t);
template int launch_gc_single<float>(const GcDesc&, const void*, const void*, void*, hipStream_t);

}  // namespace qtg
