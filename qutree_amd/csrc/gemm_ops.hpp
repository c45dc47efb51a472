// Grouped complex contractions of node tensors on MFMA (see gemm_ops.hip).
//
// Two descriptor kinds cover every dense contraction of the tree sweeps:
//   GmDesc  mode-k product   out(b,i,a) = alpha sum_l op(M)(i,l) X(b,l,a) (+ out)
//   GcDesc  hole contraction out(j,l)   = sum_{b,a} conj(X(b,j,a)) Y(b,l,a) (+ out)
// Each descriptor is split into 16 x 16 output tiles; one wavefront computes
// one tile with real-split complex MFMA (four real 16x16x4 products per
// complex k-step).  A launch walks a flat task list: descriptors are sorted by
// their first task index tile0, and a wavefront finds its descriptor from a
// per-workgroup table.  Long reductions split along K into 2 or 4 chunks
// (more wavefronts in flight) that sum in LDS inside their workgroup.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace qtg {

enum GmOp : int32_t { GM_N = 0, GM_T = 1, GM_H = 2, GM_COPY = 3 };

// Optional event pair a launcher records the kernel's own start and stop
// times into (hipExtLaunchKernelGGL); null launches plainly.
struct KernelTiming {
  hipEvent_t start, stop;
};

constexpr int kGemmWaves = 4;  // tasks (wavefronts) per workgroup of the grouped kernels
// mode-product launches whose reductions split into 8 chunks run workgroups
// of this many wavefronts (launch_gm's `waves`)
constexpr int kGemmWavesWide = 8;

struct GmDesc {
  int64_t x, m, out;                // complex-element offsets (input, matrix, output)
  int32_t bef, nin, nout, aft;      // view of X on the contracted mode
  int32_t op;                       // GmOp; GM_COPY: out = alpha X (nin == nout, m unused)
  int32_t acc;                      // 1: accumulate into out
  double are, aim;                  // alpha
  int32_t node;                     // index into the active mask (-1: always run)
  int32_t xsel;                     // X from the first (0) or second (1) input base
  int32_t ksplit;                   // K chunks per tile: 1, 2 or 4 (see order_tasks)
  int32_t pad;
  int64_t tile0;                    // first task of this descriptor
};

struct GcDesc {
  int64_t x, y, out;                // complex-element offsets (bra, ket, nx x ny output)
  int32_t bef, nx, ny, aft;         // X is (bef, nx, aft), Y is (bef, ny, aft)
  int32_t acc;                      // 1: accumulate into out
  int32_t node;                     // active-mask index (-1: always run)
  int32_t ksplit;                   // K chunks per tile: 1, 2 or 4 (see order_tasks)
  int32_t sel;                      // bit 0: X from the second input base; bit 1: Y likewise
  int64_t tile0;
};

// One hole-contraction layer in one launch (launch_hole_fused): a
// workgroup per ket tensor Z = M x_leg Psi_p (or Psi_p itself), built in
// LDS, and the hole contractions out = <X|Z> over all legs but one that
// read it; the bras X were formed before the layer.  complex128; every
// output at most 16 x 16.
struct HoleKet {
  int64_t x, m;            // Psi_p offset (state); pool offset of M (nleg x nleg, column-major), -1: Z = Psi_p
  int32_t bef, nleg, aft;  // Psi_p viewed on the ket leg
  int32_t node;            // active-mask index (-1: always run)
  int32_t it0, it1;        // this ket's items
};
struct HoleItem {
  int64_t x, out;       // bra offset (work buffer when sel & 1, else the state); n x n output in the pool
  int32_t bef, n, aft;  // the node viewed on the contracted leg
  int32_t sel;
};
int launch_hole_fused(const HoleKet* kets, int64_t nkets, const HoleItem* items, const void* psi, const void* work,
                      const void* pool, void* opool, int64_t lds_bytes, const uint8_t* active, hipStream_t s,
                      const KernelTiming* kt = nullptr);
// LDS bytes of a hole_fused workgroup whose ket tensor has N elements
inline int64_t hole_fused_lds(int64_t N) { return (4 * 256 + N) * 16; }

// number of tasks of one descriptor (16 x 16 tiles, times K chunks)
__host__ __device__ inline int64_t gm_tasks(const GmDesc& d) {
  const int64_t R = (int64_t)d.bef * d.aft;
  return ((R + 15) / 16) * ((d.nout + 15) / 16) * (d.ksplit > 1 ? d.ksplit : 1);
}
__host__ __device__ inline int64_t gc_tasks(const GcDesc& d) {
  return (int64_t)((d.nx + 15) / 16) * ((d.ny + 15) / 16) * (d.ksplit > 1 ? d.ksplit : 1);
}

// Launchers.  Input bases x0 / x1 (selected per descriptor) are cplx<T>*, the
// matrix base cplx<TM>* (converted to T on load), the output cplx<TO>*; the
// products accumulate in T.  `active` may be null.  Every alpha of a
// mode-product launch is multiplied by the complex scale (sre, sim).
// wgdesc[t]: descriptor of task t (see gemm_wg_table); null when descs holds
// one descriptor per task (descs[t]).  Either way a wavefront finds its
// descriptor without a search.  The chunks of a split tile sum in LDS inside
// their workgroup, which needs the task order of order_tasks.  No two
// descriptors of one launch may write the same output.
// lean: every descriptor of the launch passes gm_lean_ok / gc_lean_ok; the
// launch then uses a kernel without the general path (fewer registers, more
// wavefronts per SIMD).
template <typename T, typename TM, typename TO>
int launch_gm(const GmDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, const void* mbase, void* obase, const uint8_t* active, double sre, double sim,
              hipStream_t s, const KernelTiming* kt = nullptr, int waves = kGemmWaves, bool lean = false);
template <typename T, typename TO>
int launch_gc(const GcDesc* descs, const int32_t* wgdesc, int64_t ndesc, int64_t ntasks, const void* x0,
              const void* x1, void* obase, const uint8_t* active, hipStream_t s, const KernelTiming* kt = nullptr,
              bool lean = false);

// The lean forms of the grouped kernels (gemm_ops.hip) take descriptors
// whose tensors each span fewer than 2^27 elements, with a power-of-two K
// split: their element offsets fit 32 bits.
__host__ __device__ inline bool gm_lean_ok(const GmDesc& d) {
  constexpr int64_t lim = (int64_t)1 << 27;
  return (int64_t)d.bef * d.nin * d.aft < lim && (int64_t)d.bef * d.nout * d.aft < lim &&
         (int64_t)d.nin * d.nout < lim && (d.ksplit & (d.ksplit - 1)) == 0;
}
__host__ __device__ inline bool gc_lean_ok(const GcDesc& d) {
  constexpr int64_t lim = (int64_t)1 << 27;
  return (int64_t)d.bef * d.nx * d.aft < lim && (int64_t)d.bef * d.ny * d.aft < lim &&
         (d.ksplit & (d.ksplit - 1)) == 0;
}

// K split of a reduction of length K into at most maxs chunks: the largest
// power of two not above `want` whose chunks stay at least min_k long.
inline int32_t split_chunks(int64_t K, int64_t want, int64_t min_k, int32_t maxs = kGemmWaves) {
  int32_t s = 1;
  while (s < maxs && 2 * s <= want && K / (2 * s) >= min_k) s *= 2;
  return s;
}

// Orders descriptors by ksplit (largest first; stable within a class)
// and assigns tile0, so that the chunks of every split tile are consecutive
// tasks starting at a multiple of ksplit, i.e. wavefronts of one workgroup.
// Returns the launch's task count.
template <typename D>
inline int64_t order_tasks(std::vector<D>& descs, int64_t (*tasks)(const D&)) {
  std::stable_sort(descs.begin(), descs.end(),
                   [](const D& a, const D& b) { return (a.ksplit > 1 ? a.ksplit : 1) > (b.ksplit > 1 ? b.ksplit : 1); });
  int64_t t = 0;
  for (auto& d : descs) {
    d.tile0 = t;
    t += tasks(d);
  }
  return t;
}

// Appends to `out` the descriptor index of every task of a launch over
// `descs` (tile0 already assigned, ntasks in total).
template <typename D>
inline void gemm_wg_table(const D* descs, int64_t ndesc, int64_t ntasks, std::vector<int32_t>& out) {
  int64_t di = 0;
  for (int64_t t = 0; t < ntasks; ++t) {
    while (di + 1 < ndesc && descs[di + 1].tile0 <= t) ++di;
    out.push_back((int32_t)di);
  }
}
// One descriptor (tile0 ignored), no active mask; no K split.
template <typename T>
int launch_gm_single(const GmDesc& d, const void* x, const void* m, void* o, hipStream_t s);
template <typename T>
int launch_gc_single(const GcDesc& d, const void* x, const void* y, void* o, hipStream_t s);

// Real FP64 mode product (op GM_N or GM_T; one descriptor, no K split).
int launch_gm_real(const GmDesc& d, const double* x, const double* m, double* o, hipStream_t s);

// Batched FP64 solve X <- S^-1 X for n x n column-major S and X at pool
// offsets s, x; n <= nmax <= 64.  nmax > 48: Gauss-Jordan with partial
// pivoting.  nmax <= 48: Gauss-Jordan on the diagonal pivots of the positive
// definite S = A^H A, two columns per barrier, or (gram, near-identity S)
// Newton-Schulz steps.  When r >= 0 it also forms W = R X^T at w
// (R the n x n matrix at r).  When out >= 0 (nmax <= 48 only) it then writes
// the projected derivative of the node, rows L = N / n, column-major:
//   D(:, i) = f (sum_l R(i, l) Y(:, l) - sum_l W(i, l) A(:, l))
// with A at a, Y at y, D at out in their own base pointers and f the launch
// scale; an item with s < 0 only writes D = f Y (the top node).
// With gram (nmax <= 16 only) the solve forms S = A^H A and X = A^H Y itself
// instead of reading them, and a top-node item with x >= 0 writes <A|Y> and
// <A|A> (rows x n elements) to the pool at x and x + 1.
struct SolveItem {
  int64_t s, x, r, w;
  int64_t a, y, out;
  int32_t n, node, rows, pad;
};
int launch_solve(const SolveItem* items, int64_t n, cplx<double>* pool, int nmax, const uint8_t* active,
                 hipStream_t s, const KernelTiming* kt = nullptr, const cplx<double>* abase = nullptr,
                 const cplx<double>* ybase = nullptr, cplx<double>* obase = nullptr, double fre = 0.0,
                 double fim = 0.0, bool gram = false);

}  // namespace qtg
