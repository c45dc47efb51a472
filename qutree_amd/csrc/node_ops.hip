// Node-level contractions on gfx950: mode-k products (matvec/tmatvec/ctmatvec
// semantics of LA_lib.f, matrixTensorBLAS semantics of TensorBLAS.cpp) and
// hole contractions (rhomat / contractionBLAS), single and grouped.
//
// Mode product C(b,k,a) (+)= sum_l M(k,l) B(b,l,a): the effective matrix is
// staged in LDS once per workgroup; each lane owns one (b,a) row and keeps a
// chunk of MP_KCH outputs in registers, so consecutive lanes read consecutive
// b (coalesced whenever before >= 64) and every M element is an LDS broadcast.
//
// Hole contraction S(j,l) = sum_{b,a} conj(A(b,j,a)) B(b,l,a): a workgroup owns
// a 16x16 output tile and a chunk of (b,a) rows (split-K for long rows); the
// row panels of A and B are staged through LDS with lanes running along b.
// Chunks write partial tiles to a plan-owned workspace that a second kernel
// sums in chunk order (deterministic, no atomics).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace qtg {

static_assert(sizeof(qtg_mode_product_desc) == 64, "descriptor layout");
static_assert(sizeof(qtg_hole_contraction_desc) == 64, "descriptor layout");

constexpr int kBlock = 256;
constexpr int MP_KCH = 8;
constexpr int HC_T = 16;
constexpr int HC_R = 64;
constexpr int64_t HC_CHUNK = 512;

// ---------------------------------------------------------------- mode product
template <typename T>
__device__ void mode_product_rows(const qtg_mode_product_desc& d, int64_t tile, cplx<T>* __restrict__ C,
                                  const cplx<T>* __restrict__ M, const cplx<T>* __restrict__ B,
                                  cplx<T>* sM) {
  const int nB = (int)d.nB, nC = (int)d.nC;
  for (int i = threadIdx.x; i < nB * nC; i += blockDim.x) {
    const int k = i % nC, l = i / nC;
    cplx<T> v;
    if (d.op == 0) {
      v = M[d.offM + k + (int64_t)l * nC];
    } else {
      v = M[d.offM + l + (int64_t)k * nB];
      if (d.op == 2) v.im = -v.im;
    }
    sM[i] = v;  // sM[l * nC + k] = Meff(k, l)
  }
  __syncthreads();
  const int64_t rows = d.before * d.after;
  const int64_t r = tile * kBlock + threadIdx.x;
  if (r >= rows) return;
  const int64_t b = r % d.before, a = r / d.before;
  const cplx<T>* Bp = B + d.offB + b + a * d.before * nB;
  cplx<T>* Cp = C + d.offC + b + a * d.before * nC;
  const int64_t st = d.before;
  for (int kc = 0; kc < nC; kc += MP_KCH) {
    cplx<T> acc[MP_KCH];
#pragma unroll
    for (int k = 0; k < MP_KCH; ++k) {
      acc[k] = czero<T>();
      if (d.add && kc + k < nC) acc[k] = Cp[(kc + k) * st];
    }
    if (kc + MP_KCH <= nC) {
      for (int l = 0; l < nB; ++l) {
        const cplx<T> bv = Bp[l * st];
        const cplx<T>* m = sM + l * nC + kc;
#pragma unroll
        for (int k = 0; k < MP_KCH; ++k) cfma(acc[k], m[k], bv);
      }
    } else {
      for (int l = 0; l < nB; ++l) {
        const cplx<T> bv = Bp[l * st];
        const cplx<T>* m = sM + l * nC + kc;
#pragma unroll
        for (int k = 0; k < MP_KCH; ++k)
          if (kc + k < nC) cfma(acc[k], m[k], bv);
      }
    }
#pragma unroll
    for (int k = 0; k < MP_KCH; ++k)
      if (kc + k < nC) Cp[(kc + k) * st] = acc[k];
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mode_product_single(qtg_mode_product_desc d, cplx<T>* C,
                                                              const cplx<T>* M, const cplx<T>* B) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  mode_product_rows<T>(d, blockIdx.x, C, M, B, reinterpret_cast<cplx<T>*>(smem_raw));
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    mode_product_grouped(const qtg_mode_product_desc* __restrict__ descs, const int2* __restrict__ tasks,
                         cplx<T>* C, const cplx<T>* M, const cplx<T>* B) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int2 t = tasks[blockIdx.x];
  const qtg_mode_product_desc d = descs[t.x];
  mode_product_rows<T>(d, t.y, C, M, B, reinterpret_cast<cplx<T>*>(smem_raw));
}

// ----------------------------------------------------------- hole contraction
// One task: desc, 16x16 output tile (jt, lt), row chunk.  out points at the
// destination tile origin: S (+ accumulate) when the desc has one chunk, or a
// partial slab in the workspace (ld = nA).
template <typename T>
__device__ void hole_tile(const qtg_hole_contraction_desc& d, int jt, int lt, int64_t r_begin, int64_t r_end,
                          const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B, cplx<T>* out,
                          bool accumulate) {
  __shared__ cplx<T> sA[HC_R][HC_T + 1];
  __shared__ cplx<T> sB[HC_R][HC_T + 1];
  __shared__ cplx<T> red[kBlock];
  const int nA = (int)d.nA, nB = (int)d.nB;
  const int j0 = jt * HC_T, l0 = lt * HC_T;
  const int TJ = min(HC_T, nA - j0), TL = min(HC_T, nB - l0);
  const int outs = TJ * TL;
  const int G = kBlock / outs;  // row groups sharing the tile
  const int o = threadIdx.x % outs, g = threadIdx.x / outs;
  const int jj = o % TJ, ll = o / TJ;
  const int64_t before = d.before;
  cplx<T> acc = czero<T>();
  for (int64_t r0 = r_begin; r0 < r_end; r0 += HC_R) {
    const int nr = (int)min<int64_t>(HC_R, r_end - r0);
    for (int i = threadIdx.x; i < HC_R * TJ; i += kBlock) {
      const int rr = i % HC_R, c = i / HC_R;
      if (rr < nr) {
        const int64_t r = r0 + rr, b = r % before, a = r / before;
        sA[rr][c] = A[d.offA + b + (int64_t)(j0 + c) * before + a * before * nA];
      }
    }
    for (int i = threadIdx.x; i < HC_R * TL; i += kBlock) {
      const int rr = i % HC_R, c = i / HC_R;
      if (rr < nr) {
        const int64_t r = r0 + rr, b = r % before, a = r / before;
        sB[rr][c] = B[d.offB + b + (int64_t)(l0 + c) * before + a * before * nB];
      }
    }
    __syncthreads();
    if (g < G)
      for (int rr = g; rr < nr; rr += G) cfmac(acc, sA[rr][jj], sB[rr][ll]);
    __syncthreads();
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < outs) {
    cplx<T> s = red[threadIdx.x];
    for (int gg = 1; gg < G; ++gg) s = cadd(s, red[gg * outs + threadIdx.x]);
    cplx<T>* dst = out + (j0 + jj) + (int64_t)(l0 + ll) * nA;
    if (accumulate) s = cadd(s, *dst);
    *dst = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) hole_single(qtg_hole_contraction_desc d, cplx<T>* S, const cplx<T>* A,
                                                      const cplx<T>* B, int tiles_j) {
  const int jt = blockIdx.x % tiles_j, lt = blockIdx.x / tiles_j;
  hole_tile<T>(d, jt, lt, 0, d.before * d.after, A, B, S + d.offS, d.accumulate != 0);
}

struct HCTask {
  int32_t desc, jt, lt, chunk;
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
    hole_grouped(const qtg_hole_contraction_desc* __restrict__ descs, const HCTask* __restrict__ tasks,
                 const int64_t* __restrict__ part_off, const int32_t* __restrict__ nchunks, cplx<T>* S,
                 const cplx<T>* A, const cplx<T>* B, cplx<T>* work) {
  const HCTask t = tasks[blockIdx.x];
  const qtg_hole_contraction_desc d = descs[t.desc];
  const int64_t rows = d.before * d.after;
  const int64_t rb = (int64_t)t.chunk * HC_CHUNK;
  const int64_t re = min(rows, rb + HC_CHUNK);
  if (nchunks[t.desc] == 1) {
    hole_tile<T>(d, t.jt, t.lt, 0, rows, A, B, S + d.offS, d.accumulate != 0);
  } else {
    cplx<T>* slab = work + part_off[t.desc] + (int64_t)t.chunk * d.nA * d.nB;
    hole_tile<T>(d, t.jt, t.lt, rb, re, A, B, slab, false);
  }
}

// sums the partial slabs of multi-chunk descriptors in chunk order
template <typename T>
__global__ void __launch_bounds__(kBlock)
    hole_reduce(const qtg_hole_contraction_desc* __restrict__ descs, const int32_t* __restrict__ red_desc,
                const int64_t* __restrict__ red_start, int64_t nred, const int64_t* __restrict__ part_off,
                const int32_t* __restrict__ nchunks, cplx<T>* S, const cplx<T>* work, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= total) return;
  // find the descriptor whose element range holds i
  int64_t lo = 0, hi = nred - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (red_start[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int di = red_desc[lo];
  const qtg_hole_contraction_desc d = descs[di];
  const int64_t e = i - red_start[lo];
  const int64_t n = d.nA * d.nB;
  const cplx<T>* p = work + part_off[di] + e;
  cplx<T> s = p[0];
  for (int c = 1; c < nchunks[di]; ++c) s = cadd(s, p[c * n]);
  cplx<T>* dst = S + d.offS + e;
  if (d.accumulate) s = cadd(s, *dst);
  *dst = s;
}

// ------------------------------------------------------------------- host side
static int check_mp(const qtg_mode_product_desc& d) {
  if (d.before < 1 || d.after < 1 || d.nB < 1 || d.nC < 1 || d.nB > 64 || d.nC > 64 || d.op < 0 || d.op > 2) {
    set_error("mode product: need 1 <= nB, nC <= 64, before, after >= 1, op in {0,1,2}");
    return QTG_ESHAPE;
  }
  return QTG_OK;
}

static int check_hc(const qtg_hole_contraction_desc& d) {
  if (d.before < 1 || d.after < 1 || d.nA < 1 || d.nB < 1) {
    set_error("hole contraction: dimensions must be >= 1");
    return QTG_ESHAPE;
  }
  return QTG_OK;
}

template <typename T>
static int mode_product_launch(const qtg_mode_product_desc& d, cplx<T>* C, const cplx<T>* M, const cplx<T>* B,
                               hipStream_t s) {
  int rc = check_mp(d);
  if (rc) return rc;
  if (!C || !M || !B) {
    set_error("mode product: null pointer");
    return QTG_EARG;
  }
  const int64_t tiles = (d.before * d.after + kBlock - 1) / kBlock;
  const size_t lds = sizeof(cplx<T>) * d.nB * d.nC;
  hipLaunchKernelGGL(mode_product_single<T>, dim3((unsigned)tiles), dim3(kBlock), lds, s, d, C, M, B);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

template <typename T>
static int hole_launch(const qtg_hole_contraction_desc& d, cplx<T>* S, const cplx<T>* A, const cplx<T>* B,
                       hipStream_t s) {
  int rc = check_hc(d);
  if (rc) return rc;
  if (!S || !A || !B) {
    set_error("hole contraction: null pointer");
    return QTG_EARG;
  }
  const int tj = (int)((d.nA + HC_T - 1) / HC_T), tl = (int)((d.nB + HC_T - 1) / HC_T);
  hipLaunchKernelGGL(hole_single<T>, dim3(tj * tl), dim3(kBlock), 0, s, d, S, A, B, tj);
  QTG_LAUNCH_CHECK();
  return QTG_OK;
}

// ------------------------------------------------------------------------ plans
struct Plan {
  int kind = 0;  // 0 mode product, 1 hole contraction
  bool c64 = false;
  int64_t count = 0;
  void* d_descs = nullptr;
  void* d_tasks = nullptr;
  int64_t ntasks = 0;
  size_t lds = 0;
  // hole contraction extras
  int64_t* d_part_off = nullptr;
  int32_t* d_nchunks = nullptr;
  int32_t* d_red_desc = nullptr;
  int64_t* d_red_start = nullptr;
  int64_t nred = 0, red_total = 0;
  void* d_work = nullptr;
  double flops = 0, bytes = 0;
  ~Plan() {
    for (void* p : {d_descs, d_tasks, (void*)d_part_off, (void*)d_nchunks, (void*)d_red_desc,
                    (void*)d_red_start, d_work})
      if (p) (void)hipFree(p);
  }
};

template <typename X>
static int upload(void** dst, const std::vector<X>& v) {
  if (v.empty()) {
    *dst = nullptr;
    return QTG_OK;
  }
  QTG_HIP_TRY(hipMalloc(dst, sizeof(X) * v.size()));
  QTG_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
  return QTG_OK;
}

}  // namespace qtg

using namespace qtg;

extern "C" {

int qtg_mode_product_z(double* C, const double* M, const double* B, int64_t before, int64_t nB, int64_t nC,
                       int64_t after, int op, int add, qtg_stream stream) {
  qtg_mode_product_desc d{0, 0, 0, before, nB, nC, after, op, add};
  return mode_product_launch<double>(d, (cplx<double>*)C, (const cplx<double>*)M, (const cplx<double>*)B,
                                     S(stream));
}
int qtg_mode_product_c(float* C, const float* M, const float* B, int64_t before, int64_t nB, int64_t nC,
                       int64_t after, int op, int add, qtg_stream stream) {
  qtg_mode_product_desc d{0, 0, 0, before, nB, nC, after, op, add};
  return mode_product_launch<float>(d, (cplx<float>*)C, (const cplx<float>*)M, (const cplx<float>*)B, S(stream));
}
int qtg_matvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                 qtg_stream stream) {
  return qtg_mode_product_z(C, M, B, b, a, a, c, 0, add, stream);
}
int qtg_tmatvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                  qtg_stream stream) {
  return qtg_mode_product_z(C, M, B, b, a, a, c, 1, add, stream);
}
int qtg_ctmatvec_z(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                   qtg_stream stream) {
  return qtg_mode_product_z(C, M, B, b, a, a, c, 2, add, stream);
}
int qtg_hole_contraction_z(double* Sm, const double* A, const double* B, int64_t before, int64_t nA, int64_t nB,
                           int64_t after, int accumulate, qtg_stream stream) {
  qtg_hole_contraction_desc d{0, 0, 0, before, nA, nB, after, accumulate, 0};
  return hole_launch<double>(d, (cplx<double>*)Sm, (const cplx<double>*)A, (const cplx<double>*)B, S(stream));
}
int qtg_hole_contraction_c(float* Sm, const float* A, const float* B, int64_t before, int64_t nA, int64_t nB,
                           int64_t after, int accumulate, qtg_stream stream) {
  qtg_hole_contraction_desc d{0, 0, 0, before, nA, nB, after, accumulate, 0};
  return hole_launch<float>(d, (cplx<float>*)Sm, (const cplx<float>*)A, (const cplx<float>*)B, S(stream));
}
int qtg_rhomat_z(const double* bra, const double* ket, double* M, int64_t a, int64_t b, int64_t c,
                 qtg_stream stream) {
  return qtg_hole_contraction_z(M, bra, ket, b, a, a, c, 0, stream);
}

int qtg_mode_product_plan_create(qtg_plan* plan, const qtg_mode_product_desc* descs, int64_t count,
                                 int complex64) {
  if (!plan || (!descs && count > 0) || count < 0) {
    set_error("plan: bad arguments");
    return QTG_EARG;
  }
  auto* p = new Plan();
  p->kind = 0;
  p->c64 = complex64 != 0;
  p->count = count;
  std::vector<int2> tasks;
  size_t lds = 0;
  const double es = p->c64 ? 8.0 : 16.0;
  for (int64_t i = 0; i < count; ++i) {
    int rc = check_mp(descs[i]);
    if (rc) {
      delete p;
      return rc;
    }
    const auto& d = descs[i];
    const int64_t tiles = (d.before * d.after + kBlock - 1) / kBlock;
    for (int64_t t = 0; t < tiles; ++t) tasks.push_back(make_int2((int)i, (int)t));
    lds = std::max(lds, (size_t)(d.nB * d.nC) * (p->c64 ? 8 : 16));
    p->flops += 8.0 * d.before * d.after * d.nB * d.nC;
    p->bytes += es * (d.before * d.after * (d.nB + d.nC * (d.add ? 2 : 1)) + d.nB * d.nC);
  }
  std::vector<qtg_mode_product_desc> dv(descs, descs + count);
  int rc = upload(&p->d_descs, dv);
  if (!rc) rc = upload(&p->d_tasks, tasks);
  if (rc) {
    delete p;
    return rc;
  }
  p->ntasks = (int64_t)tasks.size();
  p->lds = lds;
  *plan = p;
  return QTG_OK;
}

int qtg_hole_contraction_plan_create(qtg_plan* plan, const qtg_hole_contraction_desc* descs, int64_t count,
                                     int complex64) {
  if (!plan || (!descs && count > 0) || count < 0) {
    set_error("plan: bad arguments");
    return QTG_EARG;
  }
  auto* p = new Plan();
  p->kind = 1;
  p->c64 = complex64 != 0;
  p->count = count;
  std::vector<HCTask> tasks;
  std::vector<int64_t> part_off(count, 0), red_start;
  std::vector<int32_t> nchunks(count, 1), red_desc;
  int64_t work = 0, red_total = 0;
  const double es = p->c64 ? 8.0 : 16.0;
  for (int64_t i = 0; i < count; ++i) {
    int rc = check_hc(descs[i]);
    if (rc) {
      delete p;
      return rc;
    }
    const auto& d = descs[i];
    const int64_t rows = d.before * d.after;
    const int nc = (int)((rows + HC_CHUNK - 1) / HC_CHUNK);
    const int tj = (int)((d.nA + HC_T - 1) / HC_T), tl = (int)((d.nB + HC_T - 1) / HC_T);
    nchunks[i] = nc;
    if (nc > 1) {
      part_off[i] = work;
      work += (int64_t)nc * d.nA * d.nB;
      red_desc.push_back((int32_t)i);
      red_start.push_back(red_total);
      red_total += d.nA * d.nB;
    }
    for (int c = 0; c < nc; ++c)
      for (int lt = 0; lt < tl; ++lt)
        for (int jt = 0; jt < tj; ++jt) tasks.push_back(HCTask{(int32_t)i, jt, lt, c});
    p->flops += 8.0 * rows * d.nA * d.nB;
    p->bytes += es * (rows * (d.nA + d.nB) + d.nA * d.nB * (d.accumulate ? 2 : 1));
  }
  std::vector<qtg_hole_contraction_desc> dv(descs, descs + count);
  int rc = upload(&p->d_descs, dv);
  if (!rc) rc = upload(&p->d_tasks, tasks);
  if (!rc) rc = upload((void**)&p->d_part_off, part_off);
  if (!rc) rc = upload((void**)&p->d_nchunks, nchunks);
  if (!rc) rc = upload((void**)&p->d_red_desc, red_desc);
  if (!rc) rc = upload((void**)&p->d_red_start, red_start);
  if (!rc && work > 0) {
    if (hipMalloc(&p->d_work, (size_t)work * (p->c64 ? 8 : 16)) != hipSuccess) {
      set_error("plan: workspace allocation failed");
      rc = QTG_ENOMEM;
    }
  }
  if (rc) {
    delete p;
    return rc;
  }
  p->ntasks = (int64_t)tasks.size();
  p->nred = (int64_t)red_desc.size();
  p->red_total = red_total;
  *plan = p;
  return QTG_OK;
}

int qtg_plan_execute(qtg_plan plan, void* out_base, const void* first_base, const void* second_base,
                     qtg_stream stream) {
  auto* p = static_cast<Plan*>(plan);
  if (!p || !out_base || !first_base || !second_base) {
    set_error("plan execute: null argument");
    return QTG_EARG;
  }
  if (p->ntasks == 0) return QTG_OK;
  hipStream_t s = S(stream);
  if (p->kind == 0) {
    if (p->c64)
      hipLaunchKernelGGL(mode_product_grouped<float>, dim3((unsigned)p->ntasks), dim3(kBlock), p->lds, s,
                         (const qtg_mode_product_desc*)p->d_descs, (const int2*)p->d_tasks, (cplx<float>*)out_base,
                         (const cplx<float>*)first_base, (const cplx<float>*)second_base);
    else
      hipLaunchKernelGGL(mode_product_grouped<double>, dim3((unsigned)p->ntasks), dim3(kBlock), p->lds, s,
                         (const qtg_mode_product_desc*)p->d_descs, (const int2*)p->d_tasks,
                         (cplx<double>*)out_base, (const cplx<double>*)first_base,
                         (const cplx<double>*)second_base);
    QTG_LAUNCH_CHECK();
    return QTG_OK;
  }
  const auto* descs = (const qtg_hole_contraction_desc*)p->d_descs;
  if (p->c64) {
    hipLaunchKernelGGL(hole_grouped<float>, dim3((unsigned)p->ntasks), dim3(kBlock), 0, s, descs,
                       (const HCTask*)p->d_tasks, p->d_part_off, p->d_nchunks, (cplx<float>*)out_base,
                       (const cplx<float>*)first_base, (const cplx<float>*)second_base, (cplx<float>*)p->d_work);
  } else {
    hipLaunchKernelGGL(hole_grouped<double>, dim3((unsigned)p->ntasks), dim3(kBlock), 0, s, descs,
                       (const HCTask*)p->d_tasks, p->d_part_off, p->d_nchunks, (cplx<double>*)out_base,
                       (const cplx<double>*)first_base, (const cplx<double>*)second_base,
                       (cplx<double>*)p->d_work);
  }
  QTG_LAUNCH_CHECK();
  if (p->nred > 0) {
    const unsigned blocks = (unsigned)((p->red_total + kBlock - 1) / kBlock);
    if (p->c64)
      hipLaunchKernelGGL(hole_reduce<float>, dim3(blocks), dim3(kBlock), 0, s, descs, p->d_red_desc,
                         p->d_red_start, p->nred, p->d_part_off, p->d_nchunks, (cplx<float>*)out_base,
                         (const cplx<float>*)p->d_work, p->red_total);
    else
      hipLaunchKernelGGL(hole_reduce<double>, dim3(blocks), dim3(kBlock), 0, s, descs, p->d_red_desc,
                         p->d_red_start, p->nred, p->d_part_off, p->d_nchunks, (cplx<double>*)out_base,
                         (const cplx<double>*)p->d_work, p->red_total);
    QTG_LAUNCH_CHECK();
  }
  return QTG_OK;
}

int qtg_plan_destroy(qtg_plan plan) {
  delete static_cast<Plan*>(plan);
  return QTG_OK;
}

int qtg_plan_counts(qtg_plan plan, double* flops, double* bytes) {
  auto* p = static_cast<Plan*>(plan);
  if (!p) return QTG_EARG;
  if (flops) *flops = p->flops;
  if (bytes) *bytes = p->bytes;
  return QTG_OK;
}

}  // extern "C"
