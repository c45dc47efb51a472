// Node-level contractions: mode-k products (matvec/tmatvec/ctmatvec semantics
// of LA_lib.f, matrixTensorBLAS semantics of TensorBLAS.cpp) and hole
// contractions (rhomat / contractionBLAS), single and grouped.  All of them
// run on the MFMA tile kernels of gemm_ops.hip: a mode product is a GEMM over
// the flattened (before, after) rows, a hole contraction a GEMM over the
// flattened (before, after) reduction index.  Grouped hole contractions with
// long reductions are split along K into 2 or 4 chunks that sum in LDS inside
// their workgroup, in chunk order (deterministic).
#include <algorithm>
#include <climits>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm_ops.hpp"

namespace qtg {

static_assert(sizeof(qtg_mode_product_desc) == 64, "descriptor layout");
static_assert(sizeof(qtg_hole_contraction_desc) == 64, "descriptor layout");

namespace {

constexpr int64_t kSplitK = 512;  // K per task of a split hole contraction

int to_gm(const qtg_mode_product_desc& d, GmDesc& g) {
  if (d.before < 1 || d.after < 1 || d.nB < 1 || d.nC < 1 || d.op < 0 || d.op > 2) {
    set_error("mode product: need nB, nC, before, after >= 1 and op in {0,1,2}");
    return QTG_ESHAPE;
  }
  // the tile kernels index rows of the flattened (before, after) in 64 bits
  // and matrices in 32 bits; every product is checked before it is formed
  if (d.before > INT_MAX || d.after > INT_MAX || d.nB > INT_MAX || d.nC > INT_MAX ||
      d.before * d.after > ((int64_t)1 << 40) || d.nB * d.nC >= ((int64_t)1 << 31) ||
      d.before * d.after > (((int64_t)1 << 62) / std::max(d.nB, d.nC))) {
    set_error("mode product: dimensions out of range");
    return QTG_ESHAPE;
  }
  g = GmDesc{};
  g.x = d.offB;
  g.m = d.offM;
  g.out = d.offC;
  g.bef = (int32_t)d.before;
  g.nin = (int32_t)d.nB;
  g.nout = (int32_t)d.nC;
  g.aft = (int32_t)d.after;
  g.op = d.op;
  g.acc = d.add ? 1 : 0;
  g.are = 1.0;
  g.aim = 0.0;
  g.node = -1;
  g.ksplit = 1;
  return QTG_OK;
}

int to_gc(const qtg_hole_contraction_desc& d, GcDesc& g) {
  if (d.before < 1 || d.after < 1 || d.nA < 1 || d.nB < 1) {
    set_error("hole contraction: dimensions must be >= 1");
    return QTG_ESHAPE;
  }
  // The reduction walks k = b + a before over K = before * after and steps a
  // by before * n (gc_product): both stay in 32 bits only below 2^31, with
  // 16 more for the last k-step's overrun.
  const int64_t lim = (int64_t)1 << 31;
  if (d.before >= lim || d.after >= lim || d.nA >= lim || d.nB >= lim || d.before * d.after + 16 >= lim ||
      d.before * std::max(d.nA, d.nB) >= lim || d.nA * d.nB >= lim) {
    set_error("hole contraction: before*after and before*max(nA, nB) must stay below 2^31");
    return QTG_ESHAPE;
  }
  g = GcDesc{};
  g.x = d.offA;
  g.y = d.offB;
  g.out = d.offS;
  g.bef = (int32_t)d.before;
  g.nx = (int32_t)d.nA;
  g.ny = (int32_t)d.nB;
  g.aft = (int32_t)d.after;
  g.acc = d.accumulate ? 1 : 0;
  g.node = -1;
  g.ksplit = 1;
  g.sel = 2;  // X from the first input base, Y from the second
  return QTG_OK;
}

template <typename R>
int mode_product_one(const qtg_mode_product_desc& d, cplx<R>* C, const cplx<R>* M, const cplx<R>* B, hipStream_t s) {
  if (!C || !M || !B) {
    set_error("mode product: null pointer");
    return QTG_EARG;
  }
  GmDesc g;
  int rc = to_gm(d, g);
  if (rc) return rc;
  return launch_gm_single<R>(g, B, M, C, s);
}

template <typename R>
int hole_one(const qtg_hole_contraction_desc& d, cplx<R>* Sm, const cplx<R>* A, const cplx<R>* B, hipStream_t s) {
  if (!Sm || !A || !B) {
    set_error("hole contraction: null pointer");
    return QTG_EARG;
  }
  GcDesc g;
  int rc = to_gc(d, g);
  if (rc) return rc;
  return launch_gc_single<R>(g, A, B, Sm, s);
}

struct Plan {
  int kind = 0;  // 0 mode product, 1 hole contraction
  bool c64 = false;
  int64_t ndesc = 0, ntasks = 0;
  void* d_descs = nullptr;
  void* d_wg = nullptr;  // first descriptor of every workgroup
  double flops = 0, bytes = 0;
  ~Plan() {
    for (void* p : {d_descs, d_wg})
      if (p) (void)hipFree(p);
  }
};

template <typename X>
int upload(void** dst, const std::vector<X>& v) {
  *dst = nullptr;
  if (v.empty()) return QTG_OK;
  QTG_HIP_TRY(hipMalloc(dst, sizeof(X) * v.size()));
  QTG_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
  return QTG_OK;
}

// ------------------------------------------------ no-allocation variant
// Device-resident staging of qtg_hole_contraction_grouped_noalloc_z: the
// converted descriptors, then the per-task descriptor index table.  Each GPU
// holds its own copy for the life of the process.
constexpr int64_t kNoAllocMaxDescs = 4096;
constexpr int64_t kNoAllocMaxTasks = 1 << 18;
constexpr size_t kNoAllocTableOff = (size_t)kNoAllocMaxDescs * sizeof(GcDesc);
constexpr size_t kNoAllocBytes = kNoAllocTableOff + (size_t)kNoAllocMaxTasks * sizeof(int32_t);
static_assert(kNoAllocTableOff % 256 == 0, "table alignment");

}  // namespace
}  // namespace qtg

// In host code this name is only a handle for hipGetSymbolAddress: kernels
// receive the per-device address it returns (scratch_address).
static __device__ __align__(256) unsigned char scratch[qtg::kNoAllocBytes];

namespace qtg {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device with
// that device current.
int scratch_address(void** out) {
  int dev = 0;
  QTG_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("no-allocation contraction: device index out of range");
    return QTG_EDEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("no-allocation contraction: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return QTG_EDEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return QTG_OK;
}

}  // namespace
}  // namespace qtg

using namespace qtg;

extern "C" {

int qtg_mode_product_z(double* C, const double* M, const double* B, int64_t before, int64_t nB, int64_t nC,
                       int64_t after, int op, int add, qtg_stream stream) {
  qtg_mode_product_desc d{0, 0, 0, before, nB, nC, after, op, add};
  return mode_product_one<double>(d, (cplx<double>*)C, (const cplx<double>*)M, (const cplx<double>*)B,
                                        S(stream));
}
int qtg_mode_product_c(float* C, const float* M, const float* B, int64_t before, int64_t nB, int64_t nC,
                       int64_t after, int op, int add, qtg_stream stream) {
  qtg_mode_product_desc d{0, 0, 0, before, nB, nC, after, op, add};
  return mode_product_one<float>(d, (cplx<float>*)C, (const cplx<float>*)M, (const cplx<float>*)B,
                                       S(stream));
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
  return hole_one<double>(d, (cplx<double>*)Sm, (const cplx<double>*)A, (const cplx<double>*)B,
                                S(stream));
}
int qtg_hole_contraction_c(float* Sm, const float* A, const float* B, int64_t before, int64_t nA, int64_t nB,
                           int64_t after, int accumulate, qtg_stream stream) {
  qtg_hole_contraction_desc d{0, 0, 0, before, nA, nB, after, accumulate, 0};
  return hole_one<float>(d, (cplx<float>*)Sm, (const cplx<float>*)A, (const cplx<float>*)B, S(stream));
}
int qtg_rhomat_z(const double* bra, const double* ket, double* M, int64_t a, int64_t b, int64_t c,
                 qtg_stream stream) {
  return qtg_hole_contraction_z(M, bra, ket, b, a, a, c, 0, stream);
}

int qtg_rmatvec_d(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                  qtg_stream stream) {
  if (!C || !B || !M) {
    set_error("rmatvec: null pointer");
    return QTG_EARG;
  }
  qtg_mode_product_desc d{0, 0, 0, b, a, a, c, 0, add};
  GmDesc g;
  int rc = to_gm(d, g);
  if (rc) return rc;
  return launch_gm_real(g, B, M, C, S(stream));
}
int qtg_trmatvec_d(double* C, const double* B, const double* M, int64_t a, int64_t b, int64_t c, int add,
                   qtg_stream stream) {
  if (!C || !B || !M) {
    set_error("trmatvec: null pointer");
    return QTG_EARG;
  }
  qtg_mode_product_desc d{0, 0, 0, b, a, a, c, 1, add};
  GmDesc g;
  int rc = to_gm(d, g);
  if (rc) return rc;
  return launch_gm_real(g, B, M, C, S(stream));
}

int qtg_mode_product_plan_create(qtg_plan* plan, const qtg_mode_product_desc* descs, int64_t count,
                                 int complex64) {
  if (!plan || (!descs && count > 0) || count < 0 || count > INT_MAX) {
    set_error("plan: bad arguments");
    return QTG_EARG;
  }
  auto* p = new Plan();
  p->kind = 0;
  p->c64 = complex64 != 0;
  const double es = p->c64 ? 8.0 : 16.0;
  std::vector<GmDesc> gv;
  int64_t tasks = 0;
  for (int64_t i = 0; i < count; ++i) {
    GmDesc g;
    int rc = to_gm(descs[i], g);
    if (rc) {
      delete p;
      return rc;
    }
    g.tile0 = tasks;
    tasks += gm_tasks(g);
    gv.push_back(g);
    const auto& d = descs[i];
    p->flops += 8.0 * d.before * d.after * d.nB * d.nC;
    p->bytes += es * (d.before * d.after * (d.nB + d.nC * (d.add ? 2 : 1)) + d.nB * d.nC);
  }
  std::vector<int32_t> wg;
  gemm_wg_table(gv.data(), (int64_t)gv.size(), tasks, wg);
  int rc = upload(&p->d_descs, gv);
  if (!rc) rc = upload(&p->d_wg, wg);
  if (rc) {
    delete p;
    return rc;
  }
  p->ndesc = count;
  p->ntasks = tasks;
  *plan = p;
  return QTG_OK;
}

int qtg_hole_contraction_plan_create(qtg_plan* plan, const qtg_hole_contraction_desc* descs, int64_t count,
                                     int complex64) {
  if (!plan || (!descs && count > 0) || count < 0 || count > INT_MAX) {
    set_error("plan: bad arguments");
    return QTG_EARG;
  }
  auto* p = new Plan();
  p->kind = 1;
  p->c64 = complex64 != 0;
  const double es = p->c64 ? 8.0 : 16.0;
  std::vector<GcDesc> gv;
  for (int64_t i = 0; i < count; ++i) {
    GcDesc g;
    int rc = to_gc(descs[i], g);
    if (rc) {
      delete p;
      return rc;
    }
    const int64_t K = (int64_t)g.bef * g.aft;
    if (K > 2 * kSplitK) g.ksplit = split_chunks(K, kGemmWaves, 1);
    gv.push_back(g);
    const auto& d = descs[i];
    p->flops += 8.0 * K * d.nA * d.nB;
    p->bytes += es * (K * (d.nA + d.nB) + d.nA * d.nB * (d.accumulate ? 2 : 1));
  }
  const int64_t tasks = order_tasks(gv, gc_tasks);
  std::vector<int32_t> wg;
  gemm_wg_table(gv.data(), (int64_t)gv.size(), tasks, wg);
  int rc = upload(&p->d_descs, gv);
  if (!rc) rc = upload(&p->d_wg, wg);
  if (rc) {
    delete p;
    return rc;
  }
  p->ndesc = count;
  p->ntasks = tasks;
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
    // first = M base, second = B base
    const auto* g = (const GmDesc*)p->d_descs;
    const auto* wg = (const int32_t*)p->d_wg;
    return p->c64 ? launch_gm<float, float, float>(g, wg, p->ndesc, p->ntasks, second_base, second_base,
                                                   first_base, out_base, nullptr, 1.0, 0.0, s)
                  : launch_gm<double, double, double>(g, wg, p->ndesc, p->ntasks, second_base, second_base,
                                                      first_base, out_base, nullptr, 1.0, 0.0, s);
  }
  const auto* g = (const GcDesc*)p->d_descs;
  // X (bra) from the first base, Y (ket) from the second: sel = 2 on every descriptor
  const auto* wg = (const int32_t*)p->d_wg;
  return p->c64 ? launch_gc<float, float>(g, wg, p->ndesc, p->ntasks, first_base, second_base, out_base, nullptr, s)
                : launch_gc<double, double>(g, wg, p->ndesc, p->ntasks, first_base, second_base, out_base, nullptr, s);
}

int qtg_hole_contraction_grouped_noalloc_z(const qtg_hole_contraction_desc* descs, int64_t count, double* S_base,
                                           const double* A_base, const double* B_base, qtg_stream stream) {
  if ((!descs && count > 0) || count < 0 || !S_base || !A_base || !B_base) {
    set_error("no-allocation contraction: bad arguments");
    return QTG_EARG;
  }
  if (count > kNoAllocMaxDescs) {
    set_error("no-allocation contraction: more than " + std::to_string(kNoAllocMaxDescs) + " descriptors");
    return QTG_ESHAPE;
  }
  if (count == 0) return QTG_OK;
  // the same descriptors, K splits and task order as a plan of these descs
  std::vector<GcDesc> gv;
  gv.reserve((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    GcDesc g;
    int rc = to_gc(descs[i], g);
    if (rc) return rc;
    const int64_t K = (int64_t)g.bef * g.aft;
    if (K > 2 * kSplitK) g.ksplit = split_chunks(K, kGemmWaves, 1);
    gv.push_back(g);
  }
  const int64_t tasks = order_tasks(gv, gc_tasks);
  if (tasks > kNoAllocMaxTasks) {
    set_error("no-allocation contraction: more than " + std::to_string(kNoAllocMaxTasks) + " 16 x 16 tiles");
    return QTG_ESHAPE;
  }
  std::vector<int32_t> wg;
  gemm_wg_table(gv.data(), (int64_t)gv.size(), tasks, wg);
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  hipStream_t s = S(stream);
  GcDesc* d_descs = reinterpret_cast<GcDesc*>(base);
  int32_t* d_wg = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(base) + kNoAllocTableOff);
  QTG_HIP_TRY(hipMemcpyAsync(d_descs, gv.data(), sizeof(GcDesc) * gv.size(), hipMemcpyHostToDevice, s));
  QTG_HIP_TRY(hipMemcpyAsync(d_wg, wg.data(), sizeof(int32_t) * wg.size(), hipMemcpyHostToDevice, s));
  rc = launch_gc<double, double>(d_descs, d_wg, count, tasks, A_base, B_base, S_base, nullptr, s);
  if (rc) return rc;
  // the staging tables are host temporaries and the device copy is shared by
  // every call on this GPU: the call returns with the result written
  QTG_HIP_TRY(hipStreamSynchronize(s));
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
