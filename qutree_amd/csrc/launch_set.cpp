// LaunchSet: planning and execution of an ordered list of grouped launches.
#include "launch_set.hpp"

#include <algorithm>

namespace qtg {

namespace {

constexpr int kGmSplitK = 128;  // mode products longer than 2 kGmSplitK split along K
constexpr int kGcSplitK = 512;  // hole contractions longer than 2 kGcSplitK split along K

template <typename X>
int up(X** dst, const std::vector<X>& v) {
  *dst = nullptr;
  if (v.empty()) return QTG_OK;
  QTG_HIP_TRY(hipMalloc((void**)dst, sizeof(X) * v.size()));
  QTG_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
  return QTG_OK;
}

}  // namespace

void LaunchSet::release() {
  for (void* p : {(void*)d_gm, (void*)d_gc, (void*)d_gmwg, (void*)d_gcwg})
    if (p) (void)hipFree(p);
  d_gm = nullptr;
  d_gc = nullptr;
  d_gmwg = nullptr;
  d_gcwg = nullptr;
}

void LaunchSet::push_gm(std::vector<GmDesc> descs, int x0, int x1, int m, int o) {
  if (descs.empty()) return;
  Launch L{};
  L.kind = 0;
  L.d0 = (int64_t)gm.size();
  L.nd = (int64_t)descs.size();
  for (auto& d : descs) {
    if (d.op != GM_COPY && d.nin > 2 * kGmSplitK) d.ksplit = split_chunks(d.nin, kGemmWaves, kGmSplitK / 2);
    if (d.op != GM_COPY) L.flops += 8.0 * d.bef * d.aft * (double)d.nin * d.nout;
  }
  L.ntasks = order_tasks(descs, gm_tasks);
  gm.insert(gm.end(), descs.begin(), descs.end());
  L.wg0 = (int64_t)gmwg.size();
  gemm_wg_table(gm.data() + L.d0, L.nd, L.ntasks, gmwg);
  L.x0 = x0;
  L.x1 = x1;
  L.m = m;
  L.o = o;
  flops += L.flops;
  launches.push_back(L);
}

void LaunchSet::push_gc(std::vector<GcDesc> descs, int x0, int x1, int o) {
  if (descs.empty()) return;
  Launch L{};
  L.kind = 1;
  L.d0 = (int64_t)gc.size();
  L.nd = (int64_t)descs.size();
  for (auto& d : descs) {
    const int64_t K = (int64_t)d.bef * d.aft;
    if (K > 2 * kGcSplitK) d.ksplit = split_chunks(K, kGemmWaves, 1);
    L.flops += 8.0 * d.bef * d.aft * (double)d.nx * d.ny;
  }
  L.ntasks = order_tasks(descs, gc_tasks);
  gc.insert(gc.end(), descs.begin(), descs.end());
  L.wg0 = (int64_t)gcwg.size();
  gemm_wg_table(gc.data() + L.d0, L.nd, L.ntasks, gcwg);
  L.x0 = x0;
  L.x1 = x1;
  L.m = -1;
  L.o = o;
  flops += L.flops;
  launches.push_back(L);
}

int LaunchSet::upload() {
  release();
  int rc;
  if ((rc = up(&d_gm, gm))) return rc;
  if ((rc = up(&d_gc, gc))) return rc;
  if ((rc = up(&d_gmwg, gmwg))) return rc;
  if ((rc = up(&d_gcwg, gcwg))) return rc;
  return QTG_OK;
}

int LaunchSet::run(void* const* bases, hipStream_t s) const {
  int rc;
  for (const auto& L : launches) {
    if (L.kind == 0)
      rc = launch_gm<double, double, double>(d_gm + L.d0, d_gmwg + L.wg0, L.nd, L.ntasks, bases[L.x0], bases[L.x1],
                                             bases[L.m], bases[L.o], nullptr, 1.0, 0.0, s);
    else
      rc = launch_gc<double, double>(d_gc + L.d0, d_gcwg + L.wg0, L.nd, L.ntasks, bases[L.x0], bases[L.x1],
                                     bases[L.o], nullptr, s);
    if (rc) return rc;
  }
  return QTG_OK;
}

}  // namespace qtg
