// LaunchSet: planning and execution of an ordered list of grouped launches.
#include "launch_set.hpp"

#include <algorithm>

namespace qtg {

namespace {

constexpr int kGmSplitK = 128;  // K per task of a split mode product
constexpr int kGcSplitK = 512;  // K per task of a split hole contraction

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
  for (void* p : {(void*)d_gm, (void*)d_gc, (void*)d_gmsplit, (void*)d_gcsplit, (void*)d_gmwg, (void*)d_gcwg, d_part})
    if (p) (void)hipFree(p);
  d_gm = nullptr;
  d_gc = nullptr;
  d_gmsplit = nullptr;
  d_gcsplit = nullptr;
  d_gmwg = nullptr;
  d_gcwg = nullptr;
  d_part = nullptr;
}

void LaunchSet::push_gm(std::vector<GmDesc> descs, int x0, int x1, int m, int o) {
  if (descs.empty()) return;
  Launch L{};
  L.kind = 0;
  L.d0 = (int64_t)gm.size();
  L.nd = (int64_t)descs.size();
  L.s0 = (int64_t)gmsplit.size();
  int64_t t = 0, part = 0, rblocks = 0;
  for (auto& d : descs) {
    if (d.op != GM_COPY && d.nin > 2 * kGmSplitK) {
      d.ksplit = (d.nin + kGmSplitK - 1) / kGmSplitK;
      d.part = part;
      part += (int64_t)d.ksplit * d.bef * d.aft * d.nout;
      GmDesc r = d;
      r.tile0 = rblocks;
      rblocks += ((int64_t)d.bef * d.aft * d.nout + 255) / 256;
      gmsplit.push_back(r);
    }
    d.tile0 = t;
    t += gm_tasks(d);
    if (d.op != GM_COPY) L.flops += 8.0 * d.bef * d.aft * (double)d.nin * d.nout;
    gm.push_back(d);
  }
  L.ntasks = t;
  L.ns = (int64_t)gmsplit.size() - L.s0;
  L.nrb = rblocks;
  L.wg0 = (int64_t)gmwg.size();
  gemm_wg_table(gm.data() + L.d0, L.nd, t, gmwg);
  L.x0 = x0;
  L.x1 = x1;
  L.m = m;
  L.o = o;
  partSize = std::max(partSize, part);
  flops += L.flops;
  launches.push_back(L);
}

void LaunchSet::push_gc(std::vector<GcDesc> descs, int x0, int x1, int o) {
  if (descs.empty()) return;
  Launch L{};
  L.kind = 1;
  L.d0 = (int64_t)gc.size();
  L.nd = (int64_t)descs.size();
  L.s0 = (int64_t)gcsplit.size();
  int64_t t = 0, part = 0, rblocks = 0;
  for (auto& d : descs) {
    const int64_t K = (int64_t)d.bef * d.aft;
    if (K > 2 * kGcSplitK) {
      d.ksplit = (int32_t)((K + kGcSplitK - 1) / kGcSplitK);
      d.part = part;
      part += (int64_t)d.ksplit * d.nx * d.ny;
      GcDesc r = d;
      r.tile0 = rblocks;
      rblocks += ((int64_t)d.nx * d.ny + 255) / 256;
      gcsplit.push_back(r);
    }
    d.tile0 = t;
    t += gc_tasks(d);
    L.flops += 8.0 * d.bef * d.aft * (double)d.nx * d.ny;
    gc.push_back(d);
  }
  L.ntasks = t;
  L.ns = (int64_t)gcsplit.size() - L.s0;
  L.nrb = rblocks;
  L.wg0 = (int64_t)gcwg.size();
  gemm_wg_table(gc.data() + L.d0, L.nd, t, gcwg);
  L.x0 = x0;
  L.x1 = x1;
  L.m = -1;
  L.o = o;
  partSize = std::max(partSize, part);
  flops += L.flops;
  launches.push_back(L);
}

int LaunchSet::upload() {
  release();
  int rc;
  if ((rc = up(&d_gm, gm))) return rc;
  if ((rc = up(&d_gc, gc))) return rc;
  if ((rc = up(&d_gmsplit, gmsplit))) return rc;
  if ((rc = up(&d_gcsplit, gcsplit))) return rc;
  if ((rc = up(&d_gmwg, gmwg))) return rc;
  if ((rc = up(&d_gcwg, gcwg))) return rc;
  QTG_HIP_TRY(hipMalloc(&d_part, 16 * (size_t)std::max<int64_t>(partSize, 1)));
  return QTG_OK;
}

int LaunchSet::run(void* const* bases, hipStream_t s) const {
  int rc;
  for (const auto& L : launches) {
    if (L.kind == 0) {
      if ((rc = launch_gm<double, double, double>(d_gm + L.d0, d_gmwg + L.wg0, L.nd, L.ntasks, bases[L.x0], bases[L.x1],
                                          bases[L.m], bases[L.o], d_part, nullptr, 1.0, 0.0, s)))
        return rc;
      if (L.ns > 0 &&
          (rc = launch_gm_reduce<double>(d_gmsplit + L.s0, L.ns, L.nrb, bases[L.o], d_part, nullptr, s)))
        return rc;
    } else {
      if ((rc = launch_gc<double, double>(d_gc + L.d0, d_gcwg + L.wg0, L.nd, L.ntasks, bases[L.x0], bases[L.x1],
                                  bases[L.o], d_part, nullptr, s)))
        return rc;
      if (L.ns > 0 &&
          (rc = launch_gc_reduce<double>(d_gcsplit + L.s0, L.ns, L.nrb, bases[L.o], d_part, nullptr, s)))
        return rc;
    }
  }
  return QTG_OK;
}

}  // namespace qtg
