// C ABI of the tree engine (declared in include/qutree_amd.h).
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cmf.hpp"
#include "scf.hpp"
#include "apply_fit.hpp"
#include <algorithm>
#include "ttno.hpp"
#include "engine.hpp"

using namespace qtg;

namespace {
// a NULL stream means the engine's own (non-blocking) stream
inline hipStream_t ES(qtg_engine e, qtg_stream s) {
  return s ? S(s) : static_cast<Engine*>(e)->stream();
}
}  // namespace

extern "C" {

int qtg_engine_create(const char* tree_text, int precision, qtg_engine* out) {
  if (!tree_text || !out) {
    set_error("engine_create: null argument");
    return QTG_EARG;
  }
  auto* e = new Engine();
  int rc = e->create(tree_text, precision);
  if (rc) {
    delete e;
    return rc;
  }
  *out = e;
  return QTG_OK;
}

int qtg_engine_destroy(qtg_engine e) {
  delete static_cast<Engine*>(e);
  return QTG_OK;
}

int qtg_engine_set_hamiltonian(qtg_engine e, int64_t nterms, const double* coeff, const int64_t* term_off,
                               const int64_t* fmode, const int64_t* fkind, const int64_t* fpool,
                               const double* pool) {
  if (!e || nterms < 0 || (nterms > 0 && (!coeff || !term_off || !fmode || !fkind || !fpool || !pool))) {
    set_error("engine_set_hamiltonian: bad argument");
    return QTG_EARG;
  }
  return static_cast<Engine*>(e)->set_hamiltonian(nterms, coeff, term_off, fmode, fkind, fpool, pool);
}

int64_t qtg_engine_psi_size(qtg_engine e) { return e ? static_cast<Engine*>(e)->tree().total : -1; }
int64_t qtg_engine_num_nodes(qtg_engine e) {
  return e ? (int64_t) static_cast<Engine*>(e)->tree().nodes.size() : -1;
}

int qtg_engine_node_info(qtg_engine e, int64_t addr, int64_t* info, int64_t len) {
  if (!e || !info) return QTG_EARG;
  const HTree& t = static_cast<Engine*>(e)->tree();
  if (addr < 0 || addr >= (int64_t)t.nodes.size()) return QTG_EARG;
  const HNode& n = t.nodes[addr];
  std::vector<int64_t> v = {n.parent, n.childIdx, n.bottom ? 1 : 0, n.depth, n.order(), (int64_t)n.children.size(),
                            n.bottom ? n.leaf.mode : -1, n.off};
  for (int d : n.dims) v.push_back(d);
  for (int c : n.children) v.push_back(c);
  if (len < (int64_t)v.size()) return QTG_EARG;
  std::copy(v.begin(), v.end(), info);
  return QTG_OK;
}

int qtg_engine_fill_random(qtg_engine e, uint64_t seed, int delta_lowest, double* psi_host) {
  if (!e || !psi_host) return QTG_EARG;
  std::vector<cd> psi;
  fill_random(static_cast<Engine*>(e)->tree(), seed, delta_lowest != 0, psi);
  std::memcpy(psi_host, psi.data(), sizeof(cd) * psi.size());
  return QTG_OK;
}

int qtg_engine_set_psi(qtg_engine e, const double* psi_host, qtg_stream s) {
  if (!e || !psi_host) return QTG_EARG;
  return static_cast<Engine*>(e)->set_psi(psi_host, ES(e, s));
}

int qtg_engine_set_psi_device(qtg_engine e, const double* psi_dev, qtg_stream s) {
  if (!e || !psi_dev) return QTG_EARG;
  return static_cast<Engine*>(e)->set_psi_device(psi_dev, ES(e, s));
}

int qtg_engine_build(qtg_engine e, qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->build(ES(e, s));
}

int qtg_engine_derivative(qtg_engine e, double phase_re, double phase_im, qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->derivative(ES(e, s), phase_re, phase_im);
}

int qtg_engine_layer_derivative_all(qtg_engine e, double phase_re, double phase_im, qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->layer_derivative_all(ES(e, s), phase_re, phase_im);
}

int qtg_engine_get_dpsi(qtg_engine e, double* host, int which, qtg_stream s) {
  if (!e || !host) return QTG_EARG;
  auto* en = static_cast<Engine*>(e);
  const double* src = which == 0 ? en->dpsi64() : (which == 1 ? en->apply64() : en->psi64());
  QTG_HIP_TRY(hipMemcpyAsync(host, src, sizeof(double) * 2 * en->tree().total, hipMemcpyDeviceToHost, ES(e, s)));
  QTG_HIP_TRY(hipStreamSynchronize(ES(e, s)));
  return QTG_OK;
}

int qtg_engine_energy(qtg_engine e, double* energy, qtg_stream s) {
  if (!e || !energy) return QTG_EARG;
  return static_cast<Engine*>(e)->energy(ES(e, s), energy);
}

int qtg_engine_get_matrix(qtg_engine e, int which, int64_t term, int64_t node, double* out, qtg_stream s) {
  if (!e || !out) return QTG_EARG;
  return static_cast<Engine*>(e)->get_matrix(which, (int)term, (int)node, out, ES(e, s));
}

int qtg_engine_counts(qtg_engine e, double* sweep_flops, int64_t* items) {
  if (!e) return QTG_EARG;
  auto* en = static_cast<Engine*>(e);
  if (sweep_flops) *sweep_flops = en->reference_flops();
  if (items) {
    items[0] = en->sandwich_items();
    items[1] = en->apply_terms_chunks();
    items[2] = en->apply_terms();
  }
  return QTG_OK;
}

int qtg_engine_gradient(qtg_engine e, double* grad_host, double* energy, qtg_stream s) {
  if (!e) return QTG_EARG;
  auto* en = static_cast<Engine*>(e);
  return en->gradient(ES(e, s), grad_host, energy);
}

int qtg_engine_tuning(qtg_engine e, char* buf, int64_t cap) {
  if (!e) return QTG_EARG;
  const std::string& t = static_cast<Engine*>(e)->tuned_from();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>((size_t)cap - 1, t.size());
    std::memcpy(buf, t.data(), n);
    buf[n] = 0;
  }
  return QTG_OK;
}

int qtg_engine_forward(qtg_engine e, double* energy, qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->forward(ES(e, s), energy);
}

int qtg_engine_backward(qtg_engine e, double* grad_host, qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->backward(ES(e, s), grad_host);
}

int qtg_engine_gradient_descent(qtg_engine e, int64_t steps, double eta, double* energies, double* psi_host,
                                qtg_stream s) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->gradient_descent(ES(e, s), steps, eta, energies, psi_host);
}

int qtg_engine_canonicalize(qtg_engine e, double* psi_host, qtg_stream s) {
  if (!e) return QTG_EARG;
  auto* en = static_cast<Engine*>(e);
  if (!en->planned() || en->complex64()) {
    set_error("canonicalize: needs a complex128 engine with a Hamiltonian");
    return QTG_ESTATE;
  }
  hipStream_t st = ES(e, s);
  int rc = en->canonicalize(st);
  if (rc) return rc;
  if (psi_host) {
    QTG_HIP_TRY(hipMemcpyAsync(psi_host, en->psi64(), sizeof(double) * 2 * en->tree().total, hipMemcpyDeviceToHost,
                               st));
    QTG_HIP_TRY(hipStreamSynchronize(st));
  }
  return QTG_OK;
}

int qtg_engine_plan_flops(qtg_engine e, double* executed) {
  if (!e || !executed) return QTG_EARG;
  *executed = static_cast<Engine*>(e)->sweep_flops();
  return QTG_OK;
}

// Host-only planning (no device is touched): counts of the work-item plan.
int qtg_host_plan_counts(const char* tree_text, int64_t nterms, const double* coeff, const int64_t* term_off,
                         const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool,
                         int compress, double* sweep_flops, int64_t* items) {
  if (!tree_text) return QTG_EARG;
  Engine en;
  int rc = en.create(tree_text, 0, true);
  if (rc) return rc;
  en.compress_ = compress != 0;
  // counts only: the work-item plan needs no device structures; with
  // QTG_DUMP_PLAN set the grouped plan is made too and printed
  en.path_ = std::getenv("QTG_DUMP_PLAN") ? 1 : 0;
  rc = en.set_hamiltonian(nterms, coeff, term_off, fmode, fkind, fpool, pool);
  if (rc) return rc;
  return qtg_engine_counts(&en, sweep_flops, items);
}

int qtg_host_scf_plan(const char* tree_text, int64_t nterms, const double* coeff, const int64_t* term_off,
                      const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool,
                      int64_t n_krylov, int64_t* info) {
  if (!tree_text || !info) return QTG_EARG;
  Engine en;
  int rc = en.create(tree_text, 0, true);
  if (rc) return rc;
  if ((rc = en.set_hamiltonian(nterms, coeff, term_off, fmode, fkind, fpool, pool))) return rc;
  Scf sc(en);
  if ((rc = sc.init(n_krylov))) return rc;
  info[0] = sc.sweep_length();
  info[1] = sc.launches_per_iteration();
  info[2] = sc.work_elems();
  info[3] = sc.pool_elems();
  info[4] = sc.gm_tasks_total();
  info[5] = sc.gc_tasks_total();
  return QTG_OK;
}

int qtg_host_apply_fit_plan(const char* tree_text, int64_t nterms, const double* coeff, const int64_t* term_off,
                            const int64_t* fmode, const int64_t* fkind, const int64_t* fpool, const double* pool,
                            int64_t* info) {
  if (!tree_text || !info) return QTG_EARG;
  Engine en;
  int rc = en.create(tree_text, 0, true);
  if (rc) return rc;
  if ((rc = en.set_hamiltonian(nterms, coeff, term_off, fmode, fkind, fpool, pool))) return rc;
  ApplyFit f(en);
  if ((rc = f.init())) return rc;
  info[0] = f.launches_per_iteration();
  info[1] = f.work_elems();
  info[2] = f.pool_elems();
  info[3] = f.gm_tasks_total();
  return QTG_OK;
}

int64_t qtg_host_psi_size(const char* tree_text) {
  if (!tree_text) return -1;
  Engine en;
  if (en.create(tree_text, 0, true)) return -1;
  return en.tree().total;
}

int qtg_host_fill_random(const char* tree_text, uint64_t seed, int delta_lowest, double* psi_host) {
  if (!tree_text || !psi_host) return QTG_EARG;
  Engine en;
  int rc = en.create(tree_text, 0, true);
  if (rc) return rc;
  std::vector<cd> psi;
  fill_random(en.tree(), seed, delta_lowest != 0, psi);
  std::memcpy(psi_host, psi.data(), sizeof(cd) * psi.size());
  return QTG_OK;
}

int64_t qtg_host_fill_random_stream(const char* tree_text, uint64_t seed, int64_t discard, int delta_lowest,
                                    double* psi_host) {
  if (!tree_text || !psi_host || discard < 0) return -QTG_EARG;
  Engine en;
  int rc = en.create(tree_text, 0, true);
  if (rc) return -rc;
  std::vector<cd> psi;
  const int64_t used = fill_random_stream(en.tree(), seed, discard, delta_lowest != 0, psi);
  std::memcpy(psi_host, psi.data(), sizeof(cd) * psi.size());
  return used;
}

int qtg_engine_synchronize(qtg_engine e) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->synchronize();
}

int qtg_engine_set_timing(qtg_engine e, int on) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->set_timing(on != 0);
}

int qtg_engine_read_timing(qtg_engine e, double* ms, int64_t* launches, double* flops, double* bytes) {
  if (!e) return QTG_EARG;
  return static_cast<Engine*>(e)->read_timing(ms, launches, flops, bytes);
}

int qtg_engine_set_option(qtg_engine e, const char* key, double value) {
  if (!e || !key) return QTG_EARG;
  auto* en = static_cast<Engine*>(e);
  // keys that shape the plan (buffer sizes, launch lists) are fixed by
  // set_hamiltonian: changing them afterwards would run kernels on buffers
  // sized for another plan
  static const char* const planKeys[] = {"compress",  "path",       "split_waves", "split_min_k",
                                         "gc_long_k", "gm_split_k", "tuning",      "fuse_holes"};
  for (const char* k : planKeys)
    if (!std::strcmp(key, k) && en->planned()) {
      set_error(std::string("option ") + key + ": set it before qtg_engine_set_hamiltonian");
      return QTG_ESTATE;
    }
  if (!std::strcmp(key, "rho_eps")) {
    if (!(value > 0.0)) {
      set_error("option rho_eps: must be positive");
      return QTG_EARG;
    }
    en->eps_ = value;
    // a captured sweep holds the regulariser as a launch argument
    en->invalidate_graph();
    return QTG_OK;
  }
  if (!std::strcmp(key, "compress")) {
    en->compress_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "path")) {
    if (value != 0.0 && value != 1.0) {
      set_error("option path: 0 (LDS-resident node kernels) or 1 (grouped MFMA contractions)");
      return QTG_EARG;
    }
    en->path_ = (int)value;
    return QTG_OK;
  }
  if (!std::strcmp(key, "split_waves")) {
    if (value < 1.0) {
      set_error("option split_waves: at least 1");
      return QTG_EARG;
    }
    en->splitWaves_ = (int64_t)value;
    en->userSet_ |= 1;
    return QTG_OK;
  }
  if (!std::strcmp(key, "split_min_k")) {
    if (value < 4.0) {
      set_error("option split_min_k: at least 4");
      return QTG_EARG;
    }
    en->splitMinK_ = (int)value;
    en->userSet_ |= 2;
    return QTG_OK;
  }
  if (!std::strcmp(key, "gc_long_k")) {
    if (value < 16.0) {
      set_error("option gc_long_k: at least 16");
      return QTG_EARG;
    }
    en->gcLongK_ = (int)value;
    en->userSet_ |= 4;
    return QTG_OK;
  }
  if (!std::strcmp(key, "gm_split_k")) {
    if (value < 16.0) {
      set_error("option gm_split_k: at least 16");
      return QTG_EARG;
    }
    en->gmSplitK_ = (int)value;
    en->userSet_ |= 8;
    return QTG_OK;
  }
  if (!std::strcmp(key, "tuning")) {
    en->tuning_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "grad_recompute")) {
    en->gradRecompute_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "graph")) {
    en->graph_ = value != 0.0;
    en->invalidate_graph();
    return QTG_OK;
  }
  if (!std::strcmp(key, "fuse_holes")) {
    en->fuseHoles_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "heev_side")) {
    en->heevSide_ = value != 0.0;
    en->invalidate_graph();
    return QTG_OK;
  }
  set_error(std::string("unknown option ") + key);
  return QTG_EARG;
}

// ------------------------------------------------------------------------ CMF
int qtg_cmf_create(qtg_engine e, qtg_cmf* out) {
  if (!e || !out) return QTG_EARG;
  auto* c = new Cmf(*static_cast<Engine*>(e));
  int rc = c->init();
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return QTG_OK;
}

int qtg_cmf_set_option(qtg_cmf c, const char* key, double value) {
  if (!c || !key) return QTG_EARG;
  auto* cm = static_cast<Cmf*>(c);
  if (!std::strcmp(key, "graph")) {
    cm->tickGraph_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "fuse")) {
    cm->fuseOps_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "lookahead")) {
    cm->lookahead_ = value != 0.0;
    return QTG_OK;
  }
  if (!std::strcmp(key, "device_bs")) {
    cm->deviceBs_ = value != 0.0;
    return QTG_OK;
  }
  set_error(std::string("unknown cmf option ") + key);
  return QTG_EARG;
}

int qtg_cmf_set_output_file(qtg_cmf c, const char* filename, int append) {
  if (!c) return QTG_EARG;
  auto* cm = static_cast<Cmf*>(c);
  cm->outFile_ = filename ? filename : "";
  cm->outAppend_ = append != 0;
  return QTG_OK;
}

int qtg_cmf_destroy(qtg_cmf c) {
  delete static_cast<Cmf*>(c);
  return QTG_OK;
}

int64_t qtg_cmf_integrate(qtg_cmf c, double* psi_host, double t0, double tend, double dt, double out,
                          double acc_cmf, double acc_bs, double phase_re, double phase_im, int64_t max_steps,
                          int64_t max_out, double* times, double* energies, double* autocorr, int64_t* stats) {
  if (!c || !psi_host) {
    set_error("cmf_integrate: null argument");
    return -QTG_EARG;
  }
  auto* cm = static_cast<Cmf*>(c);
  const size_t bytes = sizeof(double) * 2 * cm->size();
  if (hipMemcpyAsync(cm->state(), psi_host, bytes, hipMemcpyHostToDevice, cm->stream()) != hipSuccess ||
      hipStreamSynchronize(cm->stream()) != hipSuccess) {
    set_error("cmf_integrate: upload failed");
    return -QTG_EDEVICE;
  }
  cm->ph_re_ = phase_re;
  cm->ph_im_ = phase_im;
  CmfResult r;
  int rc = cm->integrate(t0, tend, dt, out, acc_cmf, acc_bs, max_steps, r);
  if (rc) return -rc;
  if (hipMemcpyAsync(psi_host, cm->state(), bytes, hipMemcpyDeviceToHost, cm->stream()) != hipSuccess ||
      hipStreamSynchronize(cm->stream()) != hipSuccess) {
    set_error("cmf_integrate: download failed");
    return -QTG_EDEVICE;
  }
  const int64_t n = std::min<int64_t>((int64_t)r.times.size(), max_out);
  for (int64_t i = 0; i < n; ++i) {
    if (times) times[i] = r.times[i];
    if (energies) energies[i] = r.energies[i];
    if (autocorr) {
      autocorr[2 * i] = r.autocorr[i].real();
      autocorr[2 * i + 1] = r.autocorr[i].imag();
    }
  }
  if (stats) {
    stats[0] = (int64_t)r.steps;
    stats[1] = (int64_t)r.rejected;
    stats[2] = (int64_t)r.builds;
    stats[3] = (int64_t)r.derivs;
    stats[4] = (int64_t)r.ticks;
  }
  return (int64_t)r.times.size();
}

int64_t qtg_cmf_eigenstates(qtg_cmf c, double* psi_host, double dt, double out, int64_t iterations, double acc_cmf,
                            double acc_bs, int64_t max_steps, double* energies, double* propagator,
                            double* converged, int64_t* stats, double* dt_end) {
  if (!c || !psi_host || !energies || iterations < 0 || !(out > 1e-5)) {
    set_error("cmf_eigenstates: null argument, negative iteration count or output interval <= 1e-5");
    return -QTG_EARG;
  }
  if (iterations > 0 && (!propagator || !converged)) {
    set_error("cmf_eigenstates: null output array");
    return -QTG_EARG;
  }
  auto* cm = static_cast<Cmf*>(c);
  const size_t bytes = sizeof(double) * 2 * cm->size();
  if (hipMemcpyAsync(cm->state(), psi_host, bytes, hipMemcpyHostToDevice, cm->stream()) != hipSuccess ||
      hipStreamSynchronize(cm->stream()) != hipSuccess) {
    set_error("cmf_eigenstates: upload failed");
    return -QTG_EDEVICE;
  }
  EigenResult r;
  int rc = cm->eigenstates(dt, out, iterations, acc_cmf, acc_bs, max_steps, r);
  if (rc) return -rc;
  if (hipMemcpyAsync(psi_host, cm->state(), bytes, hipMemcpyDeviceToHost, cm->stream()) != hipSuccess ||
      hipStreamSynchronize(cm->stream()) != hipSuccess) {
    set_error("cmf_eigenstates: download failed");
    return -QTG_EDEVICE;
  }
  std::copy(r.energies.begin(), r.energies.end(), energies);
  std::copy(r.propagator.begin(), r.propagator.end(), propagator);
  std::copy(r.converged.begin(), r.converged.end(), converged);
  if (stats) {
    stats[0] = (int64_t)r.steps;
    stats[1] = (int64_t)r.rejected;
    stats[2] = (int64_t)r.builds;
    stats[3] = (int64_t)r.derivs;
  }
  if (dt_end) *dt_end = r.dt_end;
  return cm->nstates();
}

// ----------------------------------------------------------------------- TTNO
int qtg_ttno_create(qtg_engine e, const int64_t* bond, const double* B, qtg_ttno* ttno) {
  if (!e || !bond || !B || !ttno) return QTG_EARG;
  auto* t = new Ttno(*static_cast<Engine*>(e));
  int rc = t->create(bond, B);
  if (rc) {
    delete t;
    return rc;
  }
  *ttno = t;
  return QTG_OK;
}

int qtg_ttno_destroy(qtg_ttno ttno) {
  delete static_cast<Ttno*>(ttno);
  return QTG_OK;
}

int qtg_ttno_apply(qtg_ttno ttno, const double* bra, const double* ket, double* out, double* expect,
                   qtg_stream stream) {
  if (!ttno || !bra || !ket) return QTG_EARG;
  return static_cast<Ttno*>(ttno)->apply(bra, ket, out, expect, S(stream));
}

int qtg_scf_create(qtg_engine e, int64_t n_krylov, qtg_scf* out) {
  if (!e || !out) return QTG_EARG;
  auto* sc = new Scf(*static_cast<Engine*>(e));
  int rc = sc->init(n_krylov);
  if (rc) {
    delete sc;
    return rc;
  }
  *out = sc;
  return QTG_OK;
}

int qtg_scf_destroy(qtg_scf s) {
  delete static_cast<Scf*>(s);
  return QTG_OK;
}

int64_t qtg_scf_sweep_length(qtg_scf s) { return s ? static_cast<Scf*>(s)->sweep_length() : -QTG_EARG; }

int qtg_scf_run(qtg_scf s, double* psi_host, int64_t n_iter, int64_t n_itp, double beta, double* energies,
                int64_t cap) {
  if (!s || !psi_host || n_iter < 0 || n_itp < 0 || cap < 0 || (cap > 0 && !energies)) {
    set_error("scf_run: bad argument");
    return QTG_EARG;
  }
  return static_cast<Scf*>(s)->run(psi_host, n_iter, n_itp, beta, energies, cap);
}

int qtg_scf_set_option(qtg_scf s, const char* key, double value) {
  if (!s || !key) return QTG_EARG;
  auto* sc = static_cast<Scf*>(s);
  if (!std::strcmp(key, "graph")) {
    sc->graph_ = value != 0.0;
    return QTG_OK;
  }
  set_error(std::string("scf_set_option: unknown key ") + key);
  return QTG_EARG;
}

int qtg_scf_info(qtg_scf s, int64_t* info) {
  if (!s || !info) return QTG_EARG;
  auto* sc = static_cast<Scf*>(s);
  info[0] = sc->sweep_length();
  info[1] = sc->launches_per_iteration();
  info[2] = sc->device_bytes();
  return QTG_OK;
}

int qtg_apply_fit_create(qtg_engine e, qtg_apply_fit* out) {
  if (!e || !out) return QTG_EARG;
  auto* f = new ApplyFit(*static_cast<Engine*>(e));
  int rc = f->init();
  if (rc) {
    delete f;
    return rc;
  }
  *out = f;
  return QTG_OK;
}

int qtg_apply_fit_destroy(qtg_apply_fit f) {
  delete static_cast<ApplyFit*>(f);
  return QTG_OK;
}

int qtg_apply_fit_run(qtg_apply_fit f, double* psi_host, uint64_t seed, int64_t max_iter, double* fidelities,
                      int64_t cap, int64_t* iterations) {
  if (!f || !psi_host || max_iter < 1 || cap < 0 || (cap > 0 && !fidelities)) {
    set_error("apply_fit_run: bad argument");
    return QTG_EARG;
  }
  return static_cast<ApplyFit*>(f)->run(psi_host, seed, max_iter, fidelities, cap, iterations);
}

int qtg_apply_fit_set_option(qtg_apply_fit f, const char* key, double value) {
  if (!f || !key) return QTG_EARG;
  if (!std::strcmp(key, "graph")) {
    static_cast<ApplyFit*>(f)->graph_ = value != 0.0;
    return QTG_OK;
  }
  set_error(std::string("apply_fit_set_option: unknown key ") + key);
  return QTG_EARG;
}

int qtg_apply_fit_info(qtg_apply_fit f, int64_t* info) {
  if (!f || !info) return QTG_EARG;
  auto* a = static_cast<ApplyFit*>(f);
  info[0] = a->launches_first();
  info[1] = a->launches_per_iteration();
  info[2] = a->device_bytes();
  return QTG_OK;
}

}  // extern "C"
