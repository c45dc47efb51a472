// Tuning table of the grouped plan: the K-split and multi-tile parameters
// that qutree_amd/autotune.py measured best for a class of plans, read when
// a Hamiltonian is planned.
//
// File: QTG_TUNING_TABLE, or tuning/gfx950.txt beside libqutree_amd.so.  One
// line per class, '#' comments:
//   class <precision> <max SPF n> <log2 state elements> <log2 SOP terms> key value ...
// A plan matches the line with the same precision whose (n, log2 N, log2 L)
// is nearest (sum of absolute differences, at most 2).  Keys: split_waves,
// split_min_k, gc_long_k, gm_split_k.  Options the caller
// sets explicitly (qtg_engine_set_option) are not overridden; option
// "tuning" 0 skips the table.
#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace qtg {

namespace {

struct TuneLine {
  int prec = 0, n = 0, logN = 0, logL = 0;
  std::vector<std::pair<std::string, double>> kv;
  std::string text;
};

std::string default_table_path() {
  Dl_info info{};
  if (dladdr(reinterpret_cast<void*>(&default_table_path), &info) && info.dli_fname) {
    std::string lib(info.dli_fname);
    const size_t slash = lib.rfind('/');
    return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tuning/gfx950.txt";
  }
  return "tuning/gfx950.txt";
}

const std::vector<TuneLine>& table() {
  static std::once_flag once;
  static std::vector<TuneLine> lines;
  std::call_once(once, [] {
    const char* env = std::getenv("QTG_TUNING_TABLE");
    std::ifstream in(env ? std::string(env) : default_table_path());
    std::string line;
    while (std::getline(in, line)) {
      const size_t hash = line.find('#');
      const std::string body = line.substr(0, hash);
      std::istringstream ss(body);
      std::string tag;
      if (!(ss >> tag) || tag != "class") continue;
      TuneLine t;
      if (!(ss >> t.prec >> t.n >> t.logN >> t.logL)) continue;
      std::string k;
      double v;
      while (ss >> k >> v) t.kv.push_back({k, v});
      t.text = body;
      lines.push_back(t);
    }
  });
  return lines;
}

int ilog2(int64_t x) {
  int r = 0;
  while (x > 1) {
    x >>= 1;
    ++r;
  }
  return r;
}

}  // namespace

void Engine::apply_tuning() {
  tunedFrom_.clear();
  if (!tuning_ || path_ != 1) return;
  const int logN = ilog2(tree_.total), logL = ilog2((int64_t)sop_.terms.size());
  const TuneLine* best = nullptr;
  int bestd = 3;
  for (const TuneLine& t : table()) {
    if (t.prec != precision_) continue;
    const int d = std::abs(t.n - nspf_) + std::abs(t.logN - logN) + std::abs(t.logL - logL);
    if (d < bestd) {
      bestd = d;
      best = &t;
    }
  }
  if (!best) return;
  for (const auto& kv : best->kv) {
    const std::string& k = kv.first;
    const double v = kv.second;
    if (k == "split_waves" && !(userSet_ & 1) && v >= 1) splitWaves_ = (int64_t)v;
    else if (k == "split_min_k" && !(userSet_ & 2) && v >= 4) splitMinK_ = (int)v;
    else if (k == "gc_long_k" && !(userSet_ & 4) && v >= 16) gcLongK_ = (int)v;
    else if (k == "gm_split_k" && !(userSet_ & 8) && v >= 16) gmSplitK_ = (int)v;
  }
  tunedFrom_ = best->text;
}

}  // namespace qtg
