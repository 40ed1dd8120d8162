// Tile-shape candidates and the autotuning cache.
//
// The cache maps a key (scope, precision, lenslet grid(s), PSF radii) to the
// fastest measured tile shape.  It lives for the process; if MLF_TUNE_CACHE
// names a file, entries are loaded from it on first use and appended to it
// when measured, one line per entry: "<key> <fwd_mt> <bwd_nt> <ms per view>".
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"
#include "rl_mfma.h"
#include "state.h"

namespace mlf {

namespace {
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tune;
bool g_tune_loaded = false;

void load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = std::getenv("MLF_TUNE_CACHE");
  if (!path || !*path) return;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key;
    int mt = 0, nt = 0;
    if (is >> key >> mt >> nt) g_tune[key] = {mt, nt};
  }
}
}  // namespace

std::vector<std::pair<int, int>> tile_candidates(Precision kp) {
  const int mt0 = kp == Precision::kF32 ? kTpzMTF32 : kTpzMTBF16;
  std::vector<std::pair<int, int>> c;
  for (int mt : {mt0, mt0 / 2})
    for (int nt : {kBwdNTiles, 3}) c.emplace_back(mt, nt);
  return c;
}

std::string tune_key(const mlf_patch_s* p, const char* scope) {
  std::ostringstream os;
  os << scope << "|" << (p->kp == Precision::kF32 ? "fp32" : "bf16") << "|" << p->n1 << "x" << p->n2
     << "|rf";
  for (int r : p->psf->rf) os << "," << r;
  return os.str();
}

bool tune_cache_get(const std::string& key, int* mt, int* nt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  load_locked();
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  *mt = it->second.first;
  *nt = it->second.second;
  return true;
}

void tune_cache_put(const std::string& key, int mt, int nt, double ms) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  load_locked();
  g_tune[key] = {mt, nt};
  const char* path = std::getenv("MLF_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << mt << " " << nt << " " << ms << "\n";
}

}  // namespace mlf
