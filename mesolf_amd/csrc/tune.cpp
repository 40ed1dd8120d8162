// Tile-shape candidates and the autotuning cache.
//
// The cache maps a key (scope, precision, lenslet grid(s), PSF radii) to the
// fastest measured tile shape.  It lives for the process; if MLF_TUNE_CACHE
// names a file, entries are loaded from it when the name is first seen and
// appended to it when measured, one line per entry:
//   "mlf-tune-v4 <key> <fwd_mt> <bwd_nt> <bwd_groups> <bwd_waves_per_simd> <ms per view>".
// Lines without the format tag (older builds) are ignored, and a cached shape
// is used only if it is one of the shapes this build has for the patch's
// precision: a stale or edited entry falls back to the default shape.
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
constexpr const char* kTuneTag = "mlf-tune-v4";
std::mutex g_tune_mu;
std::map<std::string, TileShape> g_tune;
std::string g_tune_loaded_path;  // cache file whose entries are loaded

void load_locked() {
  const char* path = std::getenv("MLF_TUNE_CACHE");
  if (!path || !*path || g_tune_loaded_path == path) return;
  g_tune_loaded_path = path;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tag, key;
    TileShape t;
    if (is >> tag >> key >> t.mt >> t.nt >> t.grp >> t.occ && tag == kTuneTag) g_tune[key] = t;
  }
}

const char* precision_name(Precision kp) {
  return kp == Precision::kF32 ? "fp32" : kp == Precision::kF16 ? "fp16" : "bf16";
}
}  // namespace

std::vector<TileShape> tile_candidates(Precision kp) {
  const int mt0 = kp == Precision::kF32 ? kTpzMTF32 : kTpzMTBF16;
  std::vector<TileShape> c;
  for (int grp : {0, 1})  // all phase groups per entry (one ratio window), or one
    for (int mt : {mt0, mt0 / 2}) {
      // BF16 first tries 4 N-tiles (phase groups of 64 phases: one 128-byte
      // line of each split-volume plane row), its default
      if (kp == Precision::kBF16) c.push_back(TileShape{mt, 4, grp, 2});
      for (int nt : {kBwdNTiles, 3}) c.push_back(TileShape{mt, nt, grp, 2});
      // 16-bit: 3 N-tiles at three waves per SIMD
      if (kp != Precision::kF32) c.push_back(TileShape{mt, 3, grp, 3});
    }
  return c;
}

int bwd_groups_per_entry(int nt, int grp) {
  const int ngrp = kPhiPadBwd / (16 * nt);
  return grp <= 0 || grp > ngrp || ngrp % grp != 0 ? ngrp : grp;
}

std::string tune_key(const mlf_patch_s* p, const char* scope) {
  std::ostringstream os;
  os << scope << "|" << precision_name(p->kp) << "|" << p->n1 << "x" << p->n2
     << "|rf";
  for (int r : p->psf->rf) os << "," << r;
  return os.str();
}

bool tune_cache_get(const std::string& key, Precision kp, TileShape* shape) {
  TileShape v;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_locked();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    v = it->second;
  }
  for (const auto& c : tile_candidates(kp))
    if (c == v) {
      *shape = v;
      return true;
    }
  return false;  // not a shape of this build: keep the default
}

void tune_cache_put(const std::string& key, const TileShape& t, double ms) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  load_locked();
  g_tune[key] = t;
  const char* path = std::getenv("MLF_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << kTuneTag << " " << key << " " << t.mt << " " << t.nt << " " << t.grp << " " << t.occ << " " << ms << "\n";
}

}  // namespace mlf
