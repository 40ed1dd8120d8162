// Batches: RL iterations of several patches of one PSF in shared launches.
//
// Each view-step kernel walks a list of work entries (slot, plane, block...)
// with a grid-stride loop, so one launch serves every attached patch and a
// captured graph of fixed grid size stays valid while patches come and go:
// only the entry lists and slot descriptors in device memory change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"
#include "rl_mfma.h"
#include "state.h"

using namespace mlf;

struct mlf_batch_s {
  mlf_psf psf = nullptr;
  Precision kp = Precision::kF32;
  int capacity = 0;
  std::vector<mlf_patch_s*> slots;
  // device tables (rebuilt when the composition changes)
  DevBuf<SlotDesc> d_slots;
  DevBuf<int32_t> d_fsched, d_bsched, d_rsched, d_counts;
  int n_f = 0, n_b = 0, n_r = 0;
  bool dirty = true;
  // launch shape (fixed for the life of the batch)
  bool tpz = false;
  int cw = 16, bm2 = 8;
  int fwd_mt = 0, bwd_nt = 0;  // tile shape shared by every attached patch
  int bwd_grp = 0, bwd_occ = 2;
  int grid_f = 0, grid_b = 0, grid_r = 0;
  int cus = 0;
  size_t smem_f = 0, smem_b = 0;
  int fwd_nbuf = 2;
  // one RL iteration's view loop as a graph
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;  // private stream for capture (other threads may use the context stream)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  std::mutex mu;
  ~mlf_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

void drop_graph(mlf_batch_s* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Grow a device table; a new allocation invalidates the captured graph.
mlf_status upload_table(mlf_batch_s* b, DevBuf<int32_t>& d, const std::vector<int32_t>& h) {
  if (h.size() > d.count) {
    drop_graph(b);
    MLF_CHECK(d.alloc(std::max(h.size(), 2 * d.count)));
  }
  if (!h.empty()) MLF_HIP(hipMemcpy(d.ptr, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return MLF_OK;
}

// Entry lists of all attached patches, merged largest radius first so the
// grid-stride loop hands every workgroup a similar mix of costs.
mlf_status rebuild(mlf_batch_s* b) {
  mlf_psf_s* ps = b->psf;
  hipStream_t st = ps->ctx->stream;
  MLF_HIP(hipStreamSynchronize(st));  // kernels in flight may read the old tables
  std::vector<SlotDesc> sd;
  struct Ent {
    int r;
    int32_t v[5];
  };
  std::vector<Ent> fe, be;
  std::vector<int32_t> rs;
  for (size_t i = 0; i < b->slots.size(); ++i) {
    mlf_patch_s* p = b->slots[i];
    sd.push_back(p->slot);
    for (size_t e = 0; e + 3 <= p->fent.size(); e += 3) {
      const int z = p->fent[e];
      fe.push_back(Ent{ps->rf[z], {(int32_t)i, z, p->fent[e + 1], p->fent[e + 2], 0}});
    }
    for (size_t e = 0; e + 4 <= p->bent.size(); e += 4) {
      const int z = p->bent[e];
      be.push_back(Ent{ps->rb[z], {(int32_t)i, z, p->bent[e + 1], p->bent[e + 2], p->bent[e + 3]}});
    }
    for (int j = 0; j < p->nblk_r; ++j) {
      rs.push_back((int32_t)i);
      rs.push_back(kRatioLens * j);
    }
  }
  auto by_r = [](const Ent& a, const Ent& c) { return a.r > c.r; };
  std::stable_sort(fe.begin(), fe.end(), by_r);
  std::stable_sort(be.begin(), be.end(), by_r);
  std::vector<int32_t> fs, bs;
  std::vector<int> fg, bg;  // XCD groups: (slot, plane) and (slot, plane, phase group)
  const int ngrp = kPhiPadBwd / (16 * std::max(1, b->bwd_nt));  // XCD groups: entries start at a phase group
  for (const Ent& e : fe) {
    fs.insert(fs.end(), e.v, e.v + 4);
    fg.push_back(e.v[0] * ps->Z + e.v[1]);
  }
  for (const Ent& e : be) {
    bs.insert(bs.end(), e.v, e.v + 5);
    bg.push_back((e.v[0] * ps->Z + e.v[1]) * ngrp + e.v[4]);
  }
  xcd_interleave(fs, 4, fg);
  xcd_interleave(bs, 5, bg);
  b->n_f = (int)fe.size();
  b->n_b = (int)be.size();
  b->n_r = (int)(rs.size() / 2);
  if (!sd.empty())
    MLF_HIP(hipMemcpy(b->d_slots.ptr, sd.data(), sd.size() * sizeof(SlotDesc), hipMemcpyHostToDevice));
  MLF_CHECK(upload_table(b, b->d_fsched, fs));
  MLF_CHECK(upload_table(b, b->d_bsched, bs));
  MLF_CHECK(upload_table(b, b->d_rsched, rs));
  const std::vector<int32_t> counts{b->n_f, b->n_b, b->n_r};
  MLF_HIP(hipMemcpy(b->d_counts.ptr, counts.data(), 3 * sizeof(int32_t), hipMemcpyHostToDevice));
  b->dirty = false;
  return MLF_OK;
}

// One RL iteration's view loop (recon:100-138) for every attached patch
// (the first `nviews` views; all when nviews <= 0).
mlf_status enqueue_views(mlf_batch_s* b, hipStream_t st, int nviews = 0) {
  mlf_psf_s* ps = b->psf;
  const PrecPack& pk = ps->pack[(int)b->kp];
  const size_t eb = elem_bytes(b->kp);
  const RatioGeo rg{b->d_rsched.ptr, b->d_counts.ptr + 2};
  const int nv = (nviews <= 0 || nviews > ps->V) ? ps->V : nviews;
  for (int k = 0; k < nv; ++k) {
    const FwdGeo g{b->d_slots.ptr, ps->d_rf.ptr, pk.d_woff.ptr, pk.d_nchunk.ptr, b->d_fsched.ptr,
                   b->d_counts.ptr + 0,
                   pk.d_finv.ptr ? pk.d_finv.ptr + (size_t)k * ps->Z : nullptr};
    const int u = ps->uv[2 * k], v = ps->uv[2 * k + 1];
    const void* Wk = pk.fwd.ptr + (size_t)k * pk.pvf * eb;
    if (b->tpz) {
      MLF_LAUNCH(launch_fwd_tpz(b->kp, b->fwd_mt, Wk, g, b->grid_f, b->smem_f, st));
      MLF_LAUNCH(launch_ratio_gather_tpz(g, rg, b->grid_r, ps->Z, b->cw, u, v, 1, nullptr, st));
    } else {
      MLF_LAUNCH(launch_fwd_mfma(b->kp, Wk, g, b->grid_f, b->smem_f, b->fwd_nbuf, st));
      MLF_LAUNCH(launch_ratio_gather(g, rg, b->grid_r, ps->Z, b->bm2, u, v, 1, nullptr, st));
    }
    const BwdGeo bg{b->d_slots.ptr, ps->d_rb.ptr, pk.d_boff.ptr, pk.d_epad.ptr, b->d_bsched.ptr,
                    b->d_counts.ptr + 1, k, bwd_groups_per_entry(b->bwd_nt, b->bwd_grp)};
    // every slot of a BF16 Toeplitz batch holds its volume in the split form
    const bool split = b->tpz && b->kp == Precision::kBF16;
    MLF_LAUNCH(launch_bwd_mfma(b->kp, b->bwd_nt, b->bwd_occ, split, pk.bwd.ptr + (size_t)k * pk.pvb * eb, bg,
                               b->grid_b, b->smem_b, (float)ps->weight[k], st));
  }
  return MLF_OK;
}

mlf_status capture(mlf_batch_s* b) {
  drop_graph(b);
  hipStream_t st = b->cap;
  MLF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const mlf_status s = enqueue_views(b, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (s != MLF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  MLF_HIP(e);
  b->graph = g;
  MLF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return MLF_OK;
}

mlf_status detach_locked(mlf_batch_s* b, mlf_patch_s* p) {
  auto it = std::find(b->slots.begin(), b->slots.end(), p);
  if (it == b->slots.end()) return set_error(MLF_ERR_INVALID_ARGUMENT, "patch is not attached");
  // work in flight reads the slot table
  MLF_HIP(hipStreamSynchronize(b->psf->ctx->stream));
  b->slots.erase(it);
  p->batch = nullptr;
  b->dirty = true;
  return MLF_OK;
}

}  // namespace

namespace mlf {
void batch_mark_dirty(mlf_batch_s* b) {
  std::lock_guard<std::mutex> lk(b->mu);
  b->dirty = true;
}

// Called by mlf_patch_destroy.
mlf_status batch_forget_patch(mlf_patch_s* p) {
  mlf_batch_s* b = p->batch;
  if (!b) return MLF_OK;
  std::lock_guard<std::mutex> lk(b->mu);
  return detach_locked(b, p);
}
}  // namespace mlf

extern "C" {

mlf_status mlf_batch_create(mlf_psf psf, mlf_precision prec, int32_t capacity, mlf_batch* out) {
  clear_error();
  if (!psf || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (capacity < 1 || capacity > 4096)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "batch capacity %d out of range", (int)capacity);
  if (prec != MLF_PRECISION_FP32 && prec != MLF_PRECISION_BF16 && prec != MLF_PRECISION_FP16)
    return set_error(MLF_ERR_UNSUPPORTED, "precision %d not supported", (int)prec);
  MLF_CHECK(ctx_activate(psf->ctx));
  {
    std::lock_guard<std::mutex> lk(g_life_mu);
    if (psf->closed) return set_error(MLF_ERR_INVALID_ARGUMENT, "psf already destroyed");
    ++psf->refs;
  }
  auto b = std::make_unique<mlf_batch_s>();
  b->psf = psf;
  struct Undo {
    mlf_batch_s* b;
    ~Undo() {
      if (b) {
        std::lock_guard<std::mutex> lk(g_life_mu);
        --b->psf->refs;
      }
    }
  } undo{b.get()};
  b->kp = kernel_precision(prec);
  b->capacity = capacity;
  MLF_CHECK(b->d_slots.alloc((size_t)capacity));
  MLF_CHECK(b->d_counts.alloc(3));
  const std::vector<int32_t> zero{0, 0, 0};
  MLF_HIP(hipMemcpy(b->d_counts.ptr, zero.data(), 3 * sizeof(int32_t), hipMemcpyHostToDevice));
  int cus = 0;
  MLF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, psf->ctx->device));
  // two workgroups per CU for the MFMA kernels (their LDS and register
  // budgets allow two), more for the light ratio kernel
  b->cus = cus;
  b->grid_f = 2 * cus;
  b->grid_b = 2 * cus;  // bwd_occ workgroups per CU once a shape is chosen
  b->grid_r = 4 * cus;
  MLF_HIP(hipEventCreate(&b->ev0));
  MLF_HIP(hipEventCreate(&b->ev1));
  MLF_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  undo.b = nullptr;
  *out = b.release();
  return MLF_OK;
}

mlf_status mlf_batch_destroy(mlf_batch b) {
  clear_error();
  if (!b) return MLF_OK;
  std::lock_guard<std::mutex> life(g_life_mu);
  mlf_psf psf = b->psf;
  (void)hipSetDevice(psf->ctx->device);
  (void)hipStreamSynchronize(psf->ctx->stream);
  {
    std::lock_guard<std::mutex> lk(b->mu);
    for (mlf_patch_s* p : b->slots) p->batch = nullptr;
    b->slots.clear();
  }
  delete b;
  if (--psf->refs == 0 && psf->closed) psf_free(psf);
  return MLF_OK;
}

mlf_status mlf_batch_attach(mlf_batch b, mlf_patch p) {
  clear_error();
  if (!b || !p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(b->mu);
  if (p->psf != b->psf) return set_error(MLF_ERR_INVALID_ARGUMENT, "patch uses another PSF");
  if (p->reference)
    return set_error(MLF_ERR_UNSUPPORTED, "reference-kernel patches cannot be batched");
  if (p->kp != b->kp) return set_error(MLF_ERR_INVALID_ARGUMENT, "patch precision differs from the batch");
  if (p->batch) return set_error(MLF_ERR_INVALID_ARGUMENT, "patch already attached to a batch");
  if ((int)b->slots.size() >= b->capacity)
    return set_error(MLF_ERR_BUDGET, "batch is full (%d patches)", b->capacity);
  MLF_CHECK(ctx_activate(b->psf->ctx));
  if (b->slots.empty()) {
    if (b->fwd_mt != p->fwd_mt || b->bwd_nt != p->bwd_nt || b->bwd_grp != p->bwd_grp || b->bwd_occ != p->bwd_occ)
      drop_graph(b);
    b->tpz = p->tpz;
    b->cw = p->cw;
    b->bm2 = p->bm2;
    b->fwd_mt = p->fwd_mt;
    b->bwd_nt = p->bwd_nt;
    b->bwd_grp = p->bwd_grp;
    b->bwd_occ = p->bwd_occ;
    b->grid_b = b->bwd_occ * b->cus;
    b->smem_f = p->smem_f;
    b->smem_b = p->smem_b;
    b->fwd_nbuf = p->fwd_nbuf;
  } else if (p->fwd_mt != b->fwd_mt || p->bwd_nt != b->bwd_nt || p->bwd_grp != b->bwd_grp ||
             p->bwd_occ != b->bwd_occ) {
    // one launch serves every patch: the patch takes the batch's tile shape
    MLF_HIP(hipStreamSynchronize(b->psf->ctx->stream));
    p->fwd_mt = b->fwd_mt;
    p->bwd_nt = b->bwd_nt;
    p->bwd_grp = b->bwd_grp;
    p->bwd_occ = b->bwd_occ;
    MLF_CHECK(patch_setup_tiles(p));
    if (p->volume_set) MLF_CHECK(patch_refresh_shadow(p));
  }
  b->slots.push_back(p);
  p->batch = b;
  b->dirty = true;
  return MLF_OK;
}

mlf_status mlf_batch_detach(mlf_batch b, mlf_patch p) {
  clear_error();
  if (!b || !p) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(b->mu);
  MLF_CHECK(ctx_activate(b->psf->ctx));
  return detach_locked(b, p);
}

mlf_status mlf_batch_size(mlf_batch b, int32_t* attached) {
  clear_error();
  if (!b || !attached) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(b->mu);
  *attached = (int32_t)b->slots.size();
  return MLF_OK;
}

mlf_status mlf_batch_set_graph(mlf_batch b, int32_t enable) {
  clear_error();
  if (!b) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(b->mu);
  b->use_graph = enable != 0;
  return MLF_OK;
}

mlf_status mlf_batch_run(mlf_batch b, int32_t iterations, int32_t filter) {
  clear_error();
  if (!b) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (iterations < 0) return set_error(MLF_ERR_INVALID_ARGUMENT, "negative iteration count");
  std::lock_guard<std::mutex> lk(b->mu);
  MLF_CHECK(ctx_activate(b->psf->ctx));
  for (mlf_patch_s* p : b->slots)
    if (!p->image_set || !p->volume_set)
      return set_error(MLF_ERR_INVALID_ARGUMENT, "image and volume must be set on every attached patch");
  hipStream_t st = b->psf->ctx->stream;
  if (b->dirty) MLF_CHECK(rebuild(b));
  MLF_HIP(hipEventRecord(b->ev0, st));
  if (!b->slots.empty()) {
    for (int it = 0; it < iterations; ++it) {
      if (b->use_graph) {
        if (!b->exec) MLF_CHECK(capture(b));
        MLF_HIP(hipGraphLaunch(b->exec, st));
      } else {
        MLF_CHECK(enqueue_views(b, st));
      }
      for (mlf_patch_s* p : b->slots) MLF_CHECK(patch_merge_split(p));
      if (filter)
        for (mlf_patch_s* p : b->slots) MLF_CHECK(patch_filter(p));
    }
  }
  MLF_HIP(hipEventRecord(b->ev1, st));
  b->timed = true;
  return MLF_OK;
}

mlf_status mlf_batch_autotune(mlf_batch b, int32_t views, mlf_tile_shape* best) {
  clear_error();
  if (!b) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(b->mu);
  if (b->slots.empty()) return set_error(MLF_ERR_INVALID_ARGUMENT, "no patch attached");
  for (mlf_patch_s* p : b->slots)
    if (!p->image_set || !p->volume_set)
      return set_error(MLF_ERR_INVALID_ARGUMENT, "image and volume must be set on every attached patch");
  MLF_CHECK(ctx_activate(b->psf->ctx));
  hipStream_t st = b->psf->ctx->stream;
  const int nv = std::max(1, std::min<int>(views <= 0 ? 8 : views, b->psf->V));
  std::vector<std::unique_ptr<DevBuf<float>>> save;
  for (mlf_patch_s* p : b->slots) {
    save.emplace_back(new DevBuf<float>());
    MLF_CHECK(save.back()->alloc(p->X.count));
    MLF_HIP(hipMemcpyAsync(save.back()->ptr, p->X.ptr, p->X.bytes(), hipMemcpyDeviceToDevice, st));
  }
  auto apply = [&](const TileShape& t) -> mlf_status {
    MLF_HIP(hipStreamSynchronize(st));
    for (mlf_patch_s* p : b->slots) {
      p->fwd_mt = t.mt;
      p->bwd_nt = t.nt;
      p->bwd_grp = t.grp;
      p->bwd_occ = t.occ;
      MLF_CHECK(patch_setup_tiles(p));
      MLF_CHECK(patch_refresh_shadow(p));
    }
    mlf_patch_s* p0 = b->slots[0];
    b->fwd_mt = t.mt;
    b->bwd_nt = t.nt;
    b->bwd_grp = t.grp;
    b->bwd_occ = t.occ;
    b->grid_b = b->bwd_occ * b->cus;
    b->cw = p0->cw;
    b->smem_f = p0->smem_f;
    b->smem_b = p0->smem_b;
    b->fwd_nbuf = p0->fwd_nbuf;
    drop_graph(b);
    b->dirty = true;
    return rebuild(b);
  };
  auto restore = [&]() -> mlf_status {
    for (size_t i = 0; i < b->slots.size(); ++i)
      MLF_HIP(hipMemcpyAsync(b->slots[i]->X.ptr, save[i]->ptr, save[i]->bytes(),
                             hipMemcpyDeviceToDevice, st));
    return MLF_OK;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MLF_HIP(hipEventCreate(&e0));
  MLF_HIP(hipEventCreate(&e1));
  double best_ms = 1e300;
  TileShape bt{b->fwd_mt, b->bwd_nt, b->bwd_grp, b->bwd_occ};
  mlf_status s = MLF_OK;
  // two rounds over the candidates, each keeping its faster round (shapes a
  // few per cent apart are otherwise decided by clock and cache noise)
  const auto cands = tile_candidates(b->kp);
  std::vector<double> cand_ms(cands.size(), 1e300);
  for (int round = 0; round < 2 && s == MLF_OK; ++round) {
    for (size_t ci = 0; ci < cands.size(); ++ci) {
      const auto& c = cands[ci];
      if (apply(c) != MLF_OK) continue;  // shape does not fit this PSF
      s = enqueue_views(b, st, 1);  // warm-up
      if (s == MLF_OK && hipEventRecord(e0, st) != hipSuccess) s = set_error(MLF_ERR_DEVICE, "event");
      if (s == MLF_OK) s = enqueue_views(b, st, nv);
      if (s == MLF_OK && hipEventRecord(e1, st) != hipSuccess) s = set_error(MLF_ERR_DEVICE, "event");
      float ms = 0.f;
      if (s == MLF_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        s = set_error(MLF_ERR_DEVICE, "event timing failed");
      if (s == MLF_OK) s = restore();
      if (s != MLF_OK) break;
      cand_ms[ci] = std::min(cand_ms[ci], (double)ms / nv);
    }
  }
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    if (cand_ms[ci] < best_ms) {
      best_ms = cand_ms[ci];
      bt = cands[ci];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MLF_CHECK(s);
  MLF_CHECK(apply(bt));
  std::vector<std::pair<int, int>> grids;
  for (mlf_patch_s* p : b->slots) grids.emplace_back(p->n1, p->n2);
  std::sort(grids.begin(), grids.end());
  std::string key = tune_key(b->slots[0], "batch");
  for (const auto& g : grids) key += "|" + std::to_string(g.first) + "x" + std::to_string(g.second);
  tune_cache_put(key, bt, best_ms);
  if (best) {
    best->fwd_columns_per_wave = bt.mt;
    best->bwd_phase_tiles = bt.nt;
    best->bwd_groups_per_entry = bt.grp;
    best->bwd_waves_per_simd = bt.occ;
    best->ms_per_view = best_ms;
  }
  return MLF_OK;
}

mlf_status mlf_batch_last_run_ms(mlf_batch b, double* ms) {
  clear_error();
  if (!b || !ms) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!b->timed) return set_error(MLF_ERR_INVALID_ARGUMENT, "no run recorded");
  MLF_CHECK(ctx_activate(b->psf->ctx));
  MLF_HIP(hipEventSynchronize(b->ev1));
  float f = 0.f;
  MLF_HIP(hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return MLF_OK;
}

}  // extern "C"
