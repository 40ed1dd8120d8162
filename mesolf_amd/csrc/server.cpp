// Serving loop: continuous batching of frame x patch deconvolution requests.
//
// One worker thread owns the device work.  Between RL iterations it admits
// queued requests into free batch slots (FIFO, skipping requests whose patch
// is busy or does not fit the memory budget), runs one iteration of every
// attached patch as one batch launch sequence, and retires the requests that
// reached their iteration count.  Patch state (volume, image, work buffers)
// stays resident per patch id so the next frame starts warm; idle patches are
// evicted least-recently-used first when a new patch needs room.  A request
// whose patch alone exceeds the budget is refused at submission; one that
// fails on the device fails alone (its patch state is dropped, so the patch's
// next frame starts cold) and the loop keeps serving.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "state.h"

using namespace mlf;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  int32_t patch_id = 0;
  int n1 = 0, n2 = 0;
  std::vector<float> image;
  float* out = nullptr;
  double t_submit = 0, t_done = 0;
  bool done = false;
  mlf_status status = MLF_OK;
  std::string error;
};

struct Resident {
  mlf_patch p = nullptr;
  int n1 = 0, n2 = 0;
  size_t bytes = 0;
  bool busy = false;
  uint64_t last_used = 0;
};

struct Active {
  std::shared_ptr<Request> req;
  int32_t patch_id = 0;
  int iters = 0;
};

}  // namespace

struct mlf_server_s {
  mlf_psf psf = nullptr;
  mlf_precision prec = MLF_PRECISION_BF16;
  mlf_server_config cfg{};
  mlf_batch batch = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
  std::map<int32_t, Resident> resident;  // worker thread only
  std::vector<Active> active;            // worker thread only
  uint64_t next_ticket = 1, clock = 0;
  bool stop = false;
  std::thread worker;
  // statistics (under mu)
  std::vector<double> latencies;
  int64_t submitted = 0, completed = 0, evictions = 0, cold_starts = 0, iterations_run = 0;
  int64_t in_flight = 0, resident_count = 0;
  int64_t waiters = 0;  // threads inside mlf_server_wait (destroy waits for them)
  uint64_t resident_bytes = 0, peak_resident_bytes = 0;
  // host-forced failures (mlf_server_inject_fault, under mu)
  int32_t fault_admit_patch = -1;
  int32_t fault_runs = 0;
  double slot_sum = 0, t_first = -1, t_last = 0;
};

namespace {

void finish(mlf_server_s* s, const std::shared_ptr<Request>& r, mlf_status st, const char* err) {
  std::lock_guard<std::mutex> lk(s->mu);
  r->done = true;
  r->status = st;
  if (err) r->error = err;
  r->t_done = now_ms();
  if (st == MLF_OK) {
    s->latencies.push_back(r->t_done - r->t_submit);
    ++s->completed;
    s->t_last = r->t_done;
  }
  --s->in_flight;
  s->cv_done.notify_all();
}

// Pick requests to admit (called with mu held).  Returns chosen requests and
// patch ids to evict first.
void choose(mlf_server_s* s, std::vector<std::shared_ptr<Request>>* admit,
            std::vector<int32_t>* evict) {
  const int free_slots = s->cfg.max_slots - (int)s->active.size();
  if (free_slots <= 0) return;
  std::vector<int32_t> chosen_ids;
  auto busy = [&](int32_t id) {
    auto it = s->resident.find(id);
    if (it != s->resident.end() && it->second.busy) return true;
    return std::find(chosen_ids.begin(), chosen_ids.end(), id) != chosen_ids.end();
  };
  uint64_t bytes = s->resident_bytes;
  for (auto it = s->pending.begin(); it != s->pending.end() && (int)admit->size() < free_slots;) {
    const std::shared_ptr<Request>& r = *it;
    if (busy(r->patch_id)) {
      ++it;
      continue;
    }
    if (!s->resident.count(r->patch_id)) {
      const size_t need = patch_bytes_estimate(s->psf, s->prec, r->n1, r->n2);
      // evict idle residents, least recently used first, until it fits
      std::vector<int32_t> victims;
      uint64_t after = bytes;
      while (after + need > s->cfg.hbm_budget_bytes) {
        int32_t victim = -1;
        uint64_t oldest = UINT64_MAX;
        for (auto& kv : s->resident) {
          const bool gone = std::find(evict->begin(), evict->end(), kv.first) != evict->end() ||
                            std::find(victims.begin(), victims.end(), kv.first) != victims.end();
          if (!gone && !busy(kv.first) && kv.second.last_used < oldest) {
            oldest = kv.second.last_used;
            victim = kv.first;
          }
        }
        if (victim < 0) break;
        victims.push_back(victim);
        after -= s->resident[victim].bytes;
      }
      if (after + need > s->cfg.hbm_budget_bytes) {
        ++it;  // cannot fit now; a later request may (its patch may be resident)
        continue;
      }
      evict->insert(evict->end(), victims.begin(), victims.end());
      bytes = after + need;
    }
    chosen_ids.push_back(r->patch_id);
    admit->push_back(r);
    it = s->pending.erase(it);
  }
}

mlf_status evict_patch(mlf_server_s* s, int32_t id) {
  auto it = s->resident.find(id);
  if (it == s->resident.end()) return MLF_OK;
  const size_t b = it->second.bytes;
  const mlf_status st = mlf_patch_destroy(it->second.p);
  s->resident.erase(it);
  std::lock_guard<std::mutex> lk(s->mu);
  s->resident_bytes -= b;
  --s->resident_count;
  ++s->evictions;
  return st;
}

mlf_status admit_one(mlf_server_s* s, const std::shared_ptr<Request>& r) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fault_admit_patch == r->patch_id) {
      s->fault_admit_patch = -1;
      return set_error(MLF_ERR_DEVICE, "injected device failure at admission of patch %d",
                       (int)r->patch_id);
    }
  }
  auto it = s->resident.find(r->patch_id);
  bool cold = false;
  if (it == s->resident.end()) {
    Resident res;
    MLF_CHECK(mlf_patch_create(s->psf, r->n1, r->n2, s->prec, &res.p));
    res.n1 = r->n1;
    res.n2 = r->n2;
    res.bytes = patch_device_bytes(res.p);
    it = s->resident.emplace(r->patch_id, res).first;
    cold = true;
    std::lock_guard<std::mutex> lk(s->mu);
    s->resident_bytes += res.bytes;
    s->peak_resident_bytes = std::max(s->peak_resident_bytes, s->resident_bytes);
    ++s->resident_count;
    ++s->cold_starts;
  } else if (it->second.n1 != r->n1 || it->second.n2 != r->n2) {
    return set_error(MLF_ERR_INVALID_ARGUMENT, "patch %d was %dx%d, request says %dx%d",
                     (int)r->patch_id, it->second.n1, it->second.n2, r->n1, r->n2);
  }
  Resident& res = it->second;
  MLF_CHECK(mlf_patch_set_image(res.p, r->image.data()));
  if (cold) MLF_CHECK(mlf_patch_init_volume(res.p));
  MLF_CHECK(mlf_batch_attach(s->batch, res.p));
  res.busy = true;
  res.last_used = ++s->clock;
  r->image.clear();
  r->image.shrink_to_fit();
  s->active.push_back(Active{r, r->patch_id, 0});
  return MLF_OK;
}

void worker_main(mlf_server_s* s) {
  (void)ctx_activate(s->psf->ctx);
  for (;;) {
    std::vector<std::shared_ptr<Request>> admit;
    std::vector<int32_t> evict;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->active.empty(); });
      if (s->stop && !s->pending.empty()) {
        // shutting down: queued requests are cancelled, those in flight finish
        std::deque<std::shared_ptr<Request>> dropped;
        dropped.swap(s->pending);
        lk.unlock();
        for (const auto& r : dropped) finish(s, r, MLF_ERR_CANCELLED, "cancelled: server destroyed");
        lk.lock();
      }
      if (s->stop && s->active.empty()) return;
      if (s->pending.empty() && s->active.empty()) continue;
      choose(s, &admit, &evict);
      if (admit.empty() && s->active.empty()) {
        // nothing fits even with every idle patch evicted: fail the oldest
        std::shared_ptr<Request> r = s->pending.front();
        s->pending.pop_front();
        lk.unlock();
        finish(s, r, MLF_ERR_BUDGET, "request does not fit the HBM budget");
        continue;
      }
    }
    for (int32_t id : evict) (void)evict_patch(s, id);
    for (const auto& r : admit) {
      const mlf_status as = admit_one(s, r);
      if (as != MLF_OK) {
        const std::string e = mlf_last_error();
        auto it = s->resident.find(r->patch_id);
        if (it != s->resident.end() && !it->second.busy) (void)evict_patch(s, r->patch_id);
        finish(s, r, as, e.c_str());
      }
    }
    if (s->active.empty()) continue;
    bool injected = false;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->fault_runs > 0) {
        --s->fault_runs;
        injected = true;
      }
    }
    mlf_status st = injected ? set_error(MLF_ERR_DEVICE, "injected device failure in a batch iteration")
                             : mlf_batch_run(s->batch, 1, s->cfg.filter);
    if (st == MLF_OK) st = mlf_context_synchronize(s->psf->ctx);
    const std::string err = st == MLF_OK ? std::string() : mlf_last_error();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      ++s->iterations_run;
      s->slot_sum += (double)s->active.size();
    }
    std::vector<Active> keep;
    std::vector<int32_t> failed;  // patches whose volume is undefined after a failed iteration
    for (Active& a : s->active) {
      ++a.iters;
      if (st == MLF_OK && a.iters < s->cfg.iterations) {
        keep.push_back(a);
        continue;
      }
      Resident& res = s->resident[a.patch_id];
      mlf_status rs = st;
      std::string re = err;
      if (mlf_batch_detach(s->batch, res.p) != MLF_OK && rs == MLF_OK) {
        rs = MLF_ERR_INTERNAL;
        re = mlf_last_error();
      }
      if (rs == MLF_OK && a.req->out && mlf_patch_get_volume(res.p, a.req->out) != MLF_OK) {
        rs = MLF_ERR_DEVICE;
        re = mlf_last_error();
      }
      res.busy = false;
      res.last_used = ++s->clock;
      if (rs != MLF_OK) failed.push_back(a.patch_id);
      finish(s, a.req, rs, rs == MLF_OK ? nullptr : re.c_str());
    }
    s->active.swap(keep);
    for (int32_t id : failed) (void)evict_patch(s, id);
  }
}

}  // namespace

extern "C" {

mlf_status mlf_server_create(mlf_psf psf, mlf_precision prec, const mlf_server_config* cfg,
                             mlf_server* out) {
  clear_error();
  if (!psf || !cfg || !out) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (cfg->max_slots < 1 || cfg->iterations < 1)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "max_slots and iterations must be >= 1");
  auto s = std::make_unique<mlf_server_s>();
  s->psf = psf;
  s->prec = prec;
  s->cfg = *cfg;
  MLF_CHECK(mlf_batch_create(psf, prec, cfg->max_slots, &s->batch));
  if (mlf_batch_set_graph(s->batch, cfg->use_graph) != MLF_OK) {
    const std::string e = mlf_last_error();
    (void)mlf_batch_destroy(s->batch);
    return set_error(MLF_ERR_INTERNAL, "%s", e.c_str());
  }
  s->worker = std::thread(worker_main, s.get());
  *out = s.release();
  return MLF_OK;
}

mlf_status mlf_server_submit(mlf_server s, int32_t patch_id, int64_t n1, int64_t n2,
                             const float* lfimg, float* volume_out, uint64_t* ticket) {
  clear_error();
  if (!s || !lfimg || !ticket) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n1 < 1 || n2 < 1 || n1 > 4096 || n2 > 4096)
    return set_error(MLF_ERR_INVALID_ARGUMENT, "bad lenslet grid");
  auto r = std::make_shared<Request>();
  r->patch_id = patch_id;
  r->n1 = (int)n1;
  r->n2 = (int)n2;
  r->image.assign(lfimg, lfimg + (size_t)n1 * n2 * kP);
  r->out = volume_out;
  const size_t need = patch_bytes_estimate(s->psf, s->prec, r->n1, r->n2);
  if (need > s->cfg.hbm_budget_bytes)
    return set_error(MLF_ERR_BUDGET, "patch %d (%lld x %lld lenslets, %zu bytes) exceeds the HBM budget of %llu bytes",
                     (int)patch_id, (long long)n1, (long long)n2, need,
                     (unsigned long long)s->cfg.hbm_budget_bytes);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return set_error(MLF_ERR_INVALID_ARGUMENT, "server is shutting down");
  r->ticket = s->next_ticket++;
  r->t_submit = now_ms();
  if (s->t_first < 0) s->t_first = r->t_submit;
  s->tickets[r->ticket] = r;
  s->pending.push_back(r);
  ++s->submitted;
  ++s->in_flight;
  *ticket = r->ticket;
  s->cv_work.notify_one();
  return MLF_OK;
}

mlf_status mlf_server_wait(mlf_server s, uint64_t ticket, double* latency_ms) {
  clear_error();
  if (!s) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return set_error(MLF_ERR_INVALID_ARGUMENT, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  ++s->waiters;
  s->cv_done.wait(lk, [&] { return r->done; });
  --s->waiters;
  s->cv_done.notify_all();
  s->tickets.erase(ticket);
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (r->status != MLF_OK) return set_error(r->status, "request failed: %s", r->error.c_str());
  return MLF_OK;
}

mlf_status mlf_server_drain(mlf_server s) {
  clear_error();
  if (!s) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  ++s->waiters;
  s->cv_done.wait(lk, [&] { return s->in_flight == 0; });
  --s->waiters;
  s->cv_done.notify_all();
  return MLF_OK;
}

mlf_status mlf_server_get_stats(mlf_server s, mlf_server_stats* st) {
  clear_error();
  if (!s || !st) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->pending = (int64_t)s->pending.size();
  st->in_flight = s->in_flight;
  st->resident_patches = s->resident_count;
  st->evictions = s->evictions;
  st->cold_starts = s->cold_starts;
  st->iterations_run = s->iterations_run;
  st->resident_bytes = s->resident_bytes;
  st->peak_resident_bytes = s->peak_resident_bytes;
  std::vector<double> l = s->latencies;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      const size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(q * l.size()) - 1);
      return l[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : l) sum += x;
    st->mean_ms = sum / l.size();
    st->max_ms = l.back();
    if (s->t_last > s->t_first) st->volumes_per_s = s->completed / ((s->t_last - s->t_first) * 1e-3);
  }
  st->mean_slots = s->iterations_run ? s->slot_sum / s->iterations_run : 0.0;
  return MLF_OK;
}

mlf_status mlf_server_inject_fault(mlf_server s, int32_t kind, int32_t patch_id) {
  clear_error();
  if (!s) return set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (kind == MLF_FAULT_ADMIT) {
    s->fault_admit_patch = patch_id;
  } else if (kind == MLF_FAULT_ITERATION) {
    ++s->fault_runs;
  } else {
    return set_error(MLF_ERR_INVALID_ARGUMENT, "unknown fault kind %d", (int)kind);
  }
  return MLF_OK;
}

mlf_status mlf_server_destroy(mlf_server s) {
  clear_error();
  if (!s) return MLF_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    // every request has finished; let threads still inside mlf_server_wait return
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  (void)ctx_activate(s->psf->ctx);
  for (auto& kv : s->resident) (void)mlf_patch_destroy(kv.second.p);
  s->resident.clear();
  (void)mlf_batch_destroy(s->batch);
  delete s;
  return MLF_OK;
}

}  // extern "C"
