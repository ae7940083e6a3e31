// Movie server: continuous batching of registration requests on one device.
//
// The reference registers one movie per process and scales by running several
// processes side by side, a subprocess per movie under the Skulk manager
// (automator/zorroSkulkManager.py:1089-1113, zorro/scripts/zorro_benchmark.py:174-248).
// Here one process keeps movie plans resident in HBM and runs several movies at
// once, each on its plan's own stream, so that the latency-bound stages of one
// movie (the one-workgroup shift solver, the logistic fit) overlap with the
// bandwidth-bound stages of another:
//
//  * requests queue in arrival order; a request is admitted when an in-flight
//    slot is free and its plan fits the HBM budget: an idle plan of the same
//    (settings, frames, shape) key is reused, otherwise one is created after
//    evicting idle plans of other keys, least recently used first;
//  * a request that fits goes ahead of a larger head of the queue that does not
//    fit yet, and every scheduling pass starts again from the oldest request;
//  * plans are created and movies run in worker threads, outside the lock, so
//    a multi-GB allocation never holds up another movie's completion;
//  * every plan is released exactly once; a request can be cancelled until it
//    is admitted.
//
// The scheduler is written against zr_plan_ops (create/run/destroy of a plan);
// the built-in backend is the zr_movie plan of this library, and host-only
// stand-ins exercise the scheduler without a GPU.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/zorro_hip.h"

int zr_capi_fail(int code, const char* msg);   // sets zr_last_error() (zr_capi.hip)

namespace {

// failures of this file's entry points and of the plan operations: the
// message goes to zr_last_error() of the calling thread, and is kept here for
// the request that a worker thread was running
thread_local std::string s_err;

int sfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  s_err = buf;
  return zr_capi_fail(code, buf);
}

double now_s() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

// ------------------------------------------------------- the HIP backend
// Upper estimate of a zr_movie plan's device bytes before it exists (the exact
// figure, zr_movie_device_bytes, replaces it once a plan of the key exists).
size_t estimate_movie_bytes(const zr_params* p, int n, int h, int w) {
  const size_t hp = p->shape_padded[0] ? p->shape_padded[0] : h;
  const size_t wp = p->shape_padded[1] ? p->shape_padded[1] : w;
  const bool crop = p->fou_crop[0] != 0 || p->fou_crop[1] != 0;
  const size_t hc = crop ? 2 * (p->fou_crop[0] / 2) : hp;
  const size_t wc = crop ? 2 * (p->fou_crop[1] / 2) : wp;
  const size_t m = p->max_shift;
  const bool masked = hp != (size_t)h || wp != (size_t)w;
  const size_t N = n, pairs = N * (N - 1) / 2;
  const size_t nspec = masked ? 2 * N + 1 : N;
  const size_t ptot = pairs + (masked ? 4 * N + 1 : 0);
  const size_t wk = wp / 2 + 1;
  bool dose = p->do_dose != 0;
  bool gauss = false;
  for (int i = 0; i < p->n_filters && i < 4; ++i) {
    dose = dose || p->filters[i] == ZR_FILTER_DOSE || p->filters[i] == ZR_FILTER_DOSENORM;
    gauss = gauss || p->filters[i] == ZR_FILTER_GAUSSLP || p->filters[i] == ZR_FILTER_BACKGROUND;
  }
  const bool spectral = dose && !masked && h == w;
  const size_t hs = spectral ? hp : hc;
  size_t b = std::min<size_t>(N, 8) * wk * hp * 8;       // row-pass scratch
  b += nspec * wk * hs * 8 + nspec * wk * 8;              // spectra
  b += ptot * (wc / 2 + 1) * 2 * m * 8;                   // pair pass 1
  b += ptot * 4 * m * m * 4;                              // lag windows
  b += 4 * (size_t)h * w * 4 + N * h * w * 4;             // sums + an owned copy of the frames
  if (gauss) b += (size_t)h * w * 4;
  if (dose) {
    b += wk * hp * 8;
    b += spectral ? 3 * N * (wk + hp) * 8 : std::min<size_t>(N, 8) * (wk + 1) * hp * 8;
  }
  if (p->do_frc) {
    const size_t box = p->frc_box > 0 ? p->frc_box : 512;
    const size_t tiles = std::max<size_t>(1, (h / (box / 2)) * (w / (box / 2)));
    b += tiles * 2 * (box / 2 + 1) * box * 8;
  }
  return (size_t)((double)b * 1.05) + (16u << 20);
}

struct HipBackend {
  int device;
};

size_t hip_estimate(void*, const zr_params* p, int n, int h, int w) { return estimate_movie_bytes(p, n, h, w); }

int hip_create(void* ctx, const zr_params* p, int n, int h, int w, void** plan, size_t* bytes) {
  const HipBackend* be = static_cast<HipBackend*>(ctx);
  zr_movie* mv = nullptr;
  int rc = zr_movie_create(be->device, p, n, h, w, &mv);
  if (rc) return sfail(rc, "%s", zr_last_error());
  size_t b = 0;
  zr_movie_device_bytes(mv, &b);
  *plan = mv;
  *bytes = b + (size_t)n * h * w * 4;       // the caller's frames are resident while the movie runs
  return ZR_OK;
}

int hip_run(void*, void* plan, const zr_serve_job* job) {
  zr_movie* mv = static_cast<zr_movie*>(plan);
  int rc = zr_movie_set_frames(mv, static_cast<const float*>(job->frames), 1);
  if (!rc) rc = zr_align(mv);
  if (!rc && (job->image_sum || job->filt_sum)) rc = zr_copy_sums_device(mv, job->image_sum, job->filt_sum, nullptr, nullptr);
  if (!rc) rc = zr_synchronize(mv);
  if (!rc && job->translations)
    rc = zr_get_solution(mv, job->translations, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (!rc && job->frc && job->frc_n && job->params->do_frc) {
    int nf = 0;
    std::vector<double> tmp(std::max(job->params->frc_box, 2) / 2 + 1);
    rc = zr_get_frc(mv, tmp.data(), &nf);
    if (!rc) {
      nf = std::min(nf, job->frc_cap);
      std::memcpy(job->frc, tmp.data(), sizeof(double) * nf);
      *job->frc_n = nf;
    }
  }
  if (rc) return sfail(rc, "%s", zr_last_error());
  return ZR_OK;
}

void hip_destroy(void*, void* plan) { zr_movie_destroy(static_cast<zr_movie*>(plan)); }

// ------------------------------------------------------------- requests
struct Request {
  long long id = 0;
  zr_params params{};
  int n = 0, h = 0, w = 0;
  const void* frames = nullptr;
  float* image_sum = nullptr;
  float* filt_sum = nullptr;
  std::string key;
  int state = ZR_REQ_QUEUED;
  int code = ZR_OK;
  std::string error;
  double arrival = 0, admitted = 0, done = 0;
  std::vector<double> translations;
  std::vector<double> frc;
  int frc_n = 0;
  bool reported = false;          // handed out by zr_server_wait_any
};

struct Plan {
  void* handle = nullptr;
  size_t bytes = 0;
};

// An admitted request and how it gets its plan: an idle one, or a new one
// for which `reserved` bytes are held.
struct Admission {
  std::shared_ptr<Request> req;
  Plan plan;
  bool create = false;
  size_t reserved = 0;
};

// Plan key: every setting field by value (not the struct's bytes, whose
// padding a C caller need not clear), then the movie dimensions.
std::string plan_key(const zr_params* p, int n, int h, int w) {
  std::string k;
  auto put = [&](const void* v, size_t nb) { k.append(static_cast<const char*>(v), nb); };
#define ZR_KEY(f) put(&p->f, sizeof(p->f));
  ZR_KEY(tri_mode) ZR_KEY(diag_width) ZR_KEY(diag_start) ZR_KEY(max_shift) ZR_KEY(b_mode) ZR_KEY(b_rad)
  ZR_KEY(sub_pix_reg) ZR_KEY(shift_method) ZR_KEY(weight_mode) ZR_KEY(peaksig_thres) ZR_KEY(logistic_k)
  ZR_KEY(logistic_nu) ZR_KEY(origin_centroid) ZR_KEY(suppress_origin) ZR_KEY(fou_crop) ZR_KEY(shape_padded)
  ZR_KEY(do_dose) ZR_KEY(dose_a) ZR_KEY(dose_b) ZR_KEY(dose_c) ZR_KEY(dose_order) ZR_KEY(dose_missing)
  ZR_KEY(pixelsize_nm) ZR_KEY(voltage_kv) ZR_KEY(do_frc) ZR_KEY(frc_box) ZR_KEY(frc_overlap) ZR_KEY(auto_max)
  ZR_KEY(shape_binned) ZR_KEY(do_hot) ZR_KEY(hot_logistic_k) ZR_KEY(hot_relax) ZR_KEY(hot_max_sigma)
  ZR_KEY(hot_psf_k2) ZR_KEY(hot_cutoff_lower) ZR_KEY(hot_cutoff_upper) ZR_KEY(hot_fixed_cutoffs)
  ZR_KEY(hot_neighbour_thres) ZR_KEY(hot_mu_sigma) ZR_KEY(pre_shift) ZR_KEY(n_filters) ZR_KEY(filters)
  ZR_KEY(k1_batch) ZR_KEY(k2_cols_per_wave) ZR_KEY(k2_pairs_per_wave) ZR_KEY(user_mask)
#undef ZR_KEY
  const int dims[3] = {n, h, w};
  put(dims, sizeof dims);
  return k;
}

}  // namespace

struct zr_server {
  zr_plan_ops ops{};
  std::unique_ptr<HipBackend> hip;    // the built-in backend, if used
  size_t budget = 0;
  int max_inflight = 1;

  std::mutex mu;
  std::condition_variable cv;         // scheduler: new requests, freed slots or plans
  std::condition_variable cv_done;    // waiters: a request finished
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<long long, std::shared_ptr<Request>> requests;
  std::deque<long long> finished;     // completion order, not yet reported
  std::map<std::string, std::vector<Plan>> idle;
  std::list<std::string> lru;         // keys with idle plans, least recently used first
  std::unordered_map<std::string, size_t> key_bytes;   // measured bytes of one plan per key
  size_t held = 0;                    // bytes of every plan (idle, running, reserved)
  int inflight = 0;
  int creating = 0;
  long long next_id = 1;
  bool closed = false;
  std::deque<Admission> runq;
  long long st_admitted = 0, st_created = 0, st_evicted = 0, st_failed = 0, st_cancelled = 0;
  size_t st_peak_held = 0;

  std::thread sched;
  std::vector<std::thread> workers;

  // ---------------------------------------------------------- helpers (lock held)
  void touch(const std::string& key) {
    lru.remove(key);
    lru.push_back(key);
  }
  void put_idle(const std::string& key, const Plan& p) {
    idle[key].push_back(p);
    touch(key);
  }
  size_t bytes_for(const Request& r) const {
    auto it = key_bytes.find(r.key);
    return it != key_bytes.end() ? it->second : ops.estimate(ops.ctx, &r.params, r.n, r.h, r.w);
  }
  // Evict idle plans (least recently used keys first) until `need` more bytes
  // fit; the evicted plans go to `dead`, to be destroyed outside the lock.
  bool evict_for(size_t need, std::vector<Plan>& dead) {
    while (held + need > budget && !lru.empty()) {
      const std::string key = lru.front();
      auto it = idle.find(key);
      Plan p = it->second.back();
      it->second.pop_back();
      if (it->second.empty()) {
        idle.erase(it);
        lru.pop_front();
      }
      held -= p.bytes;
      dead.push_back(p);
      ++st_evicted;
    }
    return held + need <= budget;
  }
  void finish(const std::shared_ptr<Request>& r, int state) {
    r->state = state;
    r->done = now_s();
    finished.push_back(r->id);
    cv_done.notify_all();
  }
  void destroy_plans(std::vector<Plan>& dead) {
    for (auto& p : dead) ops.destroy(ops.ctx, p.handle);
    dead.clear();
  }

  // ---------------------------------------------------------- scheduler
  void schedule_loop() {
    std::vector<Plan> dead;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [&] { return (closed && pending.empty()) || (!pending.empty() && inflight < max_inflight); });
      if (pending.empty()) break;      // closed, and every request admitted
      Admission chosen;
      bool found = false;
      for (auto it = pending.begin(); it != pending.end(); ++it) {
        Request& r = **it;
        auto ip = idle.find(r.key);
        if (ip != idle.end()) {
          chosen.plan = ip->second.back();
          ip->second.pop_back();
          if (ip->second.empty()) {
            idle.erase(ip);
            lru.remove(r.key);
          }
        } else {
          const bool known = key_bytes.count(r.key) != 0;
          const size_t need = bytes_for(r);
          if (!evict_for(need, dead) && (known || held > 0)) continue;
          // the first plan of a key is charged an upper estimate: with nothing
          // else held it is created and its exact size checked instead
          held += need;
          chosen.create = true;
          chosen.reserved = need;
        }
        chosen.req = *it;
        pending.erase(it);
        found = true;
        break;
      }
      if (!dead.empty()) {
        // plans evicted for this pass are destroyed outside the lock; the
        // state may change meanwhile (a movie finishing frees its plan), so a
        // pass that admitted nothing is repeated rather than judged stale
        lk.unlock();
        destroy_plans(dead);
        lk.lock();
        if (!found) continue;
      }
      if (!found) {
        if (inflight == 0 && creating == 0 && !pending.empty()) {
          // nothing running and nothing fits even after eviction
          auto r = pending.front();
          pending.pop_front();
          r->code = ZR_ERR_NOMEM;
          char buf[160];
          snprintf(buf, sizeof buf, "movie %dx%dx%d needs ~%.2f GB, more than the HBM budget %.2f GB", r->n, r->h,
                   r->w, bytes_for(*r) / 1e9, budget / 1e9);
          r->error = buf;
          ++st_failed;
          finish(r, ZR_REQ_FAILED);
          continue;
        }
        cv.wait(lk);
        continue;
      }
      chosen.req->state = ZR_REQ_RUNNING;
      chosen.req->admitted = now_s();
      ++inflight;
      ++st_admitted;
      if (chosen.create) ++creating;
      runq.push_back(std::move(chosen));
      cv.notify_all();
    }
  }

  // ---------------------------------------------------------- workers
  void worker_loop() {
    for (;;) {
      Admission a;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !runq.empty() || (closed && pending.empty() && inflight == 0); });
        if (runq.empty()) return;
        a = std::move(runq.front());
        runq.pop_front();
      }
      run_one(a);
    }
  }

  void run_one(Admission& a) {
    Request& r = *a.req;
    std::vector<Plan> dead;
    int rc = ZR_OK;
    std::string msg;
    bool have_plan = !a.create;
    Plan plan = a.plan;
    if (a.create) {
      void* h = nullptr;
      size_t actual = 0;
      s_err.clear();
      rc = ops.create(ops.ctx, &r.params, r.n, r.h, r.w, &h, &actual);
      if (rc) msg = s_err.empty() ? "plan creation failed" : s_err;
      std::unique_lock<std::mutex> lk(mu);
      --creating;
      if (rc) {
        held -= a.reserved;
      } else {
        key_bytes[r.key] = actual;
        held = held - a.reserved + actual;
        ++st_created;
        if (held > budget) evict_for(0, dead);
        if (held > budget && held == actual) {
          // alone and still over the budget: this movie can never run
          held -= actual;
          dead.push_back(Plan{h, actual});
          rc = ZR_ERR_NOMEM;
          char buf[160];
          snprintf(buf, sizeof buf, "movie %dx%dx%d needs %.2f GB, more than the HBM budget %.2f GB", r.n, r.h, r.w,
                   actual / 1e9, budget / 1e9);
          msg = buf;
        } else {
          plan = Plan{h, actual};
          have_plan = true;
          st_peak_held = std::max(st_peak_held, held);
        }
      }
    }
    if (!dead.empty()) destroy_plans(dead);
    if (!rc) {
      zr_serve_job job{};
      job.params = &r.params;
      job.frames = r.frames;
      job.n = r.n;
      job.h = r.h;
      job.w = r.w;
      job.image_sum = r.image_sum;
      job.filt_sum = r.filt_sum;
      job.translations = r.translations.data();
      job.frc = r.frc.data();
      job.frc_cap = (int)r.frc.size();
      job.frc_n = &r.frc_n;
      s_err.clear();
      rc = ops.run(ops.ctx, plan.handle, &job);
      if (rc) msg = s_err.empty() ? std::string("the movie's run failed") : s_err;
    }
    // release the plan exactly once: back to the idle pool, or destroyed after a failure
    bool broken = rc != ZR_OK && have_plan;
    {
      std::unique_lock<std::mutex> lk(mu);
      --inflight;
      if (have_plan) {
        if (broken) held -= plan.bytes;
        else put_idle(r.key, plan);
      }
      r.code = rc;
      r.error = msg;
      if (rc) ++st_failed;
      finish(a.req, rc ? ZR_REQ_FAILED : ZR_REQ_DONE);
      cv.notify_all();
    }
    if (broken) ops.destroy(ops.ctx, plan.handle);
  }

  void start() {
    sched = std::thread([this] { schedule_loop(); });
    for (int i = 0; i < max_inflight; ++i) workers.emplace_back([this] { worker_loop(); });
  }

  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu);
      closed = true;
      cv.notify_all();
    }
    if (sched.joinable()) sched.join();
    {
      std::lock_guard<std::mutex> lk(mu);
      cv.notify_all();
    }
    for (auto& t : workers)
      if (t.joinable()) t.join();
    std::vector<Plan> dead;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& kv : idle)
        for (auto& p : kv.second) dead.push_back(p);
      idle.clear();
      lru.clear();
      held = 0;
    }
    destroy_plans(dead);
  }
};

// ------------------------------------------------------------------ C ABI
static int make_server(const zr_plan_ops* ops, size_t budget, int max_inflight, zr_server** out,
                       std::unique_ptr<HipBackend> hip) {
  if (!out || !ops || !ops->create || !ops->run || !ops->destroy || !ops->estimate)
    return sfail(ZR_ERR_ARG, "NULL server argument or plan operation");
  if (max_inflight < 1 || max_inflight > 64) return sfail(ZR_ERR_ARG, "max_inflight %d outside 1..64", max_inflight);
  zr_server* s = new (std::nothrow) zr_server();
  if (!s) return sfail(ZR_ERR_NOMEM, "out of host memory");
  s->ops = *ops;
  s->hip = std::move(hip);
  if (s->hip) s->ops.ctx = s->hip.get();
  s->budget = budget;
  s->max_inflight = max_inflight;
  s->start();
  *out = s;
  return ZR_OK;
}

extern "C" int zr_server_create(int device, size_t hbm_budget, int max_inflight, zr_server** out) {
  if (hbm_budget == 0) {
    hipError_t e = hipSetDevice(device);
    size_t fr = 0, tot = 0;
    if (e == hipSuccess) e = hipMemGetInfo(&fr, &tot);
    if (e != hipSuccess) return sfail(ZR_ERR_HIP, "device %d: %s", device, hipGetErrorString(e));
    hbm_budget = (size_t)((double)fr * 0.85);
  }
  auto be = std::make_unique<HipBackend>();
  be->device = device;
  zr_plan_ops ops{};
  ops.estimate = hip_estimate;
  ops.create = hip_create;
  ops.run = hip_run;
  ops.destroy = hip_destroy;
  return make_server(&ops, hbm_budget, max_inflight, out, std::move(be));
}

extern "C" int zr_server_create_ops(const zr_plan_ops* ops, size_t hbm_budget, int max_inflight, zr_server** out) {
  if (hbm_budget == 0) return sfail(ZR_ERR_ARG, "a server with its own plan operations needs an HBM budget");
  return make_server(ops, hbm_budget, max_inflight, out, nullptr);
}

extern "C" void zr_server_destroy(zr_server* s) {
  if (!s) return;
  s->stop();
  delete s;
}

extern "C" int zr_server_submit(zr_server* s, const zr_params* p, int n, int h, int w, const void* frames,
                                float* image_sum, float* filt_sum, double arrival, long long* id) {
  if (!s || !p || !id) return sfail(ZR_ERR_ARG, "NULL argument");
  if (n < 2 || h < 16 || w < 16) return sfail(ZR_ERR_ARG, "movie %dx%dx%d: at least 2 frames of 16x16", n, h, w);
  if (!frames) return sfail(ZR_ERR_ARG, "NULL frames");
  auto r = std::make_shared<Request>();
  r->params = *p;
  r->n = n;
  r->h = h;
  r->w = w;
  r->frames = frames;
  r->image_sum = image_sum;
  r->filt_sum = filt_sum;
  r->key = plan_key(p, n, h, w);
  r->arrival = arrival > 0 ? arrival : now_s();
  r->translations.assign((size_t)n * 2, 0.0);
  r->frc.assign(2048, 0.0);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->closed) return sfail(ZR_ERR_STATE, "server is closed");
  r->id = s->next_id++;
  s->requests[r->id] = r;
  s->pending.push_back(r);
  s->cv.notify_all();
  *id = r->id;
  return ZR_OK;
}

extern "C" int zr_server_warm(zr_server* s, const zr_params* p, int n, int h, int w, int count, int* made) {
  if (!s || !p) return sfail(ZR_ERR_ARG, "NULL argument");
  Request probe;
  probe.params = *p;
  probe.n = n;
  probe.h = h;
  probe.w = w;
  probe.key = plan_key(p, n, h, w);
  int k = 0;
  for (; k < count; ++k) {
    std::vector<Plan> dead;
    size_t need;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      need = s->bytes_for(probe);
      const bool fits = s->evict_for(need, dead);
      if (!fits && s->held > 0) break;
      s->held += need;
    }
    s->destroy_plans(dead);
    void* hnd = nullptr;
    size_t actual = 0;
    int rc = s->ops.create(s->ops.ctx, p, n, h, w, &hnd, &actual);
    std::unique_lock<std::mutex> lk(s->mu);
    if (rc) {
      s->held -= need;
      if (made) *made = k;
      return rc;
    }
    s->key_bytes[probe.key] = actual;
    s->held = s->held - need + actual;
    ++s->st_created;
    s->st_peak_held = std::max(s->st_peak_held, s->held);
    s->put_idle(probe.key, Plan{hnd, actual});
    s->cv.notify_all();
  }
  if (made) *made = k;
  return ZR_OK;
}

extern "C" int zr_server_cancel(zr_server* s, long long id) {
  if (!s) return sfail(ZR_ERR_ARG, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return sfail(ZR_ERR_ARG, "unknown request %lld", id);
  auto& r = it->second;
  if (r->state != ZR_REQ_QUEUED) return sfail(ZR_ERR_STATE, "request %lld is already admitted", id);
  s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
  ++s->st_cancelled;
  s->finish(r, ZR_REQ_CANCELLED);
  s->cv.notify_all();
  return ZR_OK;
}

extern "C" int zr_server_poll(zr_server* s, long long id, int* state) {
  if (!s || !state) return sfail(ZR_ERR_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return sfail(ZR_ERR_ARG, "unknown request %lld", id);
  *state = it->second->state;
  return ZR_OK;
}

static bool is_final(int st) { return st == ZR_REQ_DONE || st == ZR_REQ_FAILED || st == ZR_REQ_CANCELLED; }

extern "C" int zr_server_wait(zr_server* s, long long id, double timeout_s, int* state) {
  if (!s) return sfail(ZR_ERR_ARG, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return sfail(ZR_ERR_ARG, "unknown request %lld", id);
  auto r = it->second;
  auto pred = [&] { return is_final(r->state); };
  if (timeout_s < 0) s->cv_done.wait(lk, pred);
  else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (state) *state = r->state;
  return ZR_OK;
}

extern "C" int zr_server_wait_any(zr_server* s, double timeout_s, long long* id, int* state) {
  if (!s || !id) return sfail(ZR_ERR_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto pred = [&] { return !s->finished.empty(); };
  if (timeout_s < 0) s->cv_done.wait(lk, pred);
  else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (s->finished.empty()) {
    *id = 0;
    if (state) *state = -1;
    return ZR_OK;
  }
  *id = s->finished.front();
  s->finished.pop_front();
  auto it = s->requests.find(*id);
  if (state) *state = it != s->requests.end() ? it->second->state : -1;
  if (it != s->requests.end()) it->second->reported = true;
  return ZR_OK;
}

extern "C" int zr_server_result(zr_server* s, long long id, double* translations, int n, double* frc, int cap,
                                int* nfrc, double* times3, char* error, int error_cap) {
  if (!s) return sfail(ZR_ERR_ARG, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return sfail(ZR_ERR_ARG, "unknown request %lld", id);
  const Request& r = *it->second;
  if (!is_final(r.state)) return sfail(ZR_ERR_STATE, "request %lld has not finished", id);
  if (translations && r.state == ZR_REQ_DONE) {
    if (n != r.n) return sfail(ZR_ERR_ARG, "request %lld has %d frames, not %d", id, r.n, n);
    std::memcpy(translations, r.translations.data(), sizeof(double) * 2 * r.n);
  }
  if (nfrc) *nfrc = r.state == ZR_REQ_DONE ? std::min(r.frc_n, std::max(cap, 0)) : 0;
  if (frc && nfrc && *nfrc) std::memcpy(frc, r.frc.data(), sizeof(double) * *nfrc);
  if (times3) {
    times3[0] = r.arrival;
    times3[1] = r.admitted;
    times3[2] = r.done;
  }
  if (error && error_cap > 0) snprintf(error, (size_t)error_cap, "%s", r.error.c_str());
  return r.state == ZR_REQ_DONE ? ZR_OK : (r.state == ZR_REQ_CANCELLED ? ZR_ERR_STATE : (r.code ? r.code : ZR_ERR_HIP));
}

extern "C" int zr_server_release(zr_server* s, long long id) {
  if (!s) return sfail(ZR_ERR_ARG, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return sfail(ZR_ERR_ARG, "unknown request %lld", id);
  if (!is_final(it->second->state)) return sfail(ZR_ERR_STATE, "request %lld has not finished", id);
  s->requests.erase(it);
  s->finished.erase(std::remove(s->finished.begin(), s->finished.end(), id), s->finished.end());
  return ZR_OK;
}

extern "C" int zr_server_stats(zr_server* s, long long* out8) {
  if (!s || !out8) return sfail(ZR_ERR_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out8[0] = s->st_admitted;
  out8[1] = s->st_created;
  out8[2] = s->st_evicted;
  out8[3] = (long long)s->st_peak_held;
  out8[4] = (long long)s->held;
  out8[5] = s->inflight;
  out8[6] = (long long)s->pending.size();
  out8[7] = s->st_failed;
  return ZR_OK;
}

extern "C" size_t zr_server_estimate_bytes(const zr_params* p, int n, int h, int w) {
  return p ? estimate_movie_bytes(p, n, h, w) : 0;
}
