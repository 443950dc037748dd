// Continuous-batching request loop over the engine (nbx_server_* in nbx.h).
//
// A request is one independent simulation: its bodies, a step count and its
// settings.  Requests are admitted first come, first served while their
// estimated device memory fits the budget (counting what the batch contexts
// actually hold instead, when that is more).  Admitted requests with the same
// settings share one batch context (one HIP stream, one host thread): a
// request joins its batch between two steps, every step advances all of the
// batch's systems with one set of kernel launches, and a request leaves its
// batch as soon as it has made its steps.  Batches with different settings
// run on their own streams, so their work overlaps on the device.  A
// system's trajectory does not depend on the other systems of its batch, so
// a request's result is the same alone and in any mix.  A batch that empties
// frees its device buffers, so only batches with work hold memory, and a
// batch left without work for kRetireIdle is retired altogether (its thread,
// stream and context end), so a stream of requests with ever new settings
// keeps only the batches it uses.  An engine call that fails mid-batch fails
// every request of that batch and resets its context: no request can receive
// another's bodies.  The engine is reached through a table of calls
// (nbx_serve_ops.h), this library's own by default.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nbx.h"
#include "nbx_serve_ops.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0, Clock::time_point t1) {
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Estimated device bytes per body (state, keys, tree nodes of both modes).
constexpr size_t kBytesPerBody = 700;
// A batch without members or newcomers for this long is retired.
constexpr auto kRetireIdle = std::chrono::milliseconds(2000);

struct Request {
  int64_t id = 0;
  nbx_request spec{};
  std::vector<double> x, y, z, vx, vy, vz, m;
  std::vector<int32_t> type;
  size_t bytes = 0;
  int32_t steps_done = 0;
  int state = 0;            // 0 queued, 1 running, 2 done, 3 failed
  nbx_status status = NBX_OK;
  std::string error;
  Clock::time_point t_submit, t_start, t_done;
};

struct BatchKey {
  int mode, precision, leaf, adaptive;
  double theta, eps, dt;
  bool operator<(const BatchKey& o) const {
    if (mode != o.mode) return mode < o.mode;
    if (precision != o.precision) return precision < o.precision;
    if (leaf != o.leaf) return leaf < o.leaf;
    if (adaptive != o.adaptive) return adaptive < o.adaptive;
    if (theta != o.theta) return theta < o.theta;
    if (eps != o.eps) return eps < o.eps;
    return dt < o.dt;
  }
};

struct Batch {
  BatchKey key{};
  nbx_ctx* ctx = nullptr;
  std::vector<Request*> members;    // index = system index in ctx
  std::deque<Request*> incoming;
  std::thread thread;
  std::atomic<uint64_t> held{0};    // device bytes of ctx after its last step
  std::atomic<bool> ended{false};   // retired: its thread has released the context and is returning
};

}  // namespace

struct nbx_server {
  nbx_engine_ops ops{};
  int device = 0;
  size_t budget = 0;
  size_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv;       // state changes: submissions, completions, stop
  std::deque<Request*> pending;
  std::map<int64_t, std::unique_ptr<Request>> requests;
  std::map<BatchKey, std::unique_ptr<Batch>> batches;
  std::vector<std::unique_ptr<Batch>> retired;   // threads still to be joined
  int64_t waiters = 0;              // threads inside nbx_server_wait
  int64_t next_id = 1;
  bool stop = false;
  std::thread scheduler;
  std::atomic<int64_t> body_steps{0};
  std::atomic<int64_t> batch_steps{0};
};

namespace {

// Caller holds the lock.  `admitted`: the request's bytes are counted in
// in_use (it left the queue).
void finish(nbx_server* s, Request* r, nbx_status st, const std::string& err, bool admitted = true) {
  r->status = st;
  r->error = err;
  r->state = st == NBX_OK ? 2 : 3;
  r->t_done = Clock::now();
  if (admitted) s->in_use -= r->bytes;
}

// An engine call failed: the context's systems no longer match the
// members, so every member fails with that error and the context is reset.
void fail_batch(nbx_server* s, Batch* b, nbx_status st, const std::string& err) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request* r : b->members) finish(s, r, st, err);
    b->members.clear();
  }
  (void)s->ops.upload_batch(b->ctx, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr);
  s->cv.notify_all();
}

// One batch: admit its newcomers, step everyone once, retire the finished.
// Free the context's buffers if no request is running or waiting to join,
// and publish what the context holds.
void settle(nbx_server* s, Batch* b) {
  if (b->members.empty()) {
    bool idle;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      idle = b->incoming.empty();
    }
    if (idle) (void)s->ops.trim(b->ctx);   // on failure the buffers stay and are reused
  }
  const uint64_t bytes = s->ops.device_bytes(b->ctx);
  {
    std::lock_guard<std::mutex> lk(s->mu);   // the scheduler admits against it
    b->held = bytes;
  }
  s->cv.notify_all();
}

void batch_loop(nbx_server* s, Batch* b) {
  for (;;) {
    std::vector<Request*> joining;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      const bool work = s->cv.wait_for(lk, kRetireIdle, [&] {
        return s->stop || !b->incoming.empty() || !b->members.empty();
      });
      if (s->stop) return;
      if (!work) {
        // Idle for kRetireIdle: leave the map under the lock, so the
        // scheduler either queued a newcomer before (and this batch stays)
        // or finds no batch for these settings and makes a new one.
        auto it = s->batches.find(b->key);
        if (it != s->batches.end() && it->second.get() == b) {
          s->retired.push_back(std::move(it->second));
          s->batches.erase(it);
          lk.unlock();
          s->ops.destroy(b->ctx);
          b->ctx = nullptr;
          b->ended = true;
          s->cv.notify_all();   // the scheduler joins this thread
          return;
        }
        continue;
      }
      joining.assign(b->incoming.begin(), b->incoming.end());
      b->incoming.clear();
    }
    for (Request* r : joining) {
      int32_t sys = -1;
      const nbx_status st = s->ops.add_system(b->ctx, r->spec.n, r->x.data(), r->y.data(), r->z.data(), r->vx.data(),
                                              r->vy.data(), r->vz.data(), r->m.data(), r->type.data(), &sys);
      if (st == NBX_OK && sys != (int32_t)b->members.size()) {
        // the context's systems and the members no longer correspond
        {
          std::lock_guard<std::mutex> lk(s->mu);
          finish(s, r, NBX_ERR_STATE, "the batch context numbered a new system out of order");
        }
        fail_batch(s, b, NBX_ERR_STATE, "the batch context numbered a new system out of order");
        continue;
      }
      std::lock_guard<std::mutex> lk(s->mu);
      if (st != NBX_OK) {
        finish(s, r, st, s->ops.last_error(b->ctx));
        s->cv.notify_all();
        continue;
      }
      r->state = 1;
      r->t_start = Clock::now();
      b->members.push_back(r);
    }
    if (b->members.empty()) {
      settle(s, b);
      continue;
    }
    int64_t bodies = 0;
    for (Request* r : b->members) bodies += r->spec.n;
    const nbx_status st = s->ops.run(b->ctx, b->key.dt, 1, nullptr);
    if (st != NBX_OK) {
      fail_batch(s, b, st, s->ops.last_error(b->ctx));
      settle(s, b);
      continue;
    }
    s->body_steps += bodies;
    s->batch_steps += 1;
    for (Request* r : b->members) r->steps_done += 1;
    // retire finished systems, highest index first so the others keep theirs
    for (int32_t k = (int32_t)b->members.size() - 1; k >= 0; --k) {
      Request* r = b->members[(size_t)k];
      if (r->steps_done < r->spec.nsteps) continue;
      const nbx_status rs = s->ops.remove_system(b->ctx, k, r->x.data(), r->y.data(), r->z.data(), r->vx.data(),
                                                 r->vy.data(), r->vz.data());
      if (rs != NBX_OK) {
        // The context may still hold the system, or hold the others in a
        // changed order: fail the whole batch (this request included).
        fail_batch(s, b, rs, s->ops.last_error(b->ctx));
        break;
      }
      std::lock_guard<std::mutex> lk(s->mu);
      finish(s, r, NBX_OK, std::string());
      b->members.erase(b->members.begin() + k);
      s->cv.notify_all();
    }
    settle(s, b);
  }
}

// Join the threads of retired batches that have ended (caller holds the
// lock; an ended thread takes the lock no more).
void reap(nbx_server* s) {
  for (size_t k = 0; k < s->retired.size();) {
    if (s->retired[k]->ended) {
      if (s->retired[k]->thread.joinable()) s->retired[k]->thread.join();
      s->retired.erase(s->retired.begin() + (long)k);
    } else {
      ++k;
    }
  }
}

// Bytes counted against the budget (caller holds the lock): the admitted
// requests' estimates, or what the batch contexts hold if that is more
// (their buffers grow with slack and stay until a batch empties).
size_t committed(const nbx_server* s) {
  uint64_t held = 0;
  for (const auto& kv : s->batches) held += kv.second->held.load();
  return std::max<size_t>(s->in_use, (size_t)held);
}

// Admission: first come, first served, while the memory fits.
void scheduler_loop(nbx_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto fits = [&] { return !s->pending.empty() && committed(s) + s->pending.front()->bytes <= s->budget; };
  auto ended = [&] {
    for (const auto& b : s->retired)
      if (b->ended) return true;
    return false;
  };
  for (;;) {
    s->cv.wait(lk, [&] { return s->stop || fits() || ended(); });
    if (s->stop) return;
    reap(s);
    while (fits()) {
      Request* r = s->pending.front();
      s->pending.pop_front();
      const BatchKey key{r->spec.mode, r->spec.precision, r->spec.leaf_size, r->spec.adaptive_dt, r->spec.theta,
                         r->spec.eps, r->spec.dt};
      auto it = s->batches.find(key);
      if (it == s->batches.end()) {
        auto b = std::make_unique<Batch>();
        b->key = key;
        nbx_config cfg;
        nbx_config_default(&cfg);
        cfg.device = s->device;
        cfg.mode = key.mode;
        cfg.precision = key.precision;
        cfg.theta = key.theta;
        cfg.eps = key.eps;
        cfg.leaf_size = key.leaf;
        cfg.adaptive_dt = key.adaptive;
        const nbx_status st = s->ops.create(&b->ctx, &cfg);
        if (st != NBX_OK) {
          s->in_use += r->bytes;
          finish(s, r, st, "could not create a context for this request's settings");
          s->cv.notify_all();
          continue;
        }
        Batch* bp = b.get();
        it = s->batches.emplace(key, std::move(b)).first;
        bp->thread = std::thread(batch_loop, s, bp);
      }
      s->in_use += r->bytes;
      it->second->incoming.push_back(r);
      s->cv.notify_all();
    }
  }
}

}  // namespace

extern "C" {

nbx_status nbx_server_create_with_ops(nbx_server** out, int device, uint64_t budget_bytes,
                                      const nbx_engine_ops* ops) {
  if (!out || budget_bytes == 0 || !ops) return NBX_ERR_ARG;
  if (!(ops->create && ops->destroy && ops->last_error && ops->add_system && ops->remove_system && ops->run &&
        ops->upload_batch && ops->trim && ops->device_bytes))
    return NBX_ERR_ARG;
  nbx_server* s = new nbx_server;
  s->ops = *ops;
  s->device = device;
  s->budget = (size_t)budget_bytes;
  s->scheduler = std::thread(scheduler_loop, s);
  *out = s;
  return NBX_OK;
}

nbx_status nbx_server_create(nbx_server** out, int device, uint64_t budget_bytes) {
  const nbx_engine_ops ops{nbx_create,     nbx_destroy,       nbx_last_error, nbx_add_system, nbx_remove_system,
                           nbx_run,        nbx_upload_batch,  nbx_trim,       nbx_device_bytes};
  return nbx_server_create_with_ops(out, device, budget_bytes, &ops);
}

void nbx_server_destroy(nbx_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->scheduler.joinable()) s->scheduler.join();
  for (auto& kv : s->batches)
    if (kv.second->thread.joinable()) kv.second->thread.join();
  for (auto& b : s->retired)
    if (b->thread.joinable()) b->thread.join();
  {
    // Requests not done fail, and threads already waiting on them return
    // before the server goes away.
    std::unique_lock<std::mutex> lk(s->mu);
    for (Request* r : s->pending) finish(s, r, NBX_ERR_STATE, "the server was stopped", false);
    s->pending.clear();
    for (auto& kv : s->requests)
      if (kv.second->state < 2) finish(s, kv.second.get(), NBX_ERR_STATE, "the server was stopped");
    s->cv.notify_all();
    s->cv.wait(lk, [&] { return s->waiters == 0; });
  }
  for (auto& kv : s->batches)
    if (kv.second->ctx) s->ops.destroy(kv.second->ctx);
  delete s;
}

nbx_status nbx_server_submit(nbx_server* s, const nbx_request* rq, int64_t* id) {
  if (!s || !rq || !id) return NBX_ERR_ARG;
  const nbx_request& q = *rq;
  if (q.n <= 0 || q.nsteps < 0 || !(q.dt > 0.0)) return NBX_ERR_ARG;
  if (!(q.x && q.y && q.z && q.vx && q.vy && q.vz && q.m && q.type)) return NBX_ERR_ARG;
  if (q.mode != NBX_MODE_REFERENCE && q.mode != NBX_MODE_PHYSICS) return NBX_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->spec = q;
  const size_t n = (size_t)q.n;
  r->x.assign(q.x, q.x + n); r->y.assign(q.y, q.y + n); r->z.assign(q.z, q.z + n);
  r->vx.assign(q.vx, q.vx + n); r->vy.assign(q.vy, q.vy + n); r->vz.assign(q.vz, q.vz + n);
  r->m.assign(q.m, q.m + n);
  r->type.assign(q.type, q.type + n);
  r->spec.x = r->spec.y = r->spec.z = r->spec.vx = r->spec.vy = r->spec.vz = r->spec.m = nullptr;
  r->spec.type = nullptr;
  r->bytes = n * kBytesPerBody;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->bytes > s->budget) return NBX_ERR_NOMEM;   // could never be admitted
  r->id = s->next_id++;
  *id = r->id;
  Request* rp = r.get();
  s->requests.emplace(r->id, std::move(r));
  if (q.nsteps == 0) {
    s->in_use += rp->bytes;
    finish(s, rp, NBX_OK, "");
  } else {
    s->pending.push_back(rp);
  }
  s->cv.notify_all();
  return NBX_OK;
}

int nbx_server_poll(nbx_server* s, int64_t id) {
  if (!s) return NBX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return NBX_ERR_ARG;
  return it->second->state;
}

nbx_status nbx_server_wait(nbx_server* s, int64_t id, double* x, double* y, double* z, double* vx, double* vy,
                           double* vz, nbx_request_timing* timing) {
  if (!s) return NBX_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  // nbx_server_destroy waits for every thread in here to leave.
  ++s->waiters;
  struct Leave {
    nbx_server* s;
    ~Leave() {
      --s->waiters;
      s->cv.notify_all();
    }
  } leave{s};
  // Look the id up again on every wake-up: another thread waiting on the
  // same id may have received the result and released the request.
  auto it = s->requests.end();
  s->cv.wait(lk, [&] {
    it = s->requests.find(id);
    return it == s->requests.end() || it->second->state >= 2;
  });
  if (it == s->requests.end()) return NBX_ERR_ARG;
  Request* r = it->second.get();
  if (timing) {
    timing->queue_ms = r->state == 2 && r->spec.nsteps > 0 ? ms_since(r->t_submit, r->t_start) : 0.0;
    timing->latency_ms = ms_since(r->t_submit, r->t_done);
  }
  const nbx_status st = r->status;
  if (r->state == 2) {
    const size_t n = (size_t)r->spec.n;
    double* outs[6] = {x, y, z, vx, vy, vz};
    const std::vector<double>* src[6] = {&r->x, &r->y, &r->z, &r->vx, &r->vy, &r->vz};
    for (int k = 0; k < 6; ++k)
      if (outs[k]) std::memcpy(outs[k], src[k]->data(), n * sizeof(double));
  }
  s->requests.erase(it);   // the result (or the failure) has been handed over
  return st;
}

nbx_status nbx_server_stats(nbx_server* s, nbx_server_info* out) {
  if (!s || !out) return NBX_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  out->queued = (int64_t)s->pending.size();
  int64_t running = 0;
  for (auto& kv : s->batches) running += (int64_t)kv.second->members.size() + (int64_t)kv.second->incoming.size();
  out->running = running;
  out->batches = (int64_t)s->batches.size();
  out->bytes_in_use = (uint64_t)s->in_use;
  out->body_steps = s->body_steps.load();
  out->batch_steps = s->batch_steps.load();
  uint64_t held = 0;
  for (auto& kv : s->batches) held += kv.second->held.load();
  out->device_bytes = held;
  return NBX_OK;
}

}  // extern "C"
