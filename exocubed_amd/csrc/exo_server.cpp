// exo_server.cpp -- continuous-batching request loop over the public C ABI.
//
// Forward-run requests (a configuration, an initial state or a seeded recipe
// for one, a step count) arrive at any time.  Prep threads build initial
// states on the host; one scheduler thread admits prepared requests in
// arrival order while their device footprint fits the memory budget and a
// slot is free, keeps their state resident in HBM, and advances all active
// requests together through one exo_batch (one launch per phase over every
// active request, steps replayed from a captured HIP graph).  The batch is
// re-formed whenever a request joins or leaves, so requests that arrive while
// others run join at the next chunk boundary and finished ones free their
// memory at once.  Each request's result is independent of the mix: the
// batched kernels compute every member exactly as it would be computed alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/exocubed.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { Q_WAIT = 0, Q_PREP, Q_READY, Q_ACTIVE, Q_DONE, Q_FAILED };

struct Request {
  uint64_t id = 0;
  exo_config cfg{};
  int kind = -1;  // >= 0: build the initial state from (kind, seed, params)
  uint64_t seed = 0;
  std::vector<double> params;
  std::vector<double> u, s;  // initial state, then the result
  int nsteps = 0, done_steps = 0;
  size_t bytes = 0;
  size_t cells = 0;
  exo_handle *h = nullptr;
  double t_submit = 0, t_admit = 0, t_done = 0;
  double dt_last = 0;
  double sums0[5] = {0, 0, 0, 0, 0}, sums1[5] = {0, 0, 0, 0, 0};
  int state = Q_WAIT;
  bool keep_state = true;
  std::string err;
};

}  // namespace

struct exo_server {
  size_t budget = 0;
  int max_active = 0;
  int chunk = 0;
  int flags = 0;
  std::mutex mu;
  std::condition_variable cv_prep, cv_sched, cv_done;
  std::deque<Request *> queue;          // submitted, not yet admitted (arrival order)
  std::vector<Request *> active;
  std::map<uint64_t, Request *> all;
  std::vector<std::thread> prep;
  std::thread sched;
  bool stop = false;
  uint64_t next_id = 1;
  size_t used = 0;
  // statistics
  double t_first = -1, t_last_done = 0, busy_s = 0;
  double cell_updates = 0;
  long long completed = 0, batches_formed = 0, chunks = 0;
  int peak_active = 0;
  size_t peak_used = 0;
  std::string err;
};

namespace {

void prep_loop(exo_server *S) {
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_prep.wait(lk, [&] {
        if (S->stop) return true;
        for (Request *q : S->queue)
          if (q->state == Q_WAIT) return true;
        return false;
      });
      if (S->stop) return;
      for (Request *q : S->queue)
        if (q->state == Q_WAIT) {
          r = q;
          break;
        }
      r->state = Q_PREP;
    }
    std::string e;
    if (r->kind >= 0) {
      size_t n = (size_t)6 * (r->cfg.eos == EXO_EOS_SHALLOW_YZ ? 4 : 5) * (r->cfg.nx + 6) *
                 (r->cfg.nx + 6) * (r->cfg.nx1 + 6);
      size_t ns = (size_t)6 * r->cfg.ntracer * (r->cfg.nx + 6) * (r->cfg.nx + 6) * (r->cfg.nx1 + 6);
      r->u.assign(n, 0.0);
      r->s.assign(ns, 0.0);
      int rc = exo_init_state_cfg(&r->cfg, r->kind, r->seed, r->params.data(),
                                  (int)r->params.size(), r->u.data(),
                                  ns ? r->s.data() : nullptr);
      if (rc != EXO_OK) e = "initial state could not be built";
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (e.empty()) {
        r->state = Q_READY;
      } else {
        r->state = Q_FAILED;
        r->err = e;
        r->t_done = now_s();
        S->queue.erase(std::find(S->queue.begin(), S->queue.end(), r));
        S->cv_done.notify_all();
      }
    }
    S->cv_sched.notify_all();
  }
}

void fail_request(exo_server *S, Request *r, const std::string &m) {
  if (r->h) exo_destroy(r->h);
  r->h = nullptr;
  r->err = m;
  r->state = Q_FAILED;
  r->t_done = now_s();
  r->u.clear();
  r->s.clear();
  r->u.shrink_to_fit();
  r->s.shrink_to_fit();
}

void sched_loop(exo_server *S) {
  exo_batch *batch = nullptr;
  bool dirty = true;
  for (;;) {
    std::vector<Request *> admit;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_sched.wait(lk, [&] {
        if (S->stop) return true;
        if (!S->active.empty()) return true;
        return !S->queue.empty() && S->queue.front()->state == Q_READY;
      });
      if (S->stop && S->active.empty()) break;
      // admission: arrival order, while the footprint fits and a slot is free
      while (!S->queue.empty() && S->queue.front()->state == Q_READY &&
             (int)(S->active.size() + admit.size()) < S->max_active &&
             S->used + S->queue.front()->bytes <= S->budget) {
        Request *r = S->queue.front();
        S->queue.pop_front();
        S->used += r->bytes;
        admit.push_back(r);
      }
      S->peak_used = std::max(S->peak_used, S->used);
    }
    for (Request *r : admit) {
      int rc = exo_create(&r->cfg, &r->h);
      if (rc == EXO_OK)
        rc = exo_upload_state(r->h, r->u.data(), r->cfg.ntracer ? r->s.data() : nullptr);
      if (rc == EXO_OK) rc = exo_conserved_sums(r->h, r->sums0);
      std::lock_guard<std::mutex> lk(S->mu);
      if (rc != EXO_OK) {
        std::string m = r->h ? exo_last_error(r->h) : "exo_create failed";
        fail_request(S, r, m);
        S->used -= r->bytes;
        S->cv_done.notify_all();
        continue;
      }
      r->t_admit = now_s();
      r->state = Q_ACTIVE;
      // the host copy of the initial state is no longer needed
      r->u.clear();
      r->s.clear();
      r->u.shrink_to_fit();
      r->s.shrink_to_fit();
      S->active.push_back(r);
      dirty = true;
    }
    if (S->active.empty()) continue;
    S->peak_active = std::max(S->peak_active, (int)S->active.size());
    if (dirty) {
      if (batch) exo_batch_destroy(batch);
      batch = nullptr;
      std::vector<exo_handle *> hs;
      for (Request *r : S->active) hs.push_back(r->h);
      if (exo_batch_create(hs.data(), (int)hs.size(), S->flags & EXO_BATCH_GRAPH, &batch) !=
          EXO_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        for (Request *r : S->active) {
          fail_request(S, r, "exo_batch_create failed");
          S->used -= r->bytes;
        }
        S->active.clear();
        S->cv_done.notify_all();
        continue;
      }
      S->batches_formed++;
      dirty = false;
    }
    // advance every active request by the same number of steps: up to the
    // chunk, and never past the request that finishes first
    int n = S->chunk;
    for (Request *r : S->active) n = std::min(n, r->nsteps - r->done_steps);
    double t0 = now_s();
    int rc = n > 0 ? exo_batch_step(batch, n, 0.0) : EXO_OK;
    double t1 = now_s();
    std::vector<Request *> finished;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->busy_s += t1 - t0;
      S->chunks++;
      if (rc != EXO_OK) {
        std::string m = exo_batch_last_error(batch);
        for (Request *r : S->active) {
          fail_request(S, r, "batch step failed: " + m);
          S->used -= r->bytes;
        }
        S->active.clear();
        dirty = true;
        S->cv_done.notify_all();
        continue;
      }
      for (Request *r : S->active) {
        r->done_steps += n;
        S->cell_updates += (double)r->cells * n;
        if (r->done_steps >= r->nsteps) finished.push_back(r);
      }
    }
    for (Request *r : finished) {
      // retire: result to the host, device memory freed
      int rc2 = EXO_OK;
      if (r->keep_state) {
        int nv = r->cfg.eos == EXO_EOS_SHALLOW_YZ ? 4 : 5;
        size_t n1 = (size_t)6 * nv * (r->cfg.nx + 6) * (r->cfg.nx + 6) * (r->cfg.nx1 + 6);
        size_t ns = (size_t)6 * r->cfg.ntracer * (r->cfg.nx + 6) * (r->cfg.nx + 6) *
                    (r->cfg.nx1 + 6);
        r->u.assign(n1, 0.0);
        r->s.assign(ns, 0.0);
        rc2 = exo_download_state(r->h, r->u.data(), ns ? r->s.data() : nullptr);
      }
      if (rc2 == EXO_OK) rc2 = exo_conserved_sums(r->h, r->sums1);
      if (rc2 == EXO_OK) rc2 = exo_new_dt(r->h, &r->dt_last);
      std::lock_guard<std::mutex> lk(S->mu);
      if (rc2 != EXO_OK) {
        fail_request(S, r, exo_last_error(r->h));
      } else {
        exo_destroy(r->h);
        r->h = nullptr;
        r->state = Q_DONE;
        r->t_done = now_s();
        S->completed++;
        S->t_last_done = r->t_done;
      }
      S->used -= r->bytes;
      S->active.erase(std::find(S->active.begin(), S->active.end(), r));
      dirty = true;
    }
    if (!finished.empty()) S->cv_done.notify_all();
  }
  if (batch) exo_batch_destroy(batch);
}

}  // namespace

extern "C" {

int exo_server_create(size_t mem_budget, int max_active, int chunk_steps, int flags,
                      int prep_threads, exo_server **out) {
  if (!out) return EXO_ERR_ARG;
  *out = nullptr;
  if (max_active <= 0 || chunk_steps <= 0) return EXO_ERR_ARG;
  if (mem_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return EXO_ERR_HIP;
    mem_budget = fr / 10 * 8;  // 80 % of what is free now
  }
  exo_server *S = new exo_server();
  S->budget = mem_budget;
  S->max_active = max_active;
  S->chunk = chunk_steps;
  S->flags = flags;
  int np = prep_threads > 0 ? prep_threads : 4;
  for (int q = 0; q < np; ++q) S->prep.emplace_back(prep_loop, S);
  S->sched = std::thread(sched_loop, S);
  *out = S;
  return EXO_OK;
}

int exo_server_submit(exo_server *S, const exo_config *cfg, const double *u0, const double *s0,
                      int kind, uint64_t seed, const double *params, int nparams, int nsteps,
                      int keep_state, uint64_t *id_out) {
  if (!S || !cfg || !id_out || nsteps <= 0) return EXO_ERR_ARG;
  if (!u0 && kind < 0) return EXO_ERR_ARG;
  size_t bytes = exo_footprint(cfg);
  if (bytes == 0) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = "invalid configuration";
    return EXO_ERR_ARG;
  }
  if (bytes > S->budget) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = "request footprint exceeds the server's memory budget";
    return EXO_ERR_NOMEM;
  }
  Request *r = new Request();
  r->cfg = *cfg;
  r->bytes = bytes;
  r->nsteps = nsteps;
  r->keep_state = keep_state != 0;
  r->cells = (size_t)6 * cfg->nx * cfg->nx * cfg->nx1;
  if (u0) {
    int nv = cfg->eos == EXO_EOS_SHALLOW_YZ ? 4 : 5;
    size_t n = (size_t)6 * nv * (cfg->nx + 6) * (cfg->nx + 6) * (cfg->nx1 + 6);
    size_t ns = (size_t)6 * cfg->ntracer * (cfg->nx + 6) * (cfg->nx + 6) * (cfg->nx1 + 6);
    if (ns && !s0) {
      delete r;
      return EXO_ERR_ARG;
    }
    r->u.assign(u0, u0 + n);
    if (ns) r->s.assign(s0, s0 + ns);
  } else {
    r->kind = kind;
    r->seed = seed;
    if (params && nparams > 0) r->params.assign(params, params + nparams);
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete r;
      return EXO_ERR_STATE;
    }
    r->id = S->next_id++;
    r->t_submit = now_s();
    if (S->t_first < 0) S->t_first = r->t_submit;
    r->state = u0 ? Q_READY : Q_WAIT;
    S->queue.push_back(r);
    S->all[r->id] = r;
    *id_out = r->id;
  }
  S->cv_prep.notify_one();
  S->cv_sched.notify_all();
  return EXO_OK;
}

// info[0..9]: latency (submit to done), queue wait (submit to admit), run
// time (admit to done), dt of the last step, cells, steps, mass before,
// mass after, energy (or momentum) total before, after
int exo_server_wait(exo_server *S, uint64_t id, double timeout_s, double *u_out, double *s_out,
                    double *info) {
  if (!S) return EXO_ERR_ARG;
  Request *r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return EXO_ERR_ARG;
    r = it->second;
    auto done = [&] { return r->state == Q_DONE || r->state == Q_FAILED; };
    if (timeout_s > 0) {
      if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return EXO_ERR_STATE;
    } else {
      S->cv_done.wait(lk, done);
    }
    S->all.erase(it);
  }
  int rc = EXO_OK;
  if (r->state == Q_FAILED) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = r->err;
    rc = EXO_ERR_STATE;
  } else {
    if (u_out && !r->u.empty()) std::memcpy(u_out, r->u.data(), r->u.size() * sizeof(double));
    if (s_out && !r->s.empty()) std::memcpy(s_out, r->s.data(), r->s.size() * sizeof(double));
  }
  if (info) {
    info[0] = r->t_done - r->t_submit;
    info[1] = r->t_admit > 0 ? r->t_admit - r->t_submit : 0;
    info[2] = r->t_admit > 0 ? r->t_done - r->t_admit : 0;
    info[3] = r->dt_last;
    info[4] = (double)r->cells;
    info[5] = r->done_steps;
    info[6] = r->sums0[0];
    info[7] = r->sums1[0];
    int ie = r->cfg.eos == EXO_EOS_SHALLOW_YZ ? 2 : 4;
    info[8] = r->sums0[ie];
    info[9] = r->sums1[ie];
  }
  delete r;
  return rc;
}

// stats[0..8]: completed, cell-updates done, seconds from the first submit to
// the last completion, seconds the scheduler spent in batch steps, batches
// formed, chunks run, peak active requests, peak bytes in use, budget bytes
int exo_server_stats(exo_server *S, double *stats, int n) {
  if (!S || !stats) return EXO_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  double v[9] = {(double)S->completed, S->cell_updates,
                 S->t_first >= 0 ? S->t_last_done - S->t_first : 0.0, S->busy_s,
                 (double)S->batches_formed, (double)S->chunks, (double)S->peak_active,
                 (double)S->peak_used, (double)S->budget};
  for (int q = 0; q < n && q < 9; ++q) stats[q] = v[q];
  return EXO_OK;
}

const char *exo_server_last_error(const exo_server *S) { return S ? S->err.c_str() : "null"; }

int exo_server_destroy(exo_server *S) {
  if (!S) return EXO_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_prep.notify_all();
  S->cv_sched.notify_all();
  for (auto &t : S->prep) t.join();
  // the scheduler finishes the requests already on the device, then exits
  S->sched.join();
  for (Request *r : S->queue) {
    if (S->all.count(r->id)) S->all.erase(r->id);
    delete r;
  }
  for (auto &kv : S->all) delete kv.second;
  delete S;
  return EXO_OK;
}

}  // extern "C"
