// tbh_serve.cpp -- continuous-batching scheduler for MIS requests
// (tbh_serve_start / tbh_submit / tbh_poll / tbh_wait / tbh_cancel).
//
// A request is one dynamic_ob_mis contraction: the maximum over its slices
// of (slice optimum + r), as maximum(contract_slices(...)) in
// src/dynamic_ob.jl:24-27.  Requests flow through two threads:
//   planner    plans each submitted request's slices (BATCHED plans, on the
//              context's plan threads) while the GPU runs earlier rounds, and
//              refuses a request alone when one of its slices' arenas cannot
//              fit the serving region of the pool;
//   scheduler  forms rounds from the planned requests -- latency-lane
//              requests (priority 1) first, then the others in arrival order,
//              each request's unscheduled slices in order -- and admits slices
//              while the round's summed arena footprint fits the region, its
//              summed estimated work (MACs of the slices' trees) stays within
//              the round bound, and at most max_resident_slices are in it (one
//              dtype per round).  Each round is one batched execution whose
//              waves mix slices of every request in it.  Rounds are pipelined:
//              round k+1 is packed and enqueued on the stream while round k
//              runs, so the GPU does not wait for the host between rounds, and
//              a request that arrives meanwhile (or a latency-lane slice)
//              waits for at most the round running and the one queued, each
//              bounded by the round work.
// A round that fails is run again request by request; only the requests
// that fail alone end FAILED.  Each slice's plan, and so its result, does not
// depend on the other slices of its round: a request's value is the same
// served alone or mixed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "tbh_runtime.hpp"

namespace tbh {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  uint64_t ticket = 0;
  int dtype = TBH_F32;
  int priority = 0;
  std::vector<SliceOwned> slices;
  std::vector<SlicePlan> plans;
  size_t next = 0;  // first slice not yet scheduled
  size_t done = 0;  // slices with a result
  uint32_t best = 0;
  bool has_best = false;
  tbh_status status = TBH_QUEUED;
  bool admitted = false;  // some slice has been scheduled (cancel no longer allowed)
  std::string err;
  double t_submit = 0, t_done = 0;
};

// One batched execution: (request, slice index) pairs and their results.
struct Round {
  int dtype = TBH_F32;
  std::vector<std::pair<std::shared_ptr<Req>, size_t>> items;
  std::vector<uint32_t> res;
};

// larger of two dtype values given as bit patterns
bool greater(int dtype, uint32_t a, uint32_t b) {
  if (dtype == TBH_F32) {
    float x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x > y;
  }
  return (int32_t)a > (int32_t)b;
}

size_t arena_of(const SlicePlan& P) { return P.empty ? 0 : (size_t)P.arena_bytes; }

}  // namespace

struct ServeState {
  tbh_ctx* ctx = nullptr;
  size_t max_resident = 0;
  double round_macs = 0;    // estimated work bound of a round
  size_t region_cap = 0;    // arena region of an idle pool: a slice's upper bound
  size_t desc_cap = 0;      // descriptor bytes a round may take in an idle pool
  std::mutex m;
  std::condition_variable cv_plan, cv_sched, cv_done;
  std::deque<std::shared_ptr<Req>> to_plan;  // submitted, not planned
  std::deque<std::shared_ptr<Req>> ready;    // planned, slices left to schedule
  std::unordered_map<uint64_t, std::shared_ptr<Req>> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread planner, scheduler;
  // counters (tbh_serve_stats)
  uint64_t rounds = 0, retried_rounds = 0, failed_requests = 0;

  void fail(const std::shared_ptr<Req>& r, const std::string& why) {
    if (r->status == TBH_DONE || r->status == TBH_FAILED) return;
    r->status = TBH_FAILED;
    r->err = why;
    r->t_done = now_s();
    failed_requests++;
    ready.erase(std::remove(ready.begin(), ready.end(), r), ready.end());
  }

  void plan_loop() {
    for (;;) {
      std::shared_ptr<Req> r;
      {
        std::unique_lock<std::mutex> g(m);
        cv_plan.wait(g, [&] { return stop || !to_plan.empty(); });
        if (stop) return;
        r = to_plan.front();
        to_plan.pop_front();
      }
      std::vector<const SliceOwned*> ptrs(r->slices.size());
      for (size_t i = 0; i < ptrs.size(); i++) ptrs[i] = &r->slices[i];
      std::vector<SlicePlan> plans;
      int rc = plan_all(ctx, ptrs, true, &plans);
      std::string err = rc ? std::string(tbh_last_error()) : std::string();
      // admission by footprint: every slice's arena must fit the region an
      // idle pool gives a round, or the request is refused alone
      for (size_t i = 0; !rc && i < plans.size(); i++)
        if (arena_of(plans[i]) > region_cap) {
          rc = TBH_ENOMEM;
          err = "slice " + std::to_string(i) + ": arena of " +
                std::to_string(arena_of(plans[i])) + " bytes exceeds the serving region (" +
                std::to_string(region_cap) + " bytes of the pool)";
        } else if (desc_bytes(plans[i]) + desc_batch_overhead(plans[i].nlevels) > desc_cap) {
          rc = TBH_ENOMEM;
          err = "slice " + std::to_string(i) + ": descriptors of " +
                std::to_string(desc_bytes(plans[i])) + " bytes exceed the pool's share for them (" +
                std::to_string(desc_cap) + " bytes)";
        }
      {
        std::lock_guard<std::mutex> g(m);
        if (r->status == TBH_FAILED) continue;  // cancelled while planning
        if (rc) {
          fail(r, (rc == TBH_ENOMEM ? "admission: " : "planning: ") + err);
          cv_done.notify_all();
          continue;
        }
        r->plans = std::move(plans);
        ready.push_back(r);
      }
      cv_sched.notify_one();
    }
  }

  // Takes slices for the next round (caller holds m); empty when nothing
  // is ready.  Slices are taken in order -- latency lane first, then arrival
  // order, each request's slices in index order -- until the next one would
  // break a bound; the first slice is always taken.
  std::shared_ptr<Round> form_round(size_t region, size_t desc_budget) {
    auto R = std::make_shared<Round>();
    if (ready.empty()) return R;
    std::vector<std::shared_ptr<Req>> order;
    for (auto& r : ready)
      if (r->priority) order.push_back(r);
    for (auto& r : ready)
      if (!r->priority) order.push_back(r);
    R->dtype = order.front()->dtype;
    const size_t cap = max_resident ? max_resident : (size_t)-1;
    size_t bytes = 0, desc = 0;
    int levels = 0;
    double work = 0;
    bool full = false;
    for (auto& r : order) {
      if (r->dtype != R->dtype) continue;
      while (r->next < r->slices.size()) {
        const SlicePlan& P = r->plans[r->next];
        const size_t a = arena_of(P), d = desc_bytes(P);
        const int lv = std::max(levels, P.nlevels);
        const double w = P.empty ? 0.0 : P.node_macs;
        if (!R->items.empty() &&
            (R->items.size() >= cap || bytes + a > region || work + w > round_macs ||
             desc + d + desc_batch_overhead(lv) > desc_budget)) {
          full = true;
          break;
        }
        bytes += a;
        desc += d;
        levels = lv;
        work += w;
        R->items.push_back({r, r->next++});
        r->admitted = true;
        r->status = TBH_RUNNING;
      }
      if (full) break;
    }
    // requests with every slice scheduled leave the ready list
    ready.erase(std::remove_if(ready.begin(), ready.end(),
                               [](const std::shared_ptr<Req>& r) {
                                 return r->next >= r->slices.size();
                               }),
                ready.end());
    R->res.assign(R->items.size(), 0);
    return R;
  }

  // Folds a finished round's results into its requests (caller holds m).
  void fold(const Round& R) {
    const double t = now_s();
    for (size_t i = 0; i < R.items.size(); i++) {
      Req& r = *R.items[i].first;
      if (r.status == TBH_FAILED) continue;
      if (!r.has_best || greater(R.dtype, R.res[i], r.best)) r.best = R.res[i];
      r.has_best = true;
      if (++r.done == r.slices.size()) {
        r.status = TBH_DONE;
        r.t_done = t;
      }
    }
  }

  // After a failed execution: each request's slices of the round run again
  // on their own (the caller holds ctx->mu, no batched run open); a request
  // that fails alone fails, the others are folded.
  void rerun_alone(const Round& R) {
    std::vector<std::shared_ptr<Req>> reqs;
    for (auto& it : R.items)
      if (std::find(reqs.begin(), reqs.end(), it.first) == reqs.end()) reqs.push_back(it.first);
    for (auto& rq : reqs) {
      Round one;
      one.dtype = R.dtype;
      for (auto& it : R.items)
        if (it.first == rq) one.items.push_back(it);
      one.res.assign(one.items.size(), 0);
      std::vector<const SliceOwned*> sl(one.items.size());
      std::vector<const SlicePlan*> pl(one.items.size());
      for (size_t i = 0; i < one.items.size(); i++) {
        sl[i] = &rq->slices[one.items[i].second];
        pl[i] = &rq->plans[one.items[i].second];
      }
      const int rc = execute_batched(ctx, R.dtype, sl,
                                     [&](size_t i, std::string*) { return pl[i]; }, true,
                                     one.res.data());
      const std::string err = rc ? std::string(tbh_last_error()) : std::string();
      std::lock_guard<std::mutex> g(m);
      if (rc)
        fail(rq, "execution: " + err);
      else
        fold(one);
    }
  }

  void sched_loop() {
    (void)hipSetDevice(ctx->device);
    BatchRunner run(ctx);
    std::unique_lock<std::mutex> held(ctx->mu, std::defer_lock);
    std::shared_ptr<Round> inflight;
    // ends the batched run and lets synchronous callers of the context in
    auto idle = [&] {
      run.close();
      if (held.owns_lock()) held.unlock();
    };
    auto finish_inflight = [&] {
      if (!inflight) return;
      void* done = nullptr;
      const int rc = run.flush(&done);
      if (rc) {
        run.close();
        {
          std::lock_guard<std::mutex> g(m);
          retried_rounds++;
        }
        rerun_alone(*inflight);
      } else {
        std::lock_guard<std::mutex> g(m);
        fold(*inflight);
      }
      inflight.reset();
      cv_done.notify_all();
    };
    for (;;) {
      bool stopping = false;
      {
        std::unique_lock<std::mutex> g(m);
        if (!inflight) cv_sched.wait(g, [&] { return stop || !ready.empty(); });
        stopping = stop;
      }
      if (stopping) {
        finish_inflight();
        idle();
        return;
      }
      if (!held.owns_lock()) {
        held.lock();
        (void)hipSetDevice(ctx->device);
      }
      if (run.region() == 0 && run.open()) {
        // no region (the pool is taken): fail what is ready, one round's worth
        const std::string err = tbh_last_error();
        idle();
        std::lock_guard<std::mutex> g(m);
        auto R = form_round((size_t)-1, (size_t)-1);
        for (auto& it : R->items) fail(it.first, "execution: " + err);
        cv_done.notify_all();
        continue;
      }
      // the next round is formed once the round in flight is issued (its
      // last launches enter the device queue near its end), so that a
      // request arriving meanwhile -- a latency-lane slice above all -- can
      // still join it
      if (inflight) run.wait_issued();
      std::shared_ptr<Round> R;
      {
        std::lock_guard<std::mutex> g(m);
        R = form_round(run.region(), run.desc_budget());
        if (!R->items.empty()) rounds++;
      }
      if (R->items.empty()) {  // nothing new: finish the round in flight, go idle
        finish_inflight();
        idle();
        continue;
      }
      std::vector<const SliceOwned*> sl(R->items.size());
      std::vector<const SlicePlan*> pl(R->items.size());
      for (size_t i = 0; i < R->items.size(); i++) {
        sl[i] = &R->items[i].first->slices[R->items[i].second];
        pl[i] = &R->items[i].first->plans[R->items[i].second];
      }
      void* done = nullptr;
      const int rc = run.enqueue(R->dtype, sl, pl, true, R->res.data(), R.get(), &done);
      if (rc) {
        // neither this round's results nor (unless already handed back)
        // those of the round in flight can be trusted: run them again
        run.close();
        {
          std::lock_guard<std::mutex> g(m);
          retried_rounds++;
        }
        if (inflight) rerun_alone(*inflight);
        rerun_alone(*R);
        inflight.reset();
        cv_done.notify_all();
        continue;
      }
      if (done) {
        std::lock_guard<std::mutex> g(m);
        fold(*inflight);
      }
      if (done) cv_done.notify_all();
      inflight = R;
    }
  }
};

}  // namespace tbh

using tbh::ServeState;
using tbh::set_error;

static std::shared_ptr<ServeState> serving(tbh_ctx* ctx) {
  if (!ctx) {
    set_error("null context");
    return nullptr;
  }
  std::lock_guard<std::mutex> g(ctx->serve_mu);
  if (!ctx->serve) set_error("serving not started (tbh_serve_start)");
  return ctx->serve;
}

extern "C" {

int tbh_serve_start_ex(tbh_ctx* ctx, const tbh_serve_config* cfg) {
  if (!ctx || !cfg) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (cfg->round_log2_macs != 0 && !(cfg->round_log2_macs >= 10 && cfg->round_log2_macs <= 60)) {
    set_error("round_log2_macs must be 0 (default) or in [10, 60]");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->serve_mu);
  if (ctx->serve) {
    set_error("serving already started");
    return TBH_EBUSY;
  }
  auto s = std::make_shared<ServeState>();
  s->ctx = ctx;
  s->max_resident = cfg->max_resident_slices;
  s->round_macs =
      std::pow(2.0, cfg->round_log2_macs ? cfg->round_log2_macs : TBH_SERVE_ROUND_LOG2_MACS);
  s->region_cap = tbh::region_bytes(ctx->pool.capacity());
  s->desc_cap = (ctx->pool.capacity() - s->region_cap) / 2;
  s->planner = std::thread([p = s.get()] { p->plan_loop(); });
  s->scheduler = std::thread([p = s.get()] { p->sched_loop(); });
  ctx->serve = s;
  return 0;
}

int tbh_serve_start(tbh_ctx* ctx, size_t max_resident_slices) {
  tbh_serve_config cfg{};
  cfg.max_resident_slices = max_resident_slices;
  return tbh_serve_start_ex(ctx, &cfg);
}

int tbh_serve_stats(tbh_ctx* ctx, uint64_t* rounds, uint64_t* retried_rounds,
                    uint64_t* failed_requests) {
  auto s = serving(ctx);
  if (!s) return TBH_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  if (rounds) *rounds = s->rounds;
  if (retried_rounds) *retried_rounds = s->retried_rounds;
  if (failed_requests) *failed_requests = s->failed_requests;
  return 0;
}

// Stops both threads after the round in flight; requests not finished by
// then end FAILED ("server stopped").
int tbh_serve_stop(tbh_ctx* ctx) {
  if (!ctx) return 0;
  std::shared_ptr<ServeState> s;
  {
    std::lock_guard<std::mutex> g(ctx->serve_mu);
    s = ctx->serve;
  }
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> g(s->m);
    s->stop = true;
  }
  s->cv_plan.notify_all();
  s->cv_sched.notify_all();
  if (s->planner.joinable()) s->planner.join();
  if (s->scheduler.joinable()) s->scheduler.join();
  {
    std::lock_guard<std::mutex> g(s->m);
    for (auto& kv : s->all)
      if (kv.second->status == TBH_QUEUED || kv.second->status == TBH_RUNNING)
        s->fail(kv.second, "server stopped");
  }
  s->cv_done.notify_all();
  std::lock_guard<std::mutex> g(ctx->serve_mu);
  ctx->serve.reset();
  return 0;
}

int tbh_submit(tbh_ctx* ctx, const tbh_request* req, uint64_t* ticket) {
  auto s = serving(ctx);
  if (!s) return TBH_EINVAL;
  if (!req || !ticket || (req->n && !req->slices)) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (req->n == 0) {
    set_error("a request needs at least one slice (maximum over zero slices)");
    return TBH_EINVAL;
  }
  if (req->dtype != TBH_F32 && req->dtype != TBH_I32) {
    set_error("unknown dtype");
    return TBH_EINVAL;
  }
  auto r = std::make_shared<tbh::Req>();
  r->dtype = req->dtype;
  r->priority = req->priority ? 1 : 0;
  r->slices.resize(req->n);
  for (size_t i = 0; i < req->n; i++) {
    int rc = tbh::convert_slice(req->slices[i], req->dtype, true, &r->slices[i]);
    if (rc) {
      set_error("slice " + std::to_string(i) + ": " + tbh_last_error());
      return rc;
    }
  }
  {
    std::lock_guard<std::mutex> g(s->m);
    if (s->stop) {
      set_error("server stopping");
      return TBH_EBUSY;
    }
    r->ticket = s->next_ticket++;
    r->t_submit = tbh::now_s();
    s->all[r->ticket] = r;
    s->to_plan.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_plan.notify_one();
  return 0;
}

static int report(ServeState& s, std::unordered_map<uint64_t, std::shared_ptr<tbh::Req>>::iterator it,
                  void* result, tbh_status* status, double* t_submit, double* t_done) {
  tbh::Req& r = *it->second;
  if (status) *status = r.status;
  if (t_submit) *t_submit = r.t_submit;
  if (t_done) *t_done = r.t_done;
  if (r.status == TBH_DONE && result) std::memcpy(result, &r.best, 4);
  if (r.status == TBH_FAILED) set_error(r.err);
  // a finished request is reported once, then forgotten
  if (r.status == TBH_DONE || r.status == TBH_FAILED) s.all.erase(it);
  return 0;
}

int tbh_poll(tbh_ctx* ctx, uint64_t ticket, void* result, tbh_status* status, double* t_submit,
             double* t_done) {
  auto s = serving(ctx);
  if (!s) return TBH_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) {
    set_error("unknown ticket " + std::to_string(ticket));
    return TBH_ENOTFOUND;
  }
  return report(*s, it, result, status, t_submit, t_done);
}

int tbh_wait(tbh_ctx* ctx, uint64_t ticket, void* result, tbh_status* status) {
  auto s = serving(ctx);
  if (!s) return TBH_EINVAL;
  std::unique_lock<std::mutex> g(s->m);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) {
    set_error("unknown ticket " + std::to_string(ticket));
    return TBH_ENOTFOUND;
  }
  std::shared_ptr<tbh::Req> r = it->second;
  s->cv_done.wait(g, [&] { return r->status == TBH_DONE || r->status == TBH_FAILED; });
  it = s->all.find(ticket);
  if (it == s->all.end()) {
    set_error("ticket " + std::to_string(ticket) + " was reported to another caller");
    return TBH_ENOTFOUND;
  }
  return report(*s, it, result, status, nullptr, nullptr);
}

int tbh_cancel(tbh_ctx* ctx, uint64_t ticket) {
  auto s = serving(ctx);
  if (!s) return TBH_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) {
    set_error("unknown ticket " + std::to_string(ticket));
    return TBH_ENOTFOUND;
  }
  auto r = it->second;
  if (r->admitted || r->status == TBH_DONE || r->status == TBH_FAILED) {
    set_error("request already admitted");
    return TBH_EBUSY;
  }
  s->to_plan.erase(std::remove(s->to_plan.begin(), s->to_plan.end(), r), s->to_plan.end());
  s->ready.erase(std::remove(s->ready.begin(), s->ready.end(), r), s->ready.end());
  r->status = TBH_FAILED;  // a planner still holding it drops it
  s->all.erase(it);
  s->cv_done.notify_all();
  return 0;
}

}  // extern "C"
