// tbh_serve.cpp -- continuous-batching scheduler for MIS requests
// (tbh_serve_start / tbh_submit / tbh_poll / tbh_wait / tbh_cancel).
//
// A request is one dynamic_ob_mis contraction: the maximum over its slices
// of (slice optimum + r), as maximum(contract_slices(...)) in
// src/dynamic_ob.jl:24-27.  Requests flow through two threads:
//   planner    plans each submitted request's slices (BATCHED plans, on the
//              context's plan threads) while the GPU runs earlier rounds;
//   scheduler  forms rounds from the planned requests in arrival order --
//              latency-lane requests (priority 1) ahead of the others, then
//              the unscheduled slices of the oldest request first, then the
//              next request's, up to max_resident_slices (and one dtype per
//              round) -- runs each round as one BATCHED execution (its waves
//              mix slices of every request in the round), and folds the
//              per-slice results into their requests.  A request joins the
//              next round after it is planned, so requests of any size are
//              admitted while others are in flight.
// Each slice's plan, and so its result, does not depend on the other slices
// of its round: a request's value is the same served alone or mixed.
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

}  // namespace

struct ServeState {
  tbh_ctx* ctx = nullptr;
  size_t max_resident = 0;
  std::mutex m;
  std::condition_variable cv_plan, cv_sched, cv_done;
  std::deque<std::shared_ptr<Req>> to_plan;  // submitted, not planned
  std::deque<std::shared_ptr<Req>> ready;    // planned, slices left to schedule
  std::unordered_map<uint64_t, std::shared_ptr<Req>> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread planner, scheduler;

  void fail(const std::shared_ptr<Req>& r, const std::string& why) {
    r->status = TBH_FAILED;
    r->err = why;
    r->t_done = now_s();
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
      {
        std::lock_guard<std::mutex> g(m);
        if (r->status == TBH_FAILED) continue;  // cancelled while planning
        if (rc) {
          fail(r, "planning: " + err);
          cv_done.notify_all();
          continue;
        }
        r->plans = std::move(plans);
        ready.push_back(r);
      }
      cv_sched.notify_one();
    }
  }

  void sched_loop() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
      // form a round
      std::vector<std::pair<std::shared_ptr<Req>, size_t>> round;
      int dtype = TBH_F32;
      {
        std::unique_lock<std::mutex> g(m);
        cv_sched.wait(g, [&] { return stop || !ready.empty(); });
        if (stop) return;
        // latency-lane requests first (in arrival order), then the rest;
        // the round's dtype is that of its first request
        std::vector<std::shared_ptr<Req>> order;
        for (auto& r : ready)
          if (r->priority) order.push_back(r);
        for (auto& r : ready)
          if (!r->priority) order.push_back(r);
        dtype = order.front()->dtype;
        const size_t cap = max_resident ? max_resident : (size_t)-1;
        for (auto& r : order) {
          if (r->dtype != dtype) continue;
          while (r->next < r->slices.size() && round.size() < cap) {
            round.push_back({r, r->next++});
            r->admitted = true;
            r->status = TBH_RUNNING;
          }
          if (round.size() >= cap) break;
        }
        // requests with every slice scheduled leave the ready list
        ready.erase(std::remove_if(ready.begin(), ready.end(),
                                   [](const std::shared_ptr<Req>& r) {
                                     return r->next >= r->slices.size();
                                   }),
                    ready.end());
      }
      // run it
      std::vector<const SliceOwned*> sl(round.size());
      std::vector<const SlicePlan*> pl(round.size());
      for (size_t i = 0; i < round.size(); i++) {
        sl[i] = &round[i].first->slices[round[i].second];
        pl[i] = &round[i].first->plans[round[i].second];
      }
      std::vector<uint32_t> res(round.size());
      int rc;
      {
        std::lock_guard<std::mutex> g(ctx->mu);
        rc = execute_batched(ctx, dtype, sl, [&](size_t i, std::string*) { return pl[i]; }, true,
                             res.data());
      }
      const std::string err = rc ? std::string(tbh_last_error()) : std::string();
      // fold results
      {
        std::lock_guard<std::mutex> g(m);
        const double t = now_s();
        for (size_t i = 0; i < round.size(); i++) {
          Req& r = *round[i].first;
          if (r.status == TBH_FAILED) continue;
          if (rc) {
            fail(round[i].first, "execution: " + err);
            ready.erase(std::remove(ready.begin(), ready.end(), round[i].first), ready.end());
            continue;
          }
          if (!r.has_best || greater(dtype, res[i], r.best)) r.best = res[i];
          r.has_best = true;
          if (++r.done == r.slices.size()) {
            r.status = TBH_DONE;
            r.t_done = t;
          }
        }
      }
      cv_done.notify_all();
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

int tbh_serve_start(tbh_ctx* ctx, size_t max_resident_slices) {
  if (!ctx) {
    set_error("null context");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->serve_mu);
  if (ctx->serve) {
    set_error("serving already started");
    return TBH_EBUSY;
  }
  auto s = std::make_shared<ServeState>();
  s->ctx = ctx;
  s->max_resident = max_resident_slices;
  s->planner = std::thread([p = s.get()] { p->plan_loop(); });
  s->scheduler = std::thread([p = s.get()] { p->sched_loop(); });
  ctx->serve = s;
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
