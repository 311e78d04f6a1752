// server.cpp -- continuous-batching request loop on one device.
//
// The reference serves pricing through MCP tool handlers that each run a
// blocking price_put_leg on a tokio worker (riga-mcp/src/main.rs:1821-1831,
// 2005-2035; price_book_async tasks at 1005-1042, 2541-2575) and plans work
// with riga-scheduler's bounded cluster x scenario work list
// (riga-scheduler/src/lib.rs:320-352).  Here one scheduler thread owns the
// device:
//   * submit() deep-copies a ticket of requests and returns at once;
//   * tickets are admitted in arrival order while the device-memory budget
//     (estimated exactly by a dry-run plan) and the in-flight limit allow;
//     everything admitted together becomes one plan (CRN scenarios share points);
//   * each plan runs on its own stream with stream-ordered memory, so a plan
//     admitted while others run overlaps with them on the device: its launches
//     are one workgroup per (item, path-chunk) unit (not persistent workgroups,
//     which would hold every slot until their queue is empty), so the hardware
//     dispatcher gives the new plan's workgroups slots as the running plans'
//     workgroups end -- the new work joins the running work within one unit's
//     time, and each plan's device span is recorded to show it;
//   * finished plans are reaped by polling their streams, results go back to
//     their tickets and the memory returns to the budget;
//   * tickets that wait longer than the queue limit are shed.
// A request's result does not depend on what it was batched with: its paths,
// scenario arithmetic and block partials are the same, and the block sums are
// double-double.
#include <math.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/riga_hip.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// A request with every borrowed array copied into owned storage.
struct OwnedRequest {
  riga_request r{};
  std::vector<double> spots, vols, q, corr, s0;
  std::vector<std::vector<double>> grid_a, grid_b, grid_c;
  std::vector<riga_implied_grid> implied;
  std::vector<riga_lv_grid> calibrated;
  std::vector<const riga_surface*> surfaces;  // resident: the handles only

  void copy_from(const riga_request& in) {
    r = in;
    const int F = in.factors > 0 ? in.factors : 0;
    spots.assign(in.spots, in.spots + F);
    vols.assign(in.vols, in.vols + F);
    q.assign(in.dividend_yields, in.dividend_yields + F);
    corr.assign(in.correlation, in.correlation + (size_t)F * F);
    s0.assign(in.initial_spots, in.initial_spots + F);
    r.spots = spots.data();
    r.vols = vols.data();
    r.dividend_yields = q.data();
    r.correlation = corr.data();
    r.initial_spots = s0.data();
    grid_a.clear();
    grid_b.clear();
    grid_c.clear();
    r.surfaces = nullptr;
    if (in.surfaces) {
      surfaces.assign(in.surfaces, in.surfaces + F);
      r.surfaces = surfaces.data();
      r.implied = nullptr;
      r.calibrated = nullptr;
    } else if (in.calibrated) {
      calibrated.resize(F);
      grid_a.resize(F);
      grid_b.resize(F);
      grid_c.resize(F);
      for (int f = 0; f < F; ++f) {
        const riga_lv_grid& g = in.calibrated[f];
        grid_a[f].assign(g.t_axis, g.t_axis + g.n_t);
        grid_b[f].assign(g.log_s, g.log_s + g.n_x);
        grid_c[f].assign(g.v, g.v + (size_t)g.n_t * g.n_x);
        calibrated[f] = riga_lv_grid{g.n_t, g.n_x, grid_a[f].data(), grid_b[f].data(), grid_c[f].data()};
      }
      r.calibrated = calibrated.data();
      r.implied = nullptr;
    } else if (in.implied) {
      implied.resize(F);
      grid_a.resize(F);
      grid_b.resize(F);
      grid_c.resize(F);
      for (int f = 0; f < F; ++f) {
        const riga_implied_grid& g = in.implied[f];
        grid_a[f].assign(g.maturities, g.maturities + g.n_maturities);
        grid_b[f].assign(g.strikes, g.strikes + g.n_strikes);
        grid_c[f].assign(g.vols, g.vols + (size_t)g.n_maturities * g.n_strikes);
        implied[f] = riga_implied_grid{g.n_maturities, g.n_strikes, grid_a[f].data(), grid_b[f].data(),
                                       grid_c[f].data()};
      }
      r.implied = implied.data();
      r.calibrated = nullptr;
    }
  }
};

enum TicketState { kQueued = 0, kRunning = 1, kDone = 2, kShed = 3, kFailed = 4 };

struct Ticket {
  uint64_t id = 0;
  int precision = RIGA_FP64;
  std::vector<std::unique_ptr<OwnedRequest>> reqs;
  int state = kQueued;
  double t_submit = 0, t_start = 0, t_end = 0;
  int64_t bytes = 0;
  int batch_tickets = 0;
  double path_steps = 0;
  double dev_start = -1.0, dev_end = -1.0;  // its plan's device span (seconds, context clock)
  std::vector<riga_result> results;
  riga_error err{};
};

struct InFlight {
  riga_plan* plan = nullptr;
  std::vector<Ticket*> tickets;
  std::vector<int> offsets;
  int64_t bytes = 0;
  std::vector<riga_result> results;
};

}  // namespace

struct riga_server {
  riga_ctx* ctx = nullptr;
  riga_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Ticket*> queue;
  std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
  std::vector<InFlight> inflight;
  int64_t bytes_in_use = 0;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  riga_server_stats stats{};
  double t0 = 0;
  int waiters = 0;  // callers inside riga_server_wait (destroy waits for them to leave)
  std::deque<std::pair<double, double>> spans;  // device spans of the latest finished plans

  riga_launch_config launch_config() const {
    riga_launch_config lc{};
    lc.paths_per_thread = cfg.paths_per_thread;  // 0: the autotuned shape, if any
    lc.normal_table = cfg.normal_table;
    lc.stream_ordered_alloc = 1;
    lc.persistent = 0;  // one workgroup per unit: plans share the device's slots
    return lc;
  }

  // reap finished plans; called with mu held.  The plans themselves are
  // destroyed by the caller with mu released.
  std::vector<riga_plan*> done_plans;
  void reap() {
    for (size_t i = 0; i < inflight.size();) {
      InFlight& f = inflight[i];
      const int ready = riga_plan_ready(f.plan);
      if (ready == 0) {
        ++i;
        continue;
      }
      riga_error err{};
      const int rc = riga_plan_wait(f.plan, f.results.data(), &err);
      const double t = now_s();
      double ds = -1.0, de = -1.0;
      if (rc == 0 && riga_plan_device_times(f.plan, &ds, &de) == 0) {
        // did this plan run on the device while another one did?
        bool met = false;
        for (const auto& sp : spans) met = met || (ds < sp.second && sp.first < de);
        if (met) stats.overlapped_plans += 1;
        spans.emplace_back(ds, de);
        if (spans.size() > 64) spans.pop_front();
      } else {
        ds = de = -1.0;
      }
      for (size_t k = 0; k < f.tickets.size(); ++k) {
        Ticket* tk = f.tickets[k];
        tk->t_end = t;
        tk->dev_start = ds;
        tk->dev_end = de;
        if (rc == 0) {
          tk->results.assign(f.results.begin() + f.offsets[k], f.results.begin() + f.offsets[k] + tk->reqs.size());
          tk->state = kDone;
          stats.completed_tickets += 1;
          stats.completed_requests += (int64_t)tk->reqs.size();
          stats.path_steps += tk->path_steps;
        } else {
          tk->err = err;
          tk->state = kFailed;
          stats.failed_tickets += 1;
        }
      }
      done_plans.push_back(f.plan);
      bytes_in_use -= f.bytes;
      inflight.erase(inflight.begin() + (long)i);
      cv_done.notify_all();
    }
  }

  void shed_stale(double t) {
    if (cfg.max_queue_seconds <= 0) return;
    while (!queue.empty() && t - queue.front()->t_submit > cfg.max_queue_seconds) {
      Ticket* tk = queue.front();
      queue.pop_front();
      tk->state = kShed;
      tk->t_end = t;
      snprintf(tk->err.code, sizeof tk->err.code, "shed");
      snprintf(tk->err.message, sizeof tk->err.message, "queued longer than %.3f s", cfg.max_queue_seconds);
      stats.shed_tickets += 1;
      cv_done.notify_all();
    }
  }

  void fail_queued(Ticket* tk, const riga_error& err) {
    auto it = std::find(queue.begin(), queue.end(), tk);
    if (it != queue.end()) queue.erase(it);
    tk->err = err;
    tk->state = kFailed;
    tk->t_end = now_s();
    stats.failed_tickets += 1;
    cv_done.notify_all();
  }

  // Device bytes of each new ticket's own plan (an upper bound for its share
  // of a joint plan), estimated once and outside the lock so that submitters
  // never wait on plan building.  Only this thread removes queued tickets, so
  // the pointers stay valid while unlocked.
  bool estimate_new(std::unique_lock<std::mutex>& lock) {
    std::vector<Ticket*> todo;
    for (Ticket* tk : queue) {
      if (tk->bytes == 0) todo.push_back(tk);
      if (todo.size() >= 32) break;
    }
    if (todo.empty()) return false;
    const riga_launch_config lc = launch_config();
    std::vector<int64_t> bytes(todo.size(), 0);
    std::vector<riga_error> errs(todo.size());
    std::vector<int> rcs(todo.size(), 0);
    lock.unlock();
    for (size_t i = 0; i < todo.size(); ++i) {
      std::vector<riga_request> own;
      for (auto& o : todo[i]->reqs) own.push_back(o->r);
      errs[i] = riga_error{};
      rcs[i] = riga_plan_estimate(ctx, own.data(), (int)own.size(), todo[i]->precision, &lc, &bytes[i], &errs[i]);
    }
    lock.lock();
    const int64_t budget = budget_bytes();
    for (size_t i = 0; i < todo.size(); ++i) {
      Ticket* tk = todo[i];
      if (rcs[i] != 0) {
        fail_queued(tk, errs[i]);  // invalid request
      } else if (bytes[i] > budget) {
        riga_error e{};
        snprintf(e.code, sizeof e.code, "memory_budget");
        snprintf(e.message, sizeof e.message, "request needs %lld bytes, budget %lld", (long long)bytes[i],
                 (long long)budget);
        fail_queued(tk, e);  // can never fit
      } else {
        tk->bytes = bytes[i] > 0 ? bytes[i] : 1;
      }
    }
    return true;
  }

  int64_t budget_bytes() const { return cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : (int64_t)64 << 30; }

  // Admit estimated tickets from the queue head into one new plan.  Called
  // with mu held; the plan is built and launched with mu released.
  bool admit(std::unique_lock<std::mutex>& lock) {
    const int max_inflight = cfg.max_inflight > 0 ? cfg.max_inflight : 2;
    if (queue.empty() || (int)inflight.size() >= max_inflight) return false;
    const int max_requests = cfg.max_batch_requests > 0 ? cfg.max_batch_requests : 4096;
    const int64_t budget = budget_bytes();
    const int precision = queue.front()->precision;
    std::vector<Ticket*> take;
    std::vector<riga_request> views;
    int64_t need = 0;
    double work = 0;
    const riga_launch_config lc = launch_config();
    while (!queue.empty()) {
      Ticket* tk = queue.front();
      if (tk->bytes == 0) break;  // not estimated yet
      if (tk->precision != precision) break;
      if (!take.empty() && (int)(views.size() + tk->reqs.size()) > max_requests) break;
      if (!take.empty() && cfg.max_batch_path_steps > 0 && work + tk->path_steps > cfg.max_batch_path_steps)
        break;  // keep one plan's run time bounded
      if (bytes_in_use + need + tk->bytes > budget) break;  // wait for memory to free up
      queue.pop_front();
      take.push_back(tk);
      for (auto& o : tk->reqs) views.push_back(o->r);
      need += tk->bytes;
      work += tk->path_steps;
    }
    if (take.empty()) return false;
    bytes_in_use += need;  // reserved while the plan is built
    InFlight f;
    f.tickets = take;
    int off = 0;
    for (Ticket* tk : take) {
      f.offsets.push_back(off);
      off += (int)tk->reqs.size();
    }
    f.results.resize(views.size());
    riga_error err{};
    lock.unlock();
    int rc = riga_plan_create(ctx, views.data(), (int)views.size(), precision, &lc, &f.plan, &err);
    if (rc == 0) rc = riga_plan_launch(f.plan, &err);
    riga_plan_info info{};
    if (rc == 0) riga_plan_query(f.plan, &info);
    lock.lock();
    bytes_in_use -= need;
    const double t = now_s();
    if (rc != 0) {
      if (f.plan) riga_plan_destroy(f.plan);
      for (Ticket* tk : take) {
        tk->err = err;
        tk->state = kFailed;
        tk->t_end = t;
        stats.failed_tickets += 1;
      }
      cv_done.notify_all();
      return true;
    }
    f.bytes = std::max<int64_t>(need, info.device_bytes);
    bytes_in_use += f.bytes;
    stats.peak_bytes = std::max(stats.peak_bytes, bytes_in_use);
    stats.batches += 1;
    stats.max_concurrent = std::max<int32_t>(stats.max_concurrent, (int32_t)inflight.size() + 1);
    for (Ticket* tk : take) {
      tk->state = kRunning;
      tk->t_start = t;
      tk->bytes = f.bytes;
      tk->batch_tickets = (int)take.size();
    }
    inflight.push_back(std::move(f));
    return true;
  }

  void destroy_done(std::unique_lock<std::mutex>& lock) {
    if (done_plans.empty()) return;
    std::vector<riga_plan*> d;
    d.swap(done_plans);
    lock.unlock();
    for (riga_plan* p : d) riga_plan_destroy(p);
    lock.lock();
  }

  void loop() {
    std::unique_lock<std::mutex> lock(mu);
    while (!stop) {
      reap();
      destroy_done(lock);
      shed_stale(now_s());
      bool progressed = estimate_new(lock);
      if (stop) break;
      reap();
      while (admit(lock)) progressed = true;
      if (!progressed) {
        // poll running plans every 100 us; sleep until work arrives otherwise
        if (!inflight.empty())
          cv_work.wait_for(lock, std::chrono::microseconds(100));
        else
          cv_work.wait_for(lock, std::chrono::milliseconds(20));
      }
    }
    // drain
    while (!inflight.empty() || !done_plans.empty()) {
      reap();
      destroy_done(lock);
      if (!inflight.empty()) {
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(200));
        lock.lock();
      }
    }
  }
};

extern "C" {

int riga_server_create(riga_ctx* ctx, const riga_server_config* cfg, riga_server** out, riga_error* err) {
  if (!ctx || !out) {
    if (err) {
      snprintf(err->code, sizeof err->code, "invalid_argument");
      snprintf(err->message, sizeof err->message, "ctx and out are required");
    }
    return -1;
  }
  std::unique_ptr<riga_server> s(new riga_server());
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  s->t0 = now_s();
  riga_server* raw = s.get();
  s->worker = std::thread([raw] { raw->loop(); });
  *out = s.release();
  return 0;
}

int riga_server_submit(riga_server* s, const riga_request* reqs, int n, int precision, uint64_t* ticket,
                       riga_error* err) {
  if (!s || !reqs || n <= 0 || !ticket) {
    if (err) {
      snprintf(err->code, sizeof err->code, "invalid_argument");
      snprintf(err->message, sizeof err->message, "server, requests and ticket are required");
    }
    return -1;
  }
  std::unique_ptr<Ticket> tk(new Ticket());
  tk->precision = precision;
  for (int i = 0; i < n; ++i) {
    std::unique_ptr<OwnedRequest> o(new OwnedRequest());
    o->copy_from(reqs[i]);
    tk->path_steps += (double)reqs[i].paths * (double)reqs[i].steps;
    tk->reqs.push_back(std::move(o));
  }
  tk->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lock(s->mu);
    tk->id = s->next_id++;
    *ticket = tk->id;
    s->queue.push_back(tk.get());
    s->stats.submitted_tickets += 1;
    s->tickets[tk->id] = std::move(tk);
  }
  s->cv_work.notify_one();
  return 0;
}

int riga_server_wait(riga_server* s, uint64_t ticket, double timeout_s, riga_result* results,
                     riga_ticket_info* info, riga_error* err) {
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) {
    if (err) {
      snprintf(err->code, sizeof err->code, "unknown_ticket");
      snprintf(err->message, sizeof err->message, "no ticket %llu", (unsigned long long)ticket);
    }
    return -1;
  }
  Ticket* tk = it->second.get();
  auto finished = [tk] { return tk->state >= kDone; };
  s->waiters += 1;
  bool got = true;
  if (timeout_s < 0)
    s->cv_done.wait(lock, finished);
  else
    got = s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), finished);
  s->waiters -= 1;
  if (s->waiters == 0) s->cv_done.notify_all();  // destroy may be waiting for the last waiter
  if (!got) return 1;  // still pending
  if (info) {
    info->state = tk->state;
    info->n_requests = (int32_t)tk->reqs.size();
    info->batch_tickets = tk->batch_tickets;
    info->submit_s = tk->t_submit - s->t0;
    info->start_s = tk->t_start > 0 ? tk->t_start - s->t0 : -1.0;
    info->end_s = tk->t_end - s->t0;
    info->latency_s = tk->t_end - tk->t_submit;
    info->device_bytes = tk->bytes;
    info->path_steps = tk->path_steps;
    info->device_start_s = tk->dev_start;
    info->device_end_s = tk->dev_end;
  }
  int rc = 0;
  if (tk->state == kDone) {
    if (results)
      for (size_t i = 0; i < tk->results.size(); ++i) results[i] = tk->results[i];
  } else {
    if (err) *err = tk->err;
    rc = tk->state == kShed ? 2 : -1;
  }
  s->tickets.erase(it);
  return rc;
}

int riga_server_stats_get(riga_server* s, riga_server_stats* out) {
  if (!s || !out) return -1;
  std::lock_guard<std::mutex> lock(s->mu);
  *out = s->stats;
  out->queued_tickets = (int32_t)s->queue.size();
  out->running_plans = (int32_t)s->inflight.size();
  out->bytes_in_use = s->bytes_in_use;
  return 0;
}

void riga_server_destroy(riga_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
    for (Ticket* tk : s->queue) {
      tk->state = kShed;
      tk->t_end = now_s();
      snprintf(tk->err.code, sizeof tk->err.code, "shed");
      snprintf(tk->err.message, sizeof tk->err.message, "the server was destroyed before the ticket ran");
      s->stats.shed_tickets += 1;
    }
    s->queue.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    // every ticket is finished now: wake the waiters and let them leave before
    // the server (and its mutex) is freed
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv_done.notify_all();
    s->cv_done.wait(lock, [s] { return s->waiters == 0; });
  }
  delete s;
}

}  // extern "C"
