// engine.cpp -- batch planner, launcher and the extern "C" entry points.
//
// A batch of requests is grouped into items: requests that share seed, path
// count, ladder, schedule, payoff terms, rate and correlation differ only in
// spots, dividend yields and vols/surfaces, i.e. they are the reference's CRN
// scenarios (riga-risk/src/lib.rs:117-186, dispatch.rs:23-70) and one lane
// prices all of them from one set of Sobol points.  Items are laid out in one
// device arena, launched per (factor count, scenario bucket) longest-first,
// merged by the finalize kernel, and turned back into PriceResult fields
// with the aggregate formulas of note.rs:262-307.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "host.h"
#include "kernels.h"

using namespace riga;

namespace {

void set_err(riga_error* e, const std::string& code, const std::string& msg) {
  if (!e) return;
  snprintf(e->code, sizeof e->code, "%s", code.c_str());
  snprintf(e->message, sizeof e->message, "%s", msg.c_str());
}

void clear_err(riga_error* e) {
  if (!e) return;
  e->code[0] = 0;
  e->message[0] = 0;
}

constexpr int RIGA_EINVAL = -1;
constexpr int RIGA_EGPU = -2;
constexpr int RIGA_ENODEV = -3;

// Device buffer.  With a stream set, memory comes from the device's
// stream-ordered pool (hipMallocAsync / hipFreeAsync): no device-wide
// synchronisation, so a serving loop can create and retire plans while other
// plans run.
// A fixed device work area handed out front to back (the allocation-free call's
// scratch array): 256-byte aligned pieces, nothing returned.
struct Scratch {
  unsigned char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > cap || bytes > cap - at) return nullptr;
    used = at + bytes;
    return base + at;
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipStream_t stream = nullptr;
  Scratch* from = nullptr;  // set: memory is taken from a scratch area, never allocated or freed
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    release();
    size_t want = std::max(bytes, (size_t)4096);
    if (from) {
      p = from->take(want);
      if (!p) return hipErrorOutOfMemory;
      cap = want;
      return hipSuccess;
    }
    hipError_t e = stream ? hipMallocAsync(&p, want, stream) : hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    else p = nullptr;
    return e;
  }
  void release() {
    if (p && !from) (void)(stream ? hipFreeAsync(p, stream) : hipFree(p));
    p = nullptr;
    cap = 0;
  }
  ~DevBuf() { release(); }
};

// Pinned host blocks recycled across plans: hipHostFree waits for the whole
// device, so a serving loop that freed each plan's download buffer would stall
// behind every plan still running.
struct PinnedPool {
  std::mutex mu;
  std::multimap<size_t, void*> idle;  // capacity -> block
  hipError_t get(size_t bytes, void** p, size_t* cap) {
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = idle.lower_bound(bytes);
      if (it != idle.end() && it->first <= 4 * bytes + 65536) {
        *p = it->second;
        *cap = it->first;
        idle.erase(it);
        return hipSuccess;
      }
    }
    size_t want = 4096;
    while (want < bytes) want <<= 1;
    hipError_t e = hipHostMalloc(p, want, hipHostMallocDefault);
    *cap = e == hipSuccess ? want : 0;
    return e;
  }
  void put(void* p, size_t cap) {
    std::lock_guard<std::mutex> g(mu);
    idle.emplace(cap, p);
  }
  ~PinnedPool() {
    for (auto& kv : idle) (void)hipHostFree(kv.second);
  }
};

struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  PinnedPool* pool = nullptr;  // plans borrow from the context's pool
  // set: the buffer is a fixed pinned block of the context (the allocation-free
  // call's download area), used as it is and never freed here
  void* fixed = nullptr;
  size_t fixed_cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    drop();
    if (fixed) {
      if (bytes > fixed_cap) return hipErrorOutOfMemory;
      p = fixed;
      cap = fixed_cap;
      return hipSuccess;
    }
    if (pool) return pool->get(std::max(bytes, (size_t)4096), &p, &cap);
    size_t want = std::max(bytes, (size_t)4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void drop() {
    if (p && !fixed) {
      if (pool)
        pool->put(p, cap);
      else
        (void)hipHostFree(p);
    }
    p = nullptr;
    cap = 0;
  }
  ~HostBuf() { drop(); }
};

}  // namespace

// Owen coordinate table of one seed: OwenEntry[dims][2^kOwenPrefixBits], the scrambled
// Sobol coordinate of every point below 2^kOwenPrefixBits in each dimension (with the
// dimension's own direction row), built on the
// device, shared by every plan of the context (plans keep a reference, so a
// table replaced by a larger one stays valid until its last plan is gone).
struct OwenTable {
  uint64_t seed = 0;
  int dims = 0;
  uint64_t last_use = 0;  // the context's request count when a plan last took it
  DevBuf buf;
  DevBuf dirs;  // the direction rows it was built from
  std::vector<uint32_t> host_dirs;
  hipEvent_t ready = nullptr;  // recorded after the build
  ~OwenTable() {
    if (ready) (void)hipEventDestroy(ready);
  }
};

struct riga_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t epoch = nullptr;  // recorded once at set-up: the origin of plans' device times
  std::string name, arch;
  size_t lds_limit = 65536;
  int n_cus = 256;  // compute units of the device
  std::mutex mu;
  std::mutex tune_mu;
  std::map<std::string, std::pair<int, int>> tuned;  // shape signature -> (paths/thread, waves)
  std::mutex owen_mu;
  std::map<uint64_t, std::shared_ptr<struct OwenTable>> owen;  // per seed, grown on demand
  uint64_t owen_uses = 0;  // requests for a table so far (least recently used ones go first)
  size_t owen_cap = (size_t)64 << 30;       // Owen table bytes the context keeps: a third of the device
  size_t owen_plan_cap = (size_t)48 << 30;  // new table bytes one plan may build: a quarter of it
  DevBuf arena, items, block_item, partials, finals, final_base, records;
  HostBuf staging, finals_host;
  PinnedPool pinned;  // plans' download buffers (plans end before their context)
  // the allocation-free call: this device's scratch array (its address from
  // hipGetSymbolAddress at init), its stream, events and pinned download block,
  // all made once here; calls through this context take noalloc_mu
  void* scratch = nullptr;
  hipStream_t noalloc_stream = nullptr;
  hipEvent_t noalloc_e0 = nullptr, noalloc_e1 = nullptr;
  void* noalloc_host = nullptr;
  std::mutex noalloc_mu;
};
constexpr size_t kNoallocHostBytes = (size_t)1 << 20;  // merged records of one request (S <= 10) + status

// A calibrated grid kept in device memory: [log_s (n_x)][v (n_t x n_x)] doubles.
// The host copy serves the tenor brackets and the LDS sizing of each plan.
struct riga_surface {
  riga_ctx* ctx = nullptr;
  LvGrid grid;
  void* dev = nullptr;
};

namespace {

#define HIP_TRY(expr, what)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      set_err(err, "gpu_error", std::string(what) + ": " + hipGetErrorString(_e));      \
      return RIGA_EGPU;                                                                  \
    }                                                                                    \
  } while (0)

// Replay options: the paths to price (all when `ids` is null) and whether to
// write their worst-of series and coupon strips.
struct ReplayOpts {
  const uint64_t* ids = nullptr;
  uint64_t n_ids = 0;
  bool ledger = false;
};

// ------------------------------------------------------------ per-request prep
struct Prep {
  const riga_request* r = nullptr;
  int F = 0;
  int ladder_dims = 0;
  std::vector<double> chol;
  std::vector<LvGrid> lv_own;              // grids copied or calibrated for this request
  std::vector<const LvGrid*> lv;           // one per factor when local vol
  std::vector<const riga_surface*> res;    // per factor: the resident surface, or null
  std::string key;                         // CRN group key
};

template <typename T>
void put_bytes(std::string& k, const T& v) {
  k.append(reinterpret_cast<const char*>(&v), sizeof v);
}

bool prepare(const riga_request& r, Prep* p, Error* e) {
  if (!validate_request(r, e)) return false;
  const int F = r.factors;
  if (F > kMaxFactors) {
    e->code = "invalid_market_shape";
    e->message = "the gfx950 kernels price at most " + std::to_string(kMaxFactors) + " underlyings per trade";
    return false;
  }
  if (r.steps > 1000000 || r.paths >= (1ull << 31)) {
    e->code = "invalid_schedule";
    e->message = "steps and paths exceed the supported range";
    return false;
  }
  p->r = &r;
  p->F = F;
  if (!cholesky_factor(F, r.correlation, &p->chol, e)) return false;
  const int cap = r.max_dimension > 0 ? r.max_dimension : 512;
  p->ladder_dims = std::max(1, std::min(r.ladder_dimension, cap));  // qmc.rs:105,249 clamp
  const bool has_lv = r.surfaces != nullptr || r.implied != nullptr || r.calibrated != nullptr;
  if (has_lv) {
    p->lv_own.resize(r.surfaces ? 0 : F);
    p->lv.assign(F, nullptr);
    p->res.assign(F, nullptr);
    for (int f = 0; f < F; ++f) {
      if (r.surfaces) {
        if (!r.surfaces[f]) {
          e->code = "invalid_local_vol";
          e->message = "resident surface missing for factor " + std::to_string(f);
          return false;
        }
        p->res[f] = r.surfaces[f];
        p->lv[f] = &r.surfaces[f]->grid;
        continue;
      }
      p->lv[f] = &p->lv_own[f];
      if (r.calibrated) {
        const riga_lv_grid& g = r.calibrated[f];
        if (g.n_t < 1 || g.n_x < 2 || g.n_x > kMaxLvX) {
          e->code = "invalid_local_vol";
          e->message = "calibrated local-vol grid must have 1+ tenors and 2.." + std::to_string(kMaxLvX) + " knots";
          return false;
        }
        p->lv_own[f].t_axis.assign(g.t_axis, g.t_axis + g.n_t);
        p->lv_own[f].log_s.assign(g.log_s, g.log_s + g.n_x);
        p->lv_own[f].v.assign(g.v, g.v + (size_t)g.n_t * g.n_x);
      } else {
        const int gx = r.grid_x > 0 ? r.grid_x : 96;
        if (gx > kMaxLvX) {
          e->code = "invalid_local_vol";
          e->message = "grid_x exceeds " + std::to_string(kMaxLvX);
          return false;
        }
        if (!dupire_calibrate(r.implied[f], r.spots[f], r.risk_free_rate, gx, &p->lv_own[f], e)) {
          e->code = "invalid_local_vol";
          return false;
        }
      }
    }
  }
  // group key: everything the scenarios of one item must share
  std::string& k = p->key;
  put_bytes(k, F);
  put_bytes(k, r.steps);
  put_bytes(k, r.maturity_years);
  put_bytes(k, r.strike);
  put_bytes(k, r.knock_in_barrier);
  put_bytes(k, r.has_knock_out);
  put_bytes(k, r.has_knock_out ? r.knock_out_barrier : 0.0);
  put_bytes(k, r.has_coupon_barrier);
  put_bytes(k, r.has_coupon_barrier ? r.coupon_barrier : 0.0);
  put_bytes(k, r.has_coupon_rate);
  put_bytes(k, r.has_coupon_rate ? r.coupon_rate : 0.0);
  put_bytes(k, r.memory_coupon);
  put_bytes(k, r.coupon_periods);
  put_bytes(k, r.fixed_first_coupon);
  put_bytes(k, r.call_start_step);
  put_bytes(k, r.paths);
  put_bytes(k, r.seed);
  put_bytes(k, r.ladder);
  put_bytes(k, p->ladder_dims);
  put_bytes(k, r.risk_free_rate);
  put_bytes(k, (int)(r.trace_level != RIGA_TRACE_OFF));
  put_bytes(k, (int)has_lv);
  for (int f = 0; f < F; ++f) put_bytes(k, r.initial_spots[f]);
  for (double v : p->chol) put_bytes(k, v);
  if (has_lv)
    for (int f = 0; f < F; ++f) {
      put_bytes(k, (int)p->lv[f]->t_axis.size());
      for (double t : p->lv[f]->t_axis) put_bytes(k, t);
    }
  return true;
}

// ------------------------------------------------------------ item building
struct Item {
  std::vector<int> members;  // request indices (scenarios in order)
  int F = 0, smax = 1;
  ItemDesc d{};
  std::vector<BridgeNode> nodes;
  std::vector<StepInfo> steps;
  std::vector<LvBracket> lvb;
  std::vector<ScenarioDesc> scen;
  std::vector<const LvGrid*> surfaces;  // distinct, in first-use order
  std::vector<const riga_surface*> surface_res;  // resident copy of each, or null
  std::vector<float> wt32;                       // FP32 tier tenor weights [step][factor]
  std::vector<int> surface_factor;
  std::vector<uint32_t> halton;
  int depth = 1;
  int stack_need = 0;  // bridge levels the kernel holds below its top at most
  double cost = 0.0;
  uint64_t span = 0;  // paths [0, span) of the request's stream that the item reads
};

// Streaming traversal of the reference bridge (qmc.rs:158-191): for every step,
// the pre-order nodes whose midpoints are first needed, and the stack depth.
void plan_bridge(int T, double dt, Item* it) {
  struct Pending {
    int hi, base;
  };
  std::vector<Pending> st;
  st.push_back({T, 1});
  int lo = 0, max_depth = 1;
  it->steps.resize(T);
  for (int t = 0; t < T; ++t) {
    StepInfo& si = it->steps[t];
    si.depth = (int)st.size();
    int c = 0;
    while (st.back().hi > lo + 1) {
      const int hi = st.back().hi, base = st.back().base;
      const int mid = (lo + hi) / 2;
      BridgeNode bn{};
      bn.node = base;
      bn.sd = sqrt((double)(hi - mid) * (double)(mid - lo) / (double)(hi - lo) * dt);
      bn.sd32 = (float)bn.sd;
      it->nodes.push_back(bn);
      st.back().base = base + (mid - lo);  // right child (mid, hi)
      st.push_back({mid, base + 1});       // left child (lo, mid)
      ++c;
      max_depth = std::max(max_depth, (int)st.size());
    }
    si.n_desc = c;
    st.pop_back();
    ++lo;
  }
  it->depth = max_depth;
  // the kernel's stack (top in registers): a step with c > 0 nodes pushes c - 1
  // of them, a step with none pops one; the levels it needs are the most pushed
  int sp = 0, need = 0;
  for (int t = 0; t < T; ++t) {
    const int c = it->steps[t].n_desc;
    if (c == 0) {
      sp = sp > 0 ? sp - 1 : 0;
    } else {
      sp += c - 1;
      need = std::max(need, sp);
    }
  }
  it->stack_need = need;
}

void plan_schedule(const riga_request& r, double dt, Item* it) {
  const int64_t T = r.steps;
  const int64_t P = r.coupon_periods < 0 ? 9 : std::max<int64_t>(r.coupon_periods, 1);  // note.rs:43
  const int64_t call_start = r.call_start_step < 0 ? 0 : r.call_start_step;            // note.rs:44
  const double coupon_rate = r.has_coupon_rate ? r.coupon_rate : 0.0;
  const double rate = r.risk_free_rate;
  int64_t prev = 0;
  for (int64_t step = 0; step < T; ++step) {
    StepInfo& si = it->steps[step];
    const int64_t period = std::min((step * P) / T, P - 1);
    const int64_t next = (step + 1 == T) ? P : std::min(((step + 1) * P) / T, P - 1);
    const int64_t a = T * (period + 1) / P, b = T * period / P;
    const int64_t period_steps = std::max<int64_t>(a > b ? a - b : 0, 1);
    const double time = (double)(step + 1) * dt;
    si.period = (int32_t)period;
    si.accrual_inc = coupon_rate / (double)period_steps;
    si.df = exp(-rate * time);
    si.flags = ((step + 1 == T || next != period) ? 1 : 0) | (step >= call_start ? 2 : 0) |
               (period != prev ? 4 : 0);
    prev = period;
  }
}

bool build_item(const std::vector<Prep>& preps, Item* it, int precision, bool records, Error* e) {
  const Prep& p0 = preps[it->members[0]];
  const riga_request& r = *p0.r;
  const int F = p0.F, T = (int)r.steps;
  const double dt = r.maturity_years / (double)r.steps;
  it->F = F;
  const int S = (int)it->members.size();
  it->smax = S <= 1 ? 1 : (S <= 4 ? 4 : 10);
  plan_bridge(T, dt, it);
  const bool mc = r.ladder == RIGA_VR_NONE;  // the Monte Carlo lane has no bridge
  if (!mc && it->depth > kBridgeStack) {
    e->code = "invalid_schedule";
    e->message = "bridge depth exceeds the supported number of steps";
    return false;
  }
  plan_schedule(r, dt, it);
  const int64_t periods = r.coupon_periods < 0 ? 9 : std::max<int64_t>(r.coupon_periods, 1);
  if (precision == RIGA_FP32 && periods > 2047) {  // the FP32 state word counts unpaid coupons in 11 bits
    e->code = "invalid_schedule";
    e->message = "the FP32 tier supports at most 2047 coupon periods";
    return false;
  }
  ItemDesc& d = it->d;
  memset(&d, 0, sizeof d);
  d.F = F;
  d.S = S;
  d.steps = T;
  d.periods = (int32_t)(r.coupon_periods < 0 ? 9 : std::max<int64_t>(r.coupon_periods, 1));
  d.call_start = (int32_t)(r.call_start_step < 0 ? 0 : r.call_start_step);
  d.flags = (r.memory_coupon ? kMemoryCoupon : 0) | (r.fixed_first_coupon ? kFixedFirst : 0) |
            (!p0.lv.empty() ? kLocalVol : 0) | (r.trace_level != RIGA_TRACE_OFF ? kTrace : 0) |
            (r.ladder == RIGA_VR_HALTON ? kHalton : 0) | (records ? kRecords : 0) | (mc ? kMonteCarlo : 0);
  if (mc) chacha_key_from_seed(r.seed, d.mc_key);
  d.ladder_dims = p0.ladder_dims;
  d.ladder_rows = (int32_t)std::min<int64_t>(p0.ladder_dims, (int64_t)F * T);
  d.paths = r.paths;
  d.seed = r.seed;
  d.path_begin = 0;
  const uint64_t max_n = d.path_begin + d.paths;
  d.nbits = 64 - __builtin_clzll(max_n);
  d.dt = dt;
  d.sqrt_dt = sqrt(dt);
  d.inv_sqrt_dt = 1.0 / sqrt(dt);
  d.root_sd = sqrt((double)T * dt);
  d.strike = r.strike;
  d.ki = r.knock_in_barrier;
  d.ko = r.has_knock_out ? r.knock_out_barrier : INFINITY;
  d.cb = r.has_coupon_barrier ? r.coupon_barrier : INFINITY;
  d.coupon_rate = r.has_coupon_rate ? r.coupon_rate : 0.0;
  d.put_df = exp(-r.risk_free_rate * r.maturity_years);
  for (int f = 0; f < F; ++f) {
    d.initial_spot[f] = r.initial_spots[f];
    d.ln_s0[f] = log(r.initial_spots[f]);
  }
  d.ln_ko = log(d.ko);  // +inf when absent
  d.ln_cb = log(d.cb);
  for (int i = 0; i < F; ++i)
    for (int j = 0; j < F; ++j) {
      d.chol[i * kMaxFactors + j] = p0.chol[i * F + j];
      d.chol32[i * kMaxFactors + j] = (float)p0.chol[i * F + j];
    }
  d.neg_half_dt = (float)(-0.5 * dt);
  d.ln_ko32 = (float)d.ln_ko;
  d.ln_cb32 = (float)d.ln_cb;
  d.root_sd32 = (float)d.root_sd;
  // scenarios
  it->scen.resize(S);
  for (int s = 0; s < S; ++s) {
    const Prep& p = preps[it->members[s]];
    const riga_request& q = *p.r;
    ScenarioDesc& sc = it->scen[s];
    memset(&sc, 0, sizeof sc);
    for (int f = 0; f < F; ++f) {
      const double vol = q.vols[f];
      sc.spot[f] = q.spots[f];
      sc.ln_rel[f] = log(q.spots[f]) - log(r.initial_spots[f]);
      sc.rq[f] = q.risk_free_rate - q.dividend_yields[f];
      sc.drift[f] = (q.risk_free_rate - q.dividend_yields[f] - 0.5 * vol * vol) * dt;  // note.rs:139-140
      sc.vsd[f] = vol * sqrt(dt);
      sc.lv[f] = -1;
      if (!p.lv.empty()) {
        const LvGrid* g = p.lv[f];
        int found = -1;
        for (size_t u = 0; u < it->surfaces.size(); ++u) {
          const LvGrid* h = it->surfaces[u];
          if (it->surface_factor[u] == f &&
              (h == g || (h->t_axis == g->t_axis && h->log_s == g->log_s && h->v == g->v))) {
            found = (int)u;
            break;
          }
        }
        if (found < 0) {
          found = (int)it->surfaces.size();
          it->surfaces.push_back(g);
          it->surface_res.push_back(p.res[f]);
          it->surface_factor.push_back(f);
        }
        sc.lv[f] = found;
      }
    }
  }
  if (!p0.lv.empty()) {
    it->lvb.resize((size_t)T * F);
    it->wt32.resize((size_t)T * F);
    for (int step = 0; step < T; ++step)
      for (int f = 0; f < F; ++f) {
        LvBracket& b = it->lvb[(size_t)step * F + f];
        lv_bracket(*p0.lv[f], (double)step * dt, &b.i0, &b.i1, &b.wt);  // note.rs:136 t = step*dt
        it->wt32[(size_t)step * F + f] = (float)b.wt;
      }
  }
  if (r.ladder == RIGA_VR_HALTON) it->halton = halton_tables(d.ladder_rows, r.seed);
  it->cost = (double)T * (double)F * (double)S * (double)r.paths;
  it->span = d.path_begin + d.paths;
  return true;
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

struct Arena {
  std::vector<unsigned char> bytes;
  template <typename T>
  int64_t add(const T* data, size_t n) {
    const size_t off = align16(bytes.size());
    bytes.resize(off + n * sizeof(T));
    if (n) memcpy(bytes.data() + off, data, n * sizeof(T));
    return (int64_t)off;
  }
};

// ------------------------------------------------------------ result assembly
// Aggregates of note.rs:262-307 from the merged double-double sums.
void fill_result(const Partial& fr, uint64_t paths, riga_result* out) {
  auto val = [](const DD& d) { return (long double)d.hi + (long double)d.lo; };
  const long double n = (long double)paths;
  const long double mean_funding = val(fr.q[kQFunding]) / n;
  const long double mean_put = val(fr.q[kQPut]) / n;
  const long double mean_cq = val(fr.q[kQCq]) / n;
  const long double mean_mq = val(fr.q[kQMq]) / n;
  auto moments = [&](int k1, int k2, uint64_t c, double* mean, double* var) {
    if (c == 0) {
      *mean = 0.0;
      *var = 0.0;
      return;
    }
    const long double s1 = val(fr.q[k1]), s2 = val(fr.q[k2]);
    *mean = (double)(s1 / (long double)c);
    *var = c > 1 ? (double)((s2 - s1 * s1 / (long double)c) / (long double)(c - 1)) : 0.0;
    if (*var < 0.0) *var = 0.0;
  };
  double pv_mean, pv_var;
  moments(kQPv1, kQPv2, paths, &pv_mean, &pv_var);
  out->paths = paths;
  out->mc_error = sqrt(pv_var / (double)paths);
  out->ki_probability = (double)fr.c[kCKi] / (double)paths;
  out->leg_funding = (double)mean_funding;
  out->leg_put = (double)mean_put;
  out->leg_coupon = (double)mean_cq;
  out->leg_memory_coupon = (double)mean_mq;
  const double total = out->leg_funding + (out->leg_coupon + out->leg_memory_coupon) / 10.0 - out->leg_put;
  out->pv = total;
  out->leg_total = total;
  out->ko_count = fr.c[kCKo];
  out->ki_count = fr.c[kCKi];
  out->put_eligible_count = fr.c[kCPutEligible];
  out->coupon_path_count = fr.c[kCCouponPaths];
  moments(kQKo1, kQKo2, fr.c[kCKo], &out->ko_mean, &out->ko_variance);
  moments(kQKi1, kQKi2, fr.c[kCKi], &out->ki_mean, &out->ki_variance);
  moments(kQCp1, kQCp2, fr.c[kCCouponPaths], &out->coupon_mean, &out->coupon_variance);
  out->n_trace = fr.n_events;
  for (int e = 0; e < fr.n_events && e < RIGA_TRACE_CAP; ++e) {
    out->trace_path[e] = fr.ev_path_lo[e];
    out->trace_step[e] = fr.ev_step[e];
  }
}

// ------------------------------------------------------------ plans
struct LaunchPlan {
  int F = 0, smax = 1, blocks = 0;  // blocks: the (item, chunk) units of the launch
  int grid = 0;                     // workgroups claiming them
  double cost = 0.0;                // the longest item's path-steps x factors (launch order)
  int mode = kLadderDirect;  // how the launch's items get their normals (kernel template mode)
  int threads = kBlock;  // lanes per workgroup
  size_t lds = 0;
  int pooled = 0;           // bottom bridge levels each workgroup keeps in an HBM slot
  int pool_n = 0;           // slots: the workgroups of this launch resident at once, at most
  size_t pool_off = 0;      // byte offset of this launch's slots in the plan's stack pool
  size_t flags_off = 0;     // word offset of its claim words (pool_n flags, one failure count)
  size_t block_off = 0;     // first (item, chunk) pair of this launch in the block table
  std::vector<int> items;   // item indices, longest first
};

// Launch-shape signature of a batch: precision and, per (factors, scenario
// bucket) launch, the item count and the log2 sizes of its paths and steps.
// The autotuner's choices are kept per signature.
using PartKey = std::tuple<int, int, int>;  // (factors, scenario bucket, ladder mode)

std::string shape_signature(int precision, const std::map<PartKey, std::vector<int>>& parts,
                            const std::vector<Item>& items) {
  std::string k = precision == RIGA_FP64 ? "d" : "f";
  for (const auto& kv : parts) {
    double paths = 0, steps = 0;
    for (int i : kv.second) {
      paths += (double)items[i].d.paths * items[i].d.S;
      steps += items[i].d.steps;
    }
    steps /= (double)kv.second.size();
    char buf[96];
    const int mode = std::get<2>(kv.first);
    snprintf(buf, sizeof buf, "|%sF%d.S%d.n%zu.p%d.t%d",
             mode == kLadderMonteCarlo ? "mc." : mode == kLadderTable ? "tab." : "",
             std::get<0>(kv.first), std::get<1>(kv.first), kv.second.size(),
             (int)std::ceil(std::log2(paths + 1.0)), (int)std::lround(std::log2(steps + 1.0)));
    k += buf;
  }
  return k;
}

}  // namespace

// A batch laid out in device memory once and launched any number of times.
struct riga_plan {
  riga_ctx* ctx = nullptr;
  int precision = RIGA_FP64;
  bool records = false;
  bool graph = false;
  int n_requests = 0;
  int ppt = 1, waves = 4;  // launch shape
  std::string signature;
  std::vector<Item> items;
  std::vector<int64_t> final_base;
  int64_t final_slots = 0, record_slots = 0, partial_slots = 0;
  std::vector<LaunchPlan> launches;
  double path_steps = 0.0;
  int blocks = 0;
  size_t lds_max = 0;
  DevBuf arena, items_d, block_item, partials, finals, final_base_d, records_d;
  DevBuf queues;  // one unit counter per launch, kQueueStride words apart
  DevBuf stack_pool, pool_words;  // launches' HBM slots for bottom bridge levels and their claim words
  size_t pool_bytes = 0, pool_nwords = 0;
  // launches after the first run on their own streams, forked from and joined
  // back into the plan's stream, so the persistent kernels of an uneven book
  // share the device instead of draining one after another
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> fork_join;  // [0]: fork; [k]: launch k done
  HostBuf finals_host;
  std::vector<PathRecordDev> recs_host;
  // replay of listed paths / coupon strips
  DevBuf path_list, series, ledger_d;
  std::vector<CouponRowDev> ledger_host;
  bool ledger = false;
  struct NormalTable {
    uint64_t seed = 0;
    int dims = 0, nbits = 0;
    int64_t paths = 0;
    const OwenEntry* owen = nullptr;
    int owen_dims = 0;  // dimensions the coordinate table covers
    DevBuf buf;
  };
  std::vector<std::unique_ptr<NormalTable>> ztables;
  // Monte Carlo lane: one stream parse per (seed, draws per path), giving the
  // first stream word of paths [0, paths)
  struct McParse {
    uint64_t seed = 0, dpp = 0, paths = 0;
    int chunk_words = 0, slow_cap = 0;
    int64_t n_chunks = 0, off_key = 0;
    DevBuf chunks, slow, slow_val, entry, cnt, block_off, starts;
    void each(void (*fn)(DevBuf&, hipStream_t), hipStream_t st) {
      for (DevBuf* b : {&chunks, &slow, &slow_val, &entry, &cnt, &block_off, &starts}) fn(*b, st);
    }
  };
  std::vector<std::unique_ptr<McParse>> mc;
  DevBuf mc_status;  // [parse][2]: paths found, error flags
  std::vector<std::shared_ptr<OwenTable>> owen_tabs;  // the context's tables this plan reads
  int64_t off_dirs = 0, off_zig = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool own_stream = true;  // false: the context's allocation-free stream and events, borrowed
  bool serial = false;     // launches one after another on `stream` (no fork)
  hipGraph_t g = nullptr;
  hipGraphExec_t gexec = nullptr;
  ~riga_plan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (g) (void)hipGraphDestroy(g);
    if (!own_stream) {
      e0 = e1 = nullptr;
      stream = nullptr;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (hipEvent_t ev : fork_join) (void)hipEventDestroy(ev);
    for (hipStream_t st : aux) (void)hipStreamDestroy(st);
    // buffers go back to the pool on this stream before it is destroyed
    for (DevBuf* b : {&arena, &items_d, &block_item, &partials, &finals, &final_base_d, &records_d, &path_list,
                      &series, &ledger_d, &queues, &stack_pool, &pool_words})
      b->release();
    for (auto& zt : ztables) zt->buf.release();
    for (auto& mp : mc) mp->each([](DevBuf& b, hipStream_t) { b.release(); }, nullptr);
    mc_status.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
  void bind_stream() {
    for (DevBuf* b : {&arena, &items_d, &block_item, &partials, &finals, &final_base_d, &records_d, &mc_status,
                      &queues, &stack_pool, &pool_words})
      b->stream = stream;
  }
  void bind_scratch(Scratch* sc) {
    for (DevBuf* b : {&arena, &items_d, &block_item, &partials, &finals, &final_base_d, &records_d, &path_list,
                      &series, &ledger_d, &mc_status, &queues})
      b->from = sc;
  }
};

namespace {

// The context's table for `seed` covering at least `dims` dimensions, building a
// larger one (at least twice the old size, so growth is rare) on `stream` when
// needed; callers wait on table->ready before their kernels read it.
// The context keeps at most ctx->owen_cap bytes of Owen tables; beyond it the least
// recently used tables of other seeds are dropped (plans still holding one keep it).

hipError_t owen_table_for(riga_ctx* ctx, uint64_t seed, int dims, hipStream_t stream,
                          std::shared_ptr<OwenTable>* out) {
  std::lock_guard<std::mutex> g(ctx->owen_mu);
  const uint64_t use = ++ctx->owen_uses;
  auto f = ctx->owen.find(seed);
  if (f != ctx->owen.end() && f->second->dims >= dims) {
    f->second->last_use = use;
    *out = f->second;
    return hipSuccess;
  }
  const int want = dims;  // the caller decides how far a table grows
  std::shared_ptr<OwenTable> t(new OwenTable());
  t->seed = seed;
  t->dims = want;
  hipError_t e = t->buf.ensure(((size_t)want << kOwenPrefixBits) * sizeof(OwenEntry));  // 2^bits entries per dimension
  // every dimension's own direction row (a table serves items whose dimensions do not wrap)
  {
    const std::vector<uint32_t>& all = sobol_directions(want);
    t->host_dirs.assign(all.begin(), all.begin() + (size_t)want * 32);  // lives as long as the copy may
  }
  if (e == hipSuccess) e = t->dirs.ensure((size_t)want * 32 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(t->dirs.p, t->host_dirs.data(), (size_t)want * 32 * sizeof(uint32_t), hipMemcpyHostToDevice,
                       stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&t->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = launch_owen_table(seed, want, (const uint32_t*)t->dirs.p, (OwenEntry*)t->buf.p, stream);
  if (e == hipSuccess) e = hipEventRecord(t->ready, stream);
  if (e != hipSuccess) return e;
  t->last_use = use;
  ctx->owen[seed] = t;
  size_t held = 0;
  for (const auto& kv : ctx->owen) held += kv.second->buf.cap;
  while (held > ctx->owen_cap && ctx->owen.size() > 1) {
    auto lru = ctx->owen.end();
    for (auto it = ctx->owen.begin(); it != ctx->owen.end(); ++it)
      if (it->first != seed && (lru == ctx->owen.end() || it->second->last_use < lru->second->last_use)) lru = it;
    if (lru == ctx->owen.end()) break;
    held -= lru->second->buf.cap;
    ctx->owen.erase(lru);
  }
  *out = t;
  return hipSuccess;
}

// Which seeds of a plan read an Owen coordinate table, and how many dimensions the
// table to build for each covers (a table covers the first dimensions of its seed;
// higher ones, and the seeds left out, hash every level -- the same integers).
// A table grows to at least twice its dimensions when it grows, and never past the
// bytes one plan may build (ctx->owen_plan_cap).  mode 1: every seed; mode k > 1:
// every seed, at most k dimensions.  mode 0: a seed whose table the context holds
// already, wide enough or as wide as it can be; otherwise a new table is built when
// the plan reads at least 2^18 points per dimension of that seed (a table costs
// 2^22 coordinates of 32 hashed levels per dimension; each coordinate read through it
// saves its gray-code XOR and 32 or 24 hashed levels),
// largest users first, while the new tables stay within
// ctx->owen_plan_cap and the context's tables that plans still hold plus the new
// ones within ctx->owen_cap.  `dims` (seed -> dimensions wanted) comes back as
// seed -> dimensions to request from owen_table_for.
void owen_table_seeds(riga_ctx* ctx, int mode, const std::map<uint64_t, double>& coords,
                      std::map<uint64_t, int>* dims) {
  const size_t row = ((size_t)1 << kOwenPrefixBits) * sizeof(OwenEntry);
  const int most = (int)std::min<size_t>(ctx->owen_plan_cap / row, (size_t)1 << 24);
  std::vector<std::pair<double, uint64_t>> order;
  for (const auto& kv : *dims) order.emplace_back(coords.at(kv.first) / std::max(1, kv.second), kv.first);
  std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  std::lock_guard<std::mutex> g(ctx->owen_mu);
  size_t pinned = 0;  // tables some plan still holds
  for (const auto& kv : ctx->owen)
    if (kv.second.use_count() > 1) pinned += kv.second->buf.cap;
  size_t fresh = 0;
  std::map<uint64_t, int> keep;
  for (const auto& pr : order) {
    const uint64_t seed = pr.second;
    const int want = std::min(dims->at(seed), mode > 1 ? std::min(mode, most) : most);
    if (want <= 0) continue;
    auto f = ctx->owen.find(seed);
    const int have = f == ctx->owen.end() ? 0 : f->second->dims;
    if (have >= want) {
      keep[seed] = have;
      continue;
    }
    const int build = mode > 1 ? want : std::min(std::max(want, 2 * have), most);
    const size_t bytes = (size_t)build * row;
    if (mode == 0 && (pr.first < (double)(1 << 18) || fresh + bytes > ctx->owen_plan_cap ||
                      pinned + fresh + bytes > ctx->owen_cap))
      continue;
    fresh += bytes;
    keep[seed] = build;
  }
  dims->swap(keep);
}

// Stream parse sizing for paths [0, mp->paths) at mp->dpp draws per path:
// chunks (one lane each) are sized so a parse fills the device (3 waves per
// SIMD at the parse's 130 VGPRs: ~196k lanes) before they grow past 512 words,
// and the stream is cut 3% beyond the words the draws need (1.022 per draw on
// average) plus eight chunks.
void mc_parse_size(riga_plan::McParse* mp) {
  const double draws = (double)mp->paths * (double)mp->dpp;
  int C = 512;
  while (C < 4096 && draws * 1.03 / C > 196608.0) C *= 2;
  mp->chunk_words = C;
  mp->slow_cap = C / 16 + 32;
  mp->n_chunks = (int64_t)std::ceil((draws * 1.03 + 8.0 * C) / C) + 1;
}

double mc_parse_bytes(const riga_plan::McParse& mp) {
  const double nc = (double)mp.n_chunks;
  return nc * (sizeof(McChunk) + (double)mp.slow_cap * 12 + 5) + ((mp.n_chunks + 255) / 256) * 8.0 +
         (double)mp.paths * 8.0;
}

hipError_t mc_parse_alloc(riga_plan::McParse* mp) {
  const size_t nc = (size_t)mp->n_chunks;
  hipError_t e;
  if ((e = mp->chunks.ensure(nc * sizeof(McChunk))) != hipSuccess) return e;
  if ((e = mp->slow.ensure(nc * (size_t)mp->slow_cap * 4)) != hipSuccess) return e;
  if ((e = mp->slow_val.ensure(nc * (size_t)mp->slow_cap * 8)) != hipSuccess) return e;
  if ((e = mp->entry.ensure(nc)) != hipSuccess) return e;
  if ((e = mp->cnt.ensure(nc * 4)) != hipSuccess) return e;
  if ((e = mp->block_off.ensure((size_t)((mp->n_chunks + 255) / 256) * 8)) != hipSuccess) return e;
  return mp->starts.ensure((size_t)mp->paths * 8);
}

hipError_t mc_parse_launch(const riga_plan::McParse& mp, const uint32_t* key, const double* zig, uint32_t* status,
                           hipStream_t stream) {
  McParseArgs pa;
  pa.key = key;
  pa.zig = zig;
  pa.chunk_words = mp.chunk_words;
  pa.slow_cap = mp.slow_cap;
  pa.n_chunks = mp.n_chunks;
  pa.slow = (uint32_t*)mp.slow.p;
  pa.slow_val = (double*)mp.slow_val.p;
  pa.chunks = (McChunk*)mp.chunks.p;
  McStartsArgs sa;
  sa.chunks = pa.chunks;
  sa.slow = pa.slow;
  sa.slow_cap = mp.slow_cap;
  sa.chunk_words = mp.chunk_words;
  sa.n_chunks = mp.n_chunks;
  sa.entry = (uint8_t*)mp.entry.p;
  sa.cnt = (uint32_t*)mp.cnt.p;
  sa.block_off = (uint64_t*)mp.block_off.p;
  sa.dpp = mp.dpp;
  sa.paths = mp.paths;
  sa.starts = (uint64_t*)mp.starts.p;
  sa.status = status;
  return launch_mc_parse(pa, sa, stream);
}

// Knots at x0 + i*dx to within 1e-9 of the spacing (linspace grids, as the
// calibration writes them): the FP32 tier then locates a knot by one multiply.
bool evenly_spaced(const std::vector<double>& x) {
  const size_t n = x.size();
  if (n < 2) return false;
  const double dx = (x[n - 1] - x[0]) / (double)(n - 1);
  if (!(dx > 0)) return false;
  for (size_t i = 0; i < n; ++i)
    if (std::fabs(x[i] - (x[0] + dx * (double)i)) > 1e-9 * dx) return false;
  return true;
}

PhoenixLaunch phoenix_launch_of(const riga_plan* pl, const LaunchPlan& lp, size_t k) {
  PhoenixLaunch L;
  L.F = lp.F;
  L.smax = lp.smax;
  L.mode = lp.mode;
  L.blocks = lp.grid;
  L.lds = lp.lds;
  L.threads = lp.threads;
  L.items = (const ItemDesc*)pl->items_d.p;
  L.block_item = (const int32_t*)pl->block_item.p + 2 * lp.block_off;
  L.n_units = lp.blocks;
  L.queue = (uint32_t*)pl->queues.p + k * kQueueStride;
  L.arena = (const unsigned char*)pl->arena.p;
  L.partials = (Partial*)pl->partials.p;
  L.records = pl->records ? (PathRecordDev*)pl->records_d.p : nullptr;
  return L;
}

// Enqueue the kernels, the finalize merge and the download of the merged records.
hipError_t enqueue(riga_plan* pl, bool timed) {
  hipError_t e;
  if (timed && (e = hipEventRecord(pl->e0, pl->stream)) != hipSuccess) return e;
  for (size_t k = 0; k < pl->mc.size(); ++k) {
    const unsigned char* a = (const unsigned char*)pl->arena.p;
    e = mc_parse_launch(*pl->mc[k], (const uint32_t*)(a + pl->mc[k]->off_key), (const double*)(a + pl->off_zig),
                        (uint32_t*)pl->mc_status.p + 2 * k, pl->stream);
    if (e != hipSuccess) return e;
  }
  for (const auto& zt : pl->ztables) {
    const unsigned char* a = (const unsigned char*)pl->arena.p;
    const uint32_t* dirs = (const uint32_t*)(a + pl->off_dirs);
    const OwenEntry* ot = zt->owen;
    e = pl->precision == RIGA_FP64
            ? launch_normal_table_f64(dirs, ot, zt->owen_dims, zt->seed, zt->dims, (int)zt->paths, zt->nbits,
                                      (double*)zt->buf.p, pl->stream)
            : launch_normal_table_f32(dirs, ot, zt->owen_dims, zt->seed, zt->dims, (int)zt->paths, zt->nbits,
                                      (float*)zt->buf.p, pl->stream);
    if (e != hipSuccess) return e;
  }
  e = launch_queue_reset((uint32_t*)pl->queues.p, (int)pl->launches.size(), pl->stream);
  if (e != hipSuccess) return e;
  const size_t nl = pl->launches.size();
  const bool fork = nl > 1 && !pl->serial;
  if (fork && (e = hipEventRecord(pl->fork_join[0], pl->stream)) != hipSuccess) return e;
  for (size_t k = 0; k < nl; ++k) {
    const LaunchPlan& lp = pl->launches[k];
    hipStream_t st = k == 0 || !fork ? pl->stream : pl->aux[k - 1];
    if (k > 0 && fork && (e = hipStreamWaitEvent(st, pl->fork_join[0], 0)) != hipSuccess) return e;
    const PhoenixLaunch L = phoenix_launch_of(pl, lp, k);
    e = pl->precision == RIGA_FP64 ? launch_phoenix_f64(L, st) : launch_phoenix_f32(L, st);
    if (e != hipSuccess) return e;
    if (k > 0 && fork) {
      if ((e = hipEventRecord(pl->fork_join[k], st)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(pl->stream, pl->fork_join[k], 0)) != hipSuccess) return e;
    }
  }
  if (pl->ledger) {
    e = launch_ledger((const ItemDesc*)pl->items_d.p, 0, pl->items[0].d.paths, (const unsigned char*)pl->arena.p,
                      (CouponRowDev*)pl->ledger_d.p, pl->stream);
    if (e != hipSuccess) return e;
  }
  if (timed && (e = hipEventRecord(pl->e1, pl->stream)) != hipSuccess) return e;
  e = launch_finalize((const ItemDesc*)pl->items_d.p, (int)pl->items.size(), (const Partial*)pl->partials.p,
                      (Partial*)pl->finals.p, (const int64_t*)pl->final_base_d.p, pl->stream);
  if (e != hipSuccess) return e;
  const size_t fin = (size_t)pl->final_slots * sizeof(Partial);
  e = hipMemcpyAsync(pl->finals_host.p, pl->finals.p, fin, hipMemcpyDeviceToHost, pl->stream);
  if (e != hipSuccess) return e;
  if (!pl->mc.empty()) {
    e = hipMemcpyAsync((unsigned char*)pl->finals_host.p + fin, pl->mc_status.p, 8 * pl->mc.size(),
                       hipMemcpyDeviceToHost, pl->stream);
    if (e != hipSuccess) return e;
  }
  // each pooled launch's count of workgroups that found no free slot (zero)
  unsigned char* failed = (unsigned char*)pl->finals_host.p + fin + 8 * pl->mc.size();
  for (size_t k = 0; k < pl->launches.size(); ++k) {
    const LaunchPlan& lp = pl->launches[k];
    if (lp.pooled <= 0) continue;
    e = hipMemcpyAsync(failed + 4 * k, (uint32_t*)pl->pool_words.p + lp.flags_off + lp.pool_n, 4,
                       hipMemcpyDeviceToHost, pl->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Static LDS of a path kernel, bounded above: the scenario-sum slots, the
// per-scenario trace counters, the FP32 tier's Euler constants and the unit id.
constexpr size_t kStaticLdsBound = sizeof(BlockShared) + 10 * 4 + 10 * kMaxFactors * sizeof(SfF32) + 64;

// Bridge stack levels (the least used positions, from both ends of the stack)
// that a launch keeps in HBM slots rather than LDS.  `split` < 0: none; > 0: that many; 0: the
// fewest that give the launch the most resident workgroups per CU -- a deep
// stack in LDS (levels x factors x lanes x 8 bytes in FP64) can otherwise hold
// occupancy below what the kernel's registers allow.
static int stack_pooled_levels(int precision, const LaunchPlan& lp, int levels, size_t lv_bytes, int real_size,
                                int split) {
  const int most = std::min(levels, kStackPooled);
  if (split < 0 || most <= 0) return 0;
  if (split > 0) return std::min(split, most);
  PhoenixLaunch L{};
  L.F = lp.F;
  L.smax = lp.smax;
  L.mode = lp.mode;
  L.threads = lp.threads;
  auto blocks_per_cu = [&](int p) {
    L.lds = (size_t)(levels - p) * lp.F * lp.threads * real_size + lv_bytes;
    int nb = 0;
    const hipError_t e = precision == RIGA_FP64 ? phoenix_blocks_per_cu_f64(L, &nb) : phoenix_blocks_per_cu_f32(L, &nb);
    return e == hipSuccess ? nb : 0;
  };
  int best = blocks_per_cu(0), best_p = 0;
  for (int p = 1; p <= most; ++p) {
    const int nb = blocks_per_cu(p);
    if (nb > best) {
      best = nb;
      best_p = p;
    }
  }
  return best_p;
}

// With `estimate` set, only the device bytes the plan would hold are computed
// (no stream, no allocation): the serving loop admits requests with it.
// With `scratch` set, every device buffer is taken from that fixed area and the
// context's allocation-free stream, events and pinned block are used: nothing
// is allocated (no shared normal tables, no Owen coordinate tables, no graph).
int plan_create(riga_ctx* ctx, const riga_request* reqs, int n, int precision, const riga_launch_config* cfg,
                bool records, riga_plan** out, riga_error* err, int64_t* estimate = nullptr,
                const ReplayOpts* replay = nullptr, Scratch* scratch = nullptr) {
  clear_err(err);
  *out = nullptr;
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n <= 0 || !reqs) {
    set_err(err, "invalid_argument", "a plan needs at least one request");
    return RIGA_EINVAL;
  }
  if (precision != RIGA_FP64 && precision != RIGA_FP32) {
    set_err(err, "invalid_precision", "precision must be RIGA_FP64 or RIGA_FP32");
    return RIGA_EINVAL;
  }
  int ppt = cfg && cfg->paths_per_thread > 0 ? cfg->paths_per_thread : 0;
  int waves = cfg && cfg->waves_per_block > 0 ? cfg->waves_per_block : 0;
  if (ppt != 0 && ppt != 1 && ppt != 2 && ppt != 4 && ppt != 8) {
    set_err(err, "invalid_launch_config", "paths_per_thread must be 1, 2, 4 or 8");
    return RIGA_EINVAL;
  }
  if (waves < 0 || waves > 4) {
    set_err(err, "invalid_launch_config", "waves_per_block must be 1, 2, 3 or 4");
    return RIGA_EINVAL;
  }
  if (scratch && (records || replay)) {
    set_err(err, "invalid_argument", "the allocation-free call prices without path records");
    return RIGA_EINVAL;
  }
  std::unique_ptr<riga_plan> pl(new riga_plan());
  pl->ctx = ctx;
  pl->finals_host.pool = &ctx->pinned;
  if (scratch) {
    pl->finals_host.fixed = ctx->noalloc_host;
    pl->finals_host.fixed_cap = kNoallocHostBytes;
    pl->bind_scratch(scratch);
  }
  pl->precision = precision;
  pl->records = records;
  pl->graph = cfg && cfg->use_graph != 0 && !records && !scratch;
  pl->n_requests = n;
  std::vector<Prep> preps(n);
  for (int i = 0; i < n; ++i) {
    Error e;
    if (reqs[i].steps >= (int64_t)1 << 20) {  // KO steps are packed in 20 bits on the device
      set_err(err, "invalid_schedule", "request " + std::to_string(i) + ": the device takes at most 1,048,575 steps");
      return RIGA_EINVAL;
    }
    // Sobol and Halton point indices n = path + 1 are 32-bit on the device (the
    // reference's are u64, qmc.rs:269-282) and trace-event path ids are int32:
    // beyond 2^31 - 1 paths the device would reuse points, so refuse
    if (reqs[i].paths > (uint64_t)INT32_MAX) {
      set_err(err, "invalid_paths", "request " + std::to_string(i) + ": the device takes at most 2,147,483,647 paths");
      return RIGA_EINVAL;
    }
    if (!prepare(reqs[i], &preps[i], &e)) {
      set_err(err, e.code, "request " + std::to_string(i) + ": " + e.message);
      return RIGA_EINVAL;
    }
  }
  // CRN grouping (records keep one scenario per item so path order is per request)
  std::vector<Item>& items = pl->items;
  {
    std::map<std::string, std::vector<int>> groups;
    std::vector<std::string> order;
    for (int i = 0; i < n; ++i) {
      std::string key = preps[i].key;
      if (records) put_bytes(key, i);
      auto itg = groups.find(key);
      if (itg == groups.end()) {
        order.push_back(key);
        groups[key] = {i};
      } else {
        itg->second.push_back(i);
      }
    }
    for (const std::string& k : order) {
      const std::vector<int>& m = groups[k];
      for (size_t s = 0; s < m.size(); s += 10) {
        Item it;
        it.members.assign(m.begin() + s, m.begin() + std::min(m.size(), s + 10));
        items.push_back(std::move(it));
      }
    }
  }
  for (Item& it : items) {
    Error e;
    if (!build_item(preps, &it, precision, records, &e)) {
      set_err(err, e.code, e.message);
      return RIGA_EINVAL;
    }
  }
  std::vector<uint64_t> ids;
  if (replay && replay->ids) {  // explain.rs:319-334: listed paths of the one request
    if (items.size() != 1 || !records) {
      set_err(err, "invalid_argument", "a path list replays one request");
      return RIGA_EINVAL;
    }
    ItemDesc& d = items[0].d;
    ids.assign(replay->ids, replay->ids + replay->n_ids);
    uint64_t top = 0;
    for (uint64_t id : ids) {
      if (id >= d.paths) {
        set_err(err, "path_out_of_range", "path_id " + std::to_string(id) +
                                              " is outside the configured path count " + std::to_string(d.paths));
        return RIGA_EINVAL;
      }
      top = std::max(top, id);
    }
    d.paths = ids.size();
    d.nbits = 64 - __builtin_clzll(top + 1);
    items[0].span = top + 1;
    items[0].cost = (double)d.steps * d.F * (double)d.paths;
  }
  pl->ledger = replay && replay->ledger;
  if (pl->ledger && (items.size() != 1 || !records)) {
    set_err(err, "invalid_argument", "coupon strips are written for one replayed request");
    return RIGA_EINVAL;
  }
  for (Item& it : items) pl->path_steps += (double)it.d.S * (double)it.d.paths * (double)it.d.steps;
  // shared normal tables: same-seed Sobol trades without wrapped direction rows
  // read their normals from one [dim][path] table when they need many more
  // normals than it holds (an uneven book), instead of hashing them per trade.
  // Decided here, since the items reading a table launch as their own kernel mode.
  const bool full_ladder = precision == RIGA_FP32 && cfg && cfg->fp32_full_ladder != 0;
  const int zmode = scratch || full_ladder ? -1 : (cfg ? cfg->normal_table : 0);
  if (full_ladder)
    for (Item& it : items)
      if (!(it.d.flags & (kHalton | kMonteCarlo))) it.d.flags |= kFullLadder;
  std::map<uint64_t, std::vector<int>> table_items;  // seed -> items reading its table
  if (zmode >= 0 && !records) {
    std::map<uint64_t, std::vector<int>> by_seed;
    for (size_t i = 0; i < items.size(); ++i) {
      const ItemDesc& d = items[i].d;
      if (!(d.flags & (kHalton | kMonteCarlo)) && d.ladder_dims >= d.F * d.steps && d.path_begin == 0)
        by_seed[d.seed].push_back((int)i);
    }
    const size_t real = precision == RIGA_FP64 ? 8 : 4;
    for (auto& kv : by_seed) {
      int dims = 0;
      uint64_t paths = 0;
      double demand = 0.0;
      for (int i : kv.second) {
        const ItemDesc& d = items[i].d;
        dims = std::max(dims, d.F * d.steps);
        paths = std::max<uint64_t>(paths, d.paths);
        demand += (double)d.paths * d.F * d.steps;
      }
      const double size = (double)dims * (double)paths;
      if (size * real > 32e9 || (zmode == 0 && demand < 4.0 * size)) continue;
      table_items[kv.first] = kv.second;
    }
  }
  std::vector<int> item_mode(items.size(), kLadderDirect);
  for (size_t i = 0; i < items.size(); ++i)
    if (items[i].d.flags & kMonteCarlo) item_mode[i] = kLadderMonteCarlo;
  for (auto& kv : table_items)
    for (int i : kv.second) item_mode[i] = kLadderTable;
  // launches: one per (factors, scenario bucket, ladder mode), blocks longest-first
  const int real_size = precision == RIGA_FP64 ? 8 : 4;
  std::map<PartKey, std::vector<int>> parts;
  for (size_t i = 0; i < items.size(); ++i)
    parts[PartKey(items[i].F, items[i].smax, item_mode[i])].push_back((int)i);
  pl->signature = shape_signature(precision, parts, items) + (full_ladder ? "|full" : "");
  if (ppt == 0 || waves == 0) {  // the autotuned shape for this signature, else 1 path x 4 waves
    int tp = 1, tw = 4;
    {
      std::lock_guard<std::mutex> g(ctx->tune_mu);
      auto f = ctx->tuned.find(pl->signature);
      if (f != ctx->tuned.end()) {
        tp = f->second.first;
        tw = f->second.second;
      }
    }
    if (ppt == 0) ppt = tp;
    if (waves == 0) waves = tw;
  }
  pl->ppt = ppt;
  pl->waves = waves;
  const int threads = 64 * waves;
  Arena arena;
  int sobol_rows = 1;
  for (Item& it : items)
    if (!(it.d.flags & (kHalton | kMonteCarlo))) sobol_rows = std::max(sobol_rows, it.d.ladder_rows);
  const std::vector<uint32_t>& dirs = sobol_directions(sobol_rows);
  const int64_t off_dirs = arena.add(dirs.data(), (size_t)sobol_rows * 32);
  bool any_mc = false;
  for (Item& it : items) any_mc = any_mc || (it.d.flags & kMonteCarlo) != 0;
  const int64_t off_zig = any_mc ? arena.add(ziggurat_tables().data(), 2 * kZigLen) : 0;
  // one Owen coordinate table per seed (built on the device at creation), covering
  // the largest raw dimension used -- for the seeds owen_table_seeds picks
  const int owen_mode = scratch ? -1 : (cfg ? cfg->owen_table : 0);
  const bool owen_table = owen_mode >= 0;
  std::map<uint64_t, int> owen_dims;
  if (owen_table) {
    std::map<uint64_t, double> coords;
    for (Item& it : items)
      if (!(it.d.flags & (kHalton | kMonteCarlo))) {
        int& m = owen_dims[it.d.seed];
        m = std::max(m, it.d.F * it.d.steps);
        coords[it.d.seed] += (double)it.d.paths * (double)it.d.F * (double)it.d.steps;
      }
    owen_table_seeds(ctx, owen_mode, coords, &owen_dims);
  }
  pl->final_base.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) {
    pl->final_base[i] = pl->final_slots;
    pl->final_slots += items[i].d.S;
  }
  std::vector<int32_t> all_blocks;
  for (auto& kv : parts) {
    std::vector<int>& idx = kv.second;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return items[a].cost > items[b].cost; });
    LaunchPlan lp;
    lp.F = std::get<0>(kv.first);
    lp.smax = std::get<1>(kv.first);
    lp.mode = std::get<2>(kv.first);
    lp.items = idx;
    int stack_levels = 0, lv_words = 0;
    for (int i : idx) {
      stack_levels = std::max(stack_levels, items[i].stack_need);
      int max_nx = 0;
      for (const LvGrid* g : items[i].surfaces) max_nx = std::max(max_nx, g->n_x());
      lv_words = std::max(lv_words, (int)items[i].surfaces.size() * (real_size == 4 ? 4 : 3) * max_nx);
    }
    lp.threads = threads;
    if (lp.mode == kLadderMonteCarlo) {  // stream ring and ziggurat tables instead of the bridge stack
      lp.lds = (size_t)mc_ring_slots(lp.smax) * threads * 8 + 2 * kZigLen * 8 + (size_t)lv_words * real_size;
    } else {
      // the allocation-free call keeps the whole stack in LDS (no slot pool)
      lp.pooled = stack_pooled_levels(precision, lp, stack_levels, (size_t)lv_words * real_size, real_size,
                                      scratch ? -1 : (cfg ? cfg->stack_split : 0));
      lp.lds = (size_t)(stack_levels - lp.pooled) * lp.F * threads * real_size + (size_t)lv_words * real_size;
    }
    if (lp.lds + kStaticLdsBound > ctx->lds_limit) {
      set_err(err, "invalid_local_vol",
              "the batch needs " + std::to_string(lp.lds) + " bytes of LDS per workgroup (limit " +
                  std::to_string(ctx->lds_limit) + "); use fewer distinct surfaces or knots");
      return RIGA_EINVAL;
    }
    pl->lds_max = std::max(pl->lds_max, lp.lds);
    int blocks = 0;
    for (int i : idx) {
      Item& it = items[i];
      ItemDesc& d = it.d;
      d.stack_levels = lp.mode == kLadderMonteCarlo ? 0 : stack_levels - lp.pooled;
      // the pooled positions are the least used: a stack position's pushes grow
      // binomially towards the middle (position k is the k-th left turn on the
      // way to a leaf), so the pool takes the top and the bottom in turn
      d.stack_pooled = lp.pooled;
      d.stack_pooled_low = lp.pooled / 2;
      d.paths_per_block = threads * ppt;
      d.n_blocks = (int32_t)((d.paths + d.paths_per_block - 1) / d.paths_per_block);
      blocks += d.n_blocks;
      d.partial_base = pl->partial_slots;
      pl->partial_slots += (int64_t)d.S * d.n_blocks;
      d.record_base = pl->record_slots;
      if (d.flags & kRecords) pl->record_slots += (int64_t)d.S * (int64_t)d.paths;
      d.off_dirs = off_dirs;
      if (owen_dims.count(d.seed) && !(d.flags & (kHalton | kMonteCarlo)))
        d.flags |= kOwenTable;  // table address set below
      d.off_zig = off_zig;
      d.off_nodes = arena.add(it.nodes.data(), it.nodes.size());
      d.off_steps = arena.add(it.steps.data(), it.steps.size());
      d.off_lvb = arena.add(it.lvb.data(), it.lvb.size());
      d.off_scen = arena.add(it.scen.data(), it.scen.size());
      d.off_halton = arena.add(it.halton.data(), it.halton.size());
      std::vector<LvSurface> su(it.surfaces.size());
      int max_nx = 0;
      bool uniform = real_size == 4 && !it.surfaces.empty();
      for (size_t u = 0; u < it.surfaces.size(); ++u) {
        const LvGrid& g = *it.surfaces[u];
        if (const riga_surface* rs = it.surface_res[u]) {  // resident: no copy in the arena
          su[u].off_t = su[u].off_x = su[u].off_v = -1;
          su[u].dev_x = (const double*)rs->dev;
          su[u].dev_v = (const double*)rs->dev + g.n_x();
        } else {
          su[u].off_t = arena.add(g.t_axis.data(), g.t_axis.size());
          su[u].off_x = arena.add(g.log_s.data(), g.log_s.size());
          su[u].off_v = arena.add(g.v.data(), g.v.size());
          su[u].dev_x = su[u].dev_v = nullptr;
        }
        su[u].n_t = g.n_t();
        su[u].n_x = g.n_x();
        su[u].factor = it.surface_factor[u];
        su[u].inv_dx = (double)(g.n_x() - 1) / (g.log_s[g.n_x() - 1] - g.log_s[0]);
        su[u].x0 = g.log_s[0];
        su[u].rel_off = (d.ln_s0[su[u].factor] - g.log_s[0]) * su[u].inv_dx;
        uniform = uniform && evenly_spaced(g.log_s);
        max_nx = std::max(max_nx, g.n_x());
      }
      if (uniform) d.flags |= kLvUniform;
      // FP32 tier Euler constants per (scenario, factor)
      std::vector<SfF32> sf((size_t)d.S * d.F);
      for (int sc = 0; sc < d.S; ++sc) {
        const riga_request& q = *preps[it.members[sc]].r;
        for (int f = 0; f < d.F; ++f) {
          SfF32& c = sf[(size_t)sc * d.F + f];
          memset(&c, 0, sizeof c);
          const double rq = q.risk_free_rate - q.dividend_yields[f];
          const int u = it.scen[sc].lv[f];
          if (u >= 0) {
            const LvSurface& g = su[u];
            c.a = (float)g.inv_dx;
            c.b = (float)g.rel_off;
            c.c = (float)(rq * d.dt);
            c.tmax = std::nextafter((float)(g.n_x - 1), 0.0f);
            c.lds_off = u * 4 * max_nx * 4;
          } else {
            const double vol = q.vols[f];
            c.a = (float)vol;
            c.c = (float)((rq - 0.5 * vol * vol) * d.dt);
          }
        }
      }
      d.off_sf32 = arena.add(sf.data(), sf.size());
      d.off_wt32 = arena.add(it.wt32.data(), it.wt32.size());
      d.off_surf = arena.add(su.data(), su.size());
      d.n_surf = (int32_t)su.size();
      d.max_nx = max_nx;
    }
    lp.blocks = blocks;
    lp.cost = items[idx.front()].cost;
    if (lp.pooled > 0) {  // a slot per workgroup that can be resident at once, and spares
      PhoenixLaunch L{};
      L.F = lp.F;
      L.smax = lp.smax;
      L.mode = lp.mode;
      L.threads = lp.threads;
      L.lds = lp.lds;
      int per_cu = 0;
      const hipError_t e = precision == RIGA_FP64 ? phoenix_blocks_per_cu_f64(L, &per_cu)
                                                  : phoenix_blocks_per_cu_f32(L, &per_cu);
      if (e != hipSuccess || per_cu <= 0) {
        set_err(err, "gpu_error", "occupancy of the path kernel could not be found");
        return RIGA_EGPU;
      }
      // plus one spare per CU: a workgroup the dispatcher starts as another ends
      // never waits on the ending one's hand-back
      lp.pool_n = std::min(blocks, (per_cu + 1) * ctx->n_cus);
      lp.pool_off = pl->pool_bytes;
      pl->pool_bytes += (size_t)lp.pool_n * lp.pooled * lp.F * lp.threads * real_size;
      lp.flags_off = pl->pool_nwords;
      pl->pool_nwords += (size_t)lp.pool_n + 1;
    }
    pl->blocks += blocks;
    pl->launches.push_back(lp);
  }
  // the launch holding the longest item is dispatched first (with the items of
  // each launch longest first, the device's work is claimed longest first)
  std::stable_sort(pl->launches.begin(), pl->launches.end(),
                   [](const LaunchPlan& a, const LaunchPlan& b) { return a.cost > b.cost; });
  pl->off_dirs = off_dirs;
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  if (scratch) {
    pl->own_stream = false;
    pl->stream = ctx->noalloc_stream;
    pl->e0 = ctx->noalloc_e0;
    pl->e1 = ctx->noalloc_e1;
  } else if (!estimate) {
    HIP_TRY(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "stream");
  }
  const bool stream_ordered = cfg && cfg->stream_ordered_alloc != 0 && !scratch;
  if (stream_ordered) pl->bind_stream();
  double ztable_bytes = 0.0;
  std::map<uint64_t, const OwenEntry*> owen_ptr;
  std::map<uint64_t, int> owen_cover;  // dimensions each seed's table covers
  if (!estimate)
    for (auto& kv : owen_dims) {
      std::shared_ptr<OwenTable> ot;
      HIP_TRY(owen_table_for(ctx, kv.first, kv.second, pl->stream, &ot), "owen table");
      HIP_TRY(hipStreamWaitEvent(pl->stream, ot->ready, 0), "owen table wait");
      owen_ptr[kv.first] = (const OwenEntry*)ot->buf.p;
      owen_cover[kv.first] = ot->dims;
      pl->owen_tabs.push_back(std::move(ot));
    }
  // a path kernel's item reads the table only if it covers all the item's
  // dimensions (the decision stays out of the step loop); the others hash
  for (Item& it : items)
    if (it.d.flags & kOwenTable) {
      const auto c = owen_cover.find(it.d.seed);
      // it holds the item's points and dimensions, and the item's rows do not wrap
      if (c != owen_cover.end() && c->second >= it.d.F * it.d.steps && it.d.nbits <= kOwenPrefixBits &&
          it.d.F * it.d.steps <= it.d.ladder_dims) {
        it.d.owen_tab = owen_ptr[it.d.seed];
        it.d.owen_dims = c->second;
      } else {
        it.d.flags &= ~kOwenTable;
      }
    }
  // the shared normal tables decided above
  {
    const size_t real = precision == RIGA_FP64 ? 8 : 4;
    for (auto& kv : table_items) {
      int dims = 0;
      uint64_t paths = 0;
      for (int i : kv.second) {
        dims = std::max(dims, items[i].d.F * items[i].d.steps);
        paths = std::max<uint64_t>(paths, items[i].d.paths);
      }
      const double bytes = (double)dims * (double)paths * real;
      std::unique_ptr<riga_plan::NormalTable> zt(new riga_plan::NormalTable());
      zt->seed = kv.first;
      zt->dims = dims;
      zt->paths = (int64_t)paths;
      zt->nbits = 64 - __builtin_clzll(paths);
      zt->owen = owen_table ? owen_ptr[kv.first] : nullptr;
      if (zt->nbits > kOwenPrefixBits) zt->owen = nullptr;  // its points pass the table's
      zt->owen_dims = zt->owen ? owen_cover[kv.first] : 0;
      if (stream_ordered) zt->buf.stream = pl->stream;
      ztable_bytes += bytes;
      if (!estimate) HIP_TRY(zt->buf.ensure((size_t)bytes), "normal table alloc");
      for (int i : kv.second) {
        items[i].d.flags |= kNormalTable;
        items[i].d.ztable = (uint64_t)(uintptr_t)zt->buf.p;
        items[i].d.zt_stride = zt->paths;
      }
      pl->ztables.push_back(std::move(zt));
    }
  }
  // Monte Carlo lane: one stream parse per (seed, draws per path)
  double mc_bytes = 0.0;
  {
    std::map<std::pair<uint64_t, uint64_t>, std::vector<int>> by_stream;
    for (size_t i = 0; i < items.size(); ++i)
      if (items[i].d.flags & kMonteCarlo)
        by_stream[{items[i].d.seed, (uint64_t)items[i].d.F * (uint64_t)items[i].d.steps}].push_back((int)i);
    for (auto& kv : by_stream) {
      std::unique_ptr<riga_plan::McParse> mp(new riga_plan::McParse());
      mp->seed = kv.first.first;
      mp->dpp = kv.first.second;
      for (int i : kv.second) mp->paths = std::max<uint64_t>(mp->paths, items[i].span);
      mc_parse_size(mp.get());
      uint32_t key[8];
      chacha_key_from_seed(mp->seed, key);
      mp->off_key = arena.add(key, 8);
      mc_bytes += mc_parse_bytes(*mp);
      if (stream_ordered) mp->each([](DevBuf& b, hipStream_t st) { b.stream = st; }, pl->stream);
      if (scratch)
        for (DevBuf* b : {&mp->chunks, &mp->slow, &mp->slow_val, &mp->entry, &mp->cnt, &mp->block_off, &mp->starts})
          b->from = scratch;
      if (!estimate) HIP_TRY(mc_parse_alloc(mp.get()), "stream parse alloc");
      for (int i : kv.second) {
        ItemDesc& d = items[i].d;
        d.mc_start = (const uint64_t*)mp->starts.p;
        d.mc_chunks = (const McChunk*)mp->chunks.p;
        d.mc_slow = (const uint32_t*)mp->slow.p;
        d.mc_val = (const double*)mp->slow_val.p;
        d.mc_n_chunks = mp->n_chunks;
        d.mc_chunk_words = mp->chunk_words;
        d.mc_slow_cap = mp->slow_cap;
      }
      pl->mc.push_back(std::move(mp));
    }
    if (!pl->mc.empty() && !estimate) HIP_TRY(pl->mc_status.ensure(8 * pl->mc.size()), "stream status alloc");
  }
  pl->off_zig = off_zig;
  // block table: (item, chunk) pairs per launch.  Items reading a shared normal
  // table go chunk-major (every item's chunk c before any item's chunk c+1), so
  // the blocks on the device at once read the same path columns of the table
  // and the rows stay in cache; other launches go item by item, longest first.
  for (LaunchPlan& lp : pl->launches) {
    lp.block_off = all_blocks.size() / 2;
    bool shared_table = false;
    int max_chunks = 0;
    for (int i : lp.items) {
      shared_table = shared_table || (items[i].d.flags & kNormalTable) != 0;
      max_chunks = std::max(max_chunks, items[i].d.n_blocks);
    }
    if (shared_table) {
      for (int c = 0; c < max_chunks; ++c)
        for (int i : lp.items)
          if (c < items[i].d.n_blocks) {
            all_blocks.push_back(i);
            all_blocks.push_back(c);
          }
    } else {
      for (int i : lp.items)
        for (int c = 0; c < items[i].d.n_blocks; ++c) {
          all_blocks.push_back(i);
          all_blocks.push_back(c);
        }
    }
  }
  std::vector<ItemDesc> descs(items.size());
  for (size_t i = 0; i < items.size(); ++i) descs[i] = items[i].d;
  const double device_bytes = ztable_bytes + mc_bytes + (double)arena.bytes.size() +
                              descs.size() * sizeof(ItemDesc) + all_blocks.size() * 4.0 +
                              (double)pl->partial_slots * sizeof(Partial) + (double)pl->final_slots * sizeof(Partial) +
                              pl->final_base.size() * 8.0 + (double)pl->record_slots * sizeof(PathRecordDev) +
                              (double)pl->pool_bytes + 4.0 * pl->pool_nwords + 8 * 4096.0;
  if (estimate) {
    *estimate = (int64_t)device_bytes;
    for (auto& zt : pl->ztables) zt->buf.stream = nullptr;
    for (auto& mp : pl->mc) mp->each([](DevBuf& b, hipStream_t) { b.stream = nullptr; }, nullptr);
    return 0;
  }
  if (scratch) {
    // every piece is at least 4 KiB and 256-byte aligned: 20 pieces at most
    const double need = device_bytes + 20 * (4096.0 + 256.0);
    const double host_need = (double)pl->final_slots * sizeof(Partial) + 8.0 * pl->mc.size();
    if (need > (double)(scratch->cap - scratch->used) || host_need > (double)kNoallocHostBytes) {
      char buf[200];
      snprintf(buf, sizeof buf, "the request needs %.0f bytes of device work area; the allocation-free call has %zu",
               need, scratch->cap - scratch->used);
      set_err(err, "scratch_exceeded", buf);
      return RIGA_EINVAL;
    }
  }

  if (!scratch) {
    HIP_TRY(hipEventCreate(&pl->e0), "event");
    HIP_TRY(hipEventCreate(&pl->e1), "event");
  }
  HIP_TRY(pl->arena.ensure(arena.bytes.size()), "arena alloc");
  HIP_TRY(pl->items_d.ensure(descs.size() * sizeof(ItemDesc)), "items alloc");
  HIP_TRY(pl->block_item.ensure(all_blocks.size() * sizeof(int32_t)), "block table alloc");
  HIP_TRY(pl->partials.ensure((size_t)pl->partial_slots * sizeof(Partial)), "partials alloc");
  HIP_TRY(pl->finals.ensure((size_t)pl->final_slots * sizeof(Partial)), "finals alloc");
  HIP_TRY(pl->final_base_d.ensure(pl->final_base.size() * sizeof(int64_t)), "final index alloc");
  HIP_TRY(pl->queues.ensure(pl->launches.size() * kQueueStride * sizeof(uint32_t)), "unit queue alloc");
  if (pl->pool_nwords > 0) {
    HIP_TRY(pl->stack_pool.ensure(pl->pool_bytes), "bridge stack pool alloc");
    HIP_TRY(pl->pool_words.ensure(pl->pool_nwords * sizeof(uint32_t)), "stack pool flags alloc");
    // every slot free; each workgroup hands its slot back when it ends
    HIP_TRY(hipMemsetAsync(pl->pool_words.p, 0, pl->pool_nwords * sizeof(uint32_t), pl->stream), "pool flags");
    for (const LaunchPlan& lp : pl->launches)
      if (lp.pooled > 0)
        for (int i : lp.items) {
          descs[i].stack_pool = (unsigned char*)pl->stack_pool.p + lp.pool_off;
          descs[i].pool_flags = (uint32_t*)pl->pool_words.p + lp.flags_off;
          descs[i].pool_n = lp.pool_n;
        }
  }
  // Each workgroup claims its unit from the launch's counter.  By default the grid
  // is one workgroup per unit and the hardware dispatcher starts each in a slot as
  // one frees up: the last units spread over idle CUs, and a plan launched while
  // another runs gets slots as the running one's workgroups end.  Persistent
  // launches hold as many workgroups as the device holds at once (at most one per
  // unit), each claiming units until none is left.
  const bool persistent = cfg && cfg->persistent != 0;
  for (LaunchPlan& lp : pl->launches) {
    lp.grid = lp.blocks;
    if (!persistent) continue;
    PhoenixLaunch L{};
    L.F = lp.F;
    L.smax = lp.smax;
    L.mode = lp.mode;
    L.threads = lp.threads;
    L.lds = lp.lds;
    int per_cu = 0;
    const hipError_t e = precision == RIGA_FP64 ? phoenix_blocks_per_cu_f64(L, &per_cu)
                                                : phoenix_blocks_per_cu_f32(L, &per_cu);
    if (e != hipSuccess || per_cu <= 0) {
      set_err(err, "gpu_error", "occupancy of the path kernel could not be found");
      return RIGA_EGPU;
    }
    lp.grid = std::min(lp.blocks, per_cu * ctx->n_cus);
  }
  pl->serial = cfg && cfg->serial_launches != 0;
  if (pl->launches.size() > 1 && !pl->serial) {
    pl->fork_join.assign(pl->launches.size(), nullptr);
    for (hipEvent_t& ev : pl->fork_join) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
    pl->aux.assign(pl->launches.size() - 1, nullptr);
    for (hipStream_t& st : pl->aux) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
  }
  HIP_TRY(pl->finals_host.ensure((size_t)pl->final_slots * sizeof(Partial) + 8 * pl->mc.size() +
                                     4 * pl->launches.size()),
          "host finals alloc");
  if (records) {
    HIP_TRY(pl->records_d.ensure((size_t)pl->record_slots * sizeof(PathRecordDev)), "records alloc");
    pl->recs_host.resize((size_t)pl->record_slots);
  }
  if (!ids.empty()) {
    HIP_TRY(pl->path_list.ensure(ids.size() * 8), "path list alloc");
    HIP_TRY(hipMemcpyAsync(pl->path_list.p, ids.data(), ids.size() * 8, hipMemcpyHostToDevice, pl->stream),
            "path list upload");
    descs[0].path_list = (const uint64_t*)pl->path_list.p;
  }
  if (pl->ledger) {
    const ItemDesc& d = descs[0];
    HIP_TRY(pl->series.ensure((size_t)d.paths * (size_t)d.steps * 8), "series alloc");
    HIP_TRY(pl->ledger_d.ensure((size_t)d.paths * (size_t)d.periods * sizeof(CouponRowDev)), "ledger alloc");
    pl->ledger_host.resize((size_t)d.paths * (size_t)d.periods);
    descs[0].series = (double*)pl->series.p;
  }
  HIP_TRY(hipMemcpyAsync(pl->arena.p, arena.bytes.data(), arena.bytes.size(), hipMemcpyHostToDevice, pl->stream),
          "arena upload");
  HIP_TRY(hipMemcpyAsync(pl->items_d.p, descs.data(), descs.size() * sizeof(ItemDesc), hipMemcpyHostToDevice,
                         pl->stream),
          "item upload");
  HIP_TRY(hipMemcpyAsync(pl->block_item.p, all_blocks.data(), all_blocks.size() * 4, hipMemcpyHostToDevice,
                         pl->stream),
          "block table upload");
  HIP_TRY(hipMemcpyAsync(pl->final_base_d.p, pl->final_base.data(), pl->final_base.size() * 8,
                         hipMemcpyHostToDevice, pl->stream),
          "final index upload");
  HIP_TRY(hipStreamSynchronize(pl->stream), "plan upload");
  if (pl->graph) {
    HIP_TRY(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    hipError_t e = enqueue(pl.get(), false);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(pl->stream, &g);
    HIP_TRY(e, "graph capture enqueue");
    HIP_TRY(e2, "graph capture end");
    pl->g = g;
    HIP_TRY(hipGraphInstantiate(&pl->gexec, g, nullptr, nullptr, 0), "graph instantiate");
  }
  *out = pl.release();
  return 0;
}

int plan_launch(riga_plan* pl, riga_error* err) {
  clear_err(err);
  if (!pl) {
    set_err(err, "invalid_argument", "plan is NULL");
    return RIGA_EINVAL;
  }
  HIP_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  if (pl->graph) {
    HIP_TRY(hipEventRecord(pl->e0, pl->stream), "event");
    HIP_TRY(hipGraphLaunch(pl->gexec, pl->stream), "graph launch");
    HIP_TRY(hipEventRecord(pl->e1, pl->stream), "event");
  } else {
    HIP_TRY(enqueue(pl, true), "phoenix launch");
  }
  if (pl->records)
    HIP_TRY(hipMemcpyAsync(pl->recs_host.data(), pl->records_d.p, pl->recs_host.size() * sizeof(PathRecordDev),
                           hipMemcpyDeviceToHost, pl->stream),
            "records download");
  if (pl->ledger)
    HIP_TRY(hipMemcpyAsync(pl->ledger_host.data(), pl->ledger_d.p, pl->ledger_host.size() * sizeof(CouponRowDev),
                           hipMemcpyDeviceToHost, pl->stream),
            "ledger download");
  return 0;
}

int plan_wait(riga_plan* pl, riga_result* results, riga_path_record* records, riga_error* err,
              riga_coupon_row* ledger = nullptr) {
  clear_err(err);
  HIP_TRY(hipStreamSynchronize(pl->stream), "kernel execution");
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, pl->e0, pl->e1), "event timing");
  const Partial* finals = (const Partial*)pl->finals_host.p;
  const uint32_t* mc_status = (const uint32_t*)(finals + pl->final_slots);
  for (size_t k = 0; k < pl->mc.size(); ++k)
    if (mc_status[2 * k] != pl->mc[k]->paths || mc_status[2 * k + 1] != 0) {
      set_err(err, "mc_stream_error",
              "Monte Carlo stream parse: " + std::to_string(mc_status[2 * k]) + " of " +
                  std::to_string(pl->mc[k]->paths) + " path starts, flags " + std::to_string(mc_status[2 * k + 1]));
      return RIGA_EGPU;
    }
  const uint32_t* pool_failed = mc_status + 2 * pl->mc.size();
  for (size_t k = 0; k < pl->launches.size(); ++k)
    if (pl->launches[k].pooled > 0 && pool_failed[k] != 0) {
      set_err(err, "gpu_error", "bridge stack pool: " + std::to_string(pool_failed[k]) +
                                    " workgroups found no free slot (results not valid)");
      return RIGA_EGPU;
    }
  for (size_t i = 0; i < pl->items.size(); ++i) {
    const Item& it = pl->items[i];
    for (int s = 0; s < it.d.S; ++s) {
      const int req = it.members[s];
      if (results) {
        riga_result* out = &results[req];
        memset(out, 0, sizeof *out);
        fill_result(finals[pl->final_base[i] + s], it.d.paths, out);
        out->device_ms = ms;
      }
      if (records && pl->records) {
        const PathRecordDev* src = pl->recs_host.data() + it.d.record_base + (int64_t)s * (int64_t)it.d.paths;
        for (uint64_t p = 0; p < it.d.paths; ++p) {
          riga_path_record& r = records[p];
          r.ko_step = src[p].ko_step;
          r.hit_ki = src[p].hit_ki;
          r.final_worst = src[p].final_worst;
          r.funding = src[p].funding;
          r.put = src[p].put;
          r.coupon_pv = src[p].coupon_pv;
          r.memory_pv = src[p].memory_pv;
          r.pv = src[p].pv;
        }
      }
    }
  }
  if (ledger && pl->ledger) {
    static_assert(sizeof(CouponRowDev) == sizeof(riga_coupon_row), "coupon row layout");
    memcpy(ledger, pl->ledger_host.data(), pl->ledger_host.size() * sizeof(CouponRowDev));
  }
  return 0;
}

int run_batch(riga_ctx* ctx, const riga_request* reqs, int n, int precision, const riga_launch_config* cfg,
              riga_result* results, riga_path_record* records, riga_error* err, const ReplayOpts* replay = nullptr,
              riga_coupon_row* ledger = nullptr) {
  riga_plan* pl = nullptr;
  int rc = plan_create(ctx, reqs, n, precision, cfg, records != nullptr, &pl, err, nullptr, replay);
  if (rc) return rc;
  rc = plan_launch(pl, err);
  if (!rc) rc = plan_wait(pl, results, records, err, ledger);
  delete pl;
  return rc;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int riga_hip_init(int device, riga_ctx** out, riga_error* err) {
  clear_err(err);
  if (!out) {
    set_err(err, "invalid_argument", "out is NULL");
    return RIGA_EINVAL;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    set_err(err, "gpu_unavailable", std::string("no HIP device: ") + hipGetErrorString(e));
    return RIGA_ENODEV;
  }
  if (device < 0 || device >= count) {
    set_err(err, "gpu_unavailable", "device index out of range");
    return RIGA_ENODEV;
  }
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    set_err(err, "gpu_unavailable", "cannot select the HIP device");
    return RIGA_ENODEV;
  }
  std::string arch = prop.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0) {
    set_err(err, "gpu_unavailable", "kernels are built for gfx950; device is " + arch);
    return RIGA_ENODEV;
  }
  std::unique_ptr<riga_ctx> ctx(new riga_ctx());
  ctx->device = device;
  ctx->name = prop.name;
  if (ctx->name.empty()) {
    char buf[256] = {0};
    if (hipDeviceGetName(buf, sizeof buf, device) == hipSuccess) ctx->name = buf;
  }
  if (ctx->name.empty()) ctx->name = "AMD Instinct (" + arch + ")";
  ctx->arch = arch;
  ctx->lds_limit = prop.sharedMemPerBlock;
  ctx->n_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (prop.totalGlobalMem > 0) {  // Owen table budgets scale with the device's memory
    ctx->owen_cap = (size_t)prop.totalGlobalMem / 3;
    ctx->owen_plan_cap = (size_t)prop.totalGlobalMem / 4;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    set_err(err, "gpu_error", "stream/event creation failed");
    return RIGA_EGPU;
  }
  // the allocation-free call's set-up, once per context with its device current:
  // this device's address of the scratch array, a stream, two events and the
  // pinned block its merged records come back through
  e = scratch_address(&ctx->scratch);
  if (e != hipSuccess || !ctx->scratch) {
    set_err(err, "gpu_error", std::string("scratch array address: ") + hipGetErrorString(e));
    return RIGA_EGPU;
  }
  if (hipStreamCreateWithFlags(&ctx->noalloc_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->noalloc_e0) != hipSuccess || hipEventCreate(&ctx->noalloc_e1) != hipSuccess ||
      hipHostMalloc(&ctx->noalloc_host, kNoallocHostBytes, hipHostMallocDefault) != hipSuccess) {
    set_err(err, "gpu_error", "allocation-free call set-up failed");
    return RIGA_EGPU;
  }
  // a path kernel that spills registers to private memory (the FP64 5-factor
  // 10-scenario kernels do) needs the runtime's private-memory set-up on its
  // first dispatch: dispatch each once now (no work; the unit counter is the
  // scratch array's first word), so that happens here and not inside an
  // allocation-free call
  e = warm_path_kernels_f64((uint32_t*)ctx->scratch, ctx->noalloc_stream);
  if (e == hipSuccess) e = warm_path_kernels_f32((uint32_t*)ctx->scratch, ctx->noalloc_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->noalloc_stream);
  if (e != hipSuccess) {
    set_err(err, "gpu_error", std::string("path kernel set-up: ") + hipGetErrorString(e));
    return RIGA_EGPU;
  }
  e = hipEventCreate(&ctx->epoch);
  if (e == hipSuccess) e = hipEventRecord(ctx->epoch, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ctx->epoch);
  if (e != hipSuccess) {
    set_err(err, "gpu_error", std::string("device clock origin: ") + hipGetErrorString(e));
    return RIGA_EGPU;
  }
  *out = ctx.release();
  return 0;
}

void riga_hip_destroy(riga_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->noalloc_stream) (void)hipStreamSynchronize(ctx->noalloc_stream);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->epoch) (void)hipEventDestroy(ctx->epoch);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->noalloc_e0) (void)hipEventDestroy(ctx->noalloc_e0);
  if (ctx->noalloc_e1) (void)hipEventDestroy(ctx->noalloc_e1);
  if (ctx->noalloc_stream) (void)hipStreamDestroy(ctx->noalloc_stream);
  if (ctx->noalloc_host) (void)hipHostFree(ctx->noalloc_host);
  delete ctx;
}

int riga_hip_device_name(riga_ctx* ctx, char* buf, size_t len, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  snprintf(buf, len, "%s", ctx->name.c_str());
  return 0;
}

int riga_hip_arch(riga_ctx* ctx, char* buf, size_t len, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  snprintf(buf, len, "%s", ctx->arch.c_str());
  return 0;
}

int riga_validate(const riga_request* req, riga_error* err) {
  clear_err(err);
  Prep p;
  Error e;
  if (!req || !prepare(*req, &p, &e)) {
    set_err(err, req ? e.code : "invalid_argument", req ? e.message : "request is NULL");
    return RIGA_EINVAL;
  }
  return 0;
}

int riga_plan_create(riga_ctx* ctx, const riga_request* reqs, int n, int precision,
                     const riga_launch_config* cfg, riga_plan** out, riga_error* err) {
  if (!out) {
    set_err(err, "invalid_argument", "out is NULL");
    return RIGA_EINVAL;
  }
  return plan_create(ctx, reqs, n, precision, cfg, false, out, err);
}

int riga_plan_estimate(riga_ctx* ctx, const riga_request* reqs, int n, int precision,
                       const riga_launch_config* cfg, int64_t* bytes, riga_error* err) {
  if (!bytes) {
    set_err(err, "invalid_argument", "bytes is NULL");
    return RIGA_EINVAL;
  }
  riga_plan* unused = nullptr;
  return plan_create(ctx, reqs, n, precision, cfg, false, &unused, err, bytes);
}

int riga_plan_launch(riga_plan* plan, riga_error* err) { return plan_launch(plan, err); }

int riga_plan_ready(riga_plan* plan) {
  if (!plan) return RIGA_EINVAL;
  const hipError_t e = hipStreamQuery(plan->stream);
  return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : RIGA_EGPU);
}

int riga_plan_wait(riga_plan* plan, riga_result* results, riga_error* err) {
  if (!plan) {
    set_err(err, "invalid_argument", "plan is NULL");
    return RIGA_EINVAL;
  }
  return plan_wait(plan, results, nullptr, err);
}

int riga_plan_run(riga_plan* plan, riga_result* results, riga_error* err) {
  int rc = riga_plan_launch(plan, err);
  return rc ? rc : riga_plan_wait(plan, results, err);
}

int riga_plan_device_times(riga_plan* plan, double* start_s, double* end_s) {
  if (!plan || !start_s || !end_s || !plan->e0 || !plan->e1 || !plan->ctx->epoch) return RIGA_EINVAL;
  if (hipEventQuery(plan->e1) != hipSuccess) return RIGA_EINVAL;  // not launched, or still running
  float a = 0.0f, b = 0.0f;
  if (hipEventElapsedTime(&a, plan->ctx->epoch, plan->e0) != hipSuccess ||
      hipEventElapsedTime(&b, plan->ctx->epoch, plan->e1) != hipSuccess)
    return RIGA_EGPU;
  *start_s = 1e-3 * (double)a;
  *end_s = 1e-3 * (double)b;
  return 0;
}

int riga_plan_query(const riga_plan* plan, riga_plan_info* info) {
  if (!plan || !info) return RIGA_EINVAL;
  memset(info, 0, sizeof *info);
  info->n_requests = plan->n_requests;
  info->n_items = (int32_t)plan->items.size();
  info->n_launches = (int32_t)plan->launches.size();
  info->blocks = plan->blocks;
  info->path_steps = plan->path_steps;
  info->lds_bytes = (int64_t)plan->lds_max;
  info->device_bytes = (int64_t)(plan->arena.cap + plan->items_d.cap + plan->block_item.cap + plan->partials.cap +
                                 plan->finals.cap + plan->final_base_d.cap + plan->records_d.cap);
  for (const auto& zt : plan->ztables) info->device_bytes += (int64_t)zt->buf.cap;
  info->normal_tables = (int32_t)plan->ztables.size();
  info->graph = plan->graph ? 1 : 0;
  info->paths_per_thread = plan->ppt;
  info->waves_per_block = plan->waves;
  info->owen_tables = (int32_t)plan->owen_tabs.size();
  return 0;
}

void riga_plan_destroy(riga_plan* plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  if (plan->stream) (void)hipStreamSynchronize(plan->stream);
  delete plan;
}

int riga_price_put_leg(riga_ctx* ctx, const riga_request* req, int precision, riga_result* out,
                       riga_error* err) {
  return run_batch(ctx, req, 1, precision, nullptr, out, nullptr, err);
}

int riga_price_put_leg_noalloc(riga_ctx* ctx, const riga_request* req, int precision, riga_result* out,
                               riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (!req || !out) {
    set_err(err, "invalid_argument", "request and result are required");
    return RIGA_EINVAL;
  }
  // calls through one context share its scratch array: one at a time
  std::lock_guard<std::mutex> g(ctx->noalloc_mu);
  Scratch sc;
  sc.base = (unsigned char*)ctx->scratch;
  sc.cap = kScratchBytes;
  riga_plan* pl = nullptr;
  int rc = plan_create(ctx, req, 1, precision, nullptr, false, &pl, err, nullptr, nullptr, &sc);
  if (rc) return rc;
  rc = plan_launch(pl, err);
  if (!rc) rc = plan_wait(pl, out, nullptr, err);
  if (rc) (void)hipStreamSynchronize(ctx->noalloc_stream);  // nothing of this call still runs
  delete pl;
  return rc;
}

int riga_price_batch(riga_ctx* ctx, const riga_request* reqs, int n, int precision,
                     const riga_launch_config* cfg, riga_result* results, riga_error* err) {
  return run_batch(ctx, reqs, n, precision, cfg, results, nullptr, err);
}

int riga_replay_paths(riga_ctx* ctx, const riga_request* req, int precision,
                      riga_path_record* records, riga_result* out, riga_error* err) {
  if (!records) {
    set_err(err, "invalid_argument", "records is NULL");
    return RIGA_EINVAL;
  }
  return run_batch(ctx, req, 1, precision, nullptr, out, records, err);
}

int riga_replay_ledger(riga_ctx* ctx, const riga_request* req, int precision, const uint64_t* path_ids,
                       uint64_t n_ids, riga_path_record* records, riga_coupon_row* ledger, riga_error* err) {
  clear_err(err);
  if (!records || !req) {
    set_err(err, "invalid_argument", "request and records are required");
    return RIGA_EINVAL;
  }
  if (path_ids && n_ids == 0) {
    set_err(err, "invalid_argument", "an empty path list");
    return RIGA_EINVAL;
  }
  ReplayOpts ro;
  ro.ids = path_ids;
  ro.n_ids = n_ids;
  ro.ledger = ledger != nullptr;
  return run_batch(ctx, req, 1, precision, nullptr, nullptr, records, err, &ro, ledger);
}

int riga_query_worst_percentile(riga_ctx* ctx, const riga_request* req, int precision, double percentile,
                                uint64_t max_sample, riga_worst_query* out, uint64_t* sample_ids,
                                riga_path_record* sample_records, riga_coupon_row* sample_ledger, riga_error* err) {
  clear_err(err);
  if (!req || !out) {
    set_err(err, "invalid_argument", "request and output are required");
    return RIGA_EINVAL;
  }
  if (!(percentile >= 0.0 && percentile <= 1.0)) {  // explain.rs:341-346
    set_err(err, "invalid_percentile", "percentile must be within [0, 1]");
    return RIGA_EINVAL;
  }
  // every path's record on the device, then the reference's stable ascending
  // sort by final worst-of and the first ceil(total * percentile) paths
  std::vector<riga_path_record> all((size_t)req->paths);
  int rc = run_batch(ctx, req, 1, precision, nullptr, nullptr, all.data(), err);
  if (rc) return rc;
  std::vector<uint64_t> order(all.size());
  std::iota(order.begin(), order.end(), (uint64_t)0);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint64_t a, uint64_t b) { return all[a].final_worst < all[b].final_worst; });
  const uint64_t total = all.size();
  const uint64_t selected = (uint64_t)std::ceil((double)total * percentile);
  out->percentile = percentile;
  out->total = total;
  out->selected = selected;
  out->threshold_worst_final = selected == 0 ? 0.0 : all[order[selected - 1]].final_worst;
  out->n_sample = std::min(selected, max_sample);
  if (out->n_sample == 0) return 0;
  for (uint64_t i = 0; i < out->n_sample; ++i)
    if (sample_ids) sample_ids[i] = order[i];
  if (!sample_records) return 0;
  // the sample again, with coupon strips
  ReplayOpts ro;
  ro.ids = order.data();
  ro.n_ids = out->n_sample;
  ro.ledger = sample_ledger != nullptr;
  return run_batch(ctx, req, 1, precision, nullptr, nullptr, sample_records, err, &ro, sample_ledger);
}

int riga_autotune(riga_ctx* ctx, const riga_request* reqs, int n, int precision, const riga_launch_config* cfg,
                  int reps, riga_tune_report* report, riga_error* err) {
  clear_err(err);
  if (!ctx || !reqs || n <= 0) {
    set_err(err, "invalid_argument", "context and at least one request are required");
    return RIGA_EINVAL;
  }
  reps = std::max(1, std::min(reps, 20));
  riga_launch_config base = cfg ? *cfg : riga_launch_config{};
  riga_tune_report rep{};
  std::vector<riga_result> res((size_t)n);
  std::string signature;
  static const int kPpt[4] = {1, 2, 4, 8};
  static const int kWv[4] = {4, 3, 2, 1};
  double best = 0.0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      riga_tune_trial& t = rep.trials[rep.n_trials++];
      t.paths_per_thread = kPpt[a];
      t.waves_per_block = kWv[b];
      t.device_ms = -1.0;
      riga_launch_config c = base;
      c.paths_per_thread = kPpt[a];
      c.waves_per_block = kWv[b];
      riga_plan* pl = nullptr;
      riga_error e{};
      int rc = plan_create(ctx, reqs, n, precision, &c, false, &pl, &e);
      if (rc != 0) {
        if (rc == RIGA_EINVAL && strstr(e.message, "LDS")) continue;  // shape does not fit
        if (err) *err = e;
        return rc;
      }
      signature = pl->signature;
      double ms = 1e300;
      for (int r = 0; r <= reps && rc == 0; ++r) {
        rc = riga_plan_run(pl, res.data(), &e);
        if (r > 0 && rc == 0) ms = std::min(ms, res[0].device_ms);
      }
      riga_plan_destroy(pl);
      if (rc != 0) {
        if (err) *err = e;
        return rc;
      }
      t.device_ms = ms;
      if (best == 0.0 || ms < best) {
        best = ms;
        rep.best = rep.n_trials - 1;
      }
    }
  if (best == 0.0) {
    set_err(err, "invalid_launch_config", "no launch shape fits this batch");
    return RIGA_EINVAL;
  }
  {
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    ctx->tuned[signature] = {rep.trials[rep.best].paths_per_thread, rep.trials[rep.best].waves_per_block};
  }
  if (report) *report = rep;
  return 0;
}

int riga_surface_create(riga_ctx* ctx, const riga_lv_grid* grid, riga_surface** out, riga_error* err) {
  if (!ctx || !grid || !out) {
    set_err(err, "invalid_argument", "context, grid and output are required");
    return RIGA_EINVAL;
  }
  if (grid->n_t < 1 || grid->n_x < 2 || grid->n_x > kMaxLvX || !grid->t_axis || !grid->log_s || !grid->v) {
    set_err(err, "invalid_local_vol",
            "calibrated local-vol grid must have 1+ tenors and 2.." + std::to_string(kMaxLvX) + " knots");
    return RIGA_EINVAL;
  }
  std::unique_ptr<riga_surface> s(new riga_surface());
  s->ctx = ctx;
  s->grid.t_axis.assign(grid->t_axis, grid->t_axis + grid->n_t);
  s->grid.log_s.assign(grid->log_s, grid->log_s + grid->n_x);
  s->grid.v.assign(grid->v, grid->v + (size_t)grid->n_t * grid->n_x);
  std::vector<double> packed(s->grid.log_s);
  packed.insert(packed.end(), s->grid.v.begin(), s->grid.v.end());
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(hipMalloc(&s->dev, packed.size() * sizeof(double)), "surface alloc");
  const hipError_t e = hipMemcpy(s->dev, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(s->dev);
    set_err(err, "gpu_error", std::string("surface upload: ") + hipGetErrorString(e));
    return RIGA_EGPU;
  }
  *out = s.release();
  return 0;
}

void riga_surface_destroy(riga_surface* s) {
  if (!s) return;
  if (s->dev) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipFree(s->dev);
  }
  delete s;
}

int riga_lv_calibrate(const riga_implied_grid* grid, double spot, double rate, int32_t grid_x,
                      double* t_axis, double* log_s, double* v, riga_error* err) {
  clear_err(err);
  if (!grid || !t_axis || !log_s || !v) {
    set_err(err, "invalid_argument", "NULL argument");
    return RIGA_EINVAL;
  }
  LvGrid g;
  Error e;
  if (!dupire_calibrate(*grid, spot, rate, grid_x, &g, &e)) {
    set_err(err, e.code, e.message);
    return RIGA_EINVAL;
  }
  memcpy(t_axis, g.t_axis.data(), g.t_axis.size() * sizeof(double));
  memcpy(log_s, g.log_s.data(), g.log_s.size() * sizeof(double));
  memcpy(v, g.v.data(), g.v.size() * sizeof(double));
  return 0;
}

double riga_lv_sigma(const riga_lv_grid* grid, double t, double s) {
  LvGrid g;
  g.t_axis.assign(grid->t_axis, grid->t_axis + grid->n_t);
  g.log_s.assign(grid->log_s, grid->log_s + grid->n_x);
  g.v.assign(grid->v, grid->v + (size_t)grid->n_t * grid->n_x);
  return lv_sigma(g, t, s);
}

int riga_debug_sobol(riga_ctx* ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point,
                     int32_t n_points, int32_t first_dim, int32_t n_dims, uint32_t* out_unscrambled,
                     uint32_t* out_scrambled, double* out_normal, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n_points <= 0 || n_dims <= 0 || first_dim < 0 || ladder_dims <= 0 ||
      first_point + (uint64_t)n_points >= (1ull << 32)) {
    set_err(err, "invalid_argument", "bad debug_sobol ranges");
    return RIGA_EINVAL;
  }
  const int rows = std::min<int64_t>(ladder_dims, (int64_t)first_dim + n_dims);
  const std::vector<uint32_t>& dirs = sobol_directions(rows);
  const uint64_t max_n = first_point + (uint64_t)n_points;
  const int nbits = 64 - __builtin_clzll(max_n);
  const size_t cells = (size_t)n_points * n_dims;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf d_dirs, d_u, d_s, d_n;
  HIP_TRY(d_dirs.ensure((size_t)rows * 32 * 4), "alloc");
  HIP_TRY(d_u.ensure(cells * 4), "alloc");
  HIP_TRY(d_s.ensure(cells * 4), "alloc");
  HIP_TRY(d_n.ensure(cells * 8), "alloc");
  HIP_TRY(hipMemcpyAsync(d_dirs.p, dirs.data(), (size_t)rows * 32 * 4, hipMemcpyHostToDevice, ctx->stream),
          "upload");
  HIP_TRY(launch_debug_sobol((const uint32_t*)d_dirs.p, ladder_dims, seed, first_point, n_points, first_dim,
                             n_dims, nbits, (uint32_t*)d_u.p, (uint32_t*)d_s.p, (double*)d_n.p, ctx->stream),
          "debug kernel");
  HIP_TRY(hipMemcpyAsync(out_unscrambled, d_u.p, cells * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_scrambled, d_s.p, cells * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_normal, d_n.p, cells * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "debug kernel execution");
  return 0;
}

int riga_debug_exp_log(riga_ctx* ctx, const double* x, uint64_t n, double* out_exp, double* out_log,
                       riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n == 0) return 0;
  if (!x || !out_exp || !out_log || n > ((uint64_t)1 << 28)) {
    set_err(err, "invalid_argument", "bad debug_exp_log arguments");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf dx, de, dl;
  HIP_TRY(dx.ensure(n * 8), "alloc");
  HIP_TRY(de.ensure(n * 8), "alloc");
  HIP_TRY(dl.ensure(n * 8), "alloc");
  HIP_TRY(hipMemcpyAsync(dx.p, x, n * 8, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIP_TRY(launch_debug_exp_log((const double*)dx.p, n, (double*)de.p, (double*)dl.p, ctx->stream), "debug kernel");
  HIP_TRY(hipMemcpyAsync(out_exp, de.p, n * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_log, dl.p, n * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "debug kernel execution");
  return 0;
}

int riga_debug_halton(riga_ctx* ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point, int32_t n_points,
                      int32_t first_dim, int32_t n_dims, uint32_t* out_digits, double* out_uniform,
                      double* out_normal, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n_points <= 0 || n_dims <= 0 || first_dim < 0 || ladder_dims <= 0 || ladder_dims > 7762 ||
      !out_digits || !out_uniform || !out_normal || first_point + (uint64_t)n_points >= (1ull << 32)) {
    set_err(err, "invalid_argument", "bad debug_halton ranges");
    return RIGA_EINVAL;
  }
  const std::vector<uint32_t> tab = halton_tables(ladder_dims, seed);
  const size_t cells = (size_t)n_points * n_dims;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf d_tab, d_dig, d_u, d_n;
  HIP_TRY(d_tab.ensure(tab.size() * 4), "alloc");
  HIP_TRY(d_dig.ensure(cells * 32 * 4), "alloc");
  HIP_TRY(d_u.ensure(cells * 8), "alloc");
  HIP_TRY(d_n.ensure(cells * 8), "alloc");
  HIP_TRY(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIP_TRY(launch_debug_halton((const uint32_t*)d_tab.p, ladder_dims, first_point, n_points, first_dim, n_dims,
                              (uint32_t*)d_dig.p, (double*)d_u.p, (double*)d_n.p, ctx->stream),
          "debug kernel");
  HIP_TRY(hipMemcpyAsync(out_digits, d_dig.p, cells * 32 * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_uniform, d_u.p, cells * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_normal, d_n.p, cells * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "debug kernel execution");
  return 0;
}

int riga_debug_sobol_f32(riga_ctx* ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point,
                         int32_t n_points, int32_t n_nodes, int32_t factors, int32_t owen_table,
                         uint32_t* out_scrambled, float* out_normal, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n_points <= 0 || n_nodes <= 0 || factors < 1 || factors > kMaxFactors || ladder_dims <= 0 ||
      first_point == 0 || first_point + (uint64_t)n_points >= (1ull << 32) ||
      (int64_t)n_nodes * factors > (int64_t)1 << 24) {
    set_err(err, "invalid_argument", "bad debug_sobol_f32 ranges");
    return RIGA_EINVAL;
  }
  const int dims = n_nodes * factors;
  const int rows = std::min(ladder_dims, dims);
  const std::vector<uint32_t>& dirs = sobol_directions(rows);
  const int nbits = 64 - __builtin_clzll(first_point + (uint64_t)n_points);
  const size_t cells = (size_t)n_points * dims;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  std::shared_ptr<OwenTable> ot;
  if ((owen_table & 1) && nbits <= kOwenPrefixBits && dims <= ladder_dims) {
    HIP_TRY(owen_table_for(ctx, seed, dims, ctx->stream, &ot), "owen table");
    HIP_TRY(hipStreamWaitEvent(ctx->stream, ot->ready, 0), "owen table wait");
  }
  DevBuf d_dirs, d_s, d_z;
  HIP_TRY(d_dirs.ensure((size_t)rows * 32 * 4), "alloc");
  HIP_TRY(d_s.ensure(cells * 4), "alloc");
  HIP_TRY(d_z.ensure(cells * 4), "alloc");
  HIP_TRY(hipMemcpyAsync(d_dirs.p, dirs.data(), (size_t)rows * 32 * 4, hipMemcpyHostToDevice, ctx->stream),
          "upload");
  HIP_TRY(launch_debug_sobol_f32((const uint32_t*)d_dirs.p, ladder_dims, ot ? (const OwenEntry*)ot->buf.p : nullptr,
                                 seed, first_point, n_points, n_nodes, factors, nbits, (owen_table & 2) != 0,
                                 (uint32_t*)d_s.p, (float*)d_z.p, ctx->stream),
          "debug kernel");
  HIP_TRY(hipMemcpyAsync(out_scrambled, d_s.p, cells * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_normal, d_z.p, cells * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "debug kernel execution");
  return 0;
}

int riga_path_cube(riga_ctx* ctx, uint64_t seed, uint64_t paths, int32_t times, int32_t factors,
                   const double* correlation, const double* s0, const double* vols, double rate,
                   const double* dividend_yields, double dt, const float* in_block, float* out_block,
                   float* out_cube, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (paths == 0 || times <= 0 || factors < 1 || factors > kMaxFactors) {
    set_err(err, "shape_mismatch", "paths and times must be positive and factors 1.." + std::to_string(kMaxFactors));
    return RIGA_EINVAL;
  }
  const double cells = (double)paths * (double)times * (double)factors;
  if (cells * 4.0 > 64e9) {
    set_err(err, "invalid_argument", "a PathCube above 64 GB");
    return RIGA_EINVAL;
  }
  if ((out_cube || in_block) && (!s0 || !vols || !dividend_yields)) {
    set_err(err, "shape_mismatch", "s0, vols and dividend_yields are required to evolve spots");
    return RIGA_EINVAL;
  }
  if (in_block && (!out_cube || out_block)) {
    set_err(err, "invalid_argument", "evolving a given block writes only the cube");
    return RIGA_EINVAL;
  }
  CubeArgs a{};
  a.paths = paths;
  a.times = times;
  if (out_cube || in_block) {
    const double sdt = sqrt(dt);
    for (int f = 0; f < factors; ++f) {  // pathcube.rs:170-175
      a.ln_s0[f] = log(s0[f]);
      a.drift[f] = (rate - dividend_yields[f] - 0.5 * vols[f] * vols[f]) * dt;
      a.diffusion[f] = vols[f] * sdt;
    }
  }
  const size_t bytes = (size_t)cells * 4;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf d_in, d_block, d_cube, d_key, d_zig, d_status;
  riga_plan::McParse mp;
  if (out_cube) HIP_TRY(d_cube.ensure(bytes), "cube alloc");
  a.cube = (float*)d_cube.p;
  if (in_block) {
    HIP_TRY(d_in.ensure(bytes), "block alloc");
    HIP_TRY(hipMemcpyAsync(d_in.p, in_block, bytes, hipMemcpyHostToDevice, ctx->stream), "block upload");
    a.in_block = (const float*)d_in.p;
  } else {
    std::vector<double> chol;
    Error e;
    if (!correlation || !cholesky_factor(factors, correlation, &chol, &e)) {
      set_err(err, "invalid_market", correlation ? e.message : "correlation is required");
      return RIGA_EINVAL;
    }
    for (int i = 0; i < factors; ++i)
      for (int j = 0; j < factors; ++j) a.chol[i * kMaxFactors + j] = chol[(size_t)i * factors + j];
    if (out_block) HIP_TRY(d_block.ensure(bytes), "block alloc");
    a.block = (float*)d_block.p;
    // the stream parse: path p starts at draw p * times * factors
    mp.seed = seed;
    mp.dpp = (uint64_t)times * (uint64_t)factors;
    mp.paths = paths;
    mc_parse_size(&mp);
    uint32_t key[8];
    chacha_key_from_seed(seed, key);
    HIP_TRY(d_key.ensure(32), "alloc");
    HIP_TRY(d_zig.ensure(2 * kZigLen * 8), "alloc");
    HIP_TRY(d_status.ensure(8), "alloc");
    HIP_TRY(mc_parse_alloc(&mp), "stream parse alloc");
    HIP_TRY(hipMemcpyAsync(d_key.p, key, 32, hipMemcpyHostToDevice, ctx->stream), "upload");
    HIP_TRY(hipMemcpyAsync(d_zig.p, ziggurat_tables().data(), 2 * kZigLen * 8, hipMemcpyHostToDevice, ctx->stream),
            "upload");
    HIP_TRY(mc_parse_launch(mp, (const uint32_t*)d_key.p, (const double*)d_zig.p, (uint32_t*)d_status.p,
                            ctx->stream),
            "stream parse");
    a.key = (const uint32_t*)d_key.p;
    a.zig = (const double*)d_zig.p;
    a.starts = (const uint64_t*)mp.starts.p;
    a.chunks = (const McChunk*)mp.chunks.p;
    a.slow = (const uint32_t*)mp.slow.p;
    a.slow_val = (const double*)mp.slow_val.p;
    a.n_chunks = mp.n_chunks;
    a.chunk_words = mp.chunk_words;
    a.slow_cap = mp.slow_cap;
  }
  HIP_TRY(launch_cube(a, factors, ctx->stream), "path cube");
  uint32_t status[2] = {(uint32_t)paths, 0};
  if (!in_block) HIP_TRY(hipMemcpyAsync(status, d_status.p, 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  if (out_block) HIP_TRY(hipMemcpyAsync(out_block, d_block.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
  if (out_cube) HIP_TRY(hipMemcpyAsync(out_cube, d_cube.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "path cube execution");
  if (status[0] != paths || status[1] != 0) {
    set_err(err, "mc_stream_error", "stream parse found " + std::to_string(status[0]) + " of " +
                                        std::to_string(paths) + " path starts, flags " + std::to_string(status[1]));
    return RIGA_EGPU;
  }
  return 0;
}

// ---- the reference's device module (gpu.rs:159-570)
}  // extern "C"

struct riga_resident_worst_of {
  riga_ctx* ctx = nullptr;
  DevBuf a, b, out;
  uint64_t n = 0;
};

struct riga_resident_lv {
  riga_ctx* ctx = nullptr;
  DevBuf normals, rows, out, ko;
  int paths = 0, steps = 0, gx = 0;
  float dt = 0.f;
  float ls_lo[2] = {0.f, 0.f}, ls_hi[2] = {0.f, 0.f};  // end knots of the two log-spot axes
};

namespace {

int lv_check(const float* normals, const float* rows, const float* ls_axis, int32_t gx, int64_t paths,
             int64_t steps, riga_error* err) {
  if (!normals || !rows || !ls_axis || gx < 2 || gx > kLvMaxKnots || paths < 0 || steps < 0 ||
      paths > INT32_MAX || steps > INT32_MAX || (double)paths * (double)steps * 2.0 > 8e9) {
    set_err(err, "shape_mismatch",
            "normals [paths][steps][2], rows [2][steps][gx], ls_axis [2][gx] with 2 <= gx <= " +
                std::to_string(kLvMaxKnots));
    return RIGA_EINVAL;
  }
  for (int f = 0; f < 2; ++f)
    if (!(ls_axis[(size_t)f * gx + gx - 1] > ls_axis[(size_t)f * gx])) {
      set_err(err, "invalid_local_vol", "each log-spot axis must increase from its first to its last knot");
      return RIGA_EINVAL;
    }
  return 0;
}

// The propagation's per-factor constants, formed in FP64 and rounded once: the
// knot position (gx - 1) (x - x_first) / (x_last - x_first) as one multiply-add
// of x, the drift (r - q) dt and the KO level ln(call_barrier * reference).
LvPropagateArgs lv_args(const float* normals, const float* rows, float* out, int* ko, const float ls_lo[2],
                        const float ls_hi[2], int gx, int64_t paths, int64_t steps, float dt, const float* s0,
                        const float* references, float call_barrier, float rate, const float* dividends) {
  LvPropagateArgs a{};
  a.normals = normals;
  a.rows = rows;
  a.out = out;
  a.ko_out = ko;
  for (int f = 0; f < 2; ++f) {
    const double scale = (double)(gx - 1) / ((double)ls_hi[f] - (double)ls_lo[f]);
    a.x0[f] = (float)std::log((double)s0[f]);
    a.knot_scale[f] = (float)scale;
    a.knot_offset[f] = (float)(-(double)ls_lo[f] * scale);
    a.drift[f] = (float)(((double)rate - (double)dividends[f]) * (double)dt);
    a.ln_ko[f] = (float)std::log((double)call_barrier * (double)references[f]);
  }
  a.sqrt_dt = (float)std::sqrt((double)dt);
  a.neg_half_dt = (float)(-0.5 * (double)dt);
  a.paths = (int)paths;
  a.steps = (int)steps;
  a.gx = gx;
  return a;
}

}  // namespace

extern "C" {

int riga_worst_of(riga_ctx* ctx, const float* a, const float* b, uint64_t n, float* out, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (n == 0) return 0;
  if (!a || !b || !out) {
    set_err(err, "shape_mismatch", "worst-of inputs must have equal length");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf da, db, dout;
  HIP_TRY(da.ensure(n * 4), "alloc");
  HIP_TRY(db.ensure(n * 4), "alloc");
  HIP_TRY(dout.ensure(n * 4), "alloc");
  HIP_TRY(hipMemcpyAsync(da.p, a, n * 4, hipMemcpyHostToDevice, ctx->stream), "H2D a");
  HIP_TRY(hipMemcpyAsync(db.p, b, n * 4, hipMemcpyHostToDevice, ctx->stream), "H2D b");
  HIP_TRY(launch_worst_of((const float*)da.p, (const float*)db.p, (float*)dout.p, n, ctx->stream), "launch");
  HIP_TRY(hipMemcpyAsync(out, dout.p, n * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "worst_of");
  return 0;
}

int riga_resident_worst_of_create(riga_ctx* ctx, const float* a, const float* b, uint64_t n,
                                  riga_resident_worst_of** out, riga_error* err) {
  clear_err(err);
  if (!ctx || !out) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  *out = nullptr;
  if (n > 0 && (!a || !b)) {
    set_err(err, "shape_mismatch", "worst-of inputs must have equal length");
    return RIGA_EINVAL;
  }
  std::unique_ptr<riga_resident_worst_of> h(new riga_resident_worst_of());
  h->ctx = ctx;
  h->n = n;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(h->a.ensure(std::max<uint64_t>(n, 1) * 4), "alloc");
  HIP_TRY(h->b.ensure(std::max<uint64_t>(n, 1) * 4), "alloc");
  HIP_TRY(h->out.ensure(std::max<uint64_t>(n, 1) * 4), "alloc");
  if (n) {
    HIP_TRY(hipMemcpyAsync(h->a.p, a, n * 4, hipMemcpyHostToDevice, ctx->stream), "H2D a");
    HIP_TRY(hipMemcpyAsync(h->b.p, b, n * 4, hipMemcpyHostToDevice, ctx->stream), "H2D b");
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream), "upload");
  *out = h.release();
  return 0;
}

int riga_resident_worst_of_reduce(riga_resident_worst_of* h, riga_error* err) {
  clear_err(err);
  if (!h) {
    set_err(err, "invalid_argument", "handle is NULL");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(h->ctx->mu);
  HIP_TRY(hipSetDevice(h->ctx->device), "hipSetDevice");
  HIP_TRY(launch_worst_of((const float*)h->a.p, (const float*)h->b.p, (float*)h->out.p, h->n, h->ctx->stream),
          "launch");
  return 0;
}

int riga_resident_worst_of_download(riga_resident_worst_of* h, float* out, riga_error* err) {
  clear_err(err);
  if (!h || (h->n && !out)) {
    set_err(err, "invalid_argument", "handle or output is NULL");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(h->ctx->mu);
  HIP_TRY(hipSetDevice(h->ctx->device), "hipSetDevice");
  if (h->n) HIP_TRY(hipMemcpyAsync(out, h->out.p, h->n * 4, hipMemcpyDeviceToHost, h->ctx->stream), "D2H");
  HIP_TRY(hipStreamSynchronize(h->ctx->stream), "download");
  return 0;
}

int riga_resident_worst_of_synchronize(riga_resident_worst_of* h, riga_error* err) {
  clear_err(err);
  if (!h) {
    set_err(err, "invalid_argument", "handle is NULL");
    return RIGA_EINVAL;
  }
  HIP_TRY(hipStreamSynchronize(h->ctx->stream), "synchronize");
  return 0;
}

void riga_resident_worst_of_destroy(riga_resident_worst_of* h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->ctx->stream);
  delete h;
}

int riga_lv_terminal_spots(riga_ctx* ctx, const float* normals, const float* rows, const float* ls_axis,
                           int32_t gx, const float* s0, const float* references, float call_barrier, int64_t paths,
                           int64_t steps, float dt, float rate, const float* dividends, float* out_terminal,
                           int32_t* out_ko, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  int rc = lv_check(normals, rows, ls_axis, gx, paths, steps, err);
  if (rc) return rc;
  if (!s0 || !references || !dividends || !out_terminal || !out_ko) {
    set_err(err, "shape_mismatch", "s0, references and dividends hold two values; outputs are required");
    return RIGA_EINVAL;
  }
  if (paths == 0) return 0;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf dn, dr, dout, dko;
  const size_t nn = (size_t)paths * steps * 2, nr = (size_t)2 * steps * gx;
  HIP_TRY(dn.ensure(std::max<size_t>(nn, 1) * 4), "alloc");
  HIP_TRY(dr.ensure(std::max<size_t>(nr, 1) * 4), "alloc");
  HIP_TRY(dout.ensure((size_t)paths * 8), "alloc");
  HIP_TRY(dko.ensure((size_t)paths * 4), "alloc");
  if (nn) HIP_TRY(hipMemcpyAsync(dn.p, normals, nn * 4, hipMemcpyHostToDevice, ctx->stream), "H2D normals");
  if (nr) HIP_TRY(hipMemcpyAsync(dr.p, rows, nr * 4, hipMemcpyHostToDevice, ctx->stream), "H2D rows");
  const float lo[2] = {ls_axis[0], ls_axis[gx]}, hi[2] = {ls_axis[gx - 1], ls_axis[2 * gx - 1]};
  const LvPropagateArgs a = lv_args((const float*)dn.p, (const float*)dr.p, (float*)dout.p, (int*)dko.p, lo, hi, gx,
                                    paths, steps, dt, s0, references, call_barrier, rate, dividends);
  HIP_TRY(launch_lv_propagate(a, ctx->stream), "lv launch");
  HIP_TRY(hipMemcpyAsync(out_terminal, dout.p, (size_t)paths * 8, hipMemcpyDeviceToHost, ctx->stream), "D2H out");
  HIP_TRY(hipMemcpyAsync(out_ko, dko.p, (size_t)paths * 4, hipMemcpyDeviceToHost, ctx->stream), "D2H ko");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "lv_terminal_spots");
  return 0;
}

int riga_resident_lv_create(riga_ctx* ctx, const float* normals, const float* rows, const float* ls_axis,
                            int32_t gx, int64_t paths, int64_t steps, float dt, riga_resident_lv** out,
                            riga_error* err) {
  clear_err(err);
  if (!ctx || !out) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  *out = nullptr;
  int rc = lv_check(normals, rows, ls_axis, gx, paths, steps, err);
  if (rc) return rc;
  std::unique_ptr<riga_resident_lv> h(new riga_resident_lv());
  h->ctx = ctx;
  h->paths = (int)paths;
  h->steps = (int)steps;
  h->gx = gx;
  h->dt = dt;
  const size_t nn = (size_t)paths * steps * 2, nr = (size_t)2 * steps * gx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(h->normals.ensure(std::max<size_t>(nn, 1) * 4), "alloc");
  HIP_TRY(h->rows.ensure(std::max<size_t>(nr, 1) * 4), "alloc");
  HIP_TRY(h->out.ensure(std::max<size_t>((size_t)paths, 1) * 8), "alloc");
  HIP_TRY(h->ko.ensure(std::max<size_t>((size_t)paths, 1) * 4), "alloc");
  for (int f = 0; f < 2; ++f) {
    h->ls_lo[f] = ls_axis[(size_t)f * gx];
    h->ls_hi[f] = ls_axis[(size_t)f * gx + gx - 1];
  }
  if (nn) HIP_TRY(hipMemcpyAsync(h->normals.p, normals, nn * 4, hipMemcpyHostToDevice, ctx->stream), "H2D normals");
  if (nr) HIP_TRY(hipMemcpyAsync(h->rows.p, rows, nr * 4, hipMemcpyHostToDevice, ctx->stream), "H2D rows");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "upload");
  *out = h.release();
  return 0;
}

int riga_resident_lv_run(riga_resident_lv* h, const float* s0, const float* references, float call_barrier,
                         float rate, const float* dividends, riga_error* err) {
  clear_err(err);
  if (!h || !s0 || !references || !dividends) {
    set_err(err, "invalid_argument", "handle, s0, references and dividends are required");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(h->ctx->mu);
  HIP_TRY(hipSetDevice(h->ctx->device), "hipSetDevice");
  // the spots enter as kernel arguments (their logs), so a run needs no upload
  const LvPropagateArgs a = lv_args((const float*)h->normals.p, (const float*)h->rows.p, (float*)h->out.p,
                                    (int*)h->ko.p, h->ls_lo, h->ls_hi, h->gx, h->paths, h->steps, h->dt, s0,
                                    references, call_barrier, rate, dividends);
  HIP_TRY(launch_lv_propagate(a, h->ctx->stream), "lv launch");
  return 0;
}

int riga_resident_lv_download(riga_resident_lv* h, float* out_terminal, int32_t* out_ko, riga_error* err) {
  clear_err(err);
  if (!h || (h->paths && (!out_terminal || !out_ko))) {
    set_err(err, "invalid_argument", "handle or outputs are NULL");
    return RIGA_EINVAL;
  }
  std::lock_guard<std::mutex> lock(h->ctx->mu);
  HIP_TRY(hipSetDevice(h->ctx->device), "hipSetDevice");
  if (h->paths) {
    HIP_TRY(hipMemcpyAsync(out_terminal, h->out.p, (size_t)h->paths * 8, hipMemcpyDeviceToHost, h->ctx->stream),
            "D2H out");
    HIP_TRY(hipMemcpyAsync(out_ko, h->ko.p, (size_t)h->paths * 4, hipMemcpyDeviceToHost, h->ctx->stream), "D2H ko");
  }
  HIP_TRY(hipStreamSynchronize(h->ctx->stream), "download");
  return 0;
}

int riga_resident_lv_synchronize(riga_resident_lv* h, riga_error* err) {
  clear_err(err);
  if (!h) {
    set_err(err, "invalid_argument", "handle is NULL");
    return RIGA_EINVAL;
  }
  HIP_TRY(hipStreamSynchronize(h->ctx->stream), "synchronize");
  return 0;
}

void riga_resident_lv_destroy(riga_resident_lv* h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->ctx->stream);
  delete h;
}

int riga_debug_mc_starts(riga_ctx* ctx, uint64_t seed, uint64_t draws_per_path, uint64_t n_paths,
                         uint64_t* out_starts, riga_error* err) {
  clear_err(err);
  if (!ctx) {
    set_err(err, "gpu_unavailable", "no device context");
    return RIGA_ENODEV;
  }
  if (draws_per_path == 0 || n_paths == 0 || !out_starts || (double)draws_per_path * (double)n_paths > 1e12) {
    set_err(err, "invalid_argument", "bad debug_mc_starts sizes");
    return RIGA_EINVAL;
  }
  riga_plan::McParse mp;
  mp.seed = seed;
  mp.dpp = draws_per_path;
  mp.paths = n_paths;
  mc_parse_size(&mp);
  uint32_t key[8];
  chacha_key_from_seed(seed, key);
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  DevBuf d_key, d_zig, d_status;
  HIP_TRY(d_key.ensure(32), "alloc");
  HIP_TRY(d_zig.ensure(2 * kZigLen * 8), "alloc");
  HIP_TRY(d_status.ensure(8), "alloc");
  HIP_TRY(mc_parse_alloc(&mp), "alloc");
  HIP_TRY(hipMemcpyAsync(d_key.p, key, 32, hipMemcpyHostToDevice, ctx->stream), "upload");
  HIP_TRY(hipMemcpyAsync(d_zig.p, ziggurat_tables().data(), 2 * kZigLen * 8, hipMemcpyHostToDevice, ctx->stream),
          "upload");
  HIP_TRY(mc_parse_launch(mp, (const uint32_t*)d_key.p, (const double*)d_zig.p, (uint32_t*)d_status.p, ctx->stream),
          "stream parse");
  uint32_t status[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(status, d_status.p, 8, hipMemcpyDeviceToHost, ctx->stream), "download");
  HIP_TRY(hipMemcpyAsync(out_starts, mp.starts.p, (size_t)n_paths * 8, hipMemcpyDeviceToHost, ctx->stream),
          "download");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "stream parse execution");
  if (status[0] != n_paths || status[1] != 0) {
    set_err(err, "mc_stream_error", "stream parse found " + std::to_string(status[0]) + " of " +
                                        std::to_string(n_paths) + " path starts, flags " + std::to_string(status[1]));
    return RIGA_EGPU;
  }
  return 0;
}

}  // extern "C"
