/*
 * riga_hip.h -- C ABI of the MI355X (gfx950) Phoenix path-simulation backend.
 *
 * This is the drop-in boundary for the reference engine's pricing path.  Each
 * entry point names the reference interface it replaces:
 *
 *   riga_hip_init / riga_hip_device_name / riga_hip_arch
 *       gpu.rs:490-512 ensure_available(), device_name(), arch()
 *       (process-global device initialised once; a build with this backend
 *        fails loudly with code "gpu_unavailable", never falls back: lib.rs:188-202)
 *   riga_price_put_leg
 *       lib.rs:183-186 price_put_leg(&PriceRequest) -> Result<PriceResult, RigaError>
 *       (note.rs:38-373 price_full_phoenix on the QMC lanes)
 *   riga_price_batch
 *       riga-risk/src/lib.rs:117-186 price_central_greeks and
 *       riga-risk/src/lib.rs:573-626 price_risk_batch: many requests priced in
 *       one launch; requests that share seed, schedule and correlation become
 *       CRN scenarios of one item and share their Sobol points
 *   riga_replay_paths
 *       explain.rs:90-316 run_paths: per-path records, written to device memory
 *       only for this query
 *
 * Conventions (lib.rs, riga-core/src/lib.rs):
 *   - every function returns 0 on success or a negative status and fills
 *     riga_error {code, message} (RigaError, riga-core/src/lib.rs:5-19);
 *   - inputs are caller-owned and only borrowed for the call; outputs are
 *     caller-allocated;
 *   - Option<f64> fields are expressed with has_* flags; None maps to the
 *     reference defaults (KO/coupon barrier +inf, coupon rate 0, 9 periods,
 *     call start step 0; note.rs:43-53).
 *   - the context is thread-safe: calls are serialised on the context's stream.
 */
#ifndef RIGA_HIP_H
#define RIGA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RIGA_TRACE_CAP 64

typedef struct riga_error {
  char code[48];
  char message[464];
} riga_error;

typedef struct riga_ctx riga_ctx;
typedef struct riga_surface riga_surface;

enum riga_precision { RIGA_FP64 = 0, RIGA_FP32 = 1 };
enum riga_ladder { RIGA_VR_NONE = 0, RIGA_VR_HALTON = 1, RIGA_VR_SOBOL = 2 };
enum riga_trace { RIGA_TRACE_OFF = 0, RIGA_TRACE_AGGREGATED = 1, RIGA_TRACE_SAMPLED = 2,
                  RIGA_TRACE_FULL = 3 };

/* ImpliedVolGrid (riga-market/src/lib.rs:84-91); vols row-major [n_maturities][n_strikes]. */
typedef struct riga_implied_grid {
  int32_t n_maturities, n_strikes;
  const double *maturities, *strikes, *vols;
} riga_implied_grid;

/* A calibrated DupireLocalVol surface (riga-market/src/lib.rs:119-123): v[n_t][n_x]. */
typedef struct riga_lv_grid {
  int32_t n_t, n_x;
  const double *t_axis, *log_s, *v;
} riga_lv_grid;

/* PriceRequest (riga-engine/src/lib.rs:36-46) with F = factors underlyings. */
typedef struct riga_request {
  int32_t factors;
  /* MarketSnapshot (riga-core/src/lib.rs:37-45) */
  const double *spots, *vols, *dividend_yields;
  const double *correlation; /* F x F row-major */
  double risk_free_rate;
  /* PhoenixTrade (riga-core/src/lib.rs:57-75) */
  const double *initial_spots;
  double maturity_years;
  int64_t steps;
  double strike, knock_in_barrier;
  int32_t has_knock_out;   double knock_out_barrier;
  int32_t has_coupon_barrier; double coupon_barrier;
  int32_t has_coupon_rate; double coupon_rate;
  int32_t memory_coupon;
  int64_t coupon_periods;  /* < 0: None */
  int32_t fixed_first_coupon;
  int64_t call_start_step; /* < 0: None */
  /* SimulationConfig (riga-core/src/lib.rs:77-86) */
  uint64_t paths, path_block_size, seed;
  int32_t trace_level;     /* riga_trace */
  int32_t ladder;          /* riga_ladder: VarianceReduction kind (NONE = the ChaCha8 Monte
                              Carlo lane, note.rs:58,124-126; HALTON; SOBOL) */
  int32_t ladder_dimension;/* VarianceReduction dimension */
  int32_t max_dimension;   /* ladder cap: 512 reproduces qmc.rs:9; larger lifts the wrap */
  /* LocalVolSpec (riga-engine/src/lib.rs:31-34): one per factor, or NULL for flat vols.
   * Either implied grids (calibrated here with grid_x log-spot points, as
   * build_local_vol, lib.rs:326-343) or already-calibrated grids. */
  const riga_implied_grid *implied;
  const riga_lv_grid *calibrated;
  int32_t grid_x;          /* 96 = GRID_X (riga-market/src/lib.rs:28) */
  /* Resident calibrated grids, one per factor (riga_surface_create); when set
   * they take the place of implied/calibrated and nothing grid-sized is copied
   * per call.  The surfaces must outlive every plan and ticket that uses them. */
  const struct riga_surface *const *surfaces;
} riga_request;

/* PriceResult (riga-core/src/lib.rs:243-255) in flat form.  Strings (ledger
 * conditions, evidence, the snapshot id) are formed by the host-language side
 * from these numbers, as note.rs:272-354 does. */
typedef struct riga_result {
  double pv, mc_error, ki_probability;
  /* LegAttribution: coupon and memory_coupon on the x10 quote scale */
  double leg_coupon, leg_memory_coupon, leg_put, leg_funding, leg_total;
  uint64_t paths, ko_count, ki_count, put_eligible_count, coupon_path_count;
  /* TransitionStats: daily_global_ko, final_fixing_eki, memory_coupon */
  double ko_mean, ko_variance, ki_mean, ki_variance, coupon_mean, coupon_variance;
  int32_t n_trace;
  int64_t trace_path[RIGA_TRACE_CAP];
  int64_t trace_step[RIGA_TRACE_CAP]; /* KO at time (step+1)*dt */
  double device_ms;                   /* kernel time of the batch this result came from */
} riga_result;

/* explain.rs PathRecord subset, one per path. */
typedef struct riga_path_record {
  int32_t ko_step; /* -1: survived */
  int32_t hit_ki;
  double final_worst, funding, put, coupon_pv, memory_pv, pv;
} riga_path_record;

/* Launch shape (chosen by the autotuner in riga_flow_amd/autotune.py or by hand). */
typedef struct riga_launch_config {
  int32_t paths_per_thread; /* 1, 2, 4 or 8 paths per lane; 0: the autotuned choice (riga_autotune)
                               for the batch's shape, else 1 */
  int32_t use_graph;        /* 1: capture the launch sequence in a hipGraph at plan creation */
  int32_t owen_table;       /* a seed's scrambled Sobol coordinates of the points below 2^22 from a
                               per-dimension table (items past those points, or whose dimensions
                               wrap their direction rows, hash) that the context keeps (16 MiB per dimension, covering a seed's first
                               dimensions; the same integers as hashing them): 0 = a table the
                               context holds already, or a new one when the plan reads at least
                               2^18 points per dimension of the seed and memory allows; 1 =
                               always; k > 1 = always, covering at most the first k dimensions;
                               -1 = hash all 32 levels */
  int32_t normal_table;     /* 0: auto (share one [dim][path] normal table among same-seed trades
                               when they need 4x more normals than it holds); 1: force; -1: off */
  int32_t stream_ordered_alloc; /* 1: plan memory from the stream-ordered pool (serving) */
  int32_t waves_per_block;  /* 1-4 waves of 64 lanes per workgroup; 0: autotuned, else 4 */
  int32_t serial_launches;  /* 0: a batch's launches (one per factor count, scenario bucket and
                               ladder mode) run concurrently on streams of their own; 1: one
                               after another on the plan's stream */
  int32_t fp32_full_ladder; /* FP32 tier: 1 hashes all 32 Owen levels of every coordinate (the FP64
                               tier's integers) instead of the top 24 plus the cell middle; no
                               shared normal table then.  For pricing what the 24-level rule costs */
  int32_t stack_split;      /* bridge stack levels nearest the root kept in private memory instead of
                               LDS: 0 = the fewest that let each launch reach the occupancy its
                               registers allow; -1 = none; k > 0 = k (at most 8) */
  int32_t persistent;       /* 0: one workgroup per (item, path-chunk) unit, each claiming the next
                               unit of the launch's longest-first queue as the hardware dispatches
                               it into a free slot; 1: persistent workgroups, as many as the device
                               holds at once, each claiming units until the queue is empty */
} riga_launch_config;

/* A batch laid out in device memory once (surfaces, factors, schedules, Sobol
 * tables) and launched any number of times; each plan owns a stream, so plans
 * launched back to back overlap on the device. */
typedef struct riga_plan riga_plan;

typedef struct riga_plan_info {
  int32_t n_requests, n_items, n_launches, blocks;
  double path_steps;    /* sum over requests of paths * steps */
  int64_t lds_bytes;    /* dynamic LDS of the largest launch */
  int64_t device_bytes; /* device memory held by the plan */
  int32_t graph;
  int32_t normal_tables; /* shared [dim][path] normal tables the plan fills per launch */
  int32_t paths_per_thread, waves_per_block; /* the launch shape in use */
  int32_t owen_tables;   /* seeds whose coordinates read an Owen prefix table (the rest hash) */
  int32_t reserved;
} riga_plan_info;

/* One autotuner trial: a launch shape and its best device time (< 0: the shape
 * does not fit, e.g. its LDS exceeds the workgroup limit). */
typedef struct riga_tune_trial {
  int32_t paths_per_thread, waves_per_block;
  double device_ms;
} riga_tune_trial;

#define RIGA_TUNE_MAX_TRIALS 16
typedef struct riga_tune_report {
  int32_t n_trials;
  int32_t best;          /* index of the fastest trial */
  riga_tune_trial trials[RIGA_TUNE_MAX_TRIALS];
} riga_tune_report;

/* Autotuner (no reference counterpart: the reference launches a fixed 256-lane
 * shape, gpu.rs:240-271).  Runs the batch once per launch shape -- paths per lane
 * 1/2/4/8 x waves per workgroup 4/3/2/1 -- timing the path kernels with device
 * events (best of `reps` after one warm-up), and keeps the fastest for the
 * batch's shape signature in the context: plans created afterwards with
 * paths_per_thread = 0 and waves_per_block = 0 use it.  cfg gives the other
 * launch options the plans will use.  Results are the same for every shape. */
int riga_autotune(riga_ctx *ctx, const riga_request *reqs, int n, int precision,
                  const riga_launch_config *cfg, int reps, riga_tune_report *report, riga_error *err);

/* ---- device (gpu.rs:490-512) ---- */
int riga_hip_init(int device, riga_ctx **out, riga_error *err);
void riga_hip_destroy(riga_ctx *ctx);
int riga_hip_device_name(riga_ctx *ctx, char *buf, size_t len, riga_error *err);
int riga_hip_arch(riga_ctx *ctx, char *buf, size_t len, riga_error *err);

/* ---- validation (lib.rs:209-321, F-factor form); no device work ---- */
int riga_validate(const riga_request *req, riga_error *err);

/* ---- pricing (lib.rs:183-186) ---- */
int riga_price_put_leg(riga_ctx *ctx, const riga_request *req, int precision, riga_result *out,
                       riga_error *err);

/* ---- pricing without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): riga_price_put_leg's result for one request, with the
 * request's tables, block partials, merged records and (Monte Carlo lane) stream
 * parse laid out in a fixed per-GPU scratch array of 32 MiB set up by
 * riga_hip_init.  No hipMalloc, hipFree or other allocating routine runs (the
 * host-side planning uses the C++ heap only).  The Owen prefix tables and shared
 * normal tables are not used (the ladder hashes all 32 levels; results are the
 * same integers, so the same prices), nor path records.  A request whose work
 * area exceeds the array fails with "scratch_exceeded" (about 11 million paths
 * for one scenario).  Calls through one context run one at a time; calls through
 * several contexts on one GPU share its array and race. */
int riga_price_put_leg_noalloc(riga_ctx *ctx, const riga_request *req, int precision, riga_result *out,
                               riga_error *err);

/* ---- batched pricing: n requests in one launch; results[i] for reqs[i] ---- */
int riga_price_batch(riga_ctx *ctx, const riga_request *reqs, int n, int precision,
                     const riga_launch_config *cfg, riga_result *results, riga_error *err);

/* ---- resident plans: create once, launch many times (no device allocation per launch) ---- */
int riga_plan_create(riga_ctx *ctx, const riga_request *reqs, int n, int precision,
                     const riga_launch_config *cfg, riga_plan **out, riga_error *err);
/* Device bytes a plan for these requests would hold, without allocating. */
int riga_plan_estimate(riga_ctx *ctx, const riga_request *reqs, int n, int precision,
                       const riga_launch_config *cfg, int64_t *bytes, riga_error *err);
int riga_plan_launch(riga_plan *plan, riga_error *err);  /* asynchronous */
int riga_plan_ready(riga_plan *plan);  /* 1: finished, 0: running, < 0: device error */
int riga_plan_wait(riga_plan *plan, riga_result *results, riga_error *err);
int riga_plan_run(riga_plan *plan, riga_result *results, riga_error *err); /* launch + wait */
int riga_plan_query(const riga_plan *plan, riga_plan_info *info);
/* Device clock of the plan's last finished launch: when its stream reached the
 * first kernel and when the path kernels were done, in seconds since the context
 * was created.  No reference counterpart (the reference runs one blocking call at
 * a time); the serving loop uses it to show plans overlapping on the device. */
int riga_plan_device_times(riga_plan *plan, double *start_s, double *end_s);
void riga_plan_destroy(riga_plan *plan);

/* ---- serving: continuous batching on one device ----
 * Replaces the reference's per-call blocking pricing in its MCP tool handlers
 * (riga-mcp/src/main.rs:1821-1831, 2005-2035; tasks 1005-1042) with one
 * device-owning scheduler: tickets are admitted in arrival order while the
 * device-memory budget and in-flight limit allow, batched into one plan, run on
 * their own stream (overlapping earlier plans), and shed after the queue limit. */
typedef struct riga_server riga_server;

typedef struct riga_server_config {
  int64_t hbm_budget_bytes;  /* device memory the in-flight plans may hold (0: 64 GiB) */
  int32_t max_inflight;      /* plans running at once (0: 2) */
  int32_t max_batch_requests;/* requests per plan (0: 4096) */
  double max_queue_seconds;  /* shed tickets queued longer (0: never) */
  int32_t paths_per_thread;  /* launch shape for every plan (0: the autotuned one, else 1) */
  int32_t normal_table;      /* as riga_launch_config.normal_table */
  double max_batch_path_steps; /* path-steps one plan may hold beyond its first ticket (0: no cap) */
} riga_server_config;

typedef struct riga_ticket_info {
  int32_t state;             /* 2 done, 3 shed, 4 failed */
  int32_t n_requests;
  int32_t batch_tickets;     /* tickets that shared its plan */
  int32_t reserved;
  double submit_s, start_s, end_s, latency_s; /* seconds since the server started */
  int64_t device_bytes;
  double path_steps;
  double device_start_s, device_end_s; /* its plan's device span (riga_plan_device_times); -1: none */
} riga_ticket_info;

typedef struct riga_server_stats {
  int64_t submitted_tickets, completed_tickets, completed_requests, shed_tickets, failed_tickets;
  int64_t batches, peak_bytes, bytes_in_use;
  double path_steps;
  int32_t queued_tickets, running_plans, max_concurrent;
  int32_t overlapped_plans; /* plans whose device span met that of a plan finished before them */
} riga_server_stats;

int riga_server_create(riga_ctx *ctx, const riga_server_config *cfg, riga_server **out,
                       riga_error *err);
/* Copies the requests; returns at once with a ticket id. */
int riga_server_submit(riga_server *server, const riga_request *reqs, int n, int precision,
                       uint64_t *ticket, riga_error *err);
/* Blocks up to timeout_s (< 0: forever).  0: done (results[n] filled); 1: still pending;
 * 2: shed; < 0: failed (err filled).  A finished ticket is released by this call. */
int riga_server_wait(riga_server *server, uint64_t ticket, double timeout_s, riga_result *results,
                     riga_ticket_info *info, riga_error *err);
int riga_server_stats_get(riga_server *server, riga_server_stats *out);
/* Finishes the running plans, sheds the queued tickets (code "shed"), wakes every
 * caller blocked in riga_server_wait and returns once they have left.  No call on
 * the server may start after destroy has begun. */
void riga_server_destroy(riga_server *server);

/* ---- replay (explain.rs:90-316): per-path records for paths [0, req->paths) ---- */
int riga_replay_paths(riga_ctx *ctx, const riga_request *req, int precision,
                      riga_path_record *records, riga_result *out, riga_error *err);

/* explain.rs:35-55 CouponPeriodRecord: one period of a path's coupon strip (nominal
 * amounts; pv = paid * discount_factor; explain.rs's coupon_by_period[period] == pv). */
typedef struct riga_coupon_row {
  int64_t period;
  double years;                 /* payment or period-end time; 0 after termination */
  int64_t qualified_steps, period_steps;
  double accrual, memory_paid, paid, memory_accrued, discount_factor, pv;
  int32_t knocked_out, pad;
} riga_coupon_row;

/* explain.rs:90-316 run_paths and :319-334 reconstruct_path: records of the
 * listed paths (path_ids NULL: all paths, n_ids ignored), in list order, and,
 * when ledger is not NULL, their coupon strips ledger[i * periods + period]
 * (periods = coupon_periods, 9 when unset).  A listed id >= paths gives
 * path_out_of_range. */
int riga_replay_ledger(riga_ctx *ctx, const riga_request *req, int precision, const uint64_t *path_ids,
                       uint64_t n_ids, riga_path_record *records, riga_coupon_row *ledger,
                       riga_error *err);

/* explain.rs:78-88 PathCubeQuery (the sample is returned through the arrays). */
typedef struct riga_worst_query {
  double percentile, threshold_worst_final;
  uint64_t selected, total, n_sample;
} riga_worst_query;

/* explain.rs:336-372 query_worst_percentile: every path's final worst-of, a
 * stable ascending sort, selected = ceil(total * percentile), the threshold and
 * the first min(selected, max_sample) paths: their ids, records and (when
 * sample_ledger is not NULL) coupon strips [max_sample][periods].  Percentile
 * outside [0, 1] gives invalid_percentile. */
int riga_query_worst_percentile(riga_ctx *ctx, const riga_request *req, int precision, double percentile,
                                uint64_t max_sample, riga_worst_query *out, uint64_t *sample_ids,
                                riga_path_record *sample_records, riga_coupon_row *sample_ledger,
                                riga_error *err);

/* ---- resident market surfaces (gpu.rs:280-327 resident_lv keeps the local-vol
 * rows and log-strike axes on the device across runs).  A calibrated grid is
 * copied to device memory once; requests then name it in riga_request.surfaces.
 * Destroy a surface only after every plan and ticket using it has finished. */
int riga_surface_create(riga_ctx *ctx, const riga_lv_grid *grid, riga_surface **out, riga_error *err);
void riga_surface_destroy(riga_surface *surface);

/* ---- local-vol calibration (riga-market/src/lib.rs:210-299 DupireLocalVol::calibrate)
 * Outputs: t_axis[n_maturities], log_s[grid_x], v[n_maturities * grid_x]. */
int riga_lv_calibrate(const riga_implied_grid *grid, double spot, double rate, int32_t grid_x,
                      double *t_axis, double *log_s, double *v, riga_error *err);
/* DupireLocalVol::sigma (riga-market/src/lib.rs:305-343) on a calibrated grid. */
double riga_lv_sigma(const riga_lv_grid *grid, double t, double s);

/* ---- diagnostics: Sobol/Owen integers and normals as the device computes them.
 * out_scrambled[p*n_dims + d] for points n = first_point + p (1-based Sobol index). */
int riga_debug_sobol(riga_ctx *ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point,
                     int32_t n_points, int32_t first_dim, int32_t n_dims,
                     uint32_t *out_unscrambled, uint32_t *out_scrambled, double *out_normal,
                     riga_error *err);

/* The FP64 tier's exp and log as its path kernels evaluate them (the Euler step,
 * the local-vol lookup, the inverse-normal tails, the ziggurat tests): out_exp[i] =
 * exp(x[i]), out_log[i] = log(x[i]), n <= 2^28.  No counterpart in the reference
 * (a diagnostic for the parity tests: the oracle builds the same functions). */
int riga_debug_exp_log(riga_ctx *ctx, const double *x, uint64_t n, double *out_exp, double *out_log,
                       riga_error *err);

/* The Halton ladder (qmc.rs:94-153) as the path kernels evaluate it, for points
 * n = first_point + p (1-based, qmc.rs:121) and dims first_dim + d (the base and
 * permutation of dim % ladder_dims): out_digits[(p*n_dims + d)*32 + k] is the
 * permuted k-th base-b digit perm[digit_k(n)] (least significant first,
 * 0xFFFFFFFF after the last), out_uniform the radical inverse and out_normal its
 * FP64 inverse normal.  ladder_dims <= 7762.  No counterpart in the reference (a
 * diagnostic for the parity tests). */
int riga_debug_halton(riga_ctx *ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point, int32_t n_points,
                      int32_t first_dim, int32_t n_dims, uint32_t *out_digits, double *out_uniform,
                      double *out_normal, riga_error *err);

/* The FP32 tier's Sobol ladder as its path kernels run it, for points
 * n = first_point + p (1-based) and dims node*factors + f, node < n_nodes:
 * out_scrambled[p][dim] is the integer the tier turns into its uniform -- the
 * top 24 Owen levels with the low byte 0x80, or all 32 levels where the top 10
 * scrambled bits are all 0 or all 1 (u within 2^-10 of 0 or 1) -- and
 * out_normal[p][dim] its FP32 inverse normal.  owen_table bit 0 reads the context's
 * coordinate table for the seed (points below 2^22, dimensions within ladder_dims), as priced batches do;
 * bit 1 hashes all 32 levels of every coordinate (the fp32_full_ladder option).
 * No counterpart in the reference (a diagnostic for the parity tests). */
int riga_debug_sobol_f32(riga_ctx *ctx, uint64_t seed, int32_t ladder_dims, uint64_t first_point,
                         int32_t n_points, int32_t n_nodes, int32_t factors, int32_t owen_table,
                         uint32_t *out_scrambled, float *out_normal, riga_error *err);

/* ---- PathCube on demand (pathcube.rs:47-76 generate_correlated_block, :146-187
 * gbm_evolve): a correlated block y = L z from ChaCha8Rng::seed_from_u64(seed),
 * `factors` StandardNormals per (path, time) cell in stream order, as floats, and
 * the GBM spot cube exp(log s0 + sum((r - q - vol^2/2) dt + vol sqrt(dt) y)) as
 * floats, both [paths][times][factors].  out_block / out_cube may be NULL.  With
 * in_block set, only gbm_evolve of that block runs (into out_cube). */
int riga_path_cube(riga_ctx *ctx, uint64_t seed, uint64_t paths, int32_t times, int32_t factors,
                   const double *correlation, const double *s0, const double *vols, double rate,
                   const double *dividend_yields, double dt, const float *in_block, float *out_block,
                   float *out_cube, riga_error *err);

/* ---- the reference's device module (crates/riga-engine/src/gpu.rs), for its callers
 * (examples/gpu_bench.rs, examples/gpu_lv_bench.rs).  FP32, synchronous like the
 * reference; resident handles keep their inputs in device memory. */
typedef struct riga_resident_worst_of riga_resident_worst_of;
typedef struct riga_resident_lv riga_resident_lv;

/* gpu.rs:171-205, :515-517 worst_of: out[i] = a[i] < b[i] ? a[i] : b[i]. */
int riga_worst_of(riga_ctx *ctx, const float *a, const float *b, uint64_t n, float *out, riga_error *err);
/* gpu.rs:331-357 resident_worst_of and ResidentWorstOf::{reduce, download, synchronize} (gpu.rs:362-405). */
int riga_resident_worst_of_create(riga_ctx *ctx, const float *a, const float *b, uint64_t n,
                                  riga_resident_worst_of **out, riga_error *err);
int riga_resident_worst_of_reduce(riga_resident_worst_of *h, riga_error *err);
int riga_resident_worst_of_download(riga_resident_worst_of *h, float *out, riga_error *err);
int riga_resident_worst_of_synchronize(riga_resident_worst_of *h, riga_error *err);
void riga_resident_worst_of_destroy(riga_resident_worst_of *h);

/* gpu.rs:207-278, :538-566 lv_terminal_spots (kernel gpu.rs:37-75): two factors, shocks
 * normals[paths][steps][2], local-vol rows[2][steps][gx] (2 <= gx <= 3072), evenly spaced
 * log-spot axes ls_axis[2][gx] (only their end knots are read, as in the reference);
 * outputs terminal spots [paths][2] and the first step both names
 * stand at or above call_barrier x reference (or -1). */
int riga_lv_terminal_spots(riga_ctx *ctx, const float *normals, const float *rows, const float *ls_axis,
                           int32_t gx, const float *s0, const float *references, float call_barrier,
                           int64_t paths, int64_t steps, float dt, float rate, const float *dividends,
                           float *out_terminal, int32_t *out_ko, riga_error *err);
/* gpu.rs:280-328 resident_lv and ResidentLv::{run, download, synchronize} (gpu.rs:407-489). */
int riga_resident_lv_create(riga_ctx *ctx, const float *normals, const float *rows, const float *ls_axis,
                            int32_t gx, int64_t paths, int64_t steps, float dt, riga_resident_lv **out,
                            riga_error *err);
int riga_resident_lv_run(riga_resident_lv *h, const float *s0, const float *references, float call_barrier,
                         float rate, const float *dividends, riga_error *err);
int riga_resident_lv_download(riga_resident_lv *h, float *out_terminal, int32_t *out_ko, riga_error *err);
int riga_resident_lv_synchronize(riga_resident_lv *h, riga_error *err);
void riga_resident_lv_destroy(riga_resident_lv *h);

/* ---- diagnostics: the Monte Carlo lane's stream parse.  out_starts[p] = index of
 * the u64 word of ChaCha8Rng::seed_from_u64(seed) at which the StandardNormal draw
 * number p * draws_per_path begins (the first draw of path p, note.rs:58,124-126). */
int riga_debug_mc_starts(riga_ctx *ctx, uint64_t seed, uint64_t draws_per_path, uint64_t n_paths,
                         uint64_t *out_starts, riga_error *err);

#ifdef __cplusplus
}
#endif
#endif
