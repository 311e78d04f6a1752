/*
 * riga_oracle.h -- CPU FP64 oracle for the Phoenix path-simulation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against.  It is loaded by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg, and by nothing else.  The product library
 * (riga_flow_amd/csrc) never links or calls it.
 *
 * It restates, in plain C, the reference engine's algorithm, generalised from
 * the reference's hard-wired two factors to F factors (dimension = step*F + f):
 *   crates/riga-engine/src/qmc.rs            Sobol/Owen/probit/Halton/bridge
 *   crates/riga-market/src/correlation.rs    validate/Jacobi/PSD clamp/Cholesky
 *   crates/riga-market/src/lib.rs:28-380     Dupire calibrate/sigma/row
 *   crates/riga-engine/src/note.rs:38-373    price_full_phoenix
 *   crates/riga-engine/src/lib.rs:209-343    validate, build_local_vol
 *   crates/riga-engine/src/explain.rs:90-316 per-path records
 * Every function cites the reference lines it follows in riga_oracle.c.
 *
 * Compiled with -O2 -ffp-contract=off so that no multiply-add is fused; the
 * reference (Rust) never contracts floating-point expressions.
 */
#ifndef RIGA_ORACLE_H
#define RIGA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RO_REFERENCE_MAX_DIM 512 /* qmc.rs:9 */
#define RO_TRACE_CAP 64          /* note.rs:57 */

/* ---- QMC primitives (qmc.rs) ---- */
double ro_probit(double p);
double ro_radical_inverse(uint64_t index, uint32_t base);
/* First `count` primitive polynomials of degree >= 2 (qmc.rs:347-363). Returns count written. */
int ro_primitive_polynomials(int count, uint32_t *out);
void ro_direction_numbers(uint32_t mask, uint32_t out[32]);
uint64_t ro_mix64(uint64_t z);
uint32_t ro_owen_scramble(uint32_t x, uint64_t dim, uint64_t seed);

/* Sobol ladder; `max_dim` is 512 for reference-compatible wrap (qmc.rs:9,249). */
typedef struct {
  int dimension;
  uint32_t *v; /* dimension x 32 */
  uint64_t seed;
} ro_sobol;
int ro_sobol_init(ro_sobol *s, int dimension, uint64_t seed, int max_dim);
void ro_sobol_free(ro_sobol *s);
uint32_t ro_sobol_unscrambled(const ro_sobol *s, uint64_t n, int dim);
uint32_t ro_sobol_scrambled(const ro_sobol *s, uint64_t n, int dim);
double ro_sobol_uniform(const ro_sobol *s, uint64_t n, int dim);
double ro_sobol_normal(const ro_sobol *s, uint64_t point, int dim);

typedef struct {
  int dimension;
  uint32_t *bases;
  uint32_t **perms;
} ro_halton;
int ro_halton_init(ro_halton *h, int dimension, uint64_t seed, int max_dim);
void ro_halton_free(ro_halton *h);
uint32_t ro_halton_digit_perm(const ro_halton *h, int dim_index, uint32_t digit);
double ro_halton_uniform(const ro_halton *h, uint64_t n, int dim);
double ro_halton_normal(const ro_halton *h, uint64_t point, int dim);

/* qmc.rs:158-191 */
void ro_brownian_increments(const double *normals, int n, double dt, double *out);

/* ---- correlation.rs ---- */
/* 0 = ok; otherwise an error message is written to err (code "invalid_market"). */
int ro_corr_validate(int n, const double *c, char *err, size_t errlen);
double ro_min_eigenvalue(int n, const double *c);
void ro_psd_clamp(int n, const double *c, double floor_, double *out);
int ro_cholesky(int n, const double *c, double *l_out, char *err, size_t errlen);
void ro_chol_apply(int n, const double *l, const double *z, double *y);

/* ---- riga-market/src/lib.rs ---- */
typedef struct {
  int n_maturities, n_strikes;
  const double *maturities; /* years, ascending */
  const double *strikes;    /* absolute, ascending */
  const double *vols;       /* [n_maturities][n_strikes] */
} ro_implied_grid;

typedef struct {
  int n_t, n_x;
  double *t_axis; /* n_t */
  double *log_s;  /* n_x */
  double *v;      /* [n_t][n_x] */
} ro_lv;

int ro_lv_calibrate(const ro_implied_grid *g, double spot, double rate, int grid_x,
                    ro_lv *out, char *err, size_t errlen);
/* Resample onto `n_times` rows with row(t) semantics (riga-market/src/lib.rs:346-364). */
int ro_lv_resample(const ro_lv *src, int n_times, const double *times, ro_lv *out);
int ro_lv_copy(int n_t, int n_x, const double *t_axis, const double *log_s, const double *v,
               ro_lv *out);
void ro_lv_free(ro_lv *lv);
double ro_lv_sigma(const ro_lv *lv, double t, double s);
void ro_lv_row(const ro_lv *lv, double t, double *out);
double ro_interp_linear(int n, const double *xs, const double *ys, double x);

/* ---- pricer (note.rs) ---- */
enum { RO_VR_NONE = 0, RO_VR_HALTON = 1, RO_VR_SOBOL = 2 };
enum { RO_TRACE_OFF = 0, RO_TRACE_AGGREGATED = 1, RO_TRACE_SAMPLED = 2, RO_TRACE_FULL = 3 };

typedef struct {
  int factors;
  /* market (riga-core MarketSnapshot) */
  const double *spots, *vols, *dividend_yields, *correlation; /* correlation F x F */
  double rate;
  /* trade (riga-core PhoenixTrade); has_* = 0 means the Option is None */
  const double *initial_spots;
  double maturity_years;
  int64_t steps;
  double strike, knock_in_barrier;
  int has_ko;  double knock_out_barrier;
  int has_cb;  double coupon_barrier;
  int has_cr;  double coupon_rate;
  int memory_coupon;
  int64_t coupon_periods;  /* < 0 means None (defaults to 9) */
  int fixed_first_coupon;
  int64_t call_start_step; /* < 0 means None (defaults to 0) */
  /* config (riga-core SimulationConfig) */
  uint64_t paths, path_block_size, seed;
  int trace_level;
  int vr_kind;       /* RO_VR_* */
  int vr_dimension;  /* requested QMC dimension */
  int max_dim;       /* ladder dimension cap: 512 = reference */
  /* local vol: NULL = flat vols, else `factors` calibrated surfaces */
  const ro_lv *lv;
} ro_request;

typedef struct {
  double pv, mc_error, ki_probability;
  double leg_coupon, leg_memory_coupon, leg_put, leg_funding, leg_total;
  uint64_t paths, ko_count, ki_count, put_eligible_count, coupon_path_count;
  double ko_mean, ko_var, ki_mean, ki_var, cpn_mean, cpn_var;
  double funding_sum, put_sum, coupon_sum, memory_sum; /* raw sequential sums */
  int n_trace;
  int64_t trace_path[RO_TRACE_CAP];
  int64_t trace_step[RO_TRACE_CAP]; /* KO step; time = (step+1)*dt */
} ro_result;

/* Per-path record (explain.rs:58-74 subset) for replay/event parity. */
typedef struct {
  int32_t ko_step;  /* -1 if the note survived */
  int32_t hit_ki;
  double final_worst, funding, put, coupon_pv, memory_pv, pv;
} ro_path_record;

/* Monte Carlo lane (VarianceReduction::None, note.rs:58,124-126): the reference
 * draws from rand_chacha 0.9.0 ChaCha8Rng::seed_from_u64(seed) through rand_distr
 * 0.5.1 StandardNormal (rand 0.9.5).  Those crates are not vendored; this is a
 * restatement of their published algorithms:
 *   seed_from_u64   rand_core 0.9 SeedableRng default: PCG32 expansion of the
 *                   u64 into the 32-byte key
 *   ChaCha8         8 rounds (4 double rounds), 64-bit block counter in words
 *                   12-13 starting at 0, 64-bit stream 0 in words 14-15; the
 *                   output is the blocks' words in order
 *   next_u64        two consecutive output words, low word first
 *   StandardNormal  256-layer ziggurat (Marsaglia-Tsang / Doornik) with the
 *                   tables of rand's ziggurat_tables.py (R = 3.6541528853610088,
 *                   V = 0.00492867323399), printed to 18 decimals in the crate
 *   random::<f64>   (u64 >> 11) * 2^-53;  Open01: 52-bit fraction in [1,2) - (1 - eps/2)
 * Pinned by the reference's golden PriceResult (fixtures/termsheet1.golden.json,
 * the full result within 1e-12 at 30,000 paths, seed 1729) and by the ChaCha
 * core's RFC 7539 known answers (20 rounds). */
void ro_chacha_block(const uint32_t in[16], int double_rounds, uint32_t out[16]);
void ro_seed_from_u64(uint64_t seed, uint32_t key[8]);
/* The k-th u64 of the stream of ChaCha8Rng::seed_from_u64(seed). */
uint64_t ro_rng_u64_at(uint64_t seed, uint64_t k);
void ro_zig_tables(double x[257], double f[257]);
/* Standard normals drawn from the stream starting at u64 index `start`; writes n
 * values and returns the u64 index after the last draw. */
uint64_t ro_std_normals(uint64_t seed, uint64_t start, uint64_t n, double *out);
/* u64 index at which each of n consecutive draws starts (draw 0 starts at `start`). */
void ro_draw_starts(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out);

/* One period of a path's coupon strip, explain.rs:35-55 CouponPeriodRecord. */
typedef struct {
  int64_t period;
  double years;
  int64_t qualified_steps, period_steps;
  double accrual, memory_paid, paid, memory_accrued, discount_factor, pv;
  int32_t knocked_out, pad;
} ro_coupon_row;

/* pathcube.rs:47-76 generate_correlated_block: [paths][times][factors] floats of
 * L z, z drawn `factors` at a time from ChaCha8Rng::seed_from_u64(seed). */
void ro_correlated_block(uint64_t seed, uint64_t paths, int64_t times, int factors, const double *chol,
                         float *out);
/* pathcube.rs:146-187 gbm_evolve: spot cube from a block (same layout). */
void ro_gbm_evolve(const double *s0, const double *vols, double rate, const double *dividend_yields,
                   double dt, uint64_t paths, int64_t times, int factors, const float *block, float *out);

/* Validation mirroring lib.rs:209-321 generalised to F factors.
 * Returns 0 or writes code/message. */
int ro_validate(const ro_request *r, char *code, size_t codelen, char *msg, size_t msglen);

/* Price paths [path_begin, path_end) -- the full run is [0, paths).  Aggregates
 * in `out` cover only that range (path-order sequential, as the reference).
 * `records` (optional) receives one record per path in the range. */
int ro_price(const ro_request *r, uint64_t path_begin, uint64_t path_end, ro_result *out,
             ro_path_record *records, ro_coupon_row *ledger, char *code, size_t codelen, char *msg,
             size_t msglen);
/* `ledger` (optional): periods rows per path in the range, explain.rs:146-262. */

/* Generate the correlated, bridged and 1/sqrt(dt)-scaled shock cube of one path:
 * out[f*steps + t].  Exposed for bridge/ladder parity tests.  On the Monte Carlo
 * lane: the path's correlated draws chol * z per step (draws of earlier paths
 * skipped, as the reference's single stream consumes them). */
int ro_path_shocks(const ro_request *r, uint64_t path_id, double *out);

/* exp and log inside the path: 0 = glibc (the reference's), 1 = the device FP64
 * tier's table-driven pair (riga_flow_amd/csrc/exp_log.h); per calling thread. */
void ro_set_device_math(int on);
int ro_get_device_math(void);
double ro_shared_exp(double x);
double ro_shared_log(double x);

#ifdef __cplusplus
}
#endif
#endif
