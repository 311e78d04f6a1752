// phoenix_layout.h -- plain-old-data records shared by the host planner and the
// fused Phoenix kernel.  Everything here lives in one device arena per batch.
//
// Vocabulary follows the reference engine: an *item* is one trade priced under
// S scenarios that share its seed, schedule, correlation and Sobol points (the
// reference's CRN scenario set, crates/riga-risk/src/lib.rs:117-186 and
// crates/riga-engine/src/dispatch.rs:23-70); a *path-step* is one (trade,
// scenario, path) advanced one observation date across all F factors.
#pragma once
#include <stdint.h>

namespace riga {

constexpr int kMaxFactors = 5;      // per-trade underlyings handled by the kernels
constexpr int kBridgeStack = 14;    // pending bridge intervals (host-checked per trade)
constexpr int kStackPooled = 8;     // bridge stack positions a lane may keep in its workgroup's HBM slot
constexpr int kTraceCap = 64;       // note.rs:57 TRACE_EVENT_CAP
// How a launch's items get their normals; the path kernel is compiled once per mode.
constexpr int kLadderDirect = 0;      // Sobol (Owen) or Halton points generated per lane
constexpr int kLadderTable = 1;       // read from the batch's shared [dim][path] normal table
constexpr int kLadderMonteCarlo = 2;  // ChaCha8 + ziggurat stream (VarianceReduction::None)
constexpr int kMaxLvX = 256;        // log-spot points per local-vol row staged in LDS
// 22 levels: 16 MiB per raw dimension (u32 entries).  Measured on the 3-asset 756-step
// batches: FP32 ten-scenario 38.96 / 37.06 / 35.55 ms, FP64 plain 14.26 / 13.74 / 13.01 ms
// for 16 / 18 / 20 bits; with the gathers non-temporal, 34.00 / 12.77 ms at 20 bits,
// 32.70 / 12.49 ms at 21 and 31.68 / 12.16 ms at 22; with whole-coordinate entries (indexed by prefix)
// (no level hashed below 2^22 points), 28.65 / 8.73 ms.
#ifndef RIGA_OWEN_PREFIX_BITS
#define RIGA_OWEN_PREFIX_BITS 22
#endif
constexpr int kOwenPrefixBits = RIGA_OWEN_PREFIX_BITS;  // coordinate tables hold points below 2^bits
static_assert(kOwenPrefixBits >= 8 && kOwenPrefixBits <= 24, "coordinate table width out of range");
// an entry is the scrambled 32-bit Sobol coordinate of one point in one dimension
typedef uint32_t OwenEntry;
constexpr int kBlock = 256;         // threads per workgroup (4 waves of 64)

enum ItemFlags : uint32_t {
  kMemoryCoupon = 1u << 0,   // PhoenixTrade::memory_coupon
  kFixedFirst = 1u << 1,     // PhoenixTrade::fixed_first_coupon
  kLocalVol = 1u << 2,       // PriceRequest::local_vol present
  kTrace = 1u << 3,          // trace_level != Off
  kHalton = 1u << 4,         // VarianceReduction::Halton ladder (else Sobol)
  kRecords = 1u << 5,        // write one PathRecord per (scenario, path)
  kOwenTable = 1u << 6,      // top kOwenPrefixBits Owen levels read from a per-dimension table
  kNormalTable = 1u << 7,    // ladder normals read from the batch's shared [dim][path] table
  kLvUniform = 1u << 8,      // FP32 tier: every surface has evenly spaced knots (slope rows)
  kMonteCarlo = 1u << 9,     // VarianceReduction::None: ChaCha8 + ziggurat normals (note.rs:124-126)
  kFullLadder = 1u << 10,    // FP32 tier: all 32 Owen levels for every coordinate (the FP64 tier's integers)
};

// Monte Carlo lane (ChaCha8 stream + StandardNormal ziggurat).
constexpr int kZigLen = 257;        // ziggurat table entries (256 layers + the x = 0 end)
// Stream words buffered per lane in LDS: three ChaCha blocks, or two for the
// 10-scenario lanes (whose local-vol rows and partials leave less LDS: two
// workgroups must still fit a CU).
constexpr int mc_ring_slots(int smax) { return smax >= 10 ? 16 : 24; }
constexpr int kMcEntries = 16;      // chunk entry offsets tracked by the stream parse
constexpr int kZigMaxLen = 250;     // longest draw accepted, in words (never reached in practice)

// Stream parse, one record per chunk of `chunk_words` words: for each entry
// offset e (the first word of the chunk that starts a draw, given how far the
// previous chunk's last draw reached), the number of draws that start in the
// chunk and the entry offset of the next chunk.
enum McFlags : uint32_t {
  kMcConstExit = 1u << 0,  // the next chunk's entry does not depend on this chunk's entry
  kMcErrLong = 1u << 1,    // a draw longer than kZigMaxLen words
  kMcErrSlow = 1u << 2,    // more slow draws than the chunk's list holds
  kMcErrExit = 1u << 3,    // a draw reaching kMcEntries or more words into the next chunk
  kMcErrShort = 1u << 4,   // the parsed stream ends before the last path's last draw
};
struct McChunk {
  int32_t count[kMcEntries];
  uint8_t exit[kMcEntries];
  int32_t nslow;           // slow words listed: (offset << 8) | words of the draw that
                           // would start there, in offset order, with that draw's value
  uint32_t flags;
};

// Per observation date, uniform across all paths of an item.
struct StepInfo {
  double accrual_inc;  // coupon_rate / period_steps(period), note.rs:159-162
  double df;           // exp(-r * (step+1) * dt), note.rs:189,210
  int32_t period;      // min(step*P/T, P-1), note.rs:151
  int32_t flags;       // bit0 period end (note.rs:197-202), bit1 call active (note.rs:166),
                       // bit2 new period: accrual restarts (note.rs:152-155)
  int32_t n_desc;      // bridge nodes first needed at this step (streaming order)
  int32_t depth;       // bridge stack depth before this step's descents (host check)
};

// One streaming-bridge node (qmc.rs:173-191): pre-order index and its
// conditional standard deviation sqrt((hi-mid)(mid-lo)/(hi-lo) * dt).
struct BridgeNode {
  double sd;
  int32_t node;
  float sd32;  // FP32 tier copy of sd
};

// Local-vol time bracket per (step, factor): riga-market/src/lib.rs:324-339
struct LvBracket {
  double wt;
  int32_t i0, i1;
};

// One local-vol surface (DupireLocalVol fields, riga-market/src/lib.rs:119-123)
// stored as byte offsets into the batch arena.
struct LvSurface {
  int64_t off_t;     // n_t tenors
  int64_t off_x;     // n_x log-spot knots
  int64_t off_v;     // n_t * n_x vols
  int32_t n_t, n_x;
  int32_t factor;    // which factor of the item this surface drives
  int32_t pad;
  double inv_dx;     // (n_x - 1) / (log_s[n_x-1] - log_s[0]): bracket guess (exact knot
                     // position when the item has kLvUniform)
  double x0;         // log_s[0]
  double rel_off;    // (ln initial_spot[factor] - x0) * inv_dx: knot position of ln(S/S0) = 0
  const double* dev_x;  // resident surface: knots and vols in device memory
  const double* dev_v;  // (null: at off_x / off_v in the arena)
};

// FP32 tier, per (scenario, factor): the Euler step's constants.
struct SfF32 {
  float a;          // local vol: 1/dx of the surface's evenly spaced knots; flat vol: vol
  float b;          // local vol: knot position of ln(S/S0) = 0
  float c;          // local vol: (rate - q) dt; flat vol: (rate - q - vol^2/2) dt
  float tmax;       // largest float below n_x - 1 (knot positions are clamped to [0, tmax])
  int32_t lds_off;  // byte offset of the surface's slope rows in the local-vol LDS area (the kernel adds the area's offset in dynamic LDS)
  int32_t pad[3];
};

struct ScenarioDesc {
  double spot[kMaxFactors];   // MarketSnapshot::spots (bumped per scenario)
  double rq[kMaxFactors];     // rate - dividend_yield[f]
  double drift[kMaxFactors];  // flat vol: (rate - q - 0.5 vol^2) dt
  double vsd[kMaxFactors];    // flat vol: vol * sqrt(dt)
  double ln_rel[kMaxFactors]; // FP32 log-state tier: ln(spot / initial_spot)
  int32_t lv[kMaxFactors];    // surface index per factor (local vol)
};

struct ItemDesc {
  int32_t F, S, steps, periods;
  int32_t call_start, flags, ladder_dims, nbits;
  uint64_t paths, seed;
  uint64_t path_begin;        // first path id (for sharded/partial runs)
  double dt, sqrt_dt, inv_sqrt_dt, root_sd;
  double strike, ki, ko, cb, coupon_rate, put_df;
  double initial_spot[kMaxFactors];
  double ln_s0[kMaxFactors];  // FP32 log-state tier: ln(initial_spot), ln(KO), ln(coupon barrier)
  double ln_ko, ln_cb;
  double chol[kMaxFactors * kMaxFactors];  // lower triangular, row-major F x F
  int64_t off_nodes;     // BridgeNode[steps - 1] in streaming order
  int64_t off_steps;     // StepInfo[steps]
  int64_t off_lvb;       // LvBracket[steps * F]
  int64_t off_dirs;      // uint32: Sobol direction rows, ladder_dims x 32
                         // (all off_* are byte offsets into the batch arena)
  int64_t off_scen;      // ScenarioDesc[S]
  int64_t off_halton;    // uint32: Halton [ladder_dims] bases, [ladder_dims] perm offsets, perms
  int64_t off_surf;      // LvSurface[n_surf] distinct surfaces used by the scenarios
  int32_t n_surf;
  int32_t max_nx;
  int32_t stack_levels;  // bridge levels below the top held in LDS (launch-wide)
  int32_t stack_pooled;  // further stack positions (the bottom and top ones) in an HBM slot
  int64_t partial_base;  // first Partial slot: [scenario][block_in_item]
  int64_t record_base;   // first PathRecord slot: [scenario][path]
  int32_t owen_dims;     // raw dimensions the Owen coordinate table covers (all of the item's when kOwenTable)
  int32_t n_blocks;
  int32_t paths_per_block;
  int32_t ladder_rows;   // ladder rows uploaded (min(ladder_dims, F*steps))
  const OwenEntry* owen_tab; // device table [dims][2^kOwenPrefixBits]: scrambled coordinates [dim][point]
  // FP32 tier: constants pre-rounded on the host, so the path loop converts nothing
  float chol32[kMaxFactors * kMaxFactors];
  float neg_half_dt, ln_ko32, ln_cb32, root_sd32;
  int64_t off_sf32;      // SfF32[S * F]
  int64_t off_wt32;      // float[steps * F]: local-vol tenor weight per (step, factor)
  uint64_t ztable;       // device address of the shared normal table (Real[dims][zt_stride])
  int64_t zt_stride;     // paths per table row
  // Monte Carlo lane: the ChaCha8 key of seed_from_u64(seed), the first stream
  // word of every path (from the parse) and the ziggurat tables in the arena
  uint32_t mc_key[8];
  const uint64_t* mc_start;  // [path]: stream word of the path's first draw
  int64_t off_zig;           // double[2 * kZigLen]: x_i then f(x_i)
  const McChunk* mc_chunks;  // the parse of this item's stream
  const uint32_t* mc_slow;   // [chunk][mc_slow_cap] slow words
  const double* mc_val;      // [chunk][mc_slow_cap] their draws
  int64_t mc_n_chunks;
  int32_t mc_chunk_words, mc_slow_cap;
  // replay: the paths priced (null: path_begin + [0, paths)) and the worst-of
  // per (step, path) for the coupon strips (null: not written)
  const uint64_t* path_list;
  double* series;        // [step][path]
  // bottom bridge levels (stack_pooled > 0): the launch's pool of HBM slots, one
  // per workgroup resident at a time, [slot][level][factor][lane] of Real, and
  // its claim words (pool_n flags, then a count of claims that found no slot)
  unsigned char* stack_pool;
  uint32_t* pool_flags;
  int32_t pool_n;
  int32_t stack_pooled_low;  // of the stack_pooled positions, how many are the bottom ones (the rest the top)
};

// Double-double accumulator (hi + lo).
struct DD {
  double hi, lo;
};

// Reduction record: one per (item, scenario, block) as a partial, and one per
// (item, scenario) once merged in block order by the finalize kernel.
enum Quantity {
  kQFunding = 0, kQPut, kQCq, kQMq,    // note.rs:254-257 sums (coupon_quote - memory_quote, memory_quote)
  kQPv1, kQPv2,                        // pv moments: sum x, sum x^2   (note.rs:258)
  kQKo1, kQKo2,                        // KO funding moments           (note.rs:223-225)
  kQKi1, kQKi2,                        // KI put moments               (note.rs:242-244)
  kQCp1, kQCp2,                        // coupon nominal moments       (note.rs:246-249)
  kNumQ
};
enum Count { kCN = 0, kCKo, kCKi, kCPutEligible, kCCouponPaths, kNumC };

struct Partial {
  DD q[kNumQ];
  uint64_t c[kNumC];
  int32_t n_events, pad;
  int32_t ev_path_lo[kTraceCap];      // path index within the item, first 64 KO in path order
  int32_t ev_step[kTraceCap];
};

// Static LDS of the kernel: the wave sums of one scenario at a time (the block
// merges them into its Partial before the next scenario reuses the slots).
constexpr int kWavesPerBlock = kBlock / 64;
struct BlockShared {
  DD acc[kWavesPerBlock][kNumQ];     // lane 0 of each wave writes its wave's sums
  uint64_t cnt[kWavesPerBlock][kNumC];
  int32_t wave_ko[kWavesPerBlock];   // KO events per wave (trace)
};

// One period of a path's coupon strip (explain.rs:35-55 CouponPeriodRecord);
// the layout of riga_coupon_row.
struct CouponRowDev {
  int64_t period;
  double years;
  int64_t qualified_steps, period_steps;
  double accrual, memory_paid, paid, memory_accrued, discount_factor, pv;
  int32_t knocked_out, pad;
};

// Per-path record (explain.rs:58-74 subset) written only for replay queries.
struct PathRecordDev {
  int32_t ko_step;
  int32_t hit_ki;
  double final_worst, funding, put, coupon_pv, memory_pv, pv;
};

}  // namespace riga
