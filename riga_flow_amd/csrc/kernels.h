// kernels.h -- host-callable launchers for the gfx950 kernels (one translation
// unit per precision tier: kernels_f64.hip is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include "phoenix_layout.h"

namespace riga {

// Largest number of CRN scenarios one lane carries, by kernel variant.
constexpr int kScenarioBuckets[3] = {1, 4, 10};

struct PhoenixLaunch {
  int F;       // factors (1..kMaxFactors)
  int smax;    // scenario bucket (1, 4 or 10)
  int mode;    // kLadderDirect, kLadderTable or kLadderMonteCarlo
  int blocks;  // grid size: one workgroup per unit
  int threads; // lanes per workgroup: 64, 128 or 256
  size_t lds;  // dynamic LDS bytes (bridge stack and local-vol rows)
  const ItemDesc* items;
  const int32_t* block_item;  // (item, chunk) units, longest items first
  int n_units;                // entries of block_item
  uint32_t* queue;            // the launch's unit counter, zero at launch
  const unsigned char* arena;
  Partial* partials;
  PathRecordDev* records;
};

hipError_t launch_phoenix_f64(const PhoenixLaunch& l, hipStream_t stream);
// Resident workgroups per CU of the launch's kernel at its shape and dynamic LDS.
hipError_t phoenix_blocks_per_cu_f64(const PhoenixLaunch& l, int* blocks);
hipError_t phoenix_blocks_per_cu_f32(const PhoenixLaunch& l, int* blocks);
// Every path-kernel variant dispatched once with no work (private memory set up).
hipError_t warm_path_kernels_f64(uint32_t* queue, hipStream_t stream);
hipError_t warm_path_kernels_f32(uint32_t* queue, hipStream_t stream);
// Zero n unit counters spaced kQueueStride words apart (a kernel, so a captured
// graph records no memset node).
constexpr int kQueueStride = 16;
hipError_t launch_queue_reset(uint32_t* queues, int n, hipStream_t stream);
hipError_t launch_debug_sobol_f32(const uint32_t* dirs, int ladder_dims, const OwenEntry* owen_tab, uint64_t seed,
                                  uint64_t first_point, int n_points, int n_nodes, int factors, int nbits,
                                  bool full_ladder, uint32_t* xs_out, float* z_out, hipStream_t stream);
hipError_t launch_phoenix_f32(const PhoenixLaunch& l, hipStream_t stream);

// Shared normal table z[dim][path] (Real = double for FP64, float for FP32); the
// Owen coordinate table (may be null) covers dimensions [0, owen_dims).
hipError_t launch_normal_table_f64(const uint32_t* dirs, const OwenEntry* owen_tab, int owen_dims, uint64_t seed,
                                   int dims, int n_paths, int nbits, double* out, hipStream_t stream);
hipError_t launch_normal_table_f32(const uint32_t* dirs, const OwenEntry* owen_tab, int owen_dims, uint64_t seed,
                                   int dims, int n_paths, int nbits, float* out, hipStream_t stream);

// Monte Carlo lane stream parse (kernels_mc.hip) for one (seed, draws per path).
struct McParseArgs {
  const uint32_t* key;   // ChaCha8 key words (device)
  const double* zig;     // ziggurat tables: x[kZigLen] then f[kZigLen] (device)
  int chunk_words;       // words per chunk, a multiple of 8
  int slow_cap;          // slow-word list capacity per chunk
  int64_t n_chunks;
  uint32_t* slow;        // [n_chunks][slow_cap]
  double* slow_val;      // [n_chunks][slow_cap]
  McChunk* chunks;       // [n_chunks]
};
struct McStartsArgs {
  const McChunk* chunks;
  const uint32_t* slow;
  int slow_cap, chunk_words;
  int64_t n_chunks;
  uint8_t* entry;        // [n_chunks] resolved entry offset
  uint32_t* cnt;         // [n_chunks] draws starting in the chunk
  uint64_t* block_off;   // [ceil(n_chunks / 256)] block sums, then their exclusive prefix sums
  uint64_t dpp;          // draws per path (steps * factors)
  uint64_t paths;        // paths priced from the stream: [0, paths)
  uint64_t* starts;      // [paths]
  uint32_t* status;      // [2]: paths found, error flags
};
hipError_t launch_mc_parse(const McParseArgs& p, const McStartsArgs& s, hipStream_t stream);

// PathCube generation from a parsed stream (draws per path = times * factors),
// or gbm_evolve of `in_block` when it is set.  Outputs [paths][times][factors].
struct CubeArgs {
  const uint32_t* key;
  const double* zig;
  const uint64_t* starts;
  const McChunk* chunks;
  const uint32_t* slow;
  const double* slow_val;
  int64_t n_chunks;
  int chunk_words, slow_cap;
  uint64_t paths;
  int times;
  double chol[kMaxFactors * kMaxFactors];
  double ln_s0[kMaxFactors], drift[kMaxFactors], diffusion[kMaxFactors];
  const float* in_block;
  float* block;  // null: not written
  float* cube;   // null: not written
};
hipError_t launch_cube(const CubeArgs& a, int factors, hipStream_t stream);

// The reference device module's kernels (gpu.rs:27-75), FP32.  The local-vol
// propagation's constants are formed on the host per factor: the knot position
// of a log-spot x is x * knot_scale + knot_offset on evenly spaced knots, the
// Euler drift is (rate - q) dt, and the KO level ln(call_barrier * reference).
struct LvPropagateArgs {
  const float* normals;  // [paths][steps][2]
  const float* rows;     // [2][steps][gx]
  float* out;            // [paths][2] terminal spots
  int* ko_out;           // [paths] first KO step or -1
  float x0[2];           // ln s0
  float knot_scale[2], knot_offset[2];
  float drift[2];        // (rate - q) dt
  float ln_ko[2];        // ln(call_barrier * reference)
  float sqrt_dt, neg_half_dt;
  int paths, steps, gx;
};
// Largest knot count the propagation stages in LDS.
constexpr int kLvMaxKnots = 3072;
hipError_t launch_worst_of(const float* a, const float* b, float* o, uint64_t n, hipStream_t stream);
hipError_t launch_lv_propagate(const LvPropagateArgs& a, hipStream_t stream);

// Owen coordinate table: out[dim][n] = the scrambled Sobol coordinate of point n < 2^kOwenPrefixBits
// from the dimension's direction row dirs[dim][32], dims [0, dims).
hipError_t launch_owen_table(uint64_t seed, int dims, const uint32_t* dirs, OwenEntry* out, hipStream_t stream);

// Coupon strips [path][period] of a replayed item from its worst-of series.
hipError_t launch_ledger(const ItemDesc* items, int item, uint64_t paths, const unsigned char* arena,
                         CouponRowDev* rows, hipStream_t stream);

hipError_t launch_finalize(const ItemDesc* items, int n_items, const Partial* partials,
                           Partial* finals, const int64_t* final_base, hipStream_t stream);

// Sobol diagnostics: unscrambled and scrambled integers and FP64 normals for
// points first_point .. first_point + n_points - 1 and dims first_dim ...
hipError_t launch_debug_sobol(const uint32_t* dirs, int ladder_dims, uint64_t seed,
                              uint64_t first_point, int n_points, int first_dim, int n_dims,
                              int nbits, uint32_t* unscrambled, uint32_t* scrambled,
                              double* normal, hipStream_t stream);


// The allocation-free pricing call's device work area (scratch.hip): a
// file-scope array of kScratchBytes per GPU; scratch_address gives the current
// device's address of it.
constexpr size_t kScratchBytes = (size_t)32 << 20;
hipError_t scratch_address(void** ptr);

// The FP64 tier's exp and log (exp_log.h) elementwise, for the parity tests.
hipError_t launch_debug_exp_log(const double* x, uint64_t n, double* e, double* l, hipStream_t stream);

// Halton diagnostics: permuted digits [cell][32] (0xFFFFFFFF after the last),
// radical inverses and FP64 normals for points first_point .. and dims first_dim ..
hipError_t launch_debug_halton(const uint32_t* tables, int rows, uint64_t first_point, int n_points, int first_dim,
                               int n_dims, uint32_t* digits, double* uniform, double* normal, hipStream_t stream);


}  // namespace riga
