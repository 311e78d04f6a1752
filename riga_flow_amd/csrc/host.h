// host.h -- host-side model preparation for the gfx950 backend: everything the
// kernel needs that is computed once per trade/market on the CPU in FP64
// (Cholesky factor, Dupire calibration, Sobol/Halton tables, schedules).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/riga_hip.h"

namespace riga {

struct Error {
  std::string code, message;
};

// ---- correlation (crates/riga-market/src/correlation.rs)
// Validates, PSD-clamps when the Jacobi minimum eigenvalue is below 1e-10,
// and returns the lower-triangular factor (row-major n x n).
bool cholesky_factor(int n, const double* c, std::vector<double>* l, Error* err);

// ---- local vol (crates/riga-market/src/lib.rs)
struct LvGrid {
  std::vector<double> t_axis, log_s, v;  // v[n_t][n_x]
  int n_t() const { return (int)t_axis.size(); }
  int n_x() const { return (int)log_s.size(); }
};
bool dupire_calibrate(const riga_implied_grid& g, double spot, double rate, int grid_x, LvGrid* out,
                      Error* err);
double lv_sigma(const LvGrid& lv, double t, double s);                   // lib.rs:305-343
void lv_bracket(const LvGrid& lv, double t, int* i0, int* i1, double* wt);  // lib.rs:324-339

// ---- ladders (crates/riga-engine/src/qmc.rs)
// Sobol direction rows 0..dims-1 (row 0 = van der Corput), 32 words each.
const std::vector<uint32_t>& sobol_directions(int dims);
// Halton: first `dims` primes and seed-keyed digit permutations, packed as
// [bases(dims)][perm offsets(dims)][perms...] in uint32 words.
std::vector<uint32_t> halton_tables(int dims, uint64_t seed);

// ---- Monte Carlo lane: rand_core 0.9 seed_from_u64 (PCG32 expansion of the
// seed into the ChaCha8 key) and rand_distr 0.5 ziggurat tables
// (ziggurat_tables.py: R = 3.6541528853610088, V = 0.00492867323399, each
// entry as printed to 18 decimals in the crate): x[257] then f(x)[257].
void chacha_key_from_seed(uint64_t seed, uint32_t key[8]);
const std::vector<double>& ziggurat_tables();

// ---- request validation (crates/riga-engine/src/lib.rs:209-321, F factors)
bool validate_request(const riga_request& r, Error* err);

}  // namespace riga
