// market.cpp -- FP64 host preparation: correlation factor, Dupire calibration,
// Sobol direction numbers, Halton permutations and request validation.
//
// Arithmetic is written operation-for-operation as the reference so the device
// receives the reference's exact factors and surfaces:
//   correlation.rs:74-260, riga-market/src/lib.rs:130-364, qmc.rs:12-22,128-140,
//   qmc.rs:319-430, riga-engine/src/lib.rs:209-321.
// Built with -ffp-contract=off (no fused multiply-add on any target).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <map>
#include <memory>
#include <mutex>

#include "host.h"

namespace riga {

namespace {

inline bool fail(Error* err, const char* code, std::string msg) {
  if (err) {
    err->code = code;
    err->message = std::move(msg);
  }
  return false;
}

// Cyclic Jacobi eigenvalues of a symmetric matrix (correlation.rs:218-260).
std::vector<double> jacobi_eigen(std::vector<double> a, int n) {
  if (n == 1) return {a[0]};
  for (int sweep = 0; sweep < 100; ++sweep) {
    int p = 0, q = 0;
    double big = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        if (fabs(a[i * n + j]) > big) {
          big = fabs(a[i * n + j]);
          p = i;
          q = j;
        }
    if (big < 1.0e-14) break;
    const double theta = 0.5 * (a[q * n + q] - a[p * n + p]) / a[p * n + q];
    const double sign = std::signbit(theta) ? -1.0 : 1.0;  // f64::signum
    const double t = sign / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0);
    const double s = t * c;
    for (int k = 0; k < n; ++k) {
      const double kp = a[k * n + p], kq = a[k * n + q];
      a[k * n + p] = c * kp - s * kq;
      a[k * n + q] = s * kp + c * kq;
    }
    for (int k = 0; k < n; ++k) {
      const double pk = a[p * n + k], qk = a[q * n + k];
      a[p * n + k] = c * pk - s * qk;
      a[q * n + k] = s * pk + c * qk;
    }
  }
  std::vector<double> e(n);
  for (int i = 0; i < n; ++i) e[i] = a[i * n + i];
  return e;
}

double min_eigen(const double* c, int n) {
  std::vector<double> e = jacobi_eigen(std::vector<double>(c, c + (size_t)n * n), n);
  double m = std::numeric_limits<double>::infinity();
  for (double v : e) m = std::fmin(m, v);
  return m;
}

}  // namespace

bool cholesky_factor(int n, const double* c, std::vector<double>* l, Error* err) {
  if (n <= 0) return fail(err, "invalid_market", "correlation matrix must have at least one factor");
  for (int i = 0; i < n; ++i) {  // correlation.rs:74-93
    if (fabs(c[i * n + i] - 1.0) > 1.0e-9)
      return fail(err, "invalid_market", "correlation diagonal [" + std::to_string(i) + "] must be exactly one");
    for (int j = 0; j < n; ++j) {
      const double v = c[i * n + j];
      if (!std::isfinite(v)) return fail(err, "invalid_market", "correlation entries must be finite");
      if (fabs(v) > 1.0 + 1.0e-9) return fail(err, "invalid_market", "correlation entry is outside [-1, 1]");
      if (fabs(v - c[j * n + i]) > 1.0e-9) return fail(err, "invalid_market", "correlation is not symmetric");
    }
  }
  std::vector<double> src(c, c + (size_t)n * n);
  const double lam = min_eigen(c, n);
  if (!(lam >= 1.0e-10)) {  // !is_positive_semidefinite(-1e-10) -> psd_clamp(1e-8), :113-136
    const double floor_ = 1.0e-8;
    if (lam < floor_) {
      const double shift = floor_ - lam;
      for (int i = 0; i < n; ++i) src[i * n + i] += shift;
      std::vector<double> inv(n);
      for (int i = 0; i < n; ++i) inv[i] = 1.0 / sqrt(src[i * n + i]);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) src[i * n + j] *= inv[i] * inv[j];
    }
  }
  l->assign((size_t)n * n, 0.0);
  double* L = l->data();
  for (int i = 0; i < n; ++i) {  // Cholesky-Banachiewicz, :150-166
    for (int j = 0; j <= i; ++j) {
      double sum = src[i * n + j];
      for (int k = 0; k < j; ++k) sum -= L[i * n + k] * L[j * n + k];
      if (i == j) {
        L[i * n + j] = sqrt(std::fmax(sum, 0.0));
      } else {
        const double pivot = L[j * n + j];
        L[i * n + j] = pivot > 0.0 ? sum / pivot : 0.0;
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------- Dupire
namespace {

// Natural cubic spline second derivatives by the Thomas algorithm (lib.rs:130-176).
std::vector<double> spline_m(const std::vector<double>& x, const double* y) {
  const size_t n = x.size();
  std::vector<double> m(n, 0.0);
  if (n < 3) return m;
  std::vector<double> h(n - 1);
  for (size_t i = 0; i + 1 < n; ++i) h[i] = x[i + 1] - x[i];
  for (double v : h)
    if (v <= 0.0) return m;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(y[i])) return m;
  std::vector<double> lower(n, 0.0), diag(n, 0.0), upper(n, 0.0), rhs(n, 0.0);
  diag[0] = diag[n - 1] = 1.0;
  for (size_t i = 1; i + 1 < n; ++i) diag[i] = 2.0 * (h[i - 1] + h[i]);
  for (size_t i = 1; i < n; ++i) lower[i] = (i < n - 1) ? h[i] : 0.0;
  for (size_t i = 0; i + 1 < n; ++i) upper[i] = (i + 1 < n - 1) ? h[i + 1] : 0.0;
  for (size_t i = 1; i + 1 < n; ++i)
    rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
  std::vector<double> cp(n), dp(n);
  cp[0] = upper[0] / diag[0];
  dp[0] = rhs[0] / diag[0];
  for (size_t i = 1; i < n; ++i) {
    double piv = diag[i] - lower[i] * cp[i - 1];
    if (fabs(piv) < 1e-300) piv = 1e-300;
    cp[i] = upper[i] / piv;
    dp[i] = (rhs[i] - lower[i] * dp[i - 1]) / piv;
  }
  m[n - 1] = dp[n - 1];
  for (size_t i = n - 1; i-- > 0;) m[i] = dp[i] - cp[i] * m[i + 1];
  return m;
}

}  // namespace

bool dupire_calibrate(const riga_implied_grid& g, double spot, double rate, int grid_x, LvGrid* out,
                      Error* err) {
  const int nt = g.n_maturities, nk = g.n_strikes;
  if (nt <= 0 || nk <= 0) return fail(err, "invalid_local_vol", "implied vol grid has an empty axis");
  for (int i = 0; i + 1 < nt; ++i)
    if (!(g.maturities[i] < g.maturities[i + 1]))
      return fail(err, "invalid_local_vol", "implied vol maturities must be strictly increasing");
  for (int i = 0; i + 1 < nk; ++i)
    if (!(g.strikes[i] < g.strikes[i + 1]))
      return fail(err, "invalid_local_vol", "implied vol strikes must be strictly increasing");
  if (!std::isfinite(spot) || spot <= 0.0)
    return fail(err, "invalid_local_vol", "local-vol calibration needs a positive finite spot");
  if (grid_x < 2) return fail(err, "invalid_local_vol", "grid_x must be at least 2");
  const int gx = grid_x;
  const double log_lo = log(std::fmax(g.strikes[0] * 0.55, 1e-6));
  const double log_hi = log(std::fmax(g.strikes[nk - 1] * 1.05, 1e-6));
  out->log_s.resize(gx);
  std::vector<double> s_axis(gx);
  for (int i = 0; i < gx; ++i) {
    out->log_s[i] = log_lo + (log_hi - log_lo) * (double)i / (double)(gx - 1);
    s_axis[i] = exp(out->log_s[i]);
  }
  std::vector<double> w((size_t)nt * gx), wx((size_t)nt * gx), wxx((size_t)nt * gx), xg((size_t)nt * gx);
  std::vector<double> xk(nk);
  for (int k = 0; k < nt; ++k) {
    const double tk = g.maturities[k];
    const double fwd = spot * exp(rate * tk);
    for (int j = 0; j < nk; ++j) xk[j] = log(g.strikes[j] / fwd);
    const double* vol = g.vols + (size_t)k * nk;
    const std::vector<double> m = spline_m(xk, vol);
    for (int i = 0; i < gx; ++i) {  // spline_eval at the clamped log-moneyness, lib.rs:179-203
      const double p = std::min(std::max(log(s_axis[i] / fwd), xk[0]), xk[nk - 1]);
      int seg = 0;
      while (seg + 1 < nk - 1 && p > xk[seg + 1]) ++seg;
      const double h = xk[seg + 1] - xk[seg];
      double t = (p - xk[seg]) / h;
      t = std::min(std::max(t, 0.0), 1.0);
      const double u = 1.0 - t;
      const double sig = u * vol[seg] + t * vol[seg + 1] +
                         (h * h / 6.0) * ((t * t * t - t) * m[seg] + (u * (u * u) - u) * m[seg + 1]);
      const double sig1 = (vol[seg + 1] - vol[seg] +
                           (h * h / 6.0) * ((3.0 * t * t - 1.0) * m[seg] - (3.0 * u * u - 1.0) * m[seg + 1])) /
                          h;
      const double sig2 = u * m[seg + 1] + t * m[seg];
      const size_t o = (size_t)k * gx + i;
      w[o] = sig * sig * tk;
      wx[o] = 2.0 * sig * tk * sig1;
      wxx[o] = 2.0 * tk * (sig1 * sig1 + sig * sig2);
      xg[o] = p;
    }
  }
  const double* T = g.maturities;
  out->t_axis.assign(T, T + nt);
  out->v.resize((size_t)nt * gx);
  const double lo2 = 0.10 * 0.10, hi2 = 3.00 * 3.00;  // LV_CLAMP_LO/HI squared, lib.rs:30-32
  for (int k = 0; k < nt; ++k) {
    for (int i = 0; i < gx; ++i) {  // lib.rs:254-291
      double dwdt;
      if (nt == 1)
        dwdt = w[i] / std::fmax(T[0], 1e-9);
      else if (k == 0)
        dwdt = (w[(size_t)gx + i] - w[i]) / std::fmax(T[1] - T[0], 1e-9);
      else if (k == nt - 1)
        dwdt = (w[(size_t)(nt - 1) * gx + i] - w[(size_t)(nt - 2) * gx + i]) / std::fmax(T[nt - 1] - T[nt - 2], 1e-9);
      else
        dwdt = (w[(size_t)(k + 1) * gx + i] - w[(size_t)(k - 1) * gx + i]) / std::fmax(T[k + 1] - T[k - 1], 1e-9);
      const size_t o = (size_t)k * gx + i;
      const double ww = std::fmax(w[o], 1e-12);
      const double denom = 1.0 - (xg[o] / ww) * wx[o] + 0.25 * (-0.25 - 1.0 / ww) * wx[o] * wx[o] + 0.5 * wxx[o];
      double lv2 = (std::isfinite(denom) && fabs(denom) > 1e-12) ? dwdt / denom : lo2;
      if (lv2 < lo2) lv2 = lo2;
      if (lv2 > hi2) lv2 = hi2;
      out->v[o] = sqrt(lv2);
    }
  }
  return true;
}

void lv_bracket(const LvGrid& lv, double t, int* i0, int* i1, double* wt) {
  const int nt = lv.n_t();
  if (nt == 1) {
    *i0 = *i1 = 0;
    *wt = 0.0;
    return;
  }
  const double tc = std::min(std::max(t, lv.t_axis[0]), lv.t_axis[nt - 1]);
  int i = 0;
  while (i + 1 < nt - 1 && tc > lv.t_axis[i + 1]) ++i;
  const double span = lv.t_axis[i + 1] - lv.t_axis[i];
  *wt = span > 0.0 ? (tc - lv.t_axis[i]) / span : 0.0;
  *i0 = i;
  *i1 = i + 1;
}

double lv_sigma(const LvGrid& lv, double t, double s) {
  const double x = log(std::fmax(s, 1e-12));
  const int n = lv.n_x();
  int lo, hi;
  if (x <= lv.log_s[0]) {
    lo = hi = 0;
  } else if (x >= lv.log_s[n - 1]) {
    lo = hi = n - 1;
  } else {
    int j = 0;
    while (j + 1 < n && x > lv.log_s[j + 1]) ++j;
    lo = j;
    hi = j + 1;
  }
  const double wx = hi == lo ? 0.0 : (x - lv.log_s[lo]) / (lv.log_s[hi] - lv.log_s[lo]);
  int i0, i1;
  double wt;
  lv_bracket(lv, t, &i0, &i1, &wt);
  const double* r0 = &lv.v[(size_t)i0 * n];
  const double* r1 = &lv.v[(size_t)i1 * n];
  const double v0 = (1.0 - wx) * r0[lo] + wx * r0[hi];
  const double v1 = (1.0 - wx) * r1[lo] + wx * r1[hi];
  return (1.0 - wt) * v0 + wt * v1;
}

// ---------------------------------------------------------------- ladders
namespace {

// Carry-less multiply modulo the polynomial `mask` of degree s.
uint32_t gf2_mulmod(uint32_t a, uint32_t b, uint32_t mask, int s) {
  uint64_t r = 0, aa = a;
  for (uint32_t bb = b; bb; bb >>= 1, aa <<= 1)
    if (bb & 1u) r ^= aa;
  for (int bit = 63; bit >= s; --bit)
    if ((r >> bit) & 1u) r ^= (uint64_t)mask << (bit - s);
  return (uint32_t)r;
}

uint32_t gf2_powmod(uint32_t base, uint64_t e, uint32_t mask, int s) {
  uint32_t acc = 1, b = base;
  for (; e; e >>= 1) {
    if (e & 1u) acc = gf2_mulmod(acc, b, mask, s);
    b = gf2_mulmod(b, b, mask, s);
  }
  return acc;
}

// x has multiplicative order 2^s - 1 modulo `mask` (qmc.rs:397-412).
bool primitive(uint32_t mask, int s) {
  const uint64_t order = (1ull << s) - 1;
  if (gf2_powmod(2, order, mask, s) != 1) return false;
  uint64_t n = order;
  for (uint64_t d = 2; d * d <= n; ++d) {
    if (n % d) continue;
    if (gf2_powmod(2, order / d, mask, s) == 1) return false;
    while (n % d == 0) n /= d;
  }
  if (n > 1 && gf2_powmod(2, order / n, mask, s) == 1) return false;
  return true;
}

// Joe-Kuo recurrence with unit initial m_1..m_s (qmc.rs:319-343).
void directions_for(uint32_t mask, uint32_t* out) {
  const int s = 31 - __builtin_clz(mask);
  uint32_t m[33] = {0};
  for (int i = 1; i <= s; ++i) m[i] = 1;
  for (int k = s + 1; k <= 32; ++k) {
    uint32_t v = ((1u << s) * m[k - s]) ^ m[k - s];
    for (int i = 1; i < s; ++i)
      if ((mask >> (s - i)) & 1u) v ^= (1u << i) * m[k - i];
    m[k] = v;
  }
  for (int k = 0; k < 32; ++k) out[k] = m[k + 1] << (31 - k);
}

std::mutex g_dir_mu;
std::vector<uint32_t> g_dirs;    // rows x 32
uint32_t g_next_mask = 0;        // next candidate polynomial
int g_next_degree = 2;

}  // namespace

const std::vector<uint32_t>& sobol_directions(int dims) {
  std::lock_guard<std::mutex> lock(g_dir_mu);
  if (g_dirs.empty()) {
    g_dirs.resize(32);
    for (int k = 0; k < 32; ++k) g_dirs[k] = 1u << (31 - k);  // dimension 1: van der Corput
    g_next_degree = 2;
    g_next_mask = (1u << 2) | 1u;
  }
  while ((int)(g_dirs.size() / 32) < dims) {  // primitive polynomials in (degree, value) order
    if (g_next_mask >= (1u << (g_next_degree + 1))) {
      ++g_next_degree;
      g_next_mask = (1u << g_next_degree) | 1u;
      continue;
    }
    const uint32_t mask = g_next_mask;
    g_next_mask += 2;
    if (!primitive(mask, g_next_degree)) continue;
    const size_t o = g_dirs.size();
    g_dirs.resize(o + 32);
    directions_for(mask, &g_dirs[o]);
  }
  return g_dirs;
}

std::vector<uint32_t> halton_tables(int dims, uint64_t seed) {
  std::vector<uint32_t> primes;
  for (uint32_t c = 2; (int)primes.size() < dims; ++c) {
    bool ok = true;
    for (uint32_t p : primes)
      if (c % p == 0) {
        ok = false;
        break;
      }
    if (ok) primes.push_back(c);
  }
  std::vector<uint32_t> out(2 * (size_t)dims);
  for (int d = 0; d < dims; ++d) {
    const uint32_t base = primes[d];
    out[d] = base;
    out[dims + d] = (uint32_t)out.size();
    const size_t o = out.size();
    out.resize(o + base);
    uint32_t* perm = &out[o];
    for (uint32_t i = 0; i < base; ++i) perm[i] = i;
    uint64_t state = seed ^ ((uint64_t)base * 0x9E3779B97F4A7C15ull);  // qmc.rs:129-140
    for (uint64_t i = base - 1; i >= 1; --i) {
      state = state * 6364136223846793005ull + 1442695040888963407ull;
      const uint64_t j = (state >> 33) % (i + 1);
      std::swap(perm[i], perm[j]);
    }
  }
  return out;
}



// ---------------------------------------------------------------- Monte Carlo lane
void chacha_key_from_seed(uint64_t seed, uint32_t key[8]) {
  // PCG32 (XSH-RR): the state advances before each output
  uint64_t st = seed;
  for (int w = 0; w < 8; ++w) {
    st = st * 6364136223846793005ULL + 11634580027462260723ULL;
    const uint32_t v = (uint32_t)(((st >> 18) ^ st) >> 27);
    const unsigned rot = (unsigned)(st >> 59);
    key[w] = rot ? (v >> rot) | (v << (32 - rot)) : v;
  }
}

const std::vector<double>& ziggurat_tables() {
  static const std::vector<double> tab = [] {
    const double R = 3.6541528853610088, V = 0.00492867323399;
    auto pdf = [](double x) { return exp(-x * x / 2.0); };
    double x[257];
    x[0] = V / pdf(R);
    x[1] = R;
    for (int i = 2; i < 256; ++i) x[i] = sqrt(-2.0 * log(V / x[i - 1] + pdf(x[i - 1])));
    x[256] = 0.0;
    std::vector<double> out(2 * 257);
    auto stored = [](double v) {  // the crate's 18-decimal literal, read back
      char buf[64];
      snprintf(buf, sizeof buf, "%.18f", v);
      return strtod(buf, nullptr);
    };
    for (int i = 0; i < 257; ++i) {
      out[i] = stored(x[i]);
      out[257 + i] = stored(pdf(x[i]));
    }
    return out;
  }();
  return tab;
}

// ---------------------------------------------------------------- validation
bool validate_request(const riga_request& r, Error* err) {
  const int F = r.factors;
  if (r.paths == 0 || r.path_block_size == 0)
    return fail(err, "invalid_paths", "paths and path_block_size must be positive");
  if (r.steps <= 0 || !std::isfinite(r.maturity_years) || r.maturity_years <= 0.0)
    return fail(err, "invalid_schedule", "steps and maturity_years must be positive and finite");
  if (F < 1 || !r.spots || !r.vols || !r.dividend_yields || !r.initial_spots || !r.correlation)
    return fail(err, "invalid_market_shape", "market and trade vectors must have one entry per factor");
  for (int f = 0; f < F; ++f) {
    const double vals[4] = {r.spots[f], r.vols[f], r.dividend_yields[f], r.initial_spots[f]};
    for (double v : vals)
      if (!std::isfinite(v)) return fail(err, "invalid_market", "market inputs must be finite");
    for (int g = 0; g < F; ++g)
      if (!std::isfinite(r.correlation[f * F + g])) return fail(err, "invalid_market", "rho must be finite");
  }
  if (!std::isfinite(r.risk_free_rate)) return fail(err, "invalid_market", "risk_free_rate must be finite");
  if (!std::isfinite(r.strike)) return fail(err, "invalid_market", "strike must be finite");
  if (!std::isfinite(r.knock_in_barrier)) return fail(err, "invalid_market", "KI barrier must be finite");
  for (int f = 0; f < F; ++f) {
    bool bad = r.spots[f] <= 0.0 || r.initial_spots[f] <= 0.0 || r.vols[f] < 0.0;
    for (int g = 0; g < F; ++g) {
      const double v = r.correlation[f * F + g];
      if (f == g) bad |= fabs(v - 1.0) > 1.0e-12;
      else bad |= fabs(v) > 1.0 || fabs(v - r.correlation[g * F + f]) > 1.0e-12;
    }
    if (bad)
      return fail(err, "invalid_market", "spots must be positive, vols non-negative, and correlation in [-1,1]");
  }
  if (r.has_knock_out && (!std::isfinite(r.knock_out_barrier) || r.knock_out_barrier <= 0.0))
    return fail(err, "invalid_trade", "knock_out_barrier must be finite and positive when present");
  if (r.has_coupon_barrier && (!std::isfinite(r.coupon_barrier) || r.coupon_barrier <= 0.0))
    return fail(err, "invalid_trade", "coupon_barrier must be finite and positive when present");
  if (r.has_coupon_rate && (!std::isfinite(r.coupon_rate) || r.coupon_rate < 0.0))
    return fail(err, "invalid_trade", "coupon_rate must be finite and non-negative when present");
  if (r.strike <= 0.0 || r.knock_in_barrier <= 0.0)
    return fail(err, "invalid_trade", "strike must be positive and barriers finite");
  if (r.ladder != RIGA_VR_SOBOL && r.ladder != RIGA_VR_HALTON && r.ladder != RIGA_VR_NONE)
    return fail(err, "unsupported_variance_reduction", "unknown variance-reduction lane");
  return true;
}

}  // namespace riga
