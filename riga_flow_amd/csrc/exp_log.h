/* exp_log.h -- FP64 exp and log written once, for the device's FP64 tier and for
 * the CPU oracle that checks it.
 *
 * Both are built from IEEE-754 operations only (add, multiply, divide, fused
 * multiply-add, integer bit manipulation), each rounded once, with no
 * contraction: compiled for gfx950 and by gcc on the host they return the same
 * bits for every input, so a path priced on the device and in the oracle
 * follows the same numbers step by step.  Accuracy is a little over half an ulp
 * (the tests compare them with glibc's exp and log and with correctly rounded
 * values).
 *
 * exp: Cody and Waite's additive range reduction (W. J. Cody, W. Waite,
 * "Software Manual for the Elementary Functions", 1980) x = m ln2 + r with
 * ln2 in two parts, |r| <= ln2/2, exp(x) = 2^m exp(r) with exp(r) from its
 * degree-13 Taylor polynomial (truncation below 5e-18), 1 + r + r^2/2 summed
 * exactly, so one rounding ends it.  No table: the kernels that call it read no
 * memory for it.
 *
 * log: x = 2^m y with y in [sqrt(1/2), sqrt(2)], log(y) = 2 atanh(u) with
 * u = (y - 1)/(y + 1), |u| <= 0.1716, as 2u + u^3 P(u^2) (the atanh series to
 * u^21, truncation below 1e-18); u carries its division remainder u_lo, so
 * 2u + 2u_lo is accurate to about 2^-106 and m ln2 is added in two parts.
 */
#ifndef RIGA_EXP_LOG_H
#define RIGA_EXP_LOG_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define RIGA_EXP_LOG_FN __device__ static inline
#else
#define RIGA_EXP_LOG_FN static inline
#endif

RIGA_EXP_LOG_FN uint64_t riga_f64_bits(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}

RIGA_EXP_LOG_FN double riga_f64_from_bits(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof x);
  return x;
}

/* exp(x) for every double x. */
RIGA_EXP_LOG_FN double riga_exp(double x) {
  const double kInvLn2 = 0x1.71547652b82fep+0; /* 1 / ln 2 */
  const double kLn2Hi = 0x1.62e42fefa39efp-1;  /* ln 2, nearest double */
  const double kLn2Lo = 0x1.abc9e3b39803fp-56; /* ln 2 - kLn2Hi */
  const double kShift = 0x1.8p52;              /* rounds x / ln2 to an integer */
  if (x != x) return x + x;
  if (x > 710.0) return (double)INFINITY;
  if (x < -746.0) return 0.0;
  double kd = x * kInvLn2 + kShift;
  const int64_t m = (int64_t)(riga_f64_bits(kd) - riga_f64_bits(kShift));
  kd -= kShift;
  /* r = x - m ln2 as rh + c: rh exact (x and m kLn2Hi share their low bits), c tiny */
  const double rh = fma(-kd, kLn2Hi, x);
  const double c = -kd * kLn2Lo;
  const double r = rh + c; /* |r| <= ln2/2 + 2^-45, for the higher terms */
  /* exp(r) = 1 + r + r^2/2 + r^3 q(r), q(r) = 1/3! + r/4! + ... + r^10/13!; the first
   * three terms are summed exactly (r^2 as an exact pair, two-sums), so the result
   * is rounded once, and only r^3 q(r) (below 0.007) carries its own rounding */
  double q = 0x1.6124613a86d09p-33;  /* 1/13! */
  q = fma(q, r, 0x1.1eed8eff8d898p-29); /* 1/12! */
  q = fma(q, r, 0x1.ae64567f544e4p-26); /* 1/11! */
  q = fma(q, r, 0x1.27e4fb7789f5cp-22); /* 1/10! */
  q = fma(q, r, 0x1.71de3a556c734p-19); /* 1/9! */
  q = fma(q, r, 0x1.a01a01a01a01ap-16); /* 1/8! */
  q = fma(q, r, 0x1.a01a01a01a01ap-13); /* 1/7! */
  q = fma(q, r, 0x1.6c16c16c16c17p-10); /* 1/6! */
  q = fma(q, r, 0x1.1111111111111p-7);  /* 1/5! */
  q = fma(q, r, 0x1.5555555555555p-5);  /* 1/4! */
  q = fma(q, r, 0x1.5555555555555p-3);  /* 1/3! */
  const double r2 = rh * rh;
  const double r2lo = fma(rh, rh, -r2);
  const double h = 0.5 * r2, hl = fma(rh, c, 0.5 * r2lo); /* r^2/2 = h + hl */
  const double tail = fma((r * r) * r, q, hl);
  const double s1 = 1.0 + rh; /* 1 + rh exactly as (s1, e1): |rh| < 1 */
  const double e1 = ((1.0 - s1) + rh) + c;
  const double s2 = s1 + h; /* + r^2/2 exactly as (s2, e2) */
  const double bb = s2 - s1;
  const double e2 = (s1 - (s2 - bb)) + (h - bb);
  const double y = s2 + ((e1 + e2) + tail); /* exp(r), in [0.70, 1.42] */
  if (m >= -1022 && m <= 1023) return y * riga_f64_from_bits((uint64_t)(m + 1023) << 52);
  if (m > 1023) /* overflows to infinity beyond DBL_MAX */
    return (y * riga_f64_from_bits((uint64_t)(m - 1 + 1023) << 52)) * 2.0;
  /* subnormal result: the first scaling is exact, the second rounds once */
  return (y * riga_f64_from_bits((uint64_t)(m + 1000 + 1023) << 52)) * 0x1p-1000;
}

/* log(x) for every double x (NaN for x < 0, -inf at 0). */
RIGA_EXP_LOG_FN double riga_log(double x) {
  const double kLn2Hi = 0x1.62e42fefa3800p-1;  /* ln 2 to 42 bits: m * kLn2Hi is exact */
  const double kLn2Lo = 0x1.ef35793c76730p-45; /* ln 2 - kLn2Hi */
  const double kSqrt2 = 0x1.6a09e667f3bcdp+0;
  if (x != x) return x + x;
  if (x < 0.0) return (double)NAN;
  if (x == 0.0) return -(double)INFINITY;
  if (x == (double)INFINITY) return x;
  int64_t m = 0;
  if (x < 0x1p-1022) { /* subnormal: scale into the normal range */
    x *= 0x1p54;
    m = -54;
  }
  const uint64_t b = riga_f64_bits(x);
  m += (int64_t)((b >> 52) & 0x7ff) - 1023;
  double y = riga_f64_from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull); /* [1, 2) */
  if (y > kSqrt2) {
    y *= 0.5;
    m += 1;
  }
  const double n = y - 1.0; /* exact: y in [1/2, 2] */
  /* d = y + 1 as an exact pair (two-sum) */
  const double d_hi = y + 1.0;
  const double bb = d_hi - y;
  const double d_lo = (y - (d_hi - bb)) + (1.0 - bb);
  const double u = n / d_hi;
  double rho = fma(-u, d_hi, n); /* n - u d_hi exactly */
  rho = fma(-u, d_lo, rho);
  /* u_lo = rho / d to ~9 bits (it is 2^-53 of u): 1/d = (1/2)(1 - h + h^2 - h^3), h = n/2 */
  const double h = 0.5 * n;
  const double inv_d = 0.5 * fma(h, fma(h, 1.0 - h, -1.0), 1.0);
  const double u_lo = rho * inv_d;
  const double v = u * u;
  double P = 0x1.8618618618618p-4; /* 2/21 */
  P = fma(v, P, 0x1.af286bca1af28p-4); /* 2/19 */
  P = fma(v, P, 0x1.e1e1e1e1e1e1ep-4); /* 2/17 */
  P = fma(v, P, 0x1.1111111111111p-3); /* 2/15 */
  P = fma(v, P, 0x1.3b13b13b13b14p-3); /* 2/13 */
  P = fma(v, P, 0x1.745d1745d1746p-3); /* 2/11 */
  P = fma(v, P, 0x1.c71c71c71c71cp-3); /* 2/9 */
  P = fma(v, P, 0x1.2492492492492p-2); /* 2/7 */
  P = fma(v, P, 0x1.999999999999ap-2); /* 2/5 */
  P = fma(v, P, 0x1.5555555555555p-1); /* 2/3 */
  const double md = (double)m;
  /* m ln2_hi + 2u exactly as a pair (two-sum), then the small terms */
  const double a = md * kLn2Hi, c = 2.0 * u;
  const double s = a + c;
  const double sb = s - a;
  const double err = (a - (s - sb)) + (c - sb);
  double tail = fma(u * v, P, 2.0 * u_lo);
  tail = fma(md, kLn2Lo, tail);
  return s + (err + tail);
}

#endif
