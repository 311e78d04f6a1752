/* exp_log.h -- FP64 exp and log written once, for the device's FP64 tier and for
 * the CPU oracle that checks it.
 *
 * Both are built from IEEE-754 operations only (add, multiply, fused
 * multiply-add, exact scaling by powers of two, integer bit manipulation), each
 * rounded once, with no contraction, and from the same two tables
 * (exp_log_tables.h, written by gen_exp_log_tables.py): compiled for gfx950 and
 * by gcc on the host they return the same bits for every input, so a path priced
 * on the device and in the oracle follows the same numbers step by step.  Both
 * stay within about 0.51 ulp (the tests compare them with correctly rounded
 * values and with glibc's exp and log).
 *
 * exp: the table-driven method of P. T. P. Tang ("Table-driven implementation of
 * the exponential function in IEEE floating-point arithmetic", ACM TOMS 15(2),
 * 1989) with 128 cells: x = (128 m + j) ln2/128 + r, |r| <= ln2/256, ln2/128 in
 * two parts; exp(x) = 2^m 2^(j/128) exp(r), 2^(j/128) = hi + lo from the table,
 * exp(r) - 1 from its degree-5 Taylor polynomial (truncation below 6e-19), and
 * y = hi + (hi (exp(r) - 1) + lo) rounded once before the exact scaling by 2^m.
 *
 * log: Tang's table-driven logarithm ("Table-driven implementation of the
 * logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990) in
 * the form x = 2^m y, y in [0.6875, 1.375), ln x = m ln2 + ln(1/invc) + ln(1 + r)
 * with r = y invc - 1 exact (invc has 8 significant bits, |r| < 2^-7), ln(1 + r)
 * from its degree-9 series and the large terms summed exactly (two-sums), so one
 * rounding ends it.  The two cells that touch 1 have invc = 1: near 1 the result
 * is r + r^2 P(r) with no cancellation.
 *
 * The table index is masked to the table in every case, NaN and out-of-range
 * inputs included; their results are selected afterwards.
 */
#ifndef RIGA_EXP_LOG_H
#define RIGA_EXP_LOG_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define RIGA_EXP_LOG_FN __device__ static inline
#define RIGA_EXP_LOG_TABLE __device__ static const
#else
#define RIGA_EXP_LOG_FN static inline
#define RIGA_EXP_LOG_TABLE static const
#endif

#include "exp_log_tables.h"

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

/* y 2^m rounded once (the exact product where it is a normal double).  The device
 * uses its ldexp instruction, the host two exact multiplies by powers of two and,
 * for a subnormal result, one rounding: the same correctly rounded value. */
RIGA_EXP_LOG_FN double riga_scale2(double y, int m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_ldexp(y, m);
#else
  if (m >= -1022 && m <= 1023) return y * riga_f64_from_bits((uint64_t)(m + 1023) << 52);
  if (m > 1023) return (y * riga_f64_from_bits((uint64_t)(m - 1 + 1023) << 52)) * 2.0;
  return (y * riga_f64_from_bits((uint64_t)(m + 1000 + 1023) << 52)) * 0x1p-1000;
#endif
}

/* exp(x) for every double x. */
RIGA_EXP_LOG_FN double riga_exp(double x) {
  const double kInvLn2N = 0x1.71547652b82fep+7; /* 128 / ln 2 */
  const double kLn2HiN = 0x1.62e42fefa39efp-8;  /* ln 2 / 128, nearest double */
  const double kLn2LoN = 0x1.abc9e3b39803fp-63; /* ln 2 / 128 - kLn2HiN */
  const double kShift = 0x1.8p52;               /* rounds x 128/ln2 to an integer */
  /* |x| <= 746 keeps k small; beyond that the result is selected below */
  const double xc = x > 746.0 ? 746.0 : (x < -746.0 ? -746.0 : x);
  double kd = xc * kInvLn2N + kShift;
  const int64_t k = (int64_t)(riga_f64_bits(kd) - riga_f64_bits(kShift)); /* kd - kShift as an integer */
  kd -= kShift;
  double r = fma(-kd, kLn2HiN, xc);
  r = fma(-kd, kLn2LoN, r); /* |r| <= ln2/256 + 2^-60 */
  const int j = (int)(k & 127);
  const int m = (int)(k >> 7); /* floor(k / 128): k = 128 m + j */
  const double hi = riga_exp_tab[2 * j], lo = riga_exp_tab[2 * j + 1];
  /* exp(r) - 1 = r + r^2 (1/2 + r/6 + r^2/24 + r^3/120) */
  double q = 0x1.1111111111111p-7;       /* 1/120 */
  q = fma(q, r, 0x1.5555555555555p-5);   /* 1/24 */
  q = fma(q, r, 0x1.5555555555555p-3);   /* 1/6 */
  q = fma(q, r, 0.5);
  const double p = fma(r * r, q, r);
  const double y = hi + fma(hi, p, lo); /* 2^(j/128) exp(r), in [0.99, 2) */
  double res = riga_scale2(y, m);
  res = x > 710.0 ? (double)INFINITY : res; /* exp(709.79) already overflows */
  res = x < -746.0 ? 0.0 : res;             /* exp(-745.14) already rounds to 0 */
  return x != x ? x + x : res;
}

/* log(x) for every double x (NaN for x < 0, -inf at 0). */
RIGA_EXP_LOG_FN double riga_log(double x) {
  const double kLn2Hi = 0x1.62e42fefa3800p-1;  /* ln 2 to 42 bits: m * kLn2Hi is exact */
  const double kLn2Lo = 0x1.ef35793c76730p-45; /* ln 2 - kLn2Hi */
  const uint64_t kOff = 0x3fe6000000000000ull; /* bits of 0.6875 */
  const int sub = x < 0x1p-1022;               /* subnormal (or not positive): scale up first */
  const double xs = sub ? x * 0x1p54 : x;
  const uint64_t b = riga_f64_bits(xs);
  const uint64_t t = b - kOff;
  const int j = (int)((t >> 45) & 127);                 /* the cell of y */
  const int m = (int)((int64_t)t >> 52) - (sub ? 54 : 0); /* x = 2^m y */
  const double y = riga_f64_from_bits(b - (t & 0xfff0000000000000ull)); /* [0.6875, 1.375) */
  const double invc = riga_log_tab[4 * j], lch = riga_log_tab[4 * j + 1], lcl = riga_log_tab[4 * j + 2];
  const double r = fma(y, invc, -1.0); /* exact */
  const double kd = (double)m;
  /* m ln2_hi + ln(1/invc)_hi + r as hi + errors (two fast two-sums: the first
   * addend is the larger, or zero) */
  const double a = kd * kLn2Hi;
  const double w = a + lch;
  const double we = (a - w) + lch;
  const double hi = w + r;
  const double hl = (w - hi) + r;
  /* ln(1 + r) - r = r^2 (-1/2 + r/3 - r^2/4 + ... + r^7/9) */
  double P = 0x1.c71c71c71c71cp-4;      /* 1/9 */
  P = fma(P, r, -0x1.0000000000000p-3); /* -1/8 */
  P = fma(P, r, 0x1.2492492492492p-3);  /* 1/7 */
  P = fma(P, r, -0x1.5555555555555p-3); /* -1/6 */
  P = fma(P, r, 0x1.999999999999ap-3);  /* 1/5 */
  P = fma(P, r, -0x1.0000000000000p-2); /* -1/4 */
  P = fma(P, r, 0x1.5555555555555p-2);  /* 1/3 */
  P = fma(P, r, -0.5);
  double lo = fma(kd, kLn2Lo, lcl) + (we + hl);
  lo = fma(r * r, P, lo);
  double res = hi + lo;
  res = x == (double)INFINITY ? x : res;
  res = x == 0.0 ? -(double)INFINITY : res;
  res = x < 0.0 ? (double)NAN : res;
  return x != x ? x + x : res;
}

#endif
