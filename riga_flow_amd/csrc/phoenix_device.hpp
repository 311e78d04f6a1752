// phoenix_device.hpp -- the fused Phoenix path kernel for gfx950.
//
// One lane owns one Sobol point (= one path of the item) and carries all S
// scenarios of the item through the whole horizon:
//
//   Sobol gray-code XOR -> Owen nested scramble -> inverse normal      (per bridge node, F dims)
//   y = L z                                                            (Cholesky row sums)
//   streaming Brownian bridge, W(t) produced in time order             (per-lane stack of F-vectors)
//   per scenario: local-vol lookup from LDS rows, Euler log-step,
//   worst-of, daily KO, day-in coupon accrual with memory, period ends (Phoenix scan)
//   path end: EKI put, funding, coupon legs -> double-double wave sums -> block partial
//
// Nothing path-sized is written to HBM unless a replay query sets kRecords.
// The reference computes the same quantities in crates/riga-engine/src/note.rs:79-259
// with the ladder in crates/riga-engine/src/qmc.rs and the surface lookup in
// crates/riga-market/src/lib.rs:305-343.
//
// Templates: Real = double (the FP64 parity tier, compiled with
// -ffp-contract=off so every expression rounds exactly as the reference's) or
// float (the fast tier: path arithmetic in FP32, payoff legs and all sums FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exp_log.h"
#include "mc_rng.hpp"
#include "phoenix_layout.h"

namespace riga {

constexpr int kWaves = kWavesPerBlock;

// ---------------------------------------------------------------- ladders
// Owen nested scramble (qmc.rs:293-314).  The reference hashes
//   mix64(seed ^ dim*C1 ^ bit*C2 ^ prefix*C3)
// and keeps bit 63 of the result.  `key` = seed ^ dim*C1 is wave-uniform; the
// prefix product is carried incrementally: P(prefix<<1 | s) = 2 P(prefix) + s*C3
// (mod 2^64), and only the top bit of the second multiply is formed, since the
// final xorshift of mix64 leaves bit 63 unchanged.
constexpr uint64_t kOwenC2 = 0xD6E8FEB86659FD93ull;
constexpr uint64_t kOwenC3 = 0xA24BAED4963EE407ull;
constexpr int kOwenTableBits = kOwenPrefixBits;  // log2 of the points per dimension of a coordinate table

// The scrambled prefix is not carried level by level: C3 is odd, so
// prefix = P * C3^-1 (mod 2^64), and its low 32 bits -- all of a prefix -- are
// P_lo * (C3_lo)^-1 (mod 2^32), one multiply after the last level.
constexpr uint32_t kOwenC3InvLo = 0xF69349B7u;  // (C3 mod 2^32)^-1 mod 2^32
static_assert((uint32_t)kOwenC3 * kOwenC3InvLo == 1u, "C3 inverse");

__device__ __forceinline__ uint32_t owen_prefix(uint64_t P) { return (uint32_t)P * kOwenC3InvLo; }

__device__ __forceinline__ void owen_level(uint32_t x, uint64_t key, int bit, uint64_t& P) {
  constexpr uint64_t C4 = 0xBF58476D1CE4E5B9ull;
  constexpr uint64_t C5 = 0x94D049BB133111EBull;
  // z = h ^ (h >> 30) for h = kb ^ P, as (kb ^ kb>>30) ^ P ^ (P>>30): the key
  // part is wave-uniform (scalar unit), the rest two three-input XORs
  const uint64_t kb = key ^ ((uint64_t)bit * kOwenC2);
  const uint64_t kk = kb ^ (kb >> 30);
  const uint64_t ps = P >> 30;
  const uint32_t zlo = __builtin_amdgcn_bitop3_b32((uint32_t)kk, (uint32_t)P, (uint32_t)ps, 0x96);
  const uint32_t zhi = __builtin_amdgcn_bitop3_b32((uint32_t)(kk >> 32), (uint32_t)(P >> 32), (uint32_t)(ps >> 32), 0x96);
  uint64_t z = ((uint64_t)zhi << 32) | zlo;
  z *= C4;
  z ^= z >> 27;
  const uint32_t top = (uint32_t)((z * C5) >> 63);
  const bool s = ((x >> bit) & 1u) != top;  // the scrambled bit
  P = (P << 1) + (s ? kOwenC3 : 0ull);      // prefix' = 2 prefix + s
}

__device__ __forceinline__ uint32_t owen_scramble(uint32_t x, uint64_t key) {
  uint64_t P = 0;
#pragma unroll 4
  for (int bit = 31; bit >= 0; --bit) owen_level(x, key, bit, P);
  return owen_prefix(P);
}

// Levels kHi..kLo of the scramble, continuing from P.
template <int kHi, int kLo>
__device__ __forceinline__ void owen_levels(uint32_t x, uint64_t key, uint64_t& P) {
#pragma unroll
  for (int bit = kHi; bit >= kLo; --bit) owen_level(x, key, bit, P);
}

// The FP32 tier's scramble.  A float carries 24 significant bits, so the tier
// hashes the top 24 levels and takes the middle of the remaining 8-bit cell
// (x = p24 << 8 | 0x80, within 2^-25 of the 32-level uniform).  Where the
// inverse normal is steep -- u < 2^-10 or u > 1 - 2^-10, i.e. the top 10 bits
// of p24 all 0 or all 1 -- the last 8 levels are hashed too, so those
// coordinates are the full 32-level scramble.  Elsewhere z moves by at most
// 2^-25 / phi(z(2^-10)) = 9.1e-6.
constexpr uint32_t kF32TailCells = 1u << 14;  // 24-bit cells below u = 2^-10

// Top 24 levels.
__device__ __forceinline__ uint32_t owen_top24(uint32_t x, uint64_t key) {
  uint64_t P = 0;
#pragma unroll 4
  for (int bit = 31; bit >= 8; --bit) owen_level(x, key, bit, P);
  return owen_prefix(P);
}

__device__ __forceinline__ bool owen_f32_tail(uint32_t p24) {
  return p24 < kF32TailCells || p24 >= (1u << 24) - kF32TailCells;
}

__device__ __forceinline__ uint32_t owen_low8(uint32_t x, uint64_t key, uint32_t p24) {
  uint64_t P = (uint64_t)p24 * kOwenC3;
  owen_levels<7, 0>(x, key, P);
  return owen_prefix(P);
}

// The FP32 rule applied to a whole scrambled coordinate.
__device__ __forceinline__ uint32_t owen_f32_of_whole(uint32_t xs, bool full_ladder) {
  return full_ladder || owen_f32_tail(xs >> 8) ? xs : ((xs & 0xFFFFFF00u) | 0x80u);
}

__device__ __forceinline__ uint32_t owen_f32(uint32_t x, uint64_t key) {
  const uint32_t p = owen_top24(x, key);
  return owen_f32_tail(p) ? owen_low8(x, key, p) : ((p << 8) | 0x80u);
}

// Unscrambled Sobol coordinate (qmc.rs:269-279): XOR of the direction numbers
// selected by the gray code of n.  `row` is wave-uniform (32 words, read as
// scalars); each bit is a sign-extended mask and one three-input logic op, with
// no branch and no per-bit load.  Bits of the gray code at or above nbits are
// zero, so a fixed 21-, 24- or 32-bit sweep gives the same XOR.
__device__ __forceinline__ uint32_t sobol_x(uint32_t gray, const uint32_t* __restrict__ row, int nbits) {
  // opaque copy: the masks are rebuilt here (one bit-field extract each)
  // instead of 32 lane masks being held across the whole path
  asm volatile("" : "+v"(gray));
  // bitop3 table 0x6c: f(a, b, c) = b ^ (a & c), i.e. x ^= mask & row[k]
  uint32_t x = 0;
  if (nbits <= 21) {
    // up to 2^21 - 1 points (2 Mi paths per item)
#pragma unroll
    for (int k = 0; k < 21; ++k)
      x = __builtin_amdgcn_bitop3_b32((uint32_t)__builtin_amdgcn_sbfe((int)gray, k, 1), x, row[k], 0x6c);
  } else if (nbits <= 24) {
#pragma unroll
    for (int k = 0; k < 24; ++k)
      x = __builtin_amdgcn_bitop3_b32((uint32_t)__builtin_amdgcn_sbfe((int)gray, k, 1), x, row[k], 0x6c);
  } else {
#pragma unroll
    for (int k = 0; k < 32; ++k)
      x = __builtin_amdgcn_bitop3_b32((uint32_t)__builtin_amdgcn_sbfe((int)gray, k, 1), x, row[k], 0x6c);
  }
  return x;
}

// Digit-scrambled radical inverse (qmc.rs:143-153) in FP64, as the reference.
__device__ __forceinline__ double halton_u(uint32_t n, uint32_t base,
                                           const uint32_t* __restrict__ perm) {
#pragma clang fp contract(off)
  const double inv_base = 1.0 / (double)base;
  double factor = inv_base;
  double result = 0.0;
  while (n > 0) {
    result += factor * (double)perm[n % base];
    n /= base;
    factor *= inv_base;
  }
  return result;
}

// Acklam inverse normal (qmc.rs:37-91) with the ladder's clamp (qmc.rs:284-289),
// as its two regions: the central rational for plow <= p <= 1 - plow, and the
// tails (4.85% of uniforms), whose clamp leaves the region test unchanged.
__device__ __forceinline__ bool probit_is_tail(double p) { return p < 0.02425 || p > 1.0 - 0.02425; }

__device__ __forceinline__ double probit_f64_central(double p) {
  constexpr double A0 = -3.969683028665376e+01, A1 = 2.209460984245205e+02,
                   A2 = -2.759285104469687e+02, A3 = 1.383577518672690e+02,
                   A4 = -3.066479806614716e+01, A5 = 2.506628277459239e+00;
  constexpr double B0 = -5.447609879822406e+01, B1 = 1.615858368580409e+02,
                   B2 = -1.556989798598866e+02, B3 = 6.680131188771972e+01,
                   B4 = -1.328068155288572e+01;
  const double q = p - 0.5;
  const double r = q * q;
  return (((((A0 * r + A1) * r + A2) * r + A3) * r + A4) * r + A5) * q /
         (((((B0 * r + B1) * r + B2) * r + B3) * r + B4) * r + 1.0);
}

__device__ __forceinline__ double probit_f64_tail(double p) {
  constexpr double C0 = -7.784894002430293e-03, C1 = -3.223964580411365e-01,
                   C2 = -2.400758277161838e+00, C3 = -2.549732539343734e+00,
                   C4 = 4.374664141464968e+00, C5 = 2.938163982698783e+00;
  constexpr double D0 = 7.784695709041462e-03, D1 = 3.224671290700398e-01,
                   D2 = 2.445134137142996e+00, D3 = 3.754408661907416e+00;
  if (p < 1.0e-12) p = 1.0e-12;
  if (p > 1.0 - 1.0e-12) p = 1.0 - 1.0e-12;
  const bool lower = p < 0.02425;
  const double q = sqrt(-2.0 * riga_log(lower ? p : 1.0 - p));
  const double num = ((((C0 * q + C1) * q + C2) * q + C3) * q + C4) * q + C5;
  const double den = (((D0 * q + D1) * q + D2) * q + D3) * q + 1.0;
  return lower ? num / den : -num / den;
}

__device__ __forceinline__ double probit_f64(double p) {
  return probit_is_tail(p) ? probit_f64_tail(p) : probit_f64_central(p);
}

// FP32 inverse normal for the fast tier: z = sqrt(2) erfinv(2u - 1) with
// M. Giles' single-precision erfinv ("Approximating the erfinv function", GPU
// Computing Gems Jade Edition, 2011): w = -ln((1 - x)(1 + x)), a polynomial in
// w - 2.5 for w < 5 and in sqrt(w) - 3 beyond.  (1 - x)(1 + x) = 4u(1 - u) and
// x = 2u - 1 are formed in FP64 from the FP64 uniform, so neither cancels; the
// logarithm is the hardware's log2 (the same instruction in every kernel).  Its
// error stays below 1.5e-6 in z for 1e-7 <= u <= 1 - 1e-7; the 2e-7 of uniforms
// outside that take the Acklam tail below, as the FP64 tier's tail.  w >= 5
// (u within 0.00168 of 0 or 1) is the rare case: the central form is
// evaluated branch-free for every coordinate and the rare ones redone after.
struct GilesArg {
  float w, x;
};
__device__ __forceinline__ GilesArg giles_arg(double p) {
  const float t = (float)(4.0 * p * (1.0 - p));
  return GilesArg{-0.693147181f * __builtin_amdgcn_logf(t), (float)(2.0 * p - 1.0)};
}
__device__ __forceinline__ bool giles_rare(GilesArg a) { return !(a.w < 5.0f); }
__device__ __forceinline__ float giles_central(GilesArg a) {
  const float w = a.w - 2.5f;
  float q = 2.81022636e-08f;
  q = fmaf(q, w, 3.43273939e-07f);
  q = fmaf(q, w, -3.5233877e-06f);
  q = fmaf(q, w, -4.39150654e-06f);
  q = fmaf(q, w, 0.00021858087f);
  q = fmaf(q, w, -0.00125372503f);
  q = fmaf(q, w, -0.00417768164f);
  q = fmaf(q, w, 0.246640727f);
  q = fmaf(q, w, 1.50140941f);
  return 1.41421356f * (q * a.x);
}
__device__ __forceinline__ float giles_outer(GilesArg a) {
  const float w = __builtin_amdgcn_sqrtf(a.w) - 3.0f;
  float q = -0.000200214257f;
  q = fmaf(q, w, 0.000100950558f);
  q = fmaf(q, w, 0.00134934322f);
  q = fmaf(q, w, -0.00367342844f);
  q = fmaf(q, w, 0.00573950773f);
  q = fmaf(q, w, -0.0076224613f);
  q = fmaf(q, w, 0.00943887047f);
  q = fmaf(q, w, 1.00167406f);
  q = fmaf(q, w, 2.83297682f);
  return 1.41421356f * (q * a.x);
}
__device__ __forceinline__ bool probit_f32_far(double p) { return p < 1e-7 || p > 1.0 - 1e-7; }

__device__ __forceinline__ float probit_f32_tail(double p) {
  constexpr float C0 = -7.784894002430293e-03f, C1 = -3.223964580411365e-01f,
                  C2 = -2.400758277161838e+00f, C3 = -2.549732539343734e+00f,
                  C4 = 4.374664141464968e+00f, C5 = 2.938163982698783e+00f;
  constexpr float D0 = 7.784695709041462e-03f, D1 = 3.224671290700398e-01f,
                  D2 = 2.445134137142996e+00f, D3 = 3.754408661907416e+00f;
  if (p < 1.0e-12) p = 1.0e-12;
  if (p > 1.0 - 1.0e-12) p = 1.0 - 1.0e-12;
  const bool lower = p < 0.5;
  const float t = (float)(lower ? p : 1.0 - p);
  // the hardware's log2, square root and reciprocal: one instruction each, the
  // same in every kernel variant (the compiler lowers the library forms by
  // context, and a path's normals must not depend on the launch it is priced in)
  const float q = __builtin_amdgcn_sqrtf(-1.386294361f * __builtin_amdgcn_logf(t));
  const float num = fmaf(fmaf(fmaf(fmaf(fmaf(C0, q, C1), q, C2), q, C3), q, C4), q, C5);
  const float den = fmaf(fmaf(fmaf(fmaf(D0, q, D1), q, D2), q, D3), q, 1.0f);
  const float z = num * __builtin_amdgcn_rcpf(den);
  return lower ? z : -z;
}

// The rare coordinates (w >= 5): the outer polynomial, or the Acklam tail
// beyond 1e-7 of 0 or 1.
__device__ __forceinline__ float probit_f32_rare(double p) {
  return probit_f32_far(p) ? probit_f32_tail(p) : giles_outer(giles_arg(p));
}

__device__ __forceinline__ float probit_f32(double p) {
  const GilesArg a = giles_arg(p);
  return giles_rare(a) ? probit_f32_rare(p) : giles_central(a);
}

__device__ __forceinline__ double probit_tail_tier(double p, double) { return probit_f64_tail(p); }
__device__ __forceinline__ float probit_tail_tier(double p, float) { return probit_f32_rare(p); }

// Inverse normals of one node's F uniforms: the main form for every coordinate,
// branch-free (FP64: the central rational; FP32: the central erfinv polynomial),
// then the rare coordinates (FP64: the 4.85% of uniforms in the tails; FP32: the
// 0.34% with w >= 5) one per lane per pass while any lane of the wave has one.  A
// branch per coordinate would run both forms in most waves and serialize the F
// coordinates' dependent chains; this way they interleave and the rare forms
// cost about one pass per node.  Each z is the same value either way.
template <typename Real, int F>
__device__ __forceinline__ void probit_node(const double (&u)[F], Real (&z)[F]) {
  uint32_t pend = 0;
#pragma unroll
  for (int f = 0; f < F; ++f) {
    if constexpr (sizeof(Real) == 8) {
      z[f] = probit_f64_central(u[f]);
      pend |= (uint32_t)probit_is_tail(u[f]) << f;
    } else {
      const GilesArg a = giles_arg(u[f]);
      z[f] = giles_central(a);
      pend |= (uint32_t)giles_rare(a) << f;
    }
  }
  while (__ballot(pend != 0u) != 0ull) {
    if (pend) {
      const int g = __builtin_ctz(pend);
      pend &= pend - 1u;
      double ug = u[0];
#pragma unroll
      for (int f = 1; f < F; ++f)
        if (g == f) ug = u[f];
      const Real zg = probit_tail_tier(ug, Real(0));
#pragma unroll
      for (int f = 0; f < F; ++f)
        if (g == f) z[f] = zg;
    }
  }
}

// ---------------------------------------------------------------- double-double
__device__ __forceinline__ DD dd_from(double x) { return DD{x, 0.0}; }

__device__ __forceinline__ DD dd_add(DD a, DD b) {
#pragma clang fp contract(off)
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b.hi - bb) + (a.lo + b.lo);
  const double hi = s + e;
  return DD{hi, e - (hi - s)};
}

// Exact square x*x = p + e (FMA error term).
__device__ __forceinline__ DD dd_sq(double x) {
  const double p = x * x;
  return DD{p, __builtin_fma(x, x, -p)};
}

__device__ __forceinline__ DD dd_shfl_xor(DD v, int m) {
  return DD{__shfl_xor(v.hi, m, 64), __shfl_xor(v.lo, m, 64)};
}

__device__ __forceinline__ DD wave_sum(DD v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = dd_add(v, dd_shfl_xor(v, m));
  return v;
}

// ---------------------------------------------------------------- bridge stack
// The streaming bridge keeps the W values of pending right endpoints, one
// F-vector per level.  The top level lives in VGPRs; the levels below it live
// in this lane's own LDS column ([level][factor][lane], so a wave touches 64
// consecutive words: no bank conflicts and no barrier).  The depth is uniform
// across the block.  A step that needs c new nodes pushes c-1 of them (the
// last one is emitted at once) and a step that needs none pops one.
//
// A launch may keep some stack positions in the workgroup's HBM slot instead,
// so that a deep stack does not hold occupancy below what registers allow: the
// least used ones.  Position k is the k-th left turn on the way to a leaf, so
// its pushes grow binomially towards the middle (756 steps: 16, 56, 112, 140,
// 112, 56, 16, 2 accesses per path): the bottom `low_levels` positions and the
// top `hbm_levels - low_levels` go to HBM, the middle stays in LDS.  A value a
// pop reads is first needed a step later.
template <typename Real>
using LdsPtr = __attribute__((address_space(3))) Real*;     // an LDS address: ds_* accesses only
template <typename Real>
using GlobalPtr = __attribute__((address_space(1))) Real*;  // an HBM address: global_* accesses only

template <typename Real, int F>
struct BridgeStack {
  Real top[F];
  LdsPtr<Real> col;     // this lane's LDS column: col[((sp - low_levels)*F + f) * stride]
  GlobalPtr<Real> hbm;  // this lane's HBM column: hbm[(slot*F + f) * stride]
  int stride;           // lanes per workgroup
  int sp;               // levels held below the top
  int low_levels;       // positions [0, low_levels) in HBM slots 0..
  int mid_end;          // positions [low_levels, mid_end) in LDS, [mid_end, ...) in HBM after the low ones

  // where position p lives: an LDS address, or (returned false) HBM slot *h
  __device__ __forceinline__ bool in_lds(int p, int* h) const {
    if (p < low_levels) {
      *h = p;
      return false;
    }
    if (p >= mid_end) {
      *h = low_levels + (p - mid_end);
      return false;
    }
    return true;
  }
  __device__ __forceinline__ void push(const Real (&v)[F]) {
    int h;
    if (in_lds(sp, &h)) {
#pragma unroll
      for (int f = 0; f < F; ++f) col[((sp - low_levels) * F + f) * stride] = top[f];
    } else {
#pragma unroll
      for (int f = 0; f < F; ++f) hbm[(h * F + f) * stride] = top[f];
    }
#pragma unroll
    for (int f = 0; f < F; ++f) top[f] = v[f];
    ++sp;
  }
  // The last step's pop finds the stack empty (its top, W(T), was the last value
  // needed): nothing is read then.
  __device__ __forceinline__ void pop() {
    if (sp == 0) return;
    --sp;
    int h;
    if (in_lds(sp, &h)) {
#pragma unroll
      for (int f = 0; f < F; ++f) top[f] = col[((sp - low_levels) * F + f) * stride];
    } else {
#pragma unroll
      for (int f = 0; f < F; ++f) top[f] = hbm[(h * F + f) * stride];
    }
  }
};

template <typename T>
__device__ __forceinline__ const T* at(const unsigned char* arena, int64_t off) {
  return reinterpret_cast<const T*>(arena + off);
}

// Tier constants: the FP64 tier reads the host's doubles, the FP32 tier their
// pre-rounded float copies (no conversion inside the path loop).
__device__ __forceinline__ double chol_of(const ItemDesc& it, int k, double) { return it.chol[k]; }
__device__ __forceinline__ float chol_of(const ItemDesc& it, int k, float) { return it.chol32[k]; }
__device__ __forceinline__ double sd_of(const BridgeNode& b, double) { return b.sd; }
__device__ __forceinline__ float sd_of(const BridgeNode& b, float) { return b.sd32; }
__device__ __forceinline__ double root_sd_of(const ItemDesc& it, double) { return it.root_sd; }
__device__ __forceinline__ float root_sd_of(const ItemDesc& it, float) { return it.root_sd32; }

// Normals of one node from the batch's shared [dim][path] table (same ladder,
// same point), and the correlation y = L z with the reference's sequential row
// sums (correlation.rs:196-206).
template <typename Real, int F>
__device__ __forceinline__ void table_normals(const ItemDesc& it, int node, uint32_t n, Real (&z)[F]) {
  const Real* zt = reinterpret_cast<const Real*>(it.ztable);
#pragma unroll
  for (int f = 0; f < F; ++f) z[f] = zt[(size_t)((uint32_t)node * F + f) * (size_t)it.zt_stride + (n - 1)];
}

// The tiers' multiply-adds, written out: both tiers are compiled without
// contraction, so the FP64 tier rounds the product and the sum apart as the
// reference does, and the FP32 tier fuses exactly where it says so -- the same
// instructions in every kernel variant, so a result does not depend on the
// launch it was priced in.
__device__ __forceinline__ double tier_madd(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float tier_madd(float a, float b, float c) { return fmaf(a, b, c); }

template <typename Real, int F>
__device__ __forceinline__ void correlate(const ItemDesc& it, const Real (&z)[F], Real (&y)[F]) {
#pragma unroll
  for (int i = 0; i < F; ++i) {
    Real acc = Real(0);
#pragma unroll
    for (int j = 0; j <= i; ++j) acc = tier_madd(chol_of(it, i * kMaxFactors + j, Real(0)), z[j], acc);
    y[i] = acc;
  }
}

// The Sobol ladder of one bridge node in two halves, so a lane can issue the
// next node's loads a step ahead: the fetch forms the unscrambled coordinates
// (dims node*F + f) or reads their scrambled values from the Owen coordinate table; the finish
// hashes the remaining levels and applies the inverse normal.
template <int F>
struct SobolFetch {
  uint32_t x[F];
  uint32_t top[F];  // coordinate-table entries (unused without a table)
};

// An item reads its scrambled coordinates from the seed's coordinate table when the
// host found one covering all its dimensions and points (kOwenTable); otherwise it
// forms the Sobol integer and hashes every Owen level -- the same integers.  The
// decision is per item, so the step loop holds one path or the other.
template <int F>
__device__ __forceinline__ void sobol_fetch(const ItemDesc& it, const unsigned char* arena, uint32_t n, uint32_t gray,
                                            int node, SobolFetch<F>& sf) {
  // the table holds points below 2^kOwenTableBits (the host sets the flag only then)
  const bool tab = (it.flags & kOwenTable) != 0 && it.nbits <= kOwenTableBits;
#pragma unroll
  for (int f = 0; f < F; ++f) {
    const uint32_t dim = (uint32_t)node * F + f;
    if (tab) {
      // a streaming read (each entry once per launch; the wave's 64 entries are
      // adjacent): non-temporal, so it does not push the shared tables out of the caches
      // lanes past the item's last path (results unused) may hold points past the
      // table: the mask keeps their reads inside the row
      sf.x[f] = 0u;
      sf.top[f] = (uint32_t)__builtin_nontemporal_load(
          &it.owen_tab[((size_t)dim << kOwenTableBits) + (n & ((1u << kOwenTableBits) - 1u))]);
    } else {
      const uint32_t d = dim % (uint32_t)it.ladder_dims;
      const uint32_t* row = at<uint32_t>(arena, it.off_dirs) + (size_t)d * 32;
      sf.x[f] = sobol_x(gray, row, it.nbits);
      sf.top[f] = 0u;
    }
  }
}

// The FP32 tier's scrambled integers of one node: 24 hashed levels per
// coordinate; the few lanes with a coordinate in a tail then finish its last 8
// levels, one coordinate per lane per pass, while any lane of the wave has one
// left.
template <int F>
__device__ __forceinline__ void sobol_ints_f32(const ItemDesc& it, int node, const SobolFetch<F>& sf,
                                               uint32_t (&xs)[F]) {
  const bool tab = (it.flags & kOwenTable) != 0 && it.nbits <= kOwenTableBits;
  if (tab) {
    // the entries are the scrambled coordinates: nothing left to hash
#pragma unroll
    for (int f = 0; f < F; ++f) xs[f] = owen_f32_of_whole(sf.top[f], (it.flags & kFullLadder) != 0);
    return;
  }
  // the full-ladder option sends every coordinate through the last 8 levels
  const uint32_t all = (it.flags & kFullLadder) ? (1u << F) - 1u : 0u;
  uint64_t keys[F];
  uint32_t pend = all;
#pragma unroll
  for (int f = 0; f < F; ++f) {
    keys[f] = it.seed ^ ((uint64_t)((uint32_t)node * F + f) * 0x9E3779B97F4A7C15ull);
    const uint32_t p = owen_top24(sf.x[f], keys[f]);
    xs[f] = (p << 8) | 0x80u;
    pend |= (uint32_t)owen_f32_tail(p) << f;
  }
  while (__ballot(pend != 0u) != 0ull) {
    if (pend) {
      const int g = __builtin_ctz(pend);
      pend &= pend - 1u;
      uint32_t xg = sf.x[0], pg = xs[0];
      uint64_t kg = keys[0];
#pragma unroll
      for (int f = 1; f < F; ++f)
        if (g == f) {
          xg = sf.x[f];
          pg = xs[f];
          kg = keys[f];
        }
      const uint32_t full = owen_low8(xg, kg, pg >> 8);
#pragma unroll
      for (int f = 0; f < F; ++f)
        if (g == f) xs[f] = full;
    }
  }
}

template <typename Real, int F, bool kUnrollOwen>
__device__ __forceinline__ void sobol_finish(const ItemDesc& it, int node, const SobolFetch<F>& sf, Real (&z)[F]) {
  const bool tab = (it.flags & kOwenTable) != 0 && it.nbits <= kOwenTableBits;
  if constexpr (sizeof(Real) == 4) {
    uint32_t xs[F];
    sobol_ints_f32<F>(it, node, sf, xs);
    double u[F];
#pragma unroll
    for (int f = 0; f < F; ++f) u[f] = ((double)xs[f] + 0.5) / 4294967296.0;
    probit_node<Real, F>(u, z);
  } else {
    double u[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const uint64_t key = it.seed ^ ((uint64_t)((uint32_t)node * F + f) * 0x9E3779B97F4A7C15ull);
      // full unroll where registers allow (the FP64 10-scenario lane keeps the loop)
      const uint32_t xs = tab ? sf.top[f] : owen_scramble(sf.x[f], key);
      u[f] = ((double)xs + 0.5) / 4294967296.0;
    }
    probit_node<Real, F>(u, z);
  }
}

// Correlated normals of one bridge node: dims node*F + f (note.rs:89-98 with F
// factors) from the batch's ladder -- Sobol with Owen scrambling, Halton, or the
// shared [dim][path] table -- and y = L z with the reference's sequential row
// sums (correlation.rs:196-206).
template <typename Real, int F, bool kUnrollOwen = true>
__device__ __forceinline__ void node_normals(const ItemDesc& it, const unsigned char* arena,
                                             uint32_t n, uint32_t gray, int node, Real (&y)[F]) {
  Real z[F];
  if (it.flags & kNormalTable) {
    table_normals<Real, F>(it, node, n, z);
  } else if (it.flags & kHalton) {
    double u[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const uint32_t d = ((uint32_t)node * F + f) % (uint32_t)it.ladder_dims;
      const uint32_t* h = at<uint32_t>(arena, it.off_halton);
      u[f] = halton_u(n, h[d], h + h[it.ladder_rows + d]);
    }
    probit_node<Real, F>(u, z);
  } else {
    SobolFetch<F> sf;
    sobol_fetch<F>(it, arena, n, gray, node, sf);
    sobol_finish<Real, F, kUnrollOwen>(it, node, sf, z);
  }
  correlate<Real, F>(it, z, y);
}


// The FP32 tier's exp of a log-state value: the hardware's exp2 (one
// instruction, the same in every kernel variant), within 2 ulp of float.
__device__ __forceinline__ float exp_f32(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

// exp/log/divide per tier: FP64 uses exp_log.h's functions (the same bits as the
// oracle's host build of them) and IEEE division; FP32 the hardware's exp2, log2
// and reciprocal, one instruction each (the library forms are lowered
// differently from one kernel variant to the next).
__device__ __forceinline__ double tier_exp(double x) { return riga_exp(x); }
__device__ __forceinline__ float tier_exp(float x) { return exp_f32(x); }
__device__ __forceinline__ double tier_log(double x) { return riga_log(x); }
__device__ __forceinline__ float tier_log(float x) { return 0.693147181f * __builtin_amdgcn_logf(x); }
__device__ __forceinline__ double tier_div(double a, double b) { return a / b; }
__device__ __forceinline__ float tier_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Local-vol lookup sigma(t, S) (riga-market/src/lib.rs:305-343) against a
// surface whose knots and two bracketing tenor rows sit in LDS.
// `arg` is the spot S (FP64 tier: x = ln max(S, 1e-12), as the reference) or,
// in the FP32 log-state tier, already x = ln S (clamped at ln 1e-12).
template <typename Real, bool kLogArg = false>
__device__ __forceinline__ Real lv_sigma(const Real* __restrict__ ls, int nx, Real inv_dx, Real wt,
                                         Real arg) {
  const Real* r0 = ls + nx;
  const Real* r1 = ls + 2 * nx;
  Real x;
  if constexpr (kLogArg)
    x = arg > Real(-27.631021115928547) ? arg : Real(-27.631021115928547);
  else
    x = tier_log(arg > Real(1e-12) ? arg : Real(1e-12));
  int lo, hi;
  if (x <= ls[0]) {
    lo = hi = 0;
  } else if (x >= ls[nx - 1]) {
    lo = hi = nx - 1;
  } else {
    int j = (int)((x - ls[0]) * inv_dx);
    j = j < 0 ? 0 : (j > nx - 2 ? nx - 2 : j);
    while (j > 0 && !(x > ls[j])) --j;
    while (j < nx - 2 && x > ls[j + 1]) ++j;
    lo = j;
    hi = j + 1;
  }
  const Real wx = (hi == lo) ? Real(0) : tier_div(x - ls[lo], ls[hi] - ls[lo]);
  const Real v0 = (Real(1) - wx) * r0[lo] + wx * r0[hi];
  const Real v1 = (Real(1) - wx) * r1[lo] + wx * r1[hi];
  return (Real(1) - wt) * v0 + wt * v1;
}

// FP32 lookup against a surface with evenly spaced knots.  LDS holds, per knot
// j, {r0[j], r0[j+1]-r0[j], r1[j], r1[j+1]-r1[j]} of the two bracketing tenor
// rows, so one 16-byte read and three FMAs give the bilinear value.  The state
// is xr = ln(S/S0); its knot position t = (xr + ln S0 - x0)/dx = xr/dx + rel_off
// is clamped to [0, n_x-1] (below the first knot and above the last the edge
// value holds, as in the general lookup).
__device__ __forceinline__ float lv_sigma_uniform(const float4* __restrict__ rows, int nx,
                                                  float inv_dx, float rel_off, float wt,
                                                  float xr) {
  float t = fmaf(xr, inv_dx, rel_off);
  t = fminf(fmaxf(t, 0.0f), (float)(nx - 1));
  int j = (int)t;
  j = j < nx - 2 ? j : nx - 2;
  const float wx = t - (float)j;
  const float4 e = rows[j];
  const float v0 = fmaf(wx, e.y, e.x);
  const float v1 = fmaf(wx, e.w, e.z);
  return fmaf(wt, v1 - v0, v0);
}

// The FP32 tier's lookup on unevenly spaced knots, out of line: a rare path
// whose search loops would otherwise widen the register budget of the kernel.
__device__ __noinline__ float lv_sigma_f32_general(const float* __restrict__ ls, int nx, float inv_dx, float wt,
                                                   float x) {
  return lv_sigma<float, true>(ls, nx, inv_dx, wt, x);
}

// ---------------------------------------------------------------- shared normal table
// Ladder normals z[dim][path] for every raw dimension a batch of same-seed
// trades uses (dims without a wrapped direction row); blockIdx.y is the
// dimension, so the direction row, the key and the Owen table row are uniform.
template <typename Real>
__global__ __launch_bounds__(256) void normal_table_kernel(const uint32_t* __restrict__ dirs,
                                                           const OwenEntry* __restrict__ owen_tab_in,
                                                           int owen_dims, uint64_t seed, int n_paths, int nbits,
                                                           Real* __restrict__ out) {
  const int path = blockIdx.x * 256 + threadIdx.x;
  const uint32_t dim = blockIdx.y;
  if (path >= n_paths) return;
  // the coordinate table covers the first owen_dims dimensions; the others hash every level
  const OwenEntry* owen_tab = (int)dim < owen_dims && nbits <= kOwenTableBits ? owen_tab_in : nullptr;
  const uint32_t n = (uint32_t)path + 1;
  const uint64_t key = seed ^ ((uint64_t)dim * 0x9E3779B97F4A7C15ull);
  uint32_t xs;
  if (owen_tab) {
    xs = owen_tab[((size_t)dim << kOwenTableBits) + (n & ((1u << kOwenTableBits) - 1u))];
  } else {
    const uint32_t gray = n ^ (n >> 1);
    const uint32_t x = sobol_x(gray, dirs + (size_t)dim * 32, nbits);
    xs = sizeof(Real) == 8 ? owen_scramble(x, key) : owen_f32(x, key);
  }
  Real z;
  if constexpr (sizeof(Real) == 8) {
    z = probit_f64(((double)xs + 0.5) / 4294967296.0);
  } else {
    // the FP32 tier's 24-level rule, as the inline ladder applies it
    if (owen_tab) xs = owen_f32_of_whole(xs, false);
    z = probit_f32(((double)xs + 0.5) / 4294967296.0);
  }
  out[(size_t)dim * n_paths + path] = z;
}

// ---------------------------------------------------------------- the kernel
// LDS-address-space views (32-bit addresses) for reads from a staged address.
typedef __attribute__((address_space(3))) const unsigned char LdsBytes;
typedef float F32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const F32x4 LdsF32x4;

// kMode: how the items of the launch get their normals (kLadderDirect: Sobol
// or Halton per lane; kLadderTable: the shared normal table; kLadderMonteCarlo:
// the ChaCha8 stream).  Each mode compiles only its own ladder code.
//
// Work queue: each workgroup claims its (item, chunk) unit from the launch's
// device counter -- the next index of the block table, which lists the longest
// items first -- so units start longest first in the order workgroups get a
// CU, whatever the dispatcher's order.  A unit's results (its block partial)
// do not depend on which workgroup ran it.  `queue` is zero when the launch
// starts.  The grid is either one workgroup per unit (each claims once) or a
// persistent grid of at most the workgroups the device holds at once (each
// claims until the table is empty).  The unit loop carries no lane-derived
// values across its back edge (the lane id enters each unit through an opaque
// copy), so it costs the kernels no registers in their hot loops.
// Resident waves per SIMD each path kernel is compiled for (its register
// budget: 512 / waves VGPRs per lane), chosen per variant from the compiler's
// spill counts and timings: more waves wherever the kernel's live state fits the
// smaller budget or spills only what the step loop rarely touches.  The
// ten-scenario FP32 lane of up to three factors runs three (its spills are the
// coupon legs, read and written at period ends and knock-outs: 49.6 -> 39.2 ms
// for the ten-scenario 3-asset batch, with its LDS, the exactly sized bridge
// stack and twelve surfaces' slope rows, admitting three workgroups per CU); the
// FP64 one-scenario direct kernel of up to three factors runs four (16.0 ->
// 14.1 ms on the 262,144-path local-vol trade).  The Monte Carlo kernels keep
// the compiler's own budget (112 VGPRs for the ten-scenario FP32 one; its word
// ring's LDS admits two workgroups per CU anyway): built for two waves, that
// kernel ran 38.8 ms instead of 34.2 on ten 1,048,576-path scenarios.
template <typename Real, int F, int SMAX, int kMode>
constexpr int path_kernel_waves() {
  if constexpr (kMode == kLadderMonteCarlo) {
    return 1;
  } else if constexpr (sizeof(Real) == 8) {
    if (SMAX >= 10) return 1;
    if (SMAX >= 4) return 2;
    if (kMode == kLadderDirect) return F <= 3 ? 4 : 3;
    return 3;
  } else {
    if (SMAX >= 10) return F <= 3 ? 3 : 2;
    if (SMAX >= 4) return 3;
    return 4;
  }
}

template <typename Real, int F, int SMAX, int kMode>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(path_kernel_waves<Real, F, SMAX, kMode>(), 8)))
void phoenix_kernel(const ItemDesc* __restrict__ items,
                                                          const int32_t* __restrict__ block_item,
                                                          const unsigned char* __restrict__ arena,
                                                          Partial* __restrict__ partials,
                                                          PathRecordDev* __restrict__ records,
                                                          uint32_t* __restrict__ queue, int n_units) {
  extern __shared__ __align__(16) unsigned char lds_dyn[];
  __shared__ BlockShared sh;      // one scenario's wave sums at a time
  __shared__ int ev_n[SMAX];      // KO events this block has recorded per scenario (trace)
  __shared__ SfF32 sf_lds[sizeof(Real) == 4 ? SMAX * F : 1];  // FP32 Euler constants, read per step
  __shared__ uint32_t unit_sh, slot_sh;

  // Each workgroup claims (item, chunk) units from the launch's counter, longest
  // items first.  With one workgroup per unit it claims once; a persistent grid
  // (at most the workgroups the device holds at once) claims until the table is
  // used up.  The counter only grows, so every workgroup reaches the end of the
  // table and leaves.
  for (;;) {
    __syncthreads();  // every lane is done with the previous unit's LDS (unit id, constants, sums)
    if (threadIdx.x == 0) unit_sh = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t unit = unit_sh;
    if (unit >= (uint32_t)n_units) return;  // the same value in every lane: the whole block leaves
    const int item = block_item[2 * unit];  // block table: (item, chunk) pairs
    const ItemDesc& it = items[item];
    const int S = it.S;
    const int steps = it.steps;
    // the lane id through an opaque copy per unit: values derived from it are then
    // formed where they are used, not hoisted out of the unit loop and carried in
    // registers through the path loops
    int tid_unit = threadIdx.x;
    asm volatile("" : "+v"(tid_unit));
    const int tid = tid_unit;
    const int nthreads = blockDim.x;  // 64, 128 or 256: the launch shape the autotuner picked
    const int nwaves = nthreads >> 6;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int chunk = block_item[2 * unit + 1];
    const bool local_vol = (it.flags & kLocalVol) != 0;
    const bool trace = (it.flags & kTrace) != 0;
    const bool want_records = (it.flags & kRecords) != 0 && records != nullptr;
    constexpr bool kLogState = sizeof(Real) == 4;  // FP32 tier carries X = ln(S / S0)
    constexpr bool kMC = kMode == kLadderMonteCarlo;
    constexpr bool kTab = kMode == kLadderTable;
    constexpr int kSurfWords = kLogState ? 4 : 3;   // LDS words per knot and surface
    const bool lv_uniform = kLogState && (it.flags & kLvUniform) != 0;
    // dynamic LDS: QMC lanes [bridge stack: (depth-1) x F x lanes][local-vol rows:
    // n_surf x kSurfWords x max_nx]; the Monte Carlo lane [stream ring: slots x
    // lanes][ziggurat x_i, f(x_i)][local-vol rows]
    uint64_t* mc_ring = reinterpret_cast<uint64_t*>(lds_dyn);
    constexpr int kRingSlots = mc_ring_slots(SMAX);
    double* zx = reinterpret_cast<double*>(mc_ring + (kMC ? kRingSlots * nthreads : 0));
    double* zf = zx + kZigLen;
    Real* lds_stack = reinterpret_cast<Real*>(lds_dyn);
    // the bottom bridge levels' HBM slot: claimed for the unit from the launch's
    // pool, which holds a slot for every workgroup that can be resident at once
    // plus one spare per CU.  The claim scans the pool at most eight times, with a
    // short sleep between passes, so it ends.
    Real* stack_low = nullptr;
    if (it.stack_pooled > 0) {
      if (tid == 0) {
        const uint32_t n = (uint32_t)it.pool_n;
        const uint32_t first = blockIdx.x % n;
        uint32_t slot = first, got = 0xFFFFFFFFu;
        for (uint32_t probe = 0; probe < 8 * n; ++probe) {
          if (atomicCAS(it.pool_flags + slot, 0u, 1u) == 0u) {
            got = slot;
            break;
          }
          slot = slot + 1 == n ? 0 : slot + 1;
          if (slot == first) __builtin_amdgcn_s_sleep(32);
        }
        if (got == 0xFFFFFFFFu) atomicAdd(it.pool_flags + n, 1u);  // the host reports the plan failed
        slot_sh = got;
      }
      __syncthreads();
      if (slot_sh == 0xFFFFFFFFu) return;  // uniform: the whole block leaves
      stack_low = reinterpret_cast<Real*>(it.stack_pool) + (size_t)slot_sh * it.stack_pooled * F * nthreads;
    }
    Real* lds_w = kMC ? reinterpret_cast<Real*>(zf + kZigLen) : lds_stack + (size_t)it.stack_levels * F * nthreads;
    const Real* lds = lds_w;
    const int nxs = it.max_nx;
    const StepInfo* stepinfo = at<StepInfo>(arena, it.off_steps);
    const BridgeNode* bnodes = at<BridgeNode>(arena, it.off_nodes);
    const LvBracket* lvb = at<LvBracket>(arena, it.off_lvb);
    const ScenarioDesc* scen = at<ScenarioDesc>(arena, it.off_scen);
    const LvSurface* surf = at<LvSurface>(arena, it.off_surf);

    for (int s = tid; s < SMAX; s += nthreads) ev_n[s] = 0;
    if constexpr (sizeof(Real) == 4) {
      const SfF32* sfg = at<SfF32>(arena, it.off_sf32);
      // surface offsets are staged as absolute LDS addresses, so a slope-row read
      // is one multiply-add of the knot index away from its address
      const int32_t lv_base = (int32_t)(uint32_t)(uintptr_t)((const LdsBytes*)lds_w);
      for (int i = tid; i < S * F; i += nthreads) {
        SfF32 c = sfg[i];
        c.lds_off += lv_base;
        sf_lds[i] = c;
      }
    }
    if constexpr (kMC) {
      const double* zg = at<double>(arena, it.off_zig);
      for (int i = tid; i < 2 * kZigLen; i += nthreads) zx[i] = zg[i];
    }
    __syncthreads();
    const Real sqrt_dt_r = (Real)it.sqrt_dt;
    // -dt/2 for the FP32 step, held in a vector register so the step loop does
    // not reload it through the scalar cache (whose wait also stalls LDS reads)
    float nhdt = it.neg_half_dt;
    asm volatile("" : "+v"(nhdt));
    // the FP32 scan's barriers and coupon rate, held the same way: reloading them
    // per scenario through the scalar cache would stall on every in-flight LDS read
    float ln_ko32 = it.ln_ko32, ln_cb32 = it.ln_cb32, cr32 = (float)it.coupon_rate;
    asm volatile("" : "+v"(ln_ko32), "+v"(ln_cb32), "+v"(cr32));
    // and the FP64 step's constants (its kernels have registers to spare: their
    // bridge stack in LDS holds them at two workgroups per CU)
    double dt64 = 0.0, sqdt64 = 0.0, ko64 = 0.0, cb64 = 0.0, cr64 = 0.0;
    double s0_64[F];
  #pragma unroll
    for (int f = 0; f < F; ++f) s0_64[f] = 1.0;
    if constexpr (!kLogState) {
      dt64 = it.dt;
      sqdt64 = it.sqrt_dt;
      ko64 = it.ko;
      cb64 = it.cb;
      cr64 = it.coupon_rate;
      asm volatile("" : "+v"(dt64), "+v"(sqdt64), "+v"(ko64), "+v"(cb64), "+v"(cr64));
  #pragma unroll
      for (int f = 0; f < F; ++f) {
        s0_64[f] = it.initial_spot[f];
        asm volatile("" : "+v"(s0_64[f]));
      }
    }

    const int ppb = it.paths_per_block;
    const int batches = ppb / nthreads;
    for (int b = 0; b < batches; ++b) {
      const uint64_t path_local = (uint64_t)chunk * ppb + (uint64_t)b * nthreads + tid;
      const bool valid = path_local < it.paths;
      const uint64_t path = it.path_list ? it.path_list[valid ? path_local : 0] : it.path_begin + path_local;
      const uint32_t n = (uint32_t)(path + 1);  // Sobol/Halton point index (qmc.rs:286, 121)
      const uint32_t gray = n ^ (n >> 1);

      // ---- path state
      BridgeStack<Real, F> bs;
      bs.col = (LdsPtr<Real>)(lds_stack + tid);
      bs.stride = nthreads;
      bs.sp = 0;
      bs.low_levels = it.stack_pooled > 0 ? it.stack_pooled_low : 0;
      bs.mid_end = bs.low_levels + it.stack_levels;
      bs.hbm = it.stack_pooled > 0 ? (GlobalPtr<Real>)(stack_low + tid) : nullptr;
      Real wlo[F];
      Real spot[SMAX][F];
      int ko_step[SMAX];
      Real accrual[SMAX], unpaid[SMAX], cpv[SMAX], mpv[SMAX];
      bool received[SMAX];
      // FP32 tier: KO step + 1 (bits 0-19, 0 = not knocked out), unpaid memory
      // coupons as a count (bits 20-30) and coupon received (bit 31) in one word
      uint32_t pk[SMAX];
      Real worst_last[SMAX];
      McRing<kRingSlots> ring;
      McCursor cur;
      uint64_t wpos = 0;  // stream index of the lane's next word
      if constexpr (kMC) {
        // the path's words start where the parse found its first draw
        wpos = it.mc_start[valid ? path : it.path_begin];
        ring.base = mc_ring + tid;
        ring.stride = nthreads;
        ring.seek(it.mc_key, wpos);
        cur.slow = it.mc_slow;
        cur.val = it.mc_val;
        cur.chunks = it.mc_chunks;
        cur.n_chunks = it.mc_n_chunks;
        cur.C = it.mc_chunk_words;
        cur.cap = it.mc_slow_cap;
        cur.seek(wpos);
      } else {
        Real y[F];
        if constexpr (kTab) {
          Real z[F];
          table_normals<Real, F>(it, 0, n, z);
          correlate<Real, F>(it, z, y);
        } else {
          node_normals<Real, F, !(sizeof(Real) == 8 && SMAX > 4)>(it, arena, n, gray, 0, y);
        }
  #pragma unroll
        for (int f = 0; f < F; ++f) {
          bs.top[f] = y[f] * root_sd_of(it, Real(0));  // W[T] = y_0 sqrt(T dt), qmc.rs:165
          wlo[f] = Real(0);
        }
      }
  #pragma unroll
      for (int s = 0; s < SMAX; ++s) {
  #pragma unroll
        for (int f = 0; f < F; ++f)
          spot[s][f] = kLogState ? (Real)scen[s < S ? s : 0].ln_rel[f] : (Real)scen[s < S ? s : 0].spot[f];
        ko_step[s] = -1;
        accrual[s] = unpaid[s] = cpv[s] = mpv[s] = Real(0);
        received[s] = false;
        pk[s] = 0u;
        worst_last[s] = kLogState ? Real(0) : Real(1);  // final_worst starts at 1.0 (note.rs:115)
      }

      int node_ptr = 0;
      // The next node's loads are issued one descent ahead, so their latency
      // overlaps the current step's scenario work: its normals from the shared
      // table, or its Sobol coordinates or Owen coordinate-table entries.
      const bool sobol = kMode == kLadderDirect && (it.flags & (kNormalTable | kHalton)) == 0;
      Real zn[F];
      SobolFetch<F> sf;
  #pragma unroll
      for (int f = 0; f < F; ++f) {
        zn[f] = Real(0);
        sf.x[f] = sf.top[f] = 0u;
      }
      if constexpr (kTab) {
        if (steps > 1) table_normals<Real, F>(it, bnodes[0].node, n, zn);
      } else {
        if (sobol && steps > 1) sobol_fetch<F>(it, arena, n, gray, bnodes[0].node, sf);
      }
      int loaded_i0[F];
  #pragma unroll
      for (int f = 0; f < F; ++f) loaded_i0[f] = -1;

      for (int step = 0; step < steps; ++step) {
        const StepInfo si = stepinfo[step];
        Real shock[F];
        Real dw[F];  // Brownian increment W[t+1] - W[t]
        if constexpr (kMC) {
          // ---- Monte Carlo lane (note.rs:123-126): F StandardNormals in stream
          // order, shocks = L z.  A wave tops up every lane's ring together when
          // any lane is short of F words, so the ChaCha blocks run converged.
          if (__ballot(ring.fill < F) != 0ull && ring.fill <= kRingSlots - 8) ring.append(it.mc_key);
          Real z[F];
  #pragma unroll
          for (int f = 0; f < F; ++f) z[f] = (Real)mc_draw(ring, cur, wpos, it.mc_key, zx);
          correlate<Real, F>(it, z, shock);
  #pragma unroll
          for (int f = 0; f < F; ++f) dw[f] = shock[f] * sqrt_dt_r;
        } else {
        // ---- streaming bridge: descend to the interval (step, step+1)
        Real wnext[F];
        if (si.n_desc == 0) {
  #pragma unroll
          for (int f = 0; f < F; ++f) wnext[f] = bs.top[f];
          bs.pop();
        } else {
          for (int d = 0; d < si.n_desc; ++d) {
            const BridgeNode bn = bnodes[node_ptr++];
            Real y[F], wmid[F];
            if constexpr (kTab) {  // prefetched a step ahead; fetch the following node's now
              correlate<Real, F>(it, zn, y);
              if (node_ptr < steps - 1) table_normals<Real, F>(it, bnodes[node_ptr].node, n, zn);
            } else if (sobol) {
              Real z[F];
              sobol_finish<Real, F, !(sizeof(Real) == 8 && SMAX > 4)>(it, bn.node, sf, z);
              correlate<Real, F>(it, z, y);
              if (node_ptr < steps - 1) sobol_fetch<F>(it, arena, n, gray, bnodes[node_ptr].node, sf);
            } else {
              node_normals<Real, F, !(sizeof(Real) == 8 && SMAX > 4)>(it, arena, n, gray, bn.node, y);
            }
  #pragma unroll
            for (int f = 0; f < F; ++f)
              wmid[f] = tier_madd(y[f], sd_of(bn, Real(0)), Real(0.5) * (wlo[f] + bs.top[f]));  // qmc.rs:185-187
            if (d + 1 < si.n_desc) {
              bs.push(wmid);
            } else {
  #pragma unroll
              for (int f = 0; f < F; ++f) wnext[f] = wmid[f];
            }
          }
        }
  #pragma unroll
        for (int f = 0; f < F; ++f) {
          dw[f] = wnext[f] - wlo[f];
          shock[f] = dw[f] * (Real)it.inv_sqrt_dt;  // qmc.rs:169, note.rs:101-105
          wlo[f] = wnext[f];
        }
        }

        // ---- local-vol rows for this step (uniform across the block)
        if (local_vol) {
          bool reload = false;
  #pragma unroll
          for (int f = 0; f < F; ++f) reload |= lvb[step * F + f].i0 != loaded_i0[f];
          if (reload) {
            __syncthreads();
            for (int u = 0; u < it.n_surf; ++u) {
              const LvSurface su = surf[u];
              const LvBracket br = lvb[step * F + su.factor];
              const double* xs = su.dev_x ? su.dev_x : at<double>(arena, su.off_x);
              const double* vv = su.dev_v ? su.dev_v : at<double>(arena, su.off_v);
              const double* v0 = vv + (size_t)br.i0 * su.n_x;
              const double* v1 = vv + (size_t)br.i1 * su.n_x;
              Real* dst = lds_w + (size_t)u * kSurfWords * nxs;
              if (lv_uniform) {
                float4* d4 = reinterpret_cast<float4*>(dst);
                for (int i = tid; i < su.n_x; i += nthreads) {
                  const int k = i + 1 < su.n_x ? i + 1 : i;
                  d4[i] = make_float4((float)v0[i], (float)(v0[k] - v0[i]), (float)v1[i],
                                      (float)(v1[k] - v1[i]));
                }
              } else {
                for (int i = tid; i < su.n_x; i += nthreads) {
                  dst[i] = (Real)xs[i];
                  dst[nxs + i] = (Real)v0[i];
                  dst[2 * nxs + i] = (Real)v1[i];
                }
              }
            }
  #pragma unroll
            for (int f = 0; f < F; ++f) loaded_i0[f] = lvb[step * F + f].i0;
            __syncthreads();
          }
        }

        const bool call_active = (si.flags & 2) != 0;
        const bool period_end = (si.flags & 1) != 0;
        const bool new_period = (si.flags & 4) != 0;
        const bool fixed_now = si.period == 0 && (it.flags & kFixedFirst) != 0;
        if constexpr (kLogState) {
          // ---- FP32 tier: X = ln(S/S0) per (scenario, factor), constants from
          // SfF32, the local-vol knot found by one FMA on evenly spaced knots.
          // The Phoenix scan below is note.rs:143-218 with the KO'd-path skip
          // turned into masks (values of a KO'd scenario are never read again).
          // staged in LDS and read per use; the opaque zero keeps the compiler from
          // hoisting these reads out of the step loop (which would pin S*F records
          // in registers for the whole path)
          int opaque = 0;
          asm volatile("" : "+s"(opaque));
          const SfF32* sf = sf_lds + opaque;
          float wt[F];
          if (local_vol) {
  #pragma unroll
            for (int f = 0; f < F; ++f) wt[f] = at<float>(arena, it.off_wt32)[step * F + f];
          }
          const float inc = fixed_now ? 0.0f : (float)si.accrual_inc;
          // Phase 1: the Euler step of every (scenario, factor).  The chains are
          // independent and the local-vol mode is chosen once per step, so the
          // factors of a scenario form one straight-line block the scheduler can
          // overlap (the LDS reads of one chain behind the arithmetic of another).
          if (lv_uniform) {
            // the Euler step of scenario slot s; slots past S (when S < SMAX) reread
            // scenario S-1's constants (in bounds) and their state is never read
  #define RIGA_LV_EULER(s)                                                                   \
    {                                                                                        \
      const int ss_ = (s) < S ? (s) : S - 1;                                                 \
      _Pragma("unroll") for (int f = 0; f < F; ++f) {                                        \
        const SfF32 c = sf[ss_ * F + f];                                                     \
        /* knot position clamped to [0, tmax] (one med3), its index and weight */            \
        const float t = __builtin_amdgcn_fmed3f(fmaf(spot[s][f], c.a, c.b), 0.0f, c.tmax);   \
        const int j = (int)t;                                                                \
        const float wx = __builtin_amdgcn_fractf(t); /* = t - j exactly, t >= 0 */           \
        const F32x4 e = *(const LdsF32x4*)(uintptr_t)(uint32_t)(c.lds_off + j * 16);     \
        const float v0 = fmaf(wx, e.y, e.x);                                                 \
        const float v1 = fmaf(wx, e.w, e.z);                                                 \
        const float vol = fmaf(wt[f], v1 - v0, v0);                                          \
        /* X += (r - q) dt + vol dW - vol^2 dt / 2 */                                        \
        spot[s][f] += fmaf(vol, fmaf(vol, nhdt, dw[f]), c.c);                                \
      }                                                                                      \
    }
            if (F * SMAX <= 30 && S == SMAX) {
              // a full bucket: one straight-line block over every scenario, so the LDS
              // reads of one (scenario, factor) chain overlap the arithmetic of others
              // (up to 30 chains; more hold too many registers)
  #pragma unroll
              for (int s = 0; s < SMAX; ++s) RIGA_LV_EULER(s)
            } else {
              // a partly filled bucket computes only its scenarios
  #pragma unroll
              for (int s = 0; s < SMAX; ++s) {
                if (s >= S) continue;
                RIGA_LV_EULER(s)
              }
            }
  #undef RIGA_LV_EULER
          } else if (local_vol) {  // knots not evenly spaced: bracketing search
  #pragma unroll
            for (int s = 0; s < SMAX; ++s) {
              if (s >= S) continue;
              const int ss = s;
  #pragma unroll
              for (int f = 0; f < F; ++f) {
                const SfF32 c = sf[ss * F + f];
                const int u = scen[ss].lv[f];
                const LvSurface& su = surf[u];
                const float vol = lv_sigma_f32_general(lds + (size_t)u * kSurfWords * nxs, su.n_x, (float)su.inv_dx,
                                                       wt[f], spot[s][f] + (float)it.ln_s0[f]);
                spot[s][f] += fmaf(vol, dw[f], fmaf(vol * vol, nhdt, c.c));
              }
            }
          } else {
  #pragma unroll
            for (int s = 0; s < SMAX; ++s) {
              if (s >= S) continue;
              const int ss = s;
  #pragma unroll
              for (int f = 0; f < F; ++f) {
                const SfF32 c = sf[ss * F + f];
                spot[s][f] += fmaf(c.a, dw[f], c.c);
              }
            }
          }
          // Phase 2: worst-of, accrual and the daily KO with its payment as
          // selects (note.rs:143-195).  The payment runs only in waves where some
          // lane of the scenario knocks out at this step (a few percent of wave
          // steps; a wave-uniform branch): for every other lane it changes nothing.
          const float cr = cr32;
          const float df = (float)si.df;
          float w_first = 0.0f;  // scenario 0's worst-of, for replay series
  #pragma unroll
          for (int s = 0; s < SMAX; ++s) {
            float w = spot[s][0];
  #pragma unroll
            for (int f = 1; f < F; ++f) w = fminf(w, spot[s][f]);
            if (s == 0) w_first = w;
            const bool alive = (pk[s] & 0xFFFFFu) == 0u;
            accrual[s] = (new_period ? 0.0f : accrual[s]) + (w >= ln_cb32 ? inc : 0.0f);
            const bool ko_now = alive & call_active & (w >= ln_ko32);  // no short-circuit branch
            if (__ballot(ko_now && s < S) != 0ull) {
              const float due = fixed_now ? cr : accrual[s];
              const float carry = due > 0.0f ? (float)((pk[s] >> 20) & 0x7FFu) * cr : 0.0f;
              const bool pay = ko_now && due + carry > 0.0f;
              cpv[s] += pay ? due * df : 0.0f;
              mpv[s] += pay ? carry * df : 0.0f;
              const uint32_t recv = pay ? 0x80000000u : (pk[s] & 0x80000000u);
              pk[s] = ko_now ? ((pk[s] & 0x7FF00000u) | recv | (uint32_t)(step + 1)) : pk[s];
            }
          }
          if (it.series != nullptr && valid)  // replay: the worst-of for the coupon strip
            it.series[(size_t)step * it.paths + path_local] = (double)exp_f32(w_first);
          // Phase 3: period end (note.rs:197-218), a step-uniform branch.
          if (period_end) {
            const bool memory = (it.flags & kMemoryCoupon) != 0;
  #pragma unroll
            for (int s = 0; s < SMAX; ++s) {
              // lanes knocked out at this very step took their coupon above
              const bool open = (pk[s] & 0xFFFFFu) == 0u;
              const float due = fixed_now ? cr : accrual[s];
              const bool pays = open && due > 0.0f;
              cpv[s] += pays ? due * df : 0.0f;
              mpv[s] += pays ? (float)((pk[s] >> 20) & 0x7FFu) * cr * df : 0.0f;
              uint32_t q = pk[s];
              q = pays ? ((q & 0x000FFFFFu) | 0x80000000u) : q;          // paid: unpaid count back to 0
              q = (open && !pays && memory) ? q + (1u << 20) : q;       // missed: one more unpaid coupon
              pk[s] = q;
            }
          }
          continue;
        }
  #pragma unroll
        for (int s = 0; s < SMAX; ++s) {
          if (s >= S) continue;
          const ScenarioDesc& sc = scen[s];
          // ---- Euler step (note.rs:134-142).  FP64: S *= exp(incr), as the
          // reference rounds it.  FP32: the state is X = ln S and X += incr, the
          // same step without an exp, a log and two divides per factor.
  #pragma unroll
          for (int f = 0; f < F; ++f) {
            Real incr;
            if (local_vol) {
              const int u = sc.lv[f];
              const LvSurface& su = surf[u];
              Real vol;
              if constexpr (kLogState) {
                if (lv_uniform)
                  vol = lv_sigma_uniform(reinterpret_cast<const float4*>(lds + (size_t)u * kSurfWords * nxs),
                                         su.n_x, (float)su.inv_dx, (float)su.rel_off,
                                         (float)lvb[step * F + f].wt, spot[s][f]);
                else
                  vol = lv_sigma<Real, true>(lds + (size_t)u * kSurfWords * nxs, su.n_x, (Real)su.inv_dx,
                                             (Real)lvb[step * F + f].wt, spot[s][f] + (Real)it.ln_s0[f]);
              } else {
                vol = lv_sigma<Real, false>(lds + (size_t)u * kSurfWords * nxs, su.n_x, (Real)su.inv_dx,
                                            (Real)lvb[step * F + f].wt, spot[s][f]);
              }
              const Real drift = ((Real)sc.rq[f] - Real(0.5) * vol * vol) * (Real)dt64;
              incr = drift + vol * (Real)sqdt64 * shock[f];
            } else {
              incr = (Real)sc.drift[f] + (Real)sc.vsd[f] * shock[f];
            }
            if constexpr (kLogState)
              spot[s][f] += incr;
            else
              spot[s][f] *= tier_exp(incr);
          }
          // ---- worst-of (note.rs:143-145); the FP32 tier keeps ln(worst)
          Real worst;
          if constexpr (kLogState) {
            worst = spot[s][0];
  #pragma unroll
            for (int f = 1; f < F; ++f) worst = fminf(worst, spot[s][f]);
          } else {
            worst = spot[s][0] / (Real)s0_64[0];
  #pragma unroll
            for (int f = 1; f < F; ++f) worst = fmin(worst, spot[s][f] / (Real)s0_64[f]);
          }
          worst_last[s] = worst;
          if (ko_step[s] >= 0) continue;  // note.rs:146-148: cash flows stop after KO
          // ---- coupon accrual (note.rs:151-163)
          if (new_period) accrual[s] = Real(0);
          const bool above_coupon = kLogState ? worst >= (Real)it.ln_cb : (double)worst >= cb64;
          if (!fixed_now && above_coupon) accrual[s] += (Real)si.accrual_inc;
          // ---- daily global KO (note.rs:166-195): every name at or above the
          // barrier, i.e. the worst-of at or above it
          bool all_above = call_active;
          if constexpr (kLogState) {
            all_above = all_above && worst >= (Real)it.ln_ko;
          } else {
  #pragma unroll
            for (int f = 0; f < F; ++f)
              all_above = all_above && ((double)(spot[s][f] / (Real)s0_64[f]) >= ko64);
          }
          if (all_above) {
            ko_step[s] = step;
            const Real due = fixed_now ? (Real)cr64 : accrual[s];
            const Real carry = due > Real(0) ? unpaid[s] : Real(0);
            if (due + carry > Real(0)) {
              cpv[s] += due * (Real)si.df;
              mpv[s] += carry * (Real)si.df;
              received[s] = true;
            }
            continue;
          }
          // ---- period end (note.rs:197-218)
          if (period_end) {
            const Real due = fixed_now ? (Real)cr64 : accrual[s];
            if (due > Real(0)) {
              cpv[s] += due * (Real)si.df;
              mpv[s] += unpaid[s] * (Real)si.df;
              received[s] = true;
              unpaid[s] = Real(0);
            } else if (it.flags & kMemoryCoupon) {
              unpaid[s] += (Real)cr64;
            }
          }
        }
        if (it.series != nullptr && valid)  // replay: the worst-of for the coupon strip
          it.series[(size_t)step * it.paths + path_local] = (double)worst_last[0];
      }

      // ---- path end (note.rs:221-258) and reductions
      if constexpr (kLogState) {
        // the FP32 tier's final worst-of, from the log spots the last step left
        // (not carried through the step loop; every path takes at least one step)
  #pragma unroll
        for (int s = 0; s < SMAX; ++s) {
          float w = spot[s][0];
  #pragma unroll
          for (int f = 1; f < F; ++f) w = fminf(w, spot[s][f]);
          worst_last[s] = w;
        }
      }
  #pragma unroll
      for (int s = 0; s < SMAX; ++s) {
        if (s >= S) continue;
        if constexpr (kLogState) {  // unpack the FP32 tier's state word
          ko_step[s] = (int)(pk[s] & 0xFFFFFu) - 1;
          received[s] = (pk[s] >> 31) != 0u;
        }
        const bool ko = ko_step[s] >= 0;
        const double funding = ko ? stepinfo[ko_step[s]].df : it.put_df;
        const double fw = kLogState ? (double)exp_f32(worst_last[s]) : (double)worst_last[s];
        const bool hit_ki = fw <= it.ki;
        const bool eligible = hit_ki && !ko;
        const double put = eligible ? fmax(it.strike - fw, 0.0) * it.put_df : 0.0;
        const double nominal = (double)cpv[s] + (double)mpv[s];
        const double coupon_quote = nominal * 10.0;
        const double memory_quote = (double)mpv[s] * 10.0;
        const double pv = funding + nominal - put;
        if (want_records && valid) {
          PathRecordDev r;
          r.ko_step = ko_step[s];
          r.hit_ki = hit_ki;
          r.final_worst = fw;
          r.funding = funding;
          r.put = put;
          r.coupon_pv = (double)cpv[s];
          r.memory_pv = (double)mpv[s];
          r.pv = pv;
          records[it.record_base + (int64_t)s * (int64_t)it.paths + (int64_t)path_local] = r;
        }
        // plain sums (funding, put, coupon and memory quotes) and (x, x^2)
        // pairs for the four moment sets, each summed over the wave in turn (the
        // scheduling barrier keeps one quantity's shuffles live at a time);
        // invalid lanes contribute zero.
        const double mx[8] = {funding, put, coupon_quote - memory_quote, memory_quote, pv,
                              ko ? funding : 0.0, hit_ki ? put : 0.0, received[s] ? nominal : 0.0};
  #pragma unroll
        for (int k = 0; k < kNumQ; ++k) {
          const double x = valid ? mx[k < kQPv1 ? k : kQPv1 + (k - kQPv1) / 2] : 0.0;
          DD v = (k >= kQPv1 && (k - kQPv1) % 2 == 1) ? dd_sq(x) : dd_from(x);
          v = wave_sum(v);
          if (lane == 0) sh.acc[wave][k] = v;
          __builtin_amdgcn_sched_barrier(0);
        }
        const uint64_t b_valid = __ballot(valid);
        const uint64_t b_ko = __ballot(valid && ko);
        const uint64_t b_ki = __ballot(valid && hit_ki);
        const uint64_t b_el = __ballot(valid && eligible);
        const uint64_t b_cp = __ballot(valid && received[s]);
        if (lane == 0) {
          sh.cnt[wave][kCN] = __popcll(b_valid);
          sh.cnt[wave][kCKo] = __popcll(b_ko);
          sh.cnt[wave][kCKi] = __popcll(b_ki);
          sh.cnt[wave][kCPutEligible] = __popcll(b_el);
          sh.cnt[wave][kCCouponPaths] = __popcll(b_cp);
          sh.wave_ko[wave] = __popcll(b_ko);
        }
        __syncthreads();
        // the block's Partial for this scenario: waves merged in order, then added
        // to the earlier batches' sums; KO events (trace) go straight to their
        // slots in path order (note.rs:174-180: the first 64)
        Partial* out = partials + it.partial_base + (int64_t)s * it.n_blocks + chunk;
        int ev_total = 0;
        if (trace) {
          int base = ev_n[s];
          for (int w = 0; w < wave; ++w) base += sh.wave_ko[w];
          const int rank = base + (int)__popcll(b_ko & ((1ull << lane) - 1ull));
          if (valid && ko && rank < kTraceCap) {  // paths_per_block <= 2048, steps < 2^20
            out->ev_path_lo[rank] = (int32_t)((int64_t)chunk * it.paths_per_block + (int64_t)(b * nthreads + tid));
            out->ev_step[rank] = (int32_t)ko_step[s];
          }
          ev_total = ev_n[s];
          for (int w = 0; w < nwaves; ++w) ev_total += sh.wave_ko[w];
          ev_total = ev_total < kTraceCap ? ev_total : kTraceCap;
        }
        if (tid < kNumQ) {
          DD a = sh.acc[0][tid];
          for (int w = 1; w < nwaves; ++w) a = dd_add(a, sh.acc[w][tid]);
          out->q[tid] = b == 0 ? a : dd_add(out->q[tid], a);
        } else if (tid >= 32 && tid < 32 + kNumC) {
          const int c = tid - 32;
          uint64_t a = 0;
          for (int w = 0; w < nwaves; ++w) a += sh.cnt[w][c];
          out->c[c] = b == 0 ? a : out->c[c] + a;
        }
        __syncthreads();  // every lane has read sh and ev_n[s]: the next scenario may reuse them
        if (tid == 48) {  // within the first wave: every launch shape has it
          ev_n[s] = ev_total;
          out->n_events = ev_total;
        }
      }
    }
    if (it.stack_pooled > 0) {  // every lane is done with the slot: hand it back
      __syncthreads();
      if (tid == 0) {
        __threadfence();
        // a returning exchange: the workgroup waits for it, so the slot is free
        // before the workgroup ends and the dispatcher starts another in its place
        // (a slot found free here was never held: counted as a failure)
        if (atomicExch(it.pool_flags + slot_sh, 0u) != 1u) atomicAdd(it.pool_flags + it.pool_n, 1u);
      }
    }
    // one workgroup per unit: this one is done (the dispatcher starts the next);
    // a persistent grid (fewer workgroups than units) claims again
    if (gridDim.x >= (uint32_t)n_units) return;
  }
}

}  // namespace riga
