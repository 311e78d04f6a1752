// mc_rng.hpp -- the Monte Carlo lane's random stream on the device:
// ChaCha8Rng::seed_from_u64(seed) (rand_chacha 0.9.0) read as u64 words, and
// rand_distr 0.5.1 StandardNormal (the 256-layer ziggurat) on top of it, as
// the reference draws them one after another (note.rs:58,124-126).
//
// The stream is one sequence per request, and a normal draw consumes a
// variable number of words (one in ~98.5% of draws).  Paths therefore start at
// data-dependent word offsets; the parse kernels (kernels_mc.hip) find them and
// compute every multi-word ("slow") draw once, and each path lane then
// regenerates its own words from its offset, taking slow draws from the parse.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exp_log.h"
#include "phoenix_layout.h"

namespace riga {

constexpr double kZigR = 3.6541528853610088;  // start of the tail (rand's ZIG_NORM_R)

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define RIGA_CHACHA_QR(a, b, c, d) \
  a += b;                          \
  d = rotl32(d ^ a, 16);           \
  c += d;                          \
  b = rotl32(b ^ c, 12);           \
  a += b;                          \
  d = rotl32(d ^ a, 8);            \
  c += d;                          \
  b = rotl32(b ^ c, 7);

// Block `block` of the ChaCha8 keystream (64-bit block counter in state words
// 12-13, stream 0 in words 14-15) as its eight u64 words, low word first.
__device__ __forceinline__ void chacha8_block(const uint32_t* __restrict__ key, uint64_t block,
                                              uint64_t (&w)[8]) {
  constexpr uint32_t c0 = 0x61707865u, c1 = 0x3320646eu, c2 = 0x79622d32u, c3 = 0x6b206574u;
  uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
  uint32_t x4 = key[0], x5 = key[1], x6 = key[2], x7 = key[3];
  uint32_t x8 = key[4], x9 = key[5], x10 = key[6], x11 = key[7];
  const uint32_t lo = (uint32_t)block, hi = (uint32_t)(block >> 32);
  uint32_t x12 = lo, x13 = hi, x14 = 0u, x15 = 0u;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    RIGA_CHACHA_QR(x0, x4, x8, x12)
    RIGA_CHACHA_QR(x1, x5, x9, x13)
    RIGA_CHACHA_QR(x2, x6, x10, x14)
    RIGA_CHACHA_QR(x3, x7, x11, x15)
    RIGA_CHACHA_QR(x0, x5, x10, x15)
    RIGA_CHACHA_QR(x1, x6, x11, x12)
    RIGA_CHACHA_QR(x2, x7, x8, x13)
    RIGA_CHACHA_QR(x3, x4, x9, x14)
  }
  auto pack = [](uint32_t a, uint32_t b) { return (uint64_t)a | ((uint64_t)b << 32); };
  w[0] = pack(x0 + c0, x1 + c1);
  w[1] = pack(x2 + c2, x3 + c3);
  w[2] = pack(x4 + key[0], x5 + key[1]);
  w[3] = pack(x6 + key[2], x7 + key[3]);
  w[4] = pack(x8 + key[4], x9 + key[5]);
  w[5] = pack(x10 + key[6], x11 + key[7]);
  w[6] = pack(x12 + lo, x13 + hi);
  w[7] = pack(x14, x15);
}
#undef RIGA_CHACHA_QR

// Word `idx` of the stream (one block computed; used on the parse's rare slow draws).
__device__ __forceinline__ uint64_t chacha8_word(const uint32_t* __restrict__ key, uint64_t idx) {
  uint64_t w[8];
  chacha8_block(key, idx >> 3, w);
  const int j = (int)(idx & 7);
  uint64_t r = w[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) r = j == k ? w[k] : r;
  return r;
}

__device__ __forceinline__ double bits_f64(uint64_t b) { return __longlong_as_double((long long)b); }

// The ziggurat's first test on a word: layer i = low 8 bits, u in [-1, 1) from
// the top 52 bits (a double in [2, 4) minus 3), x = u * x_i; the draw is x when
// |x| < x_{i+1}.
__device__ __forceinline__ bool zig_rect(uint64_t bits, const double* __restrict__ zx, double* x_out) {
  const int i = (int)(bits & 0xffu);
  const double u = bits_f64((bits >> 12) | 0x4000000000000000ull) - 3.0;
  const double x = u * zx[i];
  *x_out = x;
  return fabs(x) < zx[i + 1];
}

// The rest of a draw whose first word failed the rectangle test, fed one word
// at a time: the wedge test (one uniform), a restart, or the tail sampler for
// layer 0 (pairs of Open01 uniforms).  No multiply-add is fused here in either
// tier, so the accept/reject decisions round as the reference's do.
struct ZigSlow {
  int state;  // 1 = wedge uniform next, 0 = first word of a new attempt, 2/3 = tail x_/y_
  int i;
  double u, x, tx;

  __device__ __forceinline__ void start(uint64_t bits, const double* __restrict__ zx) {
    i = (int)(bits & 0xffu);
    u = bits_f64((bits >> 12) | 0x4000000000000000ull) - 3.0;
    x = u * zx[i];
    state = i == 0 ? 2 : 1;
  }

  // Returns true when the draw is complete (its value in *out).
  __device__ __forceinline__ bool feed(uint64_t w, const double* __restrict__ zx, const double* __restrict__ zf,
                                       double* out) {
#pragma clang fp contract(off)
    if (state == 0) {
      double xr;
      if (zig_rect(w, zx, &xr)) {
        *out = xr;
        return true;
      }
      start(w, zx);
      return false;
    }
    if (state == 1) {
      const double v = (double)(w >> 11) * (1.0 / 9007199254740992.0);  // StandardUniform f64
      if (zf[i + 1] + (zf[i] - zf[i + 1]) * v < riga_exp(-x * x / 2.0)) {
        *out = x;
        return true;
      }
      state = 0;
      return false;
    }
    // Open01: a double in [1, 2) minus (1 - eps/2)
    const double o = bits_f64((w >> 12) | 0x3FF0000000000000ull) - (1.0 - 2.220446049250313e-16 / 2.0);
    if (state == 2) {
      tx = riga_log(o) / kZigR;
      state = 3;
      return false;
    }
    const double ty = riga_log(o);
    if (-2.0 * ty < tx * tx) {
      state = 2;
      return false;
    }
    *out = u < 0.0 ? tx - kZigR : kZigR - tx;
    return true;
  }
};

// A lane's words in LDS: ring[slot * stride] for slot in [0, kSlots), `fill`
// words buffered from `head`; ChaCha blocks are appended whole.
template <int kSlots>
struct McRing {
  uint64_t* base;  // this lane's slot 0
  int stride;      // lanes per workgroup
  int head, fill;
  uint64_t next_block;

  __device__ __forceinline__ void append(const uint32_t* __restrict__ key) {
    uint64_t w[8];
    chacha8_block(key, next_block, w);
    ++next_block;
    int s = head + fill;
    s = s >= kSlots ? s - kSlots : s;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      base[s * stride] = w[j];
      s = s + 1 == kSlots ? 0 : s + 1;
    }
    fill += 8;
  }

  // Position the ring at stream word `pos`.
  __device__ __forceinline__ void seek(const uint32_t* __restrict__ key, uint64_t pos) {
    next_block = pos >> 3;
    head = 0;
    fill = 0;
    append(key);
    head = (int)(pos & 7);
    fill = 8 - head;
  }

  __device__ __forceinline__ uint64_t take(const uint32_t* __restrict__ key) {
    if (fill == 0) append(key);  // only after slow draws outran the wave's refills
    const uint64_t w = base[head * stride];
    head = head + 1 == kSlots ? 0 : head + 1;
    --fill;
    return w;
  }
};

// A lane's place in the parse's slow-word lists: the stream index of the next
// listed slow word at or after the lane's position.  A slow word that starts a
// draw carries the draw's value and length, computed once by the parse, so
// the pricing loop never runs the ziggurat's wedge or tail tests.
struct McCursor {
  const uint32_t* slow;    // [chunk][cap]: (offset << 8) | words
  const double* val;       // [chunk][cap]: the draw starting at that word
  const McChunk* chunks;
  int64_t n_chunks;
  int C, cap;
  int64_t chunk;
  int q, n;
  uint64_t next;           // ~0: past the parsed stream (no path reads there)

  __device__ __forceinline__ void advance(uint64_t pos) {
    for (;;) {
      if (q < n) {
        const uint64_t a = (uint64_t)chunk * (uint64_t)C + (slow[chunk * cap + q] >> 8);
        if (a >= pos) {
          next = a;
          return;
        }
        ++q;
        continue;
      }
      ++chunk;
      if (chunk >= n_chunks) {
        next = ~0ull;
        return;
      }
      q = 0;
      n = chunks[chunk].nslow;
    }
  }

  __device__ __forceinline__ void seek(uint64_t pos) {
    chunk = (int64_t)(pos / (uint64_t)C);
    if (chunk >= n_chunks) {
      next = ~0ull;
      n = q = 0;
      return;
    }
    n = chunks[chunk].nslow;
    const uint32_t off = (uint32_t)(pos - (uint64_t)chunk * (uint64_t)C);
    int lo = 0, hi = n;  // first entry at or after `off`
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((slow[chunk * cap + mid] >> 8) < off)
        lo = mid + 1;
      else
        hi = mid;
    }
    q = lo;
    advance(pos);
  }
};

// One StandardNormal draw: the word at stream index `wpos` either passes the
// rectangle test (x = u * x_i) or is a listed slow word whose draw the parse
// has already computed (its extra words are skipped).
template <int kSlots>
__device__ __forceinline__ double mc_draw(McRing<kSlots>& ring, McCursor& cur, uint64_t& wpos,
                                          const uint32_t* __restrict__ key, const double* __restrict__ zx) {
  const uint64_t w = ring.take(key);
  const uint64_t at = wpos++;
  if (at == cur.next) {
    const int64_t e = cur.chunk * cur.cap + cur.q;
    const double x = cur.val[e];
    const int extra = (int)(cur.slow[e] & 0xffu) - 1;
    for (int k = 0; k < extra; ++k) (void)ring.take(key);
    wpos += (uint64_t)extra;
    ++cur.q;
    cur.advance(wpos);
    return x;
  }
  const int i = (int)(w & 0xffu);
  const double u = bits_f64((w >> 12) | 0x4000000000000000ull) - 3.0;
  return u * zx[i];
}

// The draw whose first word `w0` sits at stream index `idx` and fails the
// rectangle test: its value, and the words it takes (the following words
// regenerated by index); kZigMaxLen + 1 words means too long (refused).
__device__ __forceinline__ int zig_slow_draw(const uint32_t* __restrict__ key, uint64_t idx, uint64_t w0,
                                             const double* __restrict__ zx, const double* __restrict__ zf,
                                             double* value) {
  ZigSlow s;
  s.start(w0, zx);
  for (int len = 2; len <= kZigMaxLen; ++len)
    if (s.feed(chacha8_word(key, idx + (uint64_t)(len - 1)), zx, zf, value)) return len;
  *value = 0.0;
  return kZigMaxLen + 1;
}

}  // namespace riga
