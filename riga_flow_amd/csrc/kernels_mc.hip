// kernels_mc.hip -- the Monte Carlo lane's stream parse.  Built with
// -ffp-contract=off, like the FP64 tier.
//
// The reference prices path after path from one ChaCha8Rng, drawing F
// StandardNormals per step (note.rs:58,124-126).  A draw takes one u64 word
// when the ziggurat's rectangle test accepts it and more otherwise, so the
// first word of path p (draw p * steps * F) is known only after every earlier
// draw.  The parse finds it in three passes over the stream, cut into chunks of
// `chunk_words` words:
//   1. mc_parse_kernel, one lane per chunk: the words that fail the rectangle
//      test ("slow" words, ~1.5%), and for each the draw that would start there
//      (its value and how many words it takes); from that list, for every
//      entry offset e < kMcEntries, the draws that start in the chunk and the
//      next chunk's entry offset.
//   2. mc_resolve_kernel: each chunk's true entry (from the nearest earlier
//      chunk whose exit does not depend on its entry) and its draw count;
//      mc_scan_kernel: prefix sums of the counts.
//   3. mc_starts_kernel: the stream word of every path's first draw.
// Between slow words every word is one draw, so walking a chunk's draws only
// visits its slow-word list.
#include "kernels.h"
#include "mc_rng.hpp"

namespace riga {

__device__ __forceinline__ void load_zig(const double* __restrict__ zig, double* zx, double* zf) {
  for (int i = threadIdx.x; i < kZigLen; i += blockDim.x) {
    zx[i] = zig[i];
    zf[i] = zig[kZigLen + i];
  }
}

__global__ __launch_bounds__(256) void mc_parse_kernel(McParseArgs a) {
  __shared__ double zx[kZigLen], zf[kZigLen];
  load_zig(a.zig, zx, zf);
  __syncthreads();
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= a.n_chunks) return;
  const uint32_t* key = a.key;
  const int C = a.chunk_words;
  const uint64_t base = (uint64_t)j * (uint64_t)C;
  uint32_t* slow = a.slow + (size_t)j * (size_t)a.slow_cap;
  double* val = a.slow_val + (size_t)j * (size_t)a.slow_cap;
  int ns = 0;
  uint32_t flags = 0;
  // Pass 1: the words that fail the rectangle test, with their bits (kept in
  // the value slot).  Their draws are finished in pass 2, one entry per
  // iteration, so a wave does not run the wedge and tail tests word by word.
  for (int b = 0; b < C / 8; ++b) {
    uint64_t w[8];
    chacha8_block(key, (base >> 3) + (uint64_t)b, w);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double x;
      if (!zig_rect(w[q], zx, &x)) {
        if (ns < a.slow_cap) {
          slow[ns] = (uint32_t)(b * 8 + q) << 8;
          val[ns] = bits_f64(w[q]);
        } else {
          flags |= kMcErrSlow;
        }
        ++ns;
      }
    }
  }
  const int n = ns < a.slow_cap ? ns : a.slow_cap;
  // Pass 2: each listed draw's length and value.
  for (int m = 0; m < n; ++m) {
    const uint32_t off = slow[m] >> 8;
    const uint64_t w0 = (uint64_t)__double_as_longlong(val[m]);
    double x;
    const int len = zig_slow_draw(key, base + off, w0, zx, zf, &x);
    if (len > kZigMaxLen) flags |= kMcErrLong;
    slow[m] = (off << 8) | (uint32_t)(len > 255 ? 255 : len);
    val[m] = x;
  }
  McChunk out;
  for (int e = 0; e < kMcEntries; ++e) {
    int pos = e, draws = 0;
    for (int m = 0; m < n; ++m) {
      const uint32_t s = slow[m];
      const int k = (int)(s >> 8);
      if (k < pos) continue;  // inside a draw that started earlier
      draws += k - pos + 1;   // the one-word draws before k, then the draw at k
      pos = k + (int)(s & 0xffu);
    }
    if (pos < C) {
      draws += C - pos;
      pos = C;
    }
    const int ex = pos - C;
    if (ex >= kMcEntries) flags |= kMcErrExit;
    out.count[e] = draws;
    out.exit[e] = (uint8_t)(ex < kMcEntries ? ex : kMcEntries - 1);
  }
  bool same = true;
  for (int e = 1; e < kMcEntries; ++e) same = same && out.exit[e] == out.exit[0];
  out.nslow = n;
  out.flags = flags | (same ? kMcConstExit : 0u);
  a.chunks[j] = out;
}

__global__ __launch_bounds__(256) void mc_resolve_kernel(const McChunk* __restrict__ chunks, int64_t n,
                                                         uint8_t* __restrict__ entry, uint32_t* __restrict__ cnt,
                                                         uint64_t* __restrict__ block_sum) {
  __shared__ uint64_t red[256];
  const int t = threadIdx.x;
  const int64_t j = (int64_t)blockIdx.x * 256 + t;
  uint64_t c = 0;
  if (j < n) {
    int e = 0;
    if (j > 0) {
      int64_t i = j - 1;
      while (i >= 0 && !(chunks[i].flags & kMcConstExit)) --i;
      e = i < 0 ? 0 : chunks[i].exit[0];
      for (int64_t q = i + 1; q < j; ++q) e = chunks[q].exit[e];
    }
    entry[j] = (uint8_t)e;
    c = (uint64_t)chunks[j].count[e];
    cnt[j] = (uint32_t)c;
  }
  red[t] = c;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) block_sum[blockIdx.x] = red[0];
}

// Exclusive prefix sums of the per-block draw counts, in place; also clears the
// status words the starts kernel counts into.
__global__ __launch_bounds__(1024) void mc_scan_kernel(uint64_t* __restrict__ bs, int64_t nb,
                                                       uint32_t* __restrict__ status) {
  __shared__ uint64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nb + 1023) / 1024;
  const int64_t lo = (int64_t)t * per;
  const int64_t hi = lo + per < nb ? lo + per : nb;
  uint64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += bs[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint64_t run = part[t] - s;
  for (int64_t i = lo; i < hi; ++i) {
    const uint64_t v = bs[i];
    bs[i] = run;
    run += v;
  }
  if (t < 2) status[t] = 0u;
}

__global__ __launch_bounds__(256) void mc_starts_kernel(McStartsArgs a) {
  __shared__ uint64_t sc[256];
  const int t = threadIdx.x;
  const int64_t j = (int64_t)blockIdx.x * 256 + t;
  const uint64_t c = j < a.n_chunks ? (uint64_t)a.cnt[j] : 0;
  sc[t] = c;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint64_t v = t >= off ? sc[t - off] : 0;
    __syncthreads();
    sc[t] += v;
    __syncthreads();
  }
  if (j >= a.n_chunks) return;
  const uint64_t start = a.block_off[blockIdx.x] + sc[t] - c;  // first draw of the chunk
  const uint64_t end = start + c;
  const McChunk& ch = a.chunks[j];
  if (ch.flags & (kMcErrLong | kMcErrSlow | kMcErrExit)) atomicOr(&a.status[1], ch.flags);
  const uint32_t* sl = a.slow + (size_t)j * (size_t)a.slow_cap;
  const int e = a.entry[j];
  uint32_t found = 0;
  for (uint64_t p = (start + a.dpp - 1) / a.dpp; p < a.paths && p * a.dpp < end; ++p) {
    const uint64_t m = p * a.dpp - start;  // the path's first draw, counted within the chunk
    int pos = e, at = -1;
    uint64_t idx = 0;  // draws before `pos`
    for (int q = 0; q < ch.nslow; ++q) {
      const int k = (int)(sl[q] >> 8);
      if (k < pos) continue;
      if (m < idx + (uint64_t)(k - pos)) {
        at = pos + (int)(m - idx);
        break;
      }
      idx += (uint64_t)(k - pos);
      if (m == idx) {
        at = k;
        break;
      }
      idx += 1;
      pos = k + (int)(sl[q] & 0xffu);
    }
    if (at < 0) at = pos + (int)(m - idx);
    a.starts[p] = (uint64_t)j * (uint64_t)a.chunk_words + (uint64_t)at;
    ++found;
  }
  if (found) atomicAdd(&a.status[0], found);
  // the parsed stream must hold every draw of every path (the pricing lanes
  // take slow draws from it)
  if (j == a.n_chunks - 1 && end < a.paths * a.dpp) atomicOr(&a.status[1], (uint32_t)kMcErrShort);
}

// PathCube (pathcube.rs:47-76 generate_correlated_block, :146-187 gbm_evolve):
// one lane per path regenerates its words like the pricing lanes, forms L z per
// (path, time) cell in FP64, stores it as float, and evolves
// log S += drift + diffusion * (double)(float)y, storing exp(log S) as float.
template <int F>
__global__ __launch_bounds__(256) void cube_kernel(CubeArgs a) {
  __shared__ uint64_t ring_lds[24 * 256];
  __shared__ double zx[kZigLen];
  for (int i = threadIdx.x; i < kZigLen; i += blockDim.x) zx[i] = a.zig[i];
  __syncthreads();
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = p < a.paths;
  McRing<24> ring;
  ring.base = ring_lds + threadIdx.x;
  ring.stride = blockDim.x;
  uint64_t wpos = a.starts[valid ? p : 0];
  ring.seek(a.key, wpos);
  McCursor cur;
  cur.slow = a.slow;
  cur.val = a.slow_val;
  cur.chunks = a.chunks;
  cur.n_chunks = a.n_chunks;
  cur.C = a.chunk_words;
  cur.cap = a.slow_cap;
  cur.seek(wpos);
  double log_s[F];
#pragma unroll
  for (int f = 0; f < F; ++f) log_s[f] = a.ln_s0[f];
  for (int t = 0; t < a.times; ++t) {
    if (__ballot(ring.fill < F) != 0ull && ring.fill <= 24 - 8) ring.append(a.key);
    double z[F];
#pragma unroll
    for (int f = 0; f < F; ++f) z[f] = mc_draw(ring, cur, wpos, a.key, zx);
    const size_t cell = ((size_t)p * (size_t)a.times + (size_t)t) * F;
#pragma unroll
    for (int i = 0; i < F; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j <= i; ++j) acc += a.chol[i * kMaxFactors + j] * z[j];
      const float y = (float)acc;
      if (valid && a.block) a.block[cell + i] = y;
      log_s[i] += a.drift[i] + a.diffusion[i] * (double)y;
      if (valid && a.cube) a.cube[cell + i] = (float)exp(log_s[i]);
    }
  }
}

// gbm_evolve of a given block: one lane per (path, factor).
__global__ __launch_bounds__(256) void gbm_kernel(CubeArgs a, int factors) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.paths * (uint64_t)factors) return;
  const uint64_t p = k / (uint64_t)factors;
  const int f = (int)(k % (uint64_t)factors);
  double log_s = a.ln_s0[f];
  for (int t = 0; t < a.times; ++t) {
    const size_t i = ((size_t)p * (size_t)a.times + (size_t)t) * (size_t)factors + (size_t)f;
    log_s += a.drift[f] + a.diffusion[f] * (double)a.in_block[i];
    a.cube[i] = (float)exp(log_s);
  }
}

hipError_t launch_cube(const CubeArgs& a, int factors, hipStream_t stream) {
  if (a.paths == 0 || a.times <= 0 || factors < 1 || factors > kMaxFactors) return hipErrorInvalidValue;
  if (a.in_block) {
    const uint64_t n = a.paths * (uint64_t)factors;
    hipLaunchKernelGGL(gbm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, factors);
    return hipGetLastError();
  }
  const dim3 grid((unsigned)((a.paths + 255) / 256)), block(256);
  switch (factors) {
    case 1: hipLaunchKernelGGL(cube_kernel<1>, grid, block, 0, stream, a); break;
    case 2: hipLaunchKernelGGL(cube_kernel<2>, grid, block, 0, stream, a); break;
    case 3: hipLaunchKernelGGL(cube_kernel<3>, grid, block, 0, stream, a); break;
    case 4: hipLaunchKernelGGL(cube_kernel<4>, grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL(cube_kernel<5>, grid, block, 0, stream, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_mc_parse(const McParseArgs& p, const McStartsArgs& s, hipStream_t stream) {
  if (p.chunk_words < 8 || p.chunk_words % 8 != 0 || p.chunk_words > (1 << 24) || p.n_chunks < 1 ||
      s.n_chunks != p.n_chunks)
    return hipErrorInvalidValue;
  const int64_t nb = (p.n_chunks + 255) / 256;
  hipLaunchKernelGGL(mc_parse_kernel, dim3((unsigned)nb), dim3(256), 0, stream, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc_resolve_kernel, dim3((unsigned)nb), dim3(256), 0, stream, p.chunks, p.n_chunks, s.entry,
                     s.cnt, s.block_off);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(mc_scan_kernel, dim3(1), dim3(1024), 0, stream, s.block_off, nb, s.status);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(mc_starts_kernel, dim3((unsigned)nb), dim3(256), 0, stream, s);
  return hipGetLastError();
}

}  // namespace riga
