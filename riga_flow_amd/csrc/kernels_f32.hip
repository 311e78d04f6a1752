// kernels_f32.hip -- FP32 fast tier: path arithmetic in FP32; Owen/Sobol
// integers are identical to the FP64 tier; payoff legs and sums stay FP64.
#include "kernels.h"
#include "phoenix_device.hpp"

namespace riga {

// The kernel of a launch (factor count, scenario bucket, ladder mode), passed to `op`.
template <int F, int SMAX, typename Op>
static hipError_t with_kernel(const PhoenixLaunch& l, Op op) {
  if (l.threads < 64 || l.threads > kBlock || l.threads % 64 != 0) return hipErrorInvalidValue;
  switch (l.mode) {
    case kLadderDirect: return op(phoenix_kernel<float, F, SMAX, kLadderDirect>);
    case kLadderTable: return op(phoenix_kernel<float, F, SMAX, kLadderTable>);
    case kLadderMonteCarlo: return op(phoenix_kernel<float, F, SMAX, kLadderMonteCarlo>);
    default: return hipErrorInvalidValue;
  }
}

template <typename Op>
static hipError_t with_kernel(const PhoenixLaunch& l, Op op) {
#define RIGA_F(FF)                                            \
  case FF:                                                    \
    switch (l.smax) {                                         \
      case 1: return with_kernel<FF, 1>(l, op);               \
      case 4: return with_kernel<FF, 4>(l, op);               \
      case 10: return with_kernel<FF, 10>(l, op);             \
      default: return hipErrorInvalidValue;                   \
    }
  switch (l.F) {
    RIGA_F(1)
    RIGA_F(2)
    RIGA_F(3)
    RIGA_F(4)
    RIGA_F(5)
    default: return hipErrorInvalidValue;
  }
#undef RIGA_F
}

// Workgroups of the launch's kernel that fit one CU with l.threads lanes and
// l.lds bytes of dynamic LDS (registers, LDS and wave slots all counted).
hipError_t phoenix_blocks_per_cu_f32(const PhoenixLaunch& l, int* blocks) {
  return with_kernel(l, [&](auto kernel) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kernel, l.threads, l.lds);
  });
}

// Dispatches every path-kernel variant of the tier once with no units (each
// workgroup claims past the end of an empty table and leaves): the runtime sets
// up the private memory of kernels that spill registers on `stream`'s queue now,
// not at a later call's first dispatch.  `queue` is one device word.
hipError_t warm_path_kernels_f32(uint32_t* queue, hipStream_t stream) {
  for (int F = 1; F <= kMaxFactors; ++F)
    for (int smax : kScenarioBuckets)
      for (int mode : {kLadderDirect, kLadderTable, kLadderMonteCarlo}) {
        PhoenixLaunch l{};
        l.F = F;
        l.smax = smax;
        l.mode = mode;
        l.threads = 64;
        const hipError_t e = with_kernel(l, [&](auto kernel) {
          hipLaunchKernelGGL(kernel, dim3(1), dim3(64), 0, stream, nullptr, nullptr, nullptr, nullptr, nullptr,
                             queue, 0);
          return hipGetLastError();
        });
        if (e != hipSuccess) return e;
      }
  return hipSuccess;
}

hipError_t launch_phoenix_f32(const PhoenixLaunch& l, hipStream_t stream) {
  if (l.blocks <= 0 || l.blocks > l.n_units || !l.queue) return hipErrorInvalidValue;
  return with_kernel(l, [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(l.blocks), dim3(l.threads), l.lds, stream, l.items, l.block_item, l.arena,
                       l.partials, l.records, l.queue, l.n_units);
    return hipGetLastError();
  });
}


// The FP32 tier's Sobol ladder as the path kernel runs it (coordinate-table read or
// 24 hashed levels, the wave's tail passes, the node's FP32 inverse normals) for points
// first_point + p and dims node*F + f: xs_out/z_out[p][node*F + f].
template <int F>
__global__ __launch_bounds__(256) void debug_sobol_f32_kernel(const uint32_t* __restrict__ dirs, int ladder_dims,
                                                              const OwenEntry* __restrict__ owen_tab, uint64_t seed,
                                                              uint64_t first_point, int n_points, int n_nodes,
                                                              int nbits, bool full_ladder,
                                                              uint32_t* __restrict__ xs_out,
                                                              float* __restrict__ z_out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_points) return;
  ItemDesc it{};
  it.flags = (owen_tab ? kOwenTable : 0u) | (full_ladder ? kFullLadder : 0u);
  it.ladder_dims = ladder_dims;
  it.off_dirs = 0;
  it.nbits = nbits;
  it.owen_tab = owen_tab;
  it.owen_dims = n_nodes * F;  // the host's table covers every dimension read here
  it.seed = seed;
  const uint32_t n = (uint32_t)(first_point + (uint64_t)p);
  const uint32_t gray = n ^ (n >> 1);
  const size_t row = (size_t)p * n_nodes * F;
  for (int node = 0; node < n_nodes; ++node) {
    SobolFetch<F> sf;
    sobol_fetch<F>(it, reinterpret_cast<const unsigned char*>(dirs), n, gray, node, sf);
    uint32_t xs[F];
    sobol_ints_f32<F>(it, node, sf, xs);
    double u[F];
    float z[F];
#pragma unroll
    for (int f = 0; f < F; ++f) u[f] = ((double)xs[f] + 0.5) / 4294967296.0;
    probit_node<float, F>(u, z);
#pragma unroll
    for (int f = 0; f < F; ++f) {
      xs_out[row + (size_t)node * F + f] = xs[f];
      z_out[row + (size_t)node * F + f] = z[f];
    }
  }
}

hipError_t launch_debug_sobol_f32(const uint32_t* dirs, int ladder_dims, const OwenEntry* owen_tab, uint64_t seed,
                                  uint64_t first_point, int n_points, int n_nodes, int factors, int nbits,
                                  bool full_ladder, uint32_t* xs_out, float* z_out, hipStream_t stream) {
  const dim3 grid((n_points + 255) / 256);
#define RIGA_DEBUG_F32(FF)                                                                                  \
  case FF:                                                                                                  \
    hipLaunchKernelGGL(debug_sobol_f32_kernel<FF>, grid, dim3(256), 0, stream, dirs, ladder_dims, owen_tab, \
                       seed, first_point, n_points, n_nodes, nbits, full_ladder, xs_out, z_out);            \
    break;
  switch (factors) {
    RIGA_DEBUG_F32(1)
    RIGA_DEBUG_F32(2)
    RIGA_DEBUG_F32(3)
    RIGA_DEBUG_F32(4)
    RIGA_DEBUG_F32(5)
    default:
      return hipErrorInvalidValue;
  }
#undef RIGA_DEBUG_F32
  return hipGetLastError();
}


// The reference's device module (crates/riga-engine/src/gpu.rs:27-75), for callers
// of its worst_of / lv_terminal_spots API.  Elementwise worst-of of two
// performance vectors, with the reference's `a < b ? a : b` (a NaN in a gives b).
__global__ __launch_bounds__(256) void worst_of_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                       float* __restrict__ o, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) o[i] = a[i] < b[i] ? a[i] : b[i];
}

// Two-factor FP32 local-vol propagation from given shocks (the semantics of the
// reference device module's lv_terminal_spots, gpu.rs:207-278): per path and
// step, each factor's local vol is the linear interpolation of that step's row
// at the factor's log-spot on evenly spaced knots (held at the end knots
// outside them), the log-spot takes the Euler step
//   x += (r - q) dt + vol sqrt(dt) z - vol^2 dt / 2,
// and the KO step is the first step at which both spots stand at or above
// call_barrier x reference.  Terminal spots exp(x) and KO steps are written.
//
// The shocks arrive [path][step][2], so a lane's steps are contiguous and direct
// per-lane reads would touch a different line in every lane.  A workgroup
// (256 paths) instead stages a tile of `tile` steps at a time through LDS with
// line-wide loads, together with the tile's rows as slope pairs
// (v[j], v[j+1] - v[j]): a lookup is one 8-byte LDS read and one FMA, the knot
// position one FMA of the state, and the KO test two compares of the log-spots
// against ln(call_barrier x reference).  LDS: shocks [256][2*tile + 1] floats
// (the odd row pitch keeps the per-lane reads conflict-free), rows
// [tile][2][gx] float2.
__global__ __launch_bounds__(256) void lv_propagate_kernel(LvPropagateArgs a, int tile) {
  extern __shared__ __align__(16) unsigned char lv_lds[];
  const int pitch = 2 * tile + 1;
  float* zs = reinterpret_cast<float*>(lv_lds);
  float2* slope = reinterpret_cast<float2*>(lv_lds + (((size_t)256 * pitch * 4 + 15) & ~(size_t)15));
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * 256;
  const int64_t path = p0 + tid;
  const int gx = a.gx;
  const float tmax = (float)(gx - 1);
  float x[2] = {a.x0[0], a.x0[1]};
  int ko = -1;
  const int64_t n_here = a.paths - p0 < 256 ? a.paths - p0 : 256;  // paths of this workgroup
  for (int t0 = 0; t0 < a.steps; t0 += tile) {
    const int k = a.steps - t0 < tile ? a.steps - t0 : tile;
    __syncthreads();  // the previous tile is consumed
    // shocks of paths [p0, p0 + n_here) x steps [t0, t0 + k): each path's 2k
    // floats are contiguous, consecutive lanes take consecutive floats
    const int seg = 2 * k;
    for (int i = tid; i < n_here * seg; i += 256) {
      const int pl = i / seg, e = i - pl * seg;
      zs[pl * pitch + e] = a.normals[((size_t)(p0 + pl) * a.steps + t0) * 2 + e];
    }
    for (int i = tid; i < k * 2 * gx; i += 256) {
      const int j = i % gx, sf = i / gx;  // sf = step_in_tile * 2 + factor
      const int f = sf & 1, s = sf >> 1;
      const float* row = a.rows + ((size_t)f * a.steps + (t0 + s)) * gx;
      const float v = row[j];
      slope[i] = make_float2(v, j + 1 < gx ? row[j + 1] - v : 0.0f);
    }
    __syncthreads();
    if (path < a.paths) {
      for (int s = 0; s < k; ++s) {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          const float t = __builtin_amdgcn_fmed3f(fmaf(x[f], a.knot_scale[f], a.knot_offset[f]), 0.0f, tmax);
          int j = (int)t;
          j = j < gx - 2 ? j : gx - 2;
          const float w = t - (float)j;
          const float2 e = slope[(s * 2 + f) * gx + j];
          const float vol = fmaf(w, e.y, e.x);
          const float sz = a.sqrt_dt * zs[tid * pitch + 2 * s + f];
          x[f] = x[f] + fmaf(vol, fmaf(vol, a.neg_half_dt, sz), a.drift[f]);
        }
        if (ko < 0 && x[0] >= a.ln_ko[0] && x[1] >= a.ln_ko[1]) ko = t0 + s;
      }
    }
  }
  if (path < a.paths) {
    a.out[2 * path] = expf(x[0]);
    a.out[2 * path + 1] = expf(x[1]);
    a.ko_out[path] = ko;
  }
}

// Steps per staged tile: the largest of 16, 8, 4, 2, 1 whose LDS fits 60 KiB.
static int lv_tile(int gx, size_t* lds) {
  for (int t = 16; t >= 1; t /= 2) {
    const size_t b = (((size_t)256 * (2 * t + 1) * 4 + 15) & ~(size_t)15) + (size_t)t * 2 * gx * 8;
    if (b <= 61440) {
      *lds = b;
      return t;
    }
  }
  return 0;
}

hipError_t launch_worst_of(const float* a, const float* b, float* o, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(worst_of_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, b, o, n);
  return hipGetLastError();
}

hipError_t launch_lv_propagate(const LvPropagateArgs& a, hipStream_t stream) {
  if (a.paths <= 0) return hipSuccess;
  size_t lds = 0;
  const int tile = a.gx >= 2 && a.gx <= kLvMaxKnots ? lv_tile(a.gx, &lds) : 0;
  if (tile == 0 || a.steps < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lv_propagate_kernel, dim3((unsigned)((a.paths + 255) / 256)), dim3(256), lds, stream, a, tile);
  return hipGetLastError();
}

hipError_t launch_normal_table_f32(const uint32_t* dirs, const OwenEntry* owen_tab, int owen_dims, uint64_t seed,
                                   int dims, int n_paths, int nbits, float* out, hipStream_t stream) {
  dim3 grid((n_paths + 255) / 256, dims);
  hipLaunchKernelGGL((normal_table_kernel<float>), grid, dim3(256), 0, stream, dirs, owen_tab, owen_dims, seed,
                     n_paths, nbits, out);
  return hipGetLastError();
}

}  // namespace riga
