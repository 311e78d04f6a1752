// kernels_f32.hip -- FP32 fast tier: path arithmetic in FP32; Owen/Sobol
// integers are identical to the FP64 tier; payoff legs and sums stay FP64.
#include "kernels.h"
#include "phoenix_device.hpp"

namespace riga {

template <int F, int SMAX>
static hipError_t launch_one(const PhoenixLaunch& l, hipStream_t stream) {
  if (l.threads < 64 || l.threads > kBlock || l.threads % 64 != 0) return hipErrorInvalidValue;
  switch (l.mode) {
    case kLadderDirect:
      hipLaunchKernelGGL((phoenix_kernel<float, F, SMAX, kLadderDirect>), dim3(l.blocks), dim3(l.threads), l.lds,
                         stream, l.items, l.block_item, l.arena, l.partials, l.records);
      break;
    case kLadderTable:
      hipLaunchKernelGGL((phoenix_kernel<float, F, SMAX, kLadderTable>), dim3(l.blocks), dim3(l.threads), l.lds,
                         stream, l.items, l.block_item, l.arena, l.partials, l.records);
      break;
    case kLadderMonteCarlo:
      hipLaunchKernelGGL((phoenix_kernel<float, F, SMAX, kLadderMonteCarlo>), dim3(l.blocks), dim3(l.threads),
                         l.lds, stream, l.items, l.block_item, l.arena, l.partials, l.records);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int F>
static hipError_t launch_f(const PhoenixLaunch& l, hipStream_t stream) {
  switch (l.smax) {
    case 1: return launch_one<F, 1>(l, stream);
    case 4: return launch_one<F, 4>(l, stream);
    case 10: return launch_one<F, 10>(l, stream);
    default: return hipErrorInvalidValue;
  }
}

// The FP32 tier's Sobol ladder as the path kernel runs it (prefix-table gather,
// 24 hashed levels, the wave's tail passes, the node's FP32 inverse normals) for points
// first_point + p and dims node*F + f: xs_out/z_out[p][node*F + f].
template <int F>
__global__ __launch_bounds__(256) void debug_sobol_f32_kernel(const uint32_t* __restrict__ dirs, int ladder_dims,
                                                              const uint16_t* __restrict__ owen_tab, uint64_t seed,
                                                              uint64_t first_point, int n_points, int n_nodes,
                                                              int nbits, uint32_t* __restrict__ xs_out,
                                                              float* __restrict__ z_out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_points) return;
  ItemDesc it{};
  it.flags = owen_tab ? kOwenTable : 0u;
  it.ladder_dims = ladder_dims;
  it.off_dirs = 0;
  it.nbits = nbits;
  it.owen_tab = owen_tab;
  it.seed = seed;
  const uint32_t n = (uint32_t)(first_point + (uint64_t)p);
  const uint32_t gray = n ^ (n >> 1);
  const size_t row = (size_t)p * n_nodes * F;
  for (int node = 0; node < n_nodes; ++node) {
    SobolFetch<F> sf;
    sobol_fetch<F>(it, reinterpret_cast<const unsigned char*>(dirs), gray, node, sf);
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

hipError_t launch_debug_sobol_f32(const uint32_t* dirs, int ladder_dims, const uint16_t* owen_tab, uint64_t seed,
                                  uint64_t first_point, int n_points, int n_nodes, int factors, int nbits,
                                  uint32_t* xs_out, float* z_out, hipStream_t stream) {
  const dim3 grid((n_points + 255) / 256);
#define RIGA_DEBUG_F32(FF)                                                                                  \
  case FF:                                                                                                  \
    hipLaunchKernelGGL(debug_sobol_f32_kernel<FF>, grid, dim3(256), 0, stream, dirs, ladder_dims, owen_tab, \
                       seed, first_point, n_points, n_nodes, nbits, xs_out, z_out);                         \
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

hipError_t launch_phoenix_f32(const PhoenixLaunch& l, hipStream_t stream) {
  switch (l.F) {
    case 1: return launch_f<1>(l, stream);
    case 2: return launch_f<2>(l, stream);
    case 3: return launch_f<3>(l, stream);
    case 4: return launch_f<4>(l, stream);
    case 5: return launch_f<5>(l, stream);
    default: return hipErrorInvalidValue;
  }
}

// The reference's device module (crates/riga-engine/src/gpu.rs:27-75), for callers
// of its worst_of / lv_terminal_spots API.  Elementwise worst-of of two
// performance vectors, with the reference's `a < b ? a : b` (a NaN in a gives b).
__global__ __launch_bounds__(256) void worst_of_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                       float* __restrict__ o, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) o[i] = a[i] < b[i] ? a[i] : b[i];
}

// Two-factor FP32 local-vol propagation from given shocks, one lane per path:
// normals [path][step][2], rows [factor][step][gx] (local vol per log-spot knot),
// ls_axis [factor][gx] (evenly spaced knots).  Per factor and step: the knot
// position (log S - x0) / (x_last - x0) * (gx - 1), clamped to the grid, a linear
// blend of the row, then log S += (r - q - sigma^2/2) dt + sigma sqrt(dt) z; the
// first step with both names at or above the call barrier (S / reference), or -1.
__global__ __launch_bounds__(256) void lv_propagate_kernel(LvPropagateArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.paths) return;
  const int gx = a.gx;
  float x[2], lo[2], span[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    x[f] = logf(a.s0[f]);
    lo[f] = a.ls_axis[f * gx];
    span[f] = 1.0f / (a.ls_axis[f * gx + gx - 1] - lo[f]);
  }
  const float sqdt = sqrtf(a.dt);
  const float top = (float)(gx - 1);
  const float q[2] = {a.q0, a.q1};
  const float ref[2] = {a.ref0, a.ref1};
  int ko = -1;
  for (int s = 0; s < a.steps; ++s) {
    const float* z = a.normals + ((size_t)p * a.steps + s) * 2;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      float k = (x[f] - lo[f]) * span[f] * top;
      k = k < 0.0f ? 0.0f : (k > top ? top : k);
      int j = (int)k;
      j = j > gx - 2 ? gx - 2 : j;
      const float w = k - (float)j;
      const float* row = a.rows + ((size_t)f * a.steps + s) * gx;
      const float sig = row[j] * (1.0f - w) + row[j + 1] * w;
      x[f] += (a.rate - q[f] - 0.5f * sig * sig) * a.dt + sig * sqdt * z[f];
    }
    if (ko < 0 && expf(x[0]) / ref[0] >= a.ko_bar && expf(x[1]) / ref[1] >= a.ko_bar) ko = s;
  }
  a.out[2 * p] = expf(x[0]);
  a.out[2 * p + 1] = expf(x[1]);
  a.ko_out[p] = ko;
}

hipError_t launch_worst_of(const float* a, const float* b, float* o, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(worst_of_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, b, o, n);
  return hipGetLastError();
}

hipError_t launch_lv_propagate(const LvPropagateArgs& a, hipStream_t stream) {
  if (a.paths <= 0) return hipSuccess;
  if (a.gx < 2 || a.steps < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lv_propagate_kernel, dim3((unsigned)((a.paths + 255) / 256)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_normal_table_f32(const uint32_t* dirs, const uint16_t* owen_tab, uint64_t seed,
                                   int dims, int n_paths, int nbits, float* out, hipStream_t stream) {
  dim3 grid((n_paths + 255) / 256, dims);
  hipLaunchKernelGGL((normal_table_kernel<float>), grid, dim3(256), 0, stream, dirs, owen_tab, seed,
                     n_paths, nbits, out);
  return hipGetLastError();
}

}  // namespace riga
