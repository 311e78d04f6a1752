// kernels_f64.hip -- FP64 parity tier.  Built with -ffp-contract=off so that no
// multiply-add is fused: the reference (Rust) rounds every operation separately.
#include "kernels.h"
#include "phoenix_device.hpp"

namespace riga {

template <int F, int SMAX>
static hipError_t launch_one(const PhoenixLaunch& l, hipStream_t stream) {
  if (l.threads < 64 || l.threads > kBlock || l.threads % 64 != 0) return hipErrorInvalidValue;
  switch (l.mode) {
    case kLadderDirect:
      hipLaunchKernelGGL((phoenix_kernel<double, F, SMAX, kLadderDirect>), dim3(l.blocks), dim3(l.threads), l.lds,
                         stream, l.items, l.block_item, l.arena, l.partials, l.records);
      break;
    case kLadderTable:
      hipLaunchKernelGGL((phoenix_kernel<double, F, SMAX, kLadderTable>), dim3(l.blocks), dim3(l.threads), l.lds,
                         stream, l.items, l.block_item, l.arena, l.partials, l.records);
      break;
    case kLadderMonteCarlo:
      hipLaunchKernelGGL((phoenix_kernel<double, F, SMAX, kLadderMonteCarlo>), dim3(l.blocks), dim3(l.threads),
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

hipError_t launch_phoenix_f64(const PhoenixLaunch& l, hipStream_t stream) {
  switch (l.F) {
    case 1: return launch_f<1>(l, stream);
    case 2: return launch_f<2>(l, stream);
    case 3: return launch_f<3>(l, stream);
    case 4: return launch_f<4>(l, stream);
    case 5: return launch_f<5>(l, stream);
    default: return hipErrorInvalidValue;
  }
}

// Merge the block partials of every (item, scenario) in block order.
__global__ void finalize_kernel(const ItemDesc* __restrict__ items, int n_items,
                                const Partial* __restrict__ partials, Partial* __restrict__ finals,
                                const int64_t* __restrict__ final_base) {
  // one workgroup per (item, scenario): 256 lanes each sum a fixed strided
  // subset of the block partials, then a fixed-shape tree in LDS; the order
  // never depends on timing, so the merge is deterministic.
  __shared__ DD sq[kNumQ][256];
  __shared__ uint64_t sc[kNumC][256];
  const int item = blockIdx.x;
  const int s = blockIdx.y;
  if (item >= n_items) return;
  const ItemDesc& it = items[item];
  if (s >= it.S) return;
  const Partial* p = partials + it.partial_base + (int64_t)s * it.n_blocks;
  Partial* out = finals + final_base[item] + s;
  const int t = threadIdx.x;
  DD a[kNumQ];
  uint64_t c[kNumC];
#pragma unroll
  for (int k = 0; k < kNumQ; ++k) a[k] = DD{0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kNumC; ++k) c[k] = 0;
  for (int b = t; b < it.n_blocks; b += 256) {
#pragma unroll
    for (int k = 0; k < kNumQ; ++k) a[k] = dd_add(a[k], p[b].q[k]);
#pragma unroll
    for (int k = 0; k < kNumC; ++k) c[k] += p[b].c[k];
  }
#pragma unroll
  for (int k = 0; k < kNumQ; ++k) sq[k][t] = a[k];
#pragma unroll
  for (int k = 0; k < kNumC; ++k) sc[k][t] = c[k];
  __syncthreads();
  for (int half = 128; half >= 1; half >>= 1) {
    if (t < half) {
#pragma unroll
      for (int k = 0; k < kNumQ; ++k) sq[k][t] = dd_add(sq[k][t], sq[k][t + half]);
#pragma unroll
      for (int k = 0; k < kNumC; ++k) sc[k][t] += sc[k][t + half];
    }
    __syncthreads();
  }
  if (t < kNumQ) out->q[t] = sq[t][0];
  if (t >= 32 && t < 32 + kNumC) out->c[t - 32] = sc[t - 32][0];
  if (t == 64) {  // first 64 KO events in path order: blocks in order
    int ne = 0;
    for (int b = 0; b < it.n_blocks && ne < kTraceCap; ++b) {
      for (int e = 0; e < p[b].n_events && ne < kTraceCap; ++e) {
        out->ev_path_lo[ne] = p[b].ev_path_lo[e];
        out->ev_step[ne] = p[b].ev_step[e];
        ++ne;
      }
    }
    out->n_events = ne;
  }
}


hipError_t launch_finalize(const ItemDesc* items, int n_items, const Partial* partials,
                           Partial* finals, const int64_t* final_base, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_kernel, dim3(n_items, 10), dim3(256), 0, stream, items, n_items,
                     partials, finals, final_base);
  return hipGetLastError();
}

__global__ void debug_sobol_kernel(const uint32_t* __restrict__ dirs, int ladder_dims,
                                   uint64_t seed, uint64_t first_point, int n_points,
                                   int first_dim, int n_dims, int nbits,
                                   uint32_t* __restrict__ unscrambled,
                                   uint32_t* __restrict__ scrambled, double* __restrict__ normal) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (p >= n_points || d >= n_dims) return;
  const uint32_t n = (uint32_t)(first_point + (uint64_t)p);
  const uint32_t gray = n ^ (n >> 1);
  const uint32_t dim = (uint32_t)(first_dim + d);
  const uint32_t* row = dirs + (size_t)(dim % (uint32_t)ladder_dims) * 32;
  const uint32_t x = sobol_x(gray, row, nbits);
  const uint32_t xs = owen_scramble(x, seed ^ ((uint64_t)dim * 0x9E3779B97F4A7C15ull));
  const size_t o = (size_t)p * n_dims + d;
  unscrambled[o] = x;
  scrambled[o] = xs;
  normal[o] = probit_f64(((double)xs + 0.5) / 4294967296.0);
}

hipError_t launch_debug_sobol(const uint32_t* dirs, int ladder_dims, uint64_t seed,
                              uint64_t first_point, int n_points, int first_dim, int n_dims,
                              int nbits, uint32_t* unscrambled, uint32_t* scrambled,
                              double* normal, hipStream_t stream) {
  dim3 grid((n_points + 255) / 256, n_dims);
  hipLaunchKernelGGL(debug_sobol_kernel, grid, dim3(256), 0, stream, dirs, ladder_dims, seed,
                     first_point, n_points, first_dim, n_dims, nbits, unscrambled, scrambled,
                     normal);
  return hipGetLastError();
}

// tab[dim][v] = top kOwenTableBits bits of owen_scramble(v << (32-L), key(dim)).
__global__ __launch_bounds__(256) void owen_table_kernel(uint64_t seed, int dims, uint16_t* __restrict__ out) {
  const uint32_t v = blockIdx.x * 256 + threadIdx.x;
  const uint32_t dim = blockIdx.y;
  if (v >= (1u << kOwenTableBits) || (int)dim >= dims) return;
  const uint64_t key = seed ^ ((uint64_t)dim * 0x9E3779B97F4A7C15ull);
  const uint32_t x = v << (32 - kOwenTableBits);
  uint64_t P = 0;
  for (int bit = 31; bit >= 32 - kOwenTableBits; --bit) owen_level(x, key, bit, P);
  out[(size_t)dim * (1u << kOwenTableBits) + v] = (uint16_t)owen_prefix(P);
}

hipError_t launch_owen_table(uint64_t seed, int dims, uint16_t* out, hipStream_t stream) {
  dim3 grid((1u << kOwenTableBits) / 256, dims);
  hipLaunchKernelGGL(owen_table_kernel, grid, dim3(256), 0, stream, seed, dims, out);
  return hipGetLastError();
}

// Coupon strips of replayed paths (explain.rs:146-262): one lane per path walks
// the path's worst-of series against the step table.  The daily KO is "every
// name at or above the barrier", i.e. the worst-of at or above it.
__global__ __launch_bounds__(256) void ledger_kernel(const ItemDesc* __restrict__ items, int item,
                                                     const unsigned char* __restrict__ arena,
                                                     CouponRowDev* __restrict__ rows) {
  const ItemDesc& it = items[item];
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= it.paths) return;
  const StepInfo* st = reinterpret_cast<const StepInfo*>(arena + it.off_steps);
  const int T = it.steps, P = it.periods;
  CouponRowDev* row = rows + p * (uint64_t)P;
  for (int q = 0; q < P; ++q) {
    CouponRowDev r;
    r.period = q;
    r.years = 0.0;
    r.qualified_steps = 0;
    const int64_t a = (int64_t)T * (q + 1) / P, b = (int64_t)T * q / P;
    r.period_steps = a > b ? a - b : 1;
    r.accrual = r.memory_paid = r.paid = r.memory_accrued = r.discount_factor = r.pv = 0.0;
    r.knocked_out = 0;
    r.pad = 0;
    row[q] = r;
  }
  const bool fixed_first = (it.flags & kFixedFirst) != 0;
  const bool memory = (it.flags & kMemoryCoupon) != 0;
  double accrual = 0.0, unpaid = 0.0;
  for (int step = 0; step < T; ++step) {
    const StepInfo si = st[step];
    const double worst = it.series[(size_t)step * it.paths + p];
    const int period = si.period;
    const double time = (double)(step + 1) * it.dt;
    if (si.flags & 4) accrual = 0.0;
    const bool fixed_now = period == 0 && fixed_first;
    if (!fixed_now && worst >= it.cb) {
      accrual += si.accrual_inc;
      row[period].qualified_steps += 1;
    }
    if ((si.flags & 2) && worst >= it.ko) {
      const double due = fixed_now ? it.coupon_rate : accrual;
      const double carry = due > 0.0 ? unpaid : 0.0;
      const double amount = due + carry;
      CouponRowDev& r = row[period];
      r.years = time;
      r.accrual = due;
      r.memory_paid = carry;
      r.paid = amount;
      r.discount_factor = si.df;
      r.pv = amount * si.df;
      r.knocked_out = 1;
      return;
    }
    if (si.flags & 1) {
      const double due = fixed_now ? it.coupon_rate : accrual;
      CouponRowDev& r = row[period];
      r.years = time;
      r.accrual = due;
      r.memory_paid = due > 0.0 ? unpaid : 0.0;
      r.paid = due > 0.0 ? due + unpaid : 0.0;
      r.discount_factor = si.df;
      r.pv = r.paid * si.df;
      r.memory_accrued = (due == 0.0 && memory) ? it.coupon_rate : 0.0;
      if (due > 0.0)
        unpaid = 0.0;
      else if (memory)
        unpaid += it.coupon_rate;
    }
  }
}

hipError_t launch_ledger(const ItemDesc* items, int item, uint64_t paths, const unsigned char* arena,
                         CouponRowDev* rows, hipStream_t stream) {
  hipLaunchKernelGGL(ledger_kernel, dim3((unsigned)((paths + 255) / 256)), dim3(256), 0, stream, items, item,
                     arena, rows);
  return hipGetLastError();
}

hipError_t launch_normal_table_f64(const uint32_t* dirs, const uint16_t* owen_tab, uint64_t seed,
                                   int dims, int n_paths, int nbits, double* out, hipStream_t stream) {
  dim3 grid((n_paths + 255) / 256, dims);
  hipLaunchKernelGGL((normal_table_kernel<double>), grid, dim3(256), 0, stream, dirs, owen_tab, seed,
                     n_paths, nbits, out);
  return hipGetLastError();
}

}  // namespace riga
