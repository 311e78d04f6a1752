// kernels_f64.hip -- FP64 parity tier.  Built with -ffp-contract=off so that no
// multiply-add is fused: the reference (Rust) rounds every operation separately.
#include "kernels.h"
#include "phoenix_device.hpp"

namespace riga {

// The kernel of a launch (factor count, scenario bucket, ladder mode), passed to `op`.
template <int F, int SMAX, typename Op>
static hipError_t with_kernel(const PhoenixLaunch& l, Op op) {
  if (l.threads < 64 || l.threads > kBlock || l.threads % 64 != 0) return hipErrorInvalidValue;
  switch (l.mode) {
    case kLadderDirect: return op(phoenix_kernel<double, F, SMAX, kLadderDirect>);
    case kLadderTable: return op(phoenix_kernel<double, F, SMAX, kLadderTable>);
    case kLadderMonteCarlo: return op(phoenix_kernel<double, F, SMAX, kLadderMonteCarlo>);
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
hipError_t phoenix_blocks_per_cu_f64(const PhoenixLaunch& l, int* blocks) {
  return with_kernel(l, [&](auto kernel) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kernel, l.threads, l.lds);
  });
}

// Dispatches every path-kernel variant of the tier once with no units (each
// workgroup claims past the end of an empty table and leaves): the runtime sets
// up the private memory of kernels that spill registers on `stream`'s queue now,
// not at a later call's first dispatch.  `queue` is one device word.
hipError_t warm_path_kernels_f64(uint32_t* queue, hipStream_t stream) {
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

hipError_t launch_phoenix_f64(const PhoenixLaunch& l, hipStream_t stream) {
  if (l.blocks <= 0 || l.blocks > l.n_units || !l.queue) return hipErrorInvalidValue;
  return with_kernel(l, [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(l.blocks), dim3(l.threads), l.lds, stream, l.items, l.block_item, l.arena,
                       l.partials, l.records, l.queue, l.n_units);
    return hipGetLastError();
  });
}


__global__ void queue_reset_kernel(uint32_t* __restrict__ queues, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) queues[(size_t)i * kQueueStride] = 0u;
}

hipError_t launch_queue_reset(uint32_t* queues, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(queue_reset_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, queues, n);
  return hipGetLastError();
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

// The FP64 tier's exp and log (exp_log.h) of given arguments.
__global__ void debug_exp_log_kernel(const double* __restrict__ x, uint64_t n, double* __restrict__ e,
                                     double* __restrict__ l) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  e[i] = riga_exp(x[i]);
  l[i] = riga_log(x[i]);
}

hipError_t launch_debug_exp_log(const double* x, uint64_t n, double* e, double* l, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(debug_exp_log_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x, n, e, l);
  return hipGetLastError();
}

// Halton diagnostics (qmc.rs:143-153 as the path kernels evaluate it): for
// point n = first_point + p and dim = first_dim + d, the permuted digits
// perm[digit] of n in the dimension's base, least significant first (0xFFFFFFFF
// after the last), the radical inverse from halton_u and its FP64 inverse normal.
// `tables` is the planner's Halton layout: bases[rows], offsets[rows], then the
// permutations.
__global__ void debug_halton_kernel(const uint32_t* __restrict__ tables, int rows, uint64_t first_point,
                                    int n_points, int first_dim, int n_dims, uint32_t* __restrict__ digits,
                                    double* __restrict__ uniform, double* __restrict__ normal) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (p >= n_points || d >= n_dims) return;
  const uint32_t n = (uint32_t)(first_point + (uint64_t)p);
  const int di = (int)((uint32_t)(first_dim + d) % (uint32_t)rows);
  const uint32_t base = tables[di];
  const uint32_t* perm = tables + tables[rows + di];
  const size_t o = (size_t)p * n_dims + d;
  uint32_t m = n;
  for (int k = 0; k < 32; ++k) {
    digits[o * 32 + k] = m > 0 ? perm[m % base] : 0xFFFFFFFFu;
    m /= base;
  }
  const double u = halton_u(n, base, perm);
  uniform[o] = u;
  normal[o] = probit_f64(u);
}

hipError_t launch_debug_halton(const uint32_t* tables, int rows, uint64_t first_point, int n_points, int first_dim,
                               int n_dims, uint32_t* digits, double* uniform, double* normal, hipStream_t stream) {
  dim3 grid((n_points + 255) / 256, n_dims);
  hipLaunchKernelGGL(debug_halton_kernel, grid, dim3(256), 0, stream, tables, rows, first_point, n_points,
                     first_dim, n_dims, digits, uniform, normal);
  return hipGetLastError();
}

// tab[dim][n] = owen_scramble(sobol_x(gray(n), row(dim)), key(dim)) for the points n below 2^L:
// the scrambled Sobol coordinates of a seed, one lane per entry (coalesced stores).
__global__ __launch_bounds__(256) void owen_table_kernel(uint64_t seed, int dims, const uint32_t* __restrict__ dirs,
                                                         OwenEntry* __restrict__ out) {
  const uint32_t n = blockIdx.x * 256 + threadIdx.x;
  const uint32_t dim = blockIdx.y;
  if (n >= (1u << kOwenTableBits) || (int)dim >= dims) return;
  const uint32_t x = sobol_x(n ^ (n >> 1), dirs + (size_t)dim * 32, kOwenTableBits);
  out[((size_t)dim << kOwenTableBits) + n] = owen_scramble(x, seed ^ ((uint64_t)dim * 0x9E3779B97F4A7C15ull));
}

hipError_t launch_owen_table(uint64_t seed, int dims, const uint32_t* dirs, OwenEntry* out, hipStream_t stream) {
  dim3 grid((1u << kOwenTableBits) / 256, dims);
  hipLaunchKernelGGL(owen_table_kernel, grid, dim3(256), 0, stream, seed, dims, dirs, out);
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

hipError_t launch_normal_table_f64(const uint32_t* dirs, const OwenEntry* owen_tab, int owen_dims, uint64_t seed,
                                   int dims, int n_paths, int nbits, double* out, hipStream_t stream) {
  dim3 grid((n_paths + 255) / 256, dims);
  hipLaunchKernelGGL((normal_table_kernel<double>), grid, dim3(256), 0, stream, dirs, owen_tab, owen_dims, seed,
                     n_paths, nbits, out);
  return hipGetLastError();
}

}  // namespace riga
