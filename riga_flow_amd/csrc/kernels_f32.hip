// kernels_f32.hip -- FP32 fast tier: path arithmetic in FP32; Owen/Sobol
// integers are identical to the FP64 tier; payoff legs and sums stay FP64.
#include "kernels.h"
#include "phoenix_device.hpp"

namespace riga {

template <int F, int SMAX>
static hipError_t launch_one(const PhoenixLaunch& l, hipStream_t stream) {
  if (l.threads < 64 || l.threads > kBlock || l.threads % 64 != 0) return hipErrorInvalidValue;
  if (l.mc)
    hipLaunchKernelGGL((phoenix_kernel<float, F, SMAX, true>), dim3(l.blocks), dim3(l.threads), l.lds,
                       stream, l.items, l.block_item, l.arena, l.partials, l.records);
  else
    hipLaunchKernelGGL((phoenix_kernel<float, F, SMAX, false>), dim3(l.blocks), dim3(l.threads), l.lds,
                       stream, l.items, l.block_item, l.arena, l.partials, l.records);
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

hipError_t launch_normal_table_f32(const uint32_t* dirs, const uint16_t* owen_tab, uint64_t seed,
                                   int dims, int n_paths, int nbits, float* out, hipStream_t stream) {
  dim3 grid((n_paths + 255) / 256, dims);
  hipLaunchKernelGGL((normal_table_kernel<float>), grid, dim3(256), 0, stream, dirs, owen_tab, seed,
                     n_paths, nbits, out);
  return hipGetLastError();
}

}  // namespace riga
