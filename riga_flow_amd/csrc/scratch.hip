// scratch.hip -- the fixed device work area of the allocation-free pricing call.
//
// riga_price_put_leg_noalloc lays a request's tables, block partials, merged
// records and (Monte Carlo lane) stream parse out in this array instead of
// allocating them.  Each GPU holds its own copy for the life of the process.
// The array's name is a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for the current device.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace riga {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of this device's scratch array (the caller has made the
// device current); called once per device when its context is created.
hipError_t scratch_address(void** ptr) {
  *ptr = nullptr;
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

}  // namespace riga
