// Register-tile kernels for transform type 3 (see lfx_reg_kernels.hpp).
#include "lfx_reg_kernels.hpp"

namespace lfx {
hipError_t launch_reg_3(uint32_t wavelet, const void* prog, const BatchItem* items, long long nitems,
                        const void* in, void* out, void* scr, hipStream_t st) {
  return launch_reg_tf<3>(wavelet, prog, items, nitems, in, out, scr, st);
}
}  // namespace lfx
