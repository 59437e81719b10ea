// Row-mode register-tile kernel for transform type 0 (see lfx_reg_kernels.hpp).
#include "lfx_reg_kernels.hpp"

namespace lfx {
hipError_t launch_rows_0(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st) {
  return launch_rows_tf<0>(progs, items, nitems, in, out, scr, st);
}
}  // namespace lfx
