// Across-line kernel for transform type 0 (see lfx_across_kernels.hpp).
#include "lfx_across_kernels.hpp"

namespace lfx {
hipError_t launch_across_0(const acr::Args& A, hipStream_t st) { return launch_across_tf<0>(A, st); }
}  // namespace lfx
