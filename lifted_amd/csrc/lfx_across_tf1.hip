// Across-line kernel for transform type 1 (see lfx_across_kernels.hpp).
#include "lfx_across_kernels.hpp"

namespace lfx {
hipError_t launch_across_1(const acr::Args& A, hipStream_t st) { return launch_across_tf<1>(A, st); }
}  // namespace lfx
