// Two-axis level kernel for transform type 0 (see lfx_lvl2d_kernels.hpp).
#include "lfx_lvl2d_kernels.hpp"

namespace lfx {
hipError_t launch_lvl2d_0(const l2d::Level& L, hipStream_t st) { return launch_lvl2d_tf<0>(L, st); }
}  // namespace lfx
