// Two-axis level kernels of one transform type and storage type. build.py
// compiles this file once per transform type (LFX_TF = 0..3) and storage type
// (LFX_ST: float32, bfloat16, float16), so that they compile in parallel: one
// kernel per table wavelet (its compiled step shapes), chosen on the host, so
// that each is compiled and register-allocated on its own.
#include <hip/hip_runtime.h>

#include "lfx_kernels.hpp"
#include "lfx_lvl2d.hpp"
#include "lfx_xcd.hpp"

namespace lfx {

// One level of a two-axis level-by-level transform (lfx_lvl2d.hpp): one
// wavefront per tile.
template <int TF, class W, int ST>
__global__ void __launch_bounds__(64) k_lvl2d(l2d::Level L) {
  const long long w = (long long)xcd_block(blockIdx.x, gridDim.x);
  if (w >= L.ntiles) return;
  l2d::level_tile<reg::LaneGpu, W, TF, ST>(L, w);
}

template <int TF, int ST>
hipError_t launch_lvl2d_tf(const l2d::Level& L, hipStream_t st) {
  if (L.ntiles <= 0) return hipSuccess;
  if (L.ntiles > 0x7fffffffll || L.a[0].st != ST || L.a[1].st != ST || L.a[2].st != ST) return hipErrorInvalidValue;
  switch (L.wavelet) {
#define LFX_WAVELET(nm, idv, w_)                                                                 \
  case idv:                                                                                     \
    hipLaunchKernelGGL((k_lvl2d<TF, wt::WL_##nm, ST>), dim3((unsigned)L.ntiles), dim3(64), 0, st, L); \
    break;
#define LFX_GENERIC(...)
#define LFX_UNIT(...)
#define LFX_REPEAT(...)
#define LFX_SCALE(...)
#define LFX_END(nm)
#include "wavelet_table.inc"
#undef LFX_WAVELET
#undef LFX_GENERIC
#undef LFX_UNIT
#undef LFX_REPEAT
#undef LFX_SCALE
#undef LFX_END
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define LFX_JOIN2(a, b) a##b
#define LFX_JOIN(a, b) LFX_JOIN2(a, b)
hipError_t LFX_JOIN(LFX_JOIN(LFX_JOIN(launch_lvl2d_, LFX_TF), _), LFX_ST)(const l2d::Level& L, hipStream_t st) {
  return launch_lvl2d_tf<LFX_TF, LFX_ST>(L, st);
}

}  // namespace lfx
