// Register-tile kernels of one transform type. build.py compiles this file
// once per transform type (LFX_TF = 0..3) and family (LFX_ROWS = 0: windows
// of one line, of LFX_PL = 16 or 32 samples per lane; 1: rows of four equal
// lines), so that the 34 wavelet instantiations of each compile in parallel.
// One wavefront runs one BatchItem of kind register tile.
#include <hip/hip_runtime.h>

#include "lfx_kernels.hpp"
#include "lfx_reg.hpp"
#include "lfx_xcd.hpp"

namespace lfx {

constexpr int kStageFloats = 64 * 17;

template <class W, int TF, int PL>
__device__ __forceinline__ void reg_item(const BatchItem& it, const DProgram<float>* prog, float* a0, float* a1,
                                         float* a2, float* lds) {
  if constexpr (TF == kForward || TF == kInverseAdjoint)
    reg::fwd_tile<reg::LaneGpu, W, TF, PL>(prog, it.m0, it.k, it.a, it.b, it.H, (const float*)a0, a1, a2, lds);
  else
    reg::rev_tile<reg::LaneGpu, W, TF, PL>(prog, it.m0, it.k, it.a, it.b, it.H, (const float*)a0, (const float*)a2,
                                           a1, lds);
}

// All register-tile items of a stage, any table wavelet: the item's wavelet id
// selects the compiled step shapes (a wave-uniform branch), its program index
// the coefficients in the plan's program table. PL: samples per lane of the
// plan's windows.
template <int TF, int PL>
__global__ void __launch_bounds__(64 * kRegWaves) k_reg(const BatchItem* items, long long nitems, const DProgram<float>* progs,
                                             const float* uin, float* uout, float* scr) {
  // per-wave transposition buffer: 64 rows of 17 floats (16 values per lane,
  // or 32 rows of 33 floats per round of the 32-per-lane transposes)
  __shared__ float stage[kRegWaves][kStageFloats];
  const long long w = __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * kRegWaves + (threadIdx.x >> 6)));
  if (w >= nitems) return;
  float* lds = stage[threadIdx.x >> 6];
  const BatchItem it = items[w];
  const DProgram<float>* prog = progs + it.pidx;
  // select among the kernel's pointer arguments (keeps them in the global
  // address space for the compiler, unlike indexing an array of pointers)
  auto pick = [&](int a) { return a == kArrIn ? const_cast<float*>(uin) : (a == kArrOut ? uout : scr); };
  float* a0 = pick(it.arr[0]) + it.off[0];
  float* a1 = pick(it.arr[1]) + it.off[1];
  float* a2 = pick(it.arr[2]) + it.off[2];
  switch (it.wavelet) {
#define LFX_WAVELET(nm, idv, w)                                     \
  case idv:                                                         \
    reg_item<wt::WL_##nm, TF, PL>(it, prog, a0, a1, a2, lds);       \
    return;
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
      return;
  }
}

// Row items: four equal lines of at most 512 samples per wavefront, one per
// 16-lane row, whole bands from position 0 (Zero boundary).
template <class W, int TF>
__device__ __forceinline__ void row_item(const BatchItem& it, const DProgram<float>* prog, float* a0, float* a1,
                                         float* a2, float* lds) {
  const reg::RowArgs ra{{it.rstride[0], it.rstride[1], it.rstride[2]}, it.rows};
  if constexpr (TF == kForward || TF == kInverseAdjoint)
    reg::fwd_tile<reg::RowGpu, W, TF>(prog, it.m0, it.k, 0, it.m0, 0, (const float*)a0, a1, a2, lds, ra);
  else
    reg::rev_tile<reg::RowGpu, W, TF>(prog, it.m0, it.k, 0, it.m0, 0, (const float*)a0, (const float*)a2, a1, lds, ra);
}

template <int TF>
__global__ void __launch_bounds__(64 * kRegWaves) k_reg_rows(const BatchItem* items, long long nitems,
                                                  const DProgram<float>* progs, const float* uin, float* uout,
                                                  float* scr) {
  __shared__ float stage[kRegWaves][kStageFloats];
  const long long w = __builtin_amdgcn_readfirstlane((int)(xcd_block(blockIdx.x, gridDim.x) * kRegWaves + (threadIdx.x >> 6)));
  if (w >= nitems) return;
  float* lds = stage[threadIdx.x >> 6];
  const BatchItem it = items[w];
  const DProgram<float>* prog = progs + it.pidx;
  auto pick = [&](int a) { return a == kArrIn ? const_cast<float*>(uin) : (a == kArrOut ? uout : scr); };
  float* a0 = pick(it.arr[0]) + it.off[0];
  float* a1 = pick(it.arr[1]) + it.off[1];
  float* a2 = pick(it.arr[2]) + it.off[2];
  switch (it.wavelet) {
#define LFX_WAVELET(nm, idv, w)                           \
  case idv:                                               \
    row_item<wt::WL_##nm, TF>(it, prog, a0, a1, a2, lds); \
    return;
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
      return;
  }
}

template <int TF>
hipError_t launch_rows_tf(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                          void* scr, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((nitems + kRegWaves - 1) / kRegWaves);
  hipLaunchKernelGGL((k_reg_rows<TF>), dim3(blocks), dim3(64 * kRegWaves), 0, st, items, nitems, (const DProgram<float>*)progs,
                     (const float*)in, (float*)out, (float*)scr);
  return hipGetLastError();
}

template <int TF, int PL>
hipError_t launch_reg_tf(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((nitems + kRegWaves - 1) / kRegWaves);
  hipLaunchKernelGGL((k_reg<TF, PL>), dim3(blocks), dim3(64 * kRegWaves), 0, st, items, nitems, (const DProgram<float>*)progs,
                     (const float*)in, (float*)out, (float*)scr);
  return hipGetLastError();
}

#define LFX_JOIN2(a, b) a##b
#define LFX_JOIN(a, b) LFX_JOIN2(a, b)
#if LFX_ROWS
hipError_t LFX_JOIN(launch_rows_, LFX_TF)(const void* progs, const BatchItem* items, long long nitems, const void* in,
                                          void* out, void* scr, hipStream_t st) {
  return launch_rows_tf<LFX_TF>(progs, items, nitems, in, out, scr, st);
}
#else
hipError_t LFX_JOIN(LFX_JOIN(LFX_JOIN(launch_reg_, LFX_TF), _), LFX_PL)(const void* progs, const BatchItem* items,
                                                                      long long nitems, const void* in, void* out,
                                                                      void* scr, hipStream_t st) {
  return launch_reg_tf<LFX_TF, LFX_PL>(progs, items, nitems, in, out, scr, st);
}
#endif

}  // namespace lfx
