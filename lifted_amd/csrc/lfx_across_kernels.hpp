// Across-line kernel of one transform type (included once per transform by
// lfx_across_tf{0,1,2,3}.hip, so the table's instantiations compile in
// parallel): lfx_across.hpp's blocks of lines along a strided axis.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <set>

#include "lfx_across.hpp"
#include "lfx_kernels.hpp"
#include "lfx_xcd.hpp"

namespace lfx {

extern __shared__ __align__(16) float lfx_acr_slab[];

// Row-mode lane policy whose "memory" is the workgroup's LDS slab: a pointer
// handed to the tile code is a virtual slab address (slab + virtual index),
// mapped through acr::phys on every access. The row tiles' chunk layout is
// read and written in place (no staging transposes: the slab layout keeps
// those accesses conflict-free).
struct RowLds : reg::LaneGpuT<4> {
  using V = float;
  static __device__ __forceinline__ float& at(const float* p, int off) {
    const int a = (int)(p - lfx_acr_slab) + off;
    return lfx_acr_slab[acr::phys(a)];
  }
  // x[e] of lane l = p[r rs + g], g = pos0 + E l' + e, for row r = l / 16 < nr
  // and g in [0, n); 0 otherwise
  template <int E, class S>
  static __device__ __forceinline__ void ld_raw(const S* p, long long pos0, long long n, V (&t)[E], long long rs = 0,
                                                int nr = 1) {
    const int l = lane(), r = l >> 4, lr = l & 15;
    const float* q = reinterpret_cast<const float*>(p);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int g = (int)pos0 + E * lr + e;
      t[e] = (r < nr && g >= 0 && g < n) ? at(q, r * (int)rs + g) : 0.0f;
    }
  }
  template <int E>
  static __device__ __forceinline__ void xpose_in(const V (&t)[E], V (&x)[E], float*) {
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = t[e];
  }
  template <int E, class S>
  static __device__ __forceinline__ void ld_chunks(const S* p, long long pos0, long long n, V (&x)[E], float*,
                                                   long long rs = 0, int nr = 1) {
    ld_raw<E, S>(p, pos0, n, x, rs, nr);
  }
  // p[r rs + pos0 + g] = v[e] for g = g0 + E l' + e in [lo, hi), row r < nr
  template <int E, class S>
  static __device__ __forceinline__ void st_chunks(S* p, long long pos0, const V (&v)[E], long long g0, long long lo,
                                                   long long hi, float*, long long rs = 0, int nr = 1) {
    const int l = lane(), r = l >> 4, lr = l & 15;
    const float* q = reinterpret_cast<const float*>(p);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int g = (int)g0 + E * lr + e;
      if (r < nr && g >= lo && g < hi) at(q, r * (int)rs + (int)pos0 + g) = v[e];
    }
  }
};

// One workgroup per block of acr::kLines lines: the block's lines are
// loaded (all of a thread's loads issued at once) into the slab, every level
// of four lines per wavefront runs from the slab, and the block is stored.
// One kernel per table wavelet (its compiled step shapes), chosen on the host.
template <int TF, class W>
__global__ void __launch_bounds__(acr::kThreads) k_across(acr::Args A) {
  const long long blk = (long long)xcd_block(blockIdx.x, gridDim.x);
  if (blk >= A.nblocks) return;  // whole workgroups only: before any barrier
  const int tid = (int)threadIdx.x;
  long long c0, oi, oo;
  acr::block_origin(A, blk, c0, oi, oo);
  {
    float v[acr::kPer];
    acr::fetch(A, c0, oi, tid, v);
    acr::put(A, lfx_acr_slab, tid, v);
  }
  __syncthreads();
  acr::run_lines<RowLds, W, TF>(A, lfx_acr_slab + (tid >> 6) * 4 * acr::kMaxN);
  __syncthreads();
  acr::store_block(A, lfx_acr_slab, c0, oo, tid);
}

template <int TF, class W>
static hipError_t launch_across_w(const acr::Args& A, hipStream_t st) {
  static std::mutex mu;
  static std::set<int> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const size_t lds = sizeof(float) * acr::kSlabFloats;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!done.count(dev)) {
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_across<TF, W>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      done.insert(dev);
    }
  }
  hipLaunchKernelGGL((k_across<TF, W>), dim3((unsigned)A.nblocks), dim3(acr::kThreads), lds, st, A);
  return hipGetLastError();
}

template <int TF>
hipError_t launch_across_tf(const acr::Args& A, hipStream_t st) {
  if (A.nblocks <= 0) return hipSuccess;
  if (A.nblocks > 0x7fffffffll || A.n > acr::kMaxN || A.npass < 1 || A.npass > acr::kMaxPasses)
    return hipErrorInvalidValue;
  switch (A.wavelet) {
#define LFX_WAVELET(nm, idv, w_) \
  case idv:                      \
    return launch_across_w<TF, wt::WL_##nm>(A, st);
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
}

}  // namespace lfx
