// Across-line kernel of one transform type: lfx_across.hpp's blocks of lines
// along a strided axis. build.py compiles this file once per transform type
// (LFX_TF = 0..3), so that the table's instantiations compile in parallel.
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
// those accesses at most two-way bank-conflicted).
struct RowLds : reg::LaneGpuT<4> {
  using V = float;
  static __device__ __forceinline__ float& at(const float* p, int off) {
    const int a = (int)(p - lfx_acr_slab) + off;
    return lfx_acr_slab[acr::phys(a)];
  }
  // x[e] of lane l = p[r rs + g], g = pos0 + E l' + e, for row r = l / 16 < nr
  // and g in [0, n); 0 otherwise. Every read is issued unconditionally, at a
  // row and position clamped into the wave's rows, and the values outside are
  // zeroed afterwards, so that the reads are in flight together instead of
  // each waiting behind a branch.
  //
  // A chunk whose virtual start is a multiple of E (E <= 32 divides 32) lies
  // within one 32-float block of the slab, so its physical addresses are
  // consecutive: such chunks (every row's, whole, when the virtual start of
  // row 0 is a multiple of E and the rows are 512 apart) are read from one
  // base address with immediate offsets.
  static __device__ __forceinline__ bool whole(const float* q, long long pos0, long long lo, long long hi, int E,
                                               long long rs, int nr) {
    const long long v0 = (long long)(q - lfx_acr_slab) + pos0;
    return nr == 4 && rs % acr::kMaxN == 0 && lo <= 0 && hi >= 16ll * E && v0 % E == 0;
  }
  template <int E, class S>
  static __device__ __forceinline__ void ld_raw(const S* p, long long pos0, long long n, V (&t)[E], long long rs = 0,
                                                int nr = 1) {
    const int l = lane(), r = l >> 4, lr = l & 15;
    const float* q = reinterpret_cast<const float*>(p);
    if (whole(q, pos0, -pos0, n - pos0, E, rs, nr)) {
      const float* b = &at(q, r * (int)rs + (int)pos0 + E * lr);
#pragma unroll
      for (int e = 0; e < E; ++e) t[e] = b[e];
      return;
    }
    const int rc = r < nr ? r : nr - 1, nn = (int)n;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int g = (int)pos0 + E * lr + e;
      const int gh = g < nn ? g : nn - 1, gc = gh < 0 ? 0 : gh;
      t[e] = at(q, rc * (int)rs + gc);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int g = (int)pos0 + E * lr + e;
      if (!(r < nr && g >= 0 && g < nn)) t[e] = 0.0f;
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
    if (whole(q, pos0 + g0, lo - g0, hi - g0, E, rs, nr)) {
      float* b = &at(q, r * (int)rs + (int)(pos0 + g0) + E * lr);
#pragma unroll
      for (int e = 0; e < E; ++e) b[e] = v[e];
      return;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int g = (int)g0 + E * lr + e;
      if (r < nr && g >= lo && g < hi) at(q, r * (int)rs + (int)pos0 + g) = v[e];
    }
  }
};

// One workgroup per block of acr::kLines lines: the block's lines are
// loaded (all of a thread's loads issued at once) into the slab, every level
// of each group of four lines runs from the slab (acr::kGroups / acr::kWaves
// groups per wavefront, one after the other), and the block is stored.
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
  // the wave's groups of rows of the slab: a wave-uniform base (scalar
  // registers), so the policy's address and alignment arithmetic stays scalar
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int g = wave; g < acr::kGroups; g += acr::kWaves)
    acr::run_lines<RowLds, W, TF>(A, lfx_acr_slab + g * 4 * acr::kMaxN);
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

#define LFX_JOIN2(a, b) a##b
#define LFX_JOIN(a, b) LFX_JOIN2(a, b)
hipError_t LFX_JOIN(launch_across_, LFX_TF)(const acr::Args& A, hipStream_t st) { return launch_across_tf<LFX_TF>(A, st); }

}  // namespace lfx
