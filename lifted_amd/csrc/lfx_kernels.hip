// gfx950 kernels of the lifting DWT.
//
// k_line_fwd / k_line_rev: one workgroup owns one whole 1-D line, stages it in
// LDS and runs every level there (the reference's per-line multi-level drivers,
// src/lifted-drivers-inl.hpp:75-274; single-level when L == 1, :144-179 /
// :276-311). The line is read once and written once; intermediate bands never
// touch HBM.
//   forward-stepping (Forward, InverseAdjoint): deinterleave on load, all steps,
//   write d_l to out[ns_l, ns_l + nd_l), deinterleave s_l in LDS for the next
//   level, finally write s_L to out[0, ns_L)  ->  [s_L | d_L | ... | d_1].
//   reverse-stepping (Inverse, ForwardAdjoint): read s_L, then per level read
//   d_l, run the steps, interleave into the next level's s; the finest level
//   is interleaved straight into the output.
// LDS: a band buffer of n elements plus one of ceil(n/2) (levels alternate).
#include <hip/hip_runtime.h>

#include "lfx_device.hpp"
#include "lfx_kernels.hpp"

namespace lfx {

using dev::Band;
using dev::bf16_t;
using dev::f16_t;
using dev::load_as;
using dev::store_as;

__device__ __forceinline__ void line_offsets(const LineMap& m, long long line, long long& so, long long& dof) {
  so = 0;
  dof = 0;
  for (int k = m.nod - 1; k >= 0; --k) {
    const long long q = line % m.shape[k];
    line /= m.shape[k];
    so += q * m.sst[k];
    dof += q * m.dstr[k];
  }
}

__device__ __forceinline__ long long level_len(long long n, int l) { return (n + (1ll << l) - 1) >> l; }

template <class C, class S>
__global__ void __launch_bounds__(kLineThreads) k_line_fwd(const DProgram<C>* prog, int bc, long long n, int L,
                                                           const S* src, S* dst, LineMap lm, long long line0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* P = reinterpret_cast<C*>(smem_raw);
  C* Q = P + n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  long long so, dof;
  line_offsets(lm, line0 + blockIdx.x, so, dof);
  const S* in = src + so;
  S* out = dst + dof;
  const long long ns0 = n - n / 2;
  for (long long i = tid; i < n; i += nthr) P[(i & 1) ? ns0 + (i >> 1) : (i >> 1)] = load_as<S, C>(in, i * lm.s_ax);
  __syncthreads();
  C* cur = P;
  C* oth = Q;
  long long m = n;
  for (int l = 0; l < L; ++l) {
    const long long nd = m / 2, ns = m - nd;
    const Band<C> s{cur, 0, (int)ns}, d{cur + ns, 0, (int)nd};
    dev::level_steps<C>(prog, bc, s, d, ns, nd, tid, nthr);
    for (long long i = tid; i < nd; i += nthr) store_as<S, C>(out, (ns + i) * lm.d_ax, cur[ns + i]);
    if (l + 1 < L) {
      const long long ns2 = ns - ns / 2;
      for (long long i = tid; i < ns; i += nthr) oth[(i & 1) ? ns2 + (i >> 1) : (i >> 1)] = cur[i];
      __syncthreads();
      C* t = cur;
      cur = oth;
      oth = t;
    }
    m = ns;
  }
  for (long long i = tid; i < m; i += nthr) store_as<S, C>(out, i * lm.d_ax, cur[i]);
}

template <class C, class S>
__global__ void __launch_bounds__(kLineThreads) k_line_rev(const DProgram<C>* prog, int bc, long long n, int L,
                                                           const S* src, S* dst, LineMap lm, long long line0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* buf[2];
  buf[0] = reinterpret_cast<C*>(smem_raw);
  buf[1] = buf[0] + n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  long long so, dof;
  line_offsets(lm, line0 + blockIdx.x, so, dof);
  const S* in = src + so;
  S* out = dst + dof;
  C* cur = buf[(L - 1) & 1];
  {
    const long long mL = level_len(n, L - 1);
    const long long nsL = mL - mL / 2;
    for (long long i = tid; i < nsL; i += nthr) cur[i] = load_as<S, C>(in, i * lm.s_ax);
  }
  for (int l = L - 1; l >= 0; --l) {
    const long long m = level_len(n, l);
    const long long nd = m / 2, ns = m - nd;
    for (long long i = tid; i < nd; i += nthr) cur[ns + i] = load_as<S, C>(in, (ns + i) * lm.s_ax);
    __syncthreads();
    const Band<C> s{cur, 0, (int)ns}, d{cur + ns, 0, (int)nd};
    dev::level_steps<C>(prog, bc, s, d, ns, nd, tid, nthr);
    if (l > 0) {
      C* nxt = buf[(l - 1) & 1];
      for (long long j = tid; j < m; j += nthr) nxt[j] = (j & 1) ? cur[ns + (j >> 1)] : cur[j >> 1];
      __syncthreads();
      cur = nxt;
    } else {
      for (long long j = tid; j < m; j += nthr)
        store_as<S, C>(out, j * lm.d_ax, (j & 1) ? cur[ns + (j >> 1)] : cur[j >> 1]);
    }
  }
}

// Strided elementwise copy over an ND box (used by the level-by-level inverse,
// which copies in -> out before transforming in place, drivers-inl.hpp:542-567).
template <class S>
__global__ void k_copy_nd(const S* src, S* dst, CopyMap cm, long long total) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    long long r = e, so = 0, dof = 0;
    for (int k = cm.nd - 1; k >= 0; --k) {
      const long long q = r % cm.shape[k];
      r /= cm.shape[k];
      so += q * cm.sst[k];
      dof += q * cm.dstr[k];
    }
    dst[dof] = src[so];
  }
}

// ---------------------------------------------------------------------------
// launchers

template <class C, class S>
static hipError_t launch_line_t(bool fwd, const void* prog, int bc, long long n, int L, const void* src, void* dst,
                                const LineMap& lm, long long nlines, hipStream_t st) {
  const size_t lds = line_lds_bytes(n, sizeof(C));
  auto kf = fwd ? k_line_fwd<C, S> : k_line_rev<C, S>;
  static bool attr_set[2] = {false, false};
  if (!attr_set[fwd ? 1 : 0]) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kf),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    if (e != hipSuccess) return e;
    attr_set[fwd ? 1 : 0] = true;
  }
  const long long kMaxGrid = 1ll << 30;
  for (long long l0 = 0; l0 < nlines; l0 += kMaxGrid) {
    const long long cnt = nlines - l0 < kMaxGrid ? nlines - l0 : kMaxGrid;
    hipLaunchKernelGGL(kf, dim3((unsigned)cnt), dim3(kLineThreads), lds, st,
                       reinterpret_cast<const DProgram<C>*>(prog), bc, n, L, reinterpret_cast<const S*>(src),
                       reinterpret_cast<S*>(dst), lm, l0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_line(int storage, bool fwd, const void* prog, int bc, long long n, int L, const void* src,
                       void* dst, const LineMap& lm, long long nlines, hipStream_t st) {
  switch (storage) {
    case kF32: return launch_line_t<float, float>(fwd, prog, bc, n, L, src, dst, lm, nlines, st);
    case kF64: return launch_line_t<double, double>(fwd, prog, bc, n, L, src, dst, lm, nlines, st);
    case kBF16: return launch_line_t<float, bf16_t>(fwd, prog, bc, n, L, src, dst, lm, nlines, st);
    case kF16: return launch_line_t<float, f16_t>(fwd, prog, bc, n, L, src, dst, lm, nlines, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_copy(int storage, const void* src, void* dst, const CopyMap& cm, long long total,
                       hipStream_t st) {
  if (total <= 0) return hipSuccess;
  const int thr = 256;
  long long blocks = (total + thr - 1) / thr;
  if (blocks > 65536) blocks = 65536;
  const size_t es = storage_bytes(storage);
  if (es == 8)
    hipLaunchKernelGGL(k_copy_nd<double>, dim3((unsigned)blocks), dim3(thr), 0, st,
                       reinterpret_cast<const double*>(src), reinterpret_cast<double*>(dst), cm, total);
  else if (es == 4)
    hipLaunchKernelGGL(k_copy_nd<float>, dim3((unsigned)blocks), dim3(thr), 0, st,
                       reinterpret_cast<const float*>(src), reinterpret_cast<float*>(dst), cm, total);
  else
    hipLaunchKernelGGL(k_copy_nd<uint16_t>, dim3((unsigned)blocks), dim3(thr), 0, st,
                       reinterpret_cast<const uint16_t*>(src), reinterpret_cast<uint16_t*>(dst), cm, total);
  return hipGetLastError();
}

}  // namespace lfx
