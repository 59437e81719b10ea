// gfx950 kernels of the lifting DWT: thin __global__ wrappers around the
// bodies in lfx_bodies.hpp (blockIdx -> line / tile, line base offsets from a
// LineMap3) and their launchers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <set>
#include <type_traits>
#include <utility>

#include "lfx_bodies.hpp"
#include "lfx_kernels.hpp"
#include "lfx_xcd.hpp"

namespace lfx {

using dev::bf16_t;
using dev::f16_t;

// ---------------------------------------------------------------------------
// ND wrappers: blockIdx -> (line, tile); a LineMap3 gives the line's base
// offsets in the three arrays.

__device__ __forceinline__ void line_bases(const LineMap3& m, long long line, long long* off) {
  off[0] = off[1] = off[2] = 0;
  for (int k = m.nod - 1; k >= 0; --k) {
    const long long q = line % m.shape[k];
    line /= m.shape[k];
    off[0] += q * m.str[0][k];
    off[1] += q * m.str[1][k];
    off[2] += q * m.str[2][k];
  }
}

template <class C>
__device__ __forceinline__ C* lds_base() {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  return reinterpret_cast<C*>(smem_raw);
}

static_assert(sizeof(DProgram<double>) <= kProgLdsBytes, "step program does not fit its LDS copy");

// Copy the step program into LDS once per workgroup: every lifting step reads
// its descriptor and coefficients, and from global memory each of those reads
// would be a dependent round trip (long while other kernels stream HBM).
template <class C>
__device__ __forceinline__ const DProgram<C>* stage_program(const DProgram<C>* g) {
  __shared__ __align__(16) uint32_t prog_lds[kProgLdsBytes / 4];
  constexpr int W = (int)(sizeof(DProgram<C>) / 4);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(g);
  for (int i = (int)threadIdx.x; i < W; i += (int)blockDim.x) prog_lds[i] = src[i];
  __syncthreads();
  return reinterpret_cast<const DProgram<C>*>(prog_lds);
}

// array 0: input, array 1: output, array 2: separate s_L source (reverse only)
template <class C, class SI, class SO>
__global__ void __launch_bounds__(kThreads) k_line_fwd(const DProgram<C>* prog, int bc, long long n, int L,
                                                       const SI* src, SO* dst, LineMap3 lm, long long cap0,
                                                       long long line0) {
  long long off[3];
  line_bases(lm, line0 + blockIdx.x, off);
  C* P = lds_base<C>();
  prog = stage_program<C>(prog);
  line_fwd_body<C, SI, SO>(prog, bc, n, L, src + off[0], lm.ax[0], dst + off[1], lm.ax[1], P, P + cap0);
}

template <class C, class SI, class SS, class SO>
__global__ void __launch_bounds__(kThreads) k_line_rev(const DProgram<C>* prog, int bc, long long n, int L,
                                                       const SI* src, const SS* ssrc, SO* dst, LineMap3 lm,
                                                       long long cap0, long long line0) {
  long long off[3];
  line_bases(lm, line0 + blockIdx.x, off);
  C* P = lds_base<C>();
  prog = stage_program<C>(prog);
  line_rev_body<C, SI, SS, SO>(prog, bc, n, L, src + off[0], lm.ax[0], ssrc + off[2], lm.ax[2], dst + off[1],
                               lm.ax[1], P, P + cap0);
}

// Whole lines whose two band buffers live in global scratch (gstride
// elements per workgroup) instead of LDS: the same bodies, one workgroup per
// line, its threads synchronized by workgroup barriers (which also order the
// workgroup's global-memory accesses). For lines longer than the LDS limit
// under the Periodic boundary, whose wrap reads the far end of every band.
template <class C, class S>
__global__ void __launch_bounds__(kThreads) k_gline_fwd(const DProgram<C>* prog, int bc, long long n, int L,
                                                        const S* src, S* dst, LineMap3 lm, C* gscr,
                                                        long long gstride, long long line0) {
  long long off[3];
  line_bases(lm, line0 + blockIdx.x, off);
  C* P = gscr + (long long)blockIdx.x * gstride;
  prog = stage_program<C>(prog);
  line_fwd_body<C, S, S>(prog, bc, n, L, src + off[0], lm.ax[0], dst + off[1], lm.ax[1], P, P + n);
}

template <class C, class S>
__global__ void __launch_bounds__(kThreads) k_gline_rev(const DProgram<C>* prog, int bc, long long n, int L,
                                                        const S* src, S* dst, LineMap3 lm, C* gscr,
                                                        long long gstride, long long line0) {
  long long off[3];
  line_bases(lm, line0 + blockIdx.x, off);
  C* P = gscr + (long long)blockIdx.x * gstride;
  prog = stage_program<C>(prog);
  line_rev_body<C, S, S, S>(prog, bc, n, L, src + off[0], lm.ax[0], src + off[2], lm.ax[2], dst + off[1], lm.ax[1],
                            P, P + n);
}

// forward tiles: array 0 = input band, 1 = output (d regions), 2 = s_k out
template <class C, class SI, class SO, class SK>
__global__ void __launch_bounds__(kThreads) k_tile_fwd(const DProgram<C>* prog, int bc, TileArgs ta, const SI* src,
                                                       SO* dst, SK* sdst, LineMap3 lm, long long cap0,
                                                       long long blk0) {
  const long long blk = blk0 + blockIdx.x;
  const long long line = blk / ta.ntiles, tile = blk % ta.ntiles;
  long long off[3];
  line_bases(lm, line, off);
  C* P = lds_base<C>();
  const long long a = tile * ta.T, b = lmin(a + ta.T, ta.m0);
  prog = stage_program<C>(prog);
  tile_fwd_body<C, SI, SO, SK>(prog, bc, ta.m0, ta.k, a, b, ta.H, src + off[0], lm.ax[0], dst + off[1], lm.ax[1],
                               sdst + off[2], lm.ax[2], P, P + cap0);
}

// reverse tiles: array 0 = input (d regions), 1 = output band, 2 = s_k in
template <class C, class SI, class SK, class SO>
__global__ void __launch_bounds__(kThreads) k_tile_rev(const DProgram<C>* prog, int bc, TileArgs ta, const SI* src,
                                                       const SK* ssrc, SO* dst, LineMap3 lm, long long cap0,
                                                       long long blk0) {
  const long long blk = blk0 + blockIdx.x;
  const long long line = blk / ta.ntiles, tile = blk % ta.ntiles;
  long long off[3];
  line_bases(lm, line, off);
  C* P = lds_base<C>();
  const long long a = tile * ta.T, b = lmin(a + ta.T, ta.m0);
  prog = stage_program<C>(prog);
  tile_rev_body<C, SI, SK, SO>(prog, bc, ta.m0, ta.k, a, b, ta.H, src + off[0], lm.ax[0], ssrc + off[2], lm.ax[2],
                               dst + off[1], lm.ax[1], P, P + cap0);
}

// Strided elementwise copy over an ND box (used by the level-by-level inverse,
// which copies in -> out before transforming in place, drivers-inl.hpp:542-567).
// Rows along the last dim of at least 256 elements: a workgroup per row
// (grid-stride over rows), the row's offsets computed once; otherwise one
// element per thread with the full index decomposition.
template <class S>
__global__ void k_copy_nd(const S* src, S* dst, CopyMap cm, long long total) {
  const long long inner = cm.shape[cm.nd - 1];
  if (inner >= 256) {
    const long long rows = total / inner;
    const long long si = cm.sst[cm.nd - 1], di = cm.dstr[cm.nd - 1];
    for (long long row = blockIdx.x; row < rows; row += gridDim.x) {
      long long r = row, so = 0, dof = 0;
      for (int k = cm.nd - 2; k >= 0; --k) {
        const long long q = r % cm.shape[k];
        r /= cm.shape[k];
        so += q * cm.sst[k];
        dof += q * cm.dstr[k];
      }
      for (long long i = threadIdx.x; i < inner; i += blockDim.x) dst[dof + i * di] = src[so + i * si];
    }
    return;
  }
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

// Box permutation (see PermMap): 64 x 64 tiles through LDS, so that reads run
// along the source's fast dim and writes along the destination's.
__device__ __forceinline__ void perm_batch(const PermMap& pm, long long b, long long& os, long long& od) {
  os = 0;
  od = 0;
  for (int k = pm.nb - 1; k >= 0; --k) {
    const long long q = b % pm.bshape[k];
    b /= pm.bshape[k];
    os += q * pm.bss[k];
    od += q * pm.bds[k];
  }
}

template <class SI, class SO, class C>
__global__ void __launch_bounds__(256) k_permute(const SI* __restrict__ src, SO* __restrict__ dst, PermMap pm) {
  __shared__ C tile[64][65];
  const long long i0 = (long long)blockIdx.x * 64, o0 = (long long)blockIdx.y * 64;
  const int tx = (int)(threadIdx.x & 63), ty = (int)(threadIdx.x >> 6);
  for (long long b = blockIdx.z; b < pm.nbatch; b += gridDim.z) {
    long long os, od;
    perm_batch(pm, b, os, od);
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
      const long long i = i0 + tx, o = o0 + r;
      if (i < pm.ni && o < pm.no) tile[r][tx] = dev::load_as<SI, C>(src, os + i * pm.si_i + o * pm.si_o);
    }
    __syncthreads();
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
      const long long i = i0 + r, o = o0 + tx;
      if (i < pm.ni && o < pm.no) dev::store_as<SO, C>(dst, od + i * pm.so_i + o * pm.so_o, tile[tx][r]);
    }
    __syncthreads();
  }
}

// Full tiles, both fast dims unit-stride and every row aligned to four
// elements (checked by the launcher): the same 64 x 64 tiles with 4-element
// vector loads along i and stores along o (a quarter of the memory
// instructions; float32 tiles run about 15 % faster than scalar ones on
// 512^3 boxes). 16-bit sides move 8 bytes per access and convert in registers.
template <class S>
struct alignas(4 * sizeof(S)) Vec4 {
  S v[4];
};
template <class SI, class SO>
__global__ void __launch_bounds__(256) k_permute_v4(const SI* __restrict__ src, SO* __restrict__ dst, PermMap pm) {
  __shared__ float tile[64][65];
  const long long i0 = (long long)blockIdx.x * 64, o0 = (long long)blockIdx.y * 64;
  const int q = (int)(threadIdx.x & 15), rr = (int)(threadIdx.x >> 4);  // 16 float4 per tile row, 16 rows per pass
  for (long long b = blockIdx.z; b < pm.nbatch; b += gridDim.z) {
    long long os, od;
    perm_batch(pm, b, os, od);
    const SI* s = src + os + i0 + 4 * q;
    SO* d = dst + od + o0 + 4 * q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int r = rr + 16 * k;  // o within the tile
      const Vec4<SI> v = *reinterpret_cast<const Vec4<SI>*>(s + (o0 + r) * pm.si_o);
#pragma unroll
      for (int j = 0; j < 4; ++j) tile[r][4 * q + j] = dev::load_as<SI, float>(v.v, j);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = rr + 16 * k;  // i within the tile
      Vec4<SO> v;
#pragma unroll
      for (int j = 0; j < 4; ++j) dev::store_as<SO, float>(v.v, j, tile[4 * q + j][c]);
      *reinterpret_cast<Vec4<SO>*>(d + (i0 + c) * pm.so_i) = v;
    }
    __syncthreads();
  }
}

// Whether k_permute_v4 applies: unit fast strides, whole 64 x 64 tiles, and
// every row start (tile rows, batch offsets, base pointers) aligned to four
// elements of its side.
static bool permute_v4_ok(const PermMap& pm, const void* src, size_t es_in, const void* dst, size_t es_out) {
  if (pm.same || pm.si_i != 1 || pm.so_o != 1 || pm.ni % 64 || pm.no % 64) return false;
  if (pm.si_o % 4 || pm.so_i % 4 || ((uintptr_t)src & (4 * es_in - 1)) || ((uintptr_t)dst & (4 * es_out - 1)))
    return false;
  for (int k = 0; k < pm.nb; ++k)
    if (pm.bss[k] % 4 || pm.bds[k] % 4) return false;
  return true;
}

// fi == fo: both sides run along dim i. Lines of at least 256 elements: a
// workgroup per line (grid-stride over lines), the batch offsets computed once
// per line and the threads striding along it. Shorter lines: one element per
// thread over the flattened (line, i) space.
template <class SI, class SO, class C>
__global__ void __launch_bounds__(256) k_convert(const SI* __restrict__ src, SO* __restrict__ dst, PermMap pm) {
  if (pm.ni >= 256) {
    for (long long b = blockIdx.x; b < pm.nbatch; b += gridDim.x) {
      long long os, od;
      perm_batch(pm, b, os, od);
      const SI* s = src + os;
      SO* d = dst + od;
      if (pm.si_i == 1 && pm.so_i == 1) {
        for (long long i = threadIdx.x; i < pm.ni; i += 256) dev::store_as<SO, C>(d, i, dev::load_as<SI, C>(s, i));
      } else {
        for (long long i = threadIdx.x; i < pm.ni; i += 256)
          dev::store_as<SO, C>(d, i * pm.so_i, dev::load_as<SI, C>(s, i * pm.si_i));
      }
    }
  } else {
    const long long total = pm.ni * pm.nbatch;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
      const long long b = e / pm.ni, i = e - b * pm.ni;
      long long os, od;
      perm_batch(pm, b, os, od);
      dev::store_as<SO, C>(dst, od + i * pm.so_i, dev::load_as<SI, C>(src, os + i * pm.si_i));
    }
  }
}

// fi == fo with unit stride along i on both sides, lines of at least 256
// elements, a multiple of 4 long, each starting on a 4-element boundary:
// a workgroup per line, 4 elements per thread and access.
template <class SI, class SO>
__global__ void __launch_bounds__(256) k_convert_v4(const SI* __restrict__ src, SO* __restrict__ dst, PermMap pm) {
  const long long n4 = pm.ni / 4;
  for (long long b = blockIdx.x; b < pm.nbatch; b += gridDim.x) {
    long long os, od;
    perm_batch(pm, b, os, od);
    const Vec4<SI>* s = reinterpret_cast<const Vec4<SI>*>(src + os);
    Vec4<SO>* d = reinterpret_cast<Vec4<SO>*>(dst + od);
    for (long long i = threadIdx.x; i < n4; i += 256) {
      const Vec4<SI> v = s[i];
      Vec4<SO> w;
#pragma unroll
      for (int j = 0; j < 4; ++j) dev::store_as<SO, float>(w.v, j, dev::load_as<SI, float>(v.v, j));
      d[i] = w;
    }
  }
}

static bool convert_v4_ok(const PermMap& pm, const void* src, size_t es_in, const void* dst, size_t es_out) {
  if (!pm.same || pm.si_i != 1 || pm.so_i != 1 || pm.ni < 256 || pm.ni % 4) return false;
  if (((uintptr_t)src & (4 * es_in - 1)) || ((uintptr_t)dst & (4 * es_out - 1))) return false;
  for (int k = 0; k < pm.nb; ++k)
    if (pm.bss[k] % 4 || pm.bds[k] % 4) return false;
  return true;
}

template <class SI, class SO, class C>
static hipError_t permute_t(const void* src, void* dst, const PermMap& pm, hipStream_t st) {
  const unsigned z = (unsigned)std::min<long long>(pm.nbatch, 65535);
  if (pm.same) {
    const long long work = pm.ni >= 256 ? pm.nbatch : (pm.ni * pm.nbatch + 255) / 256;
    const unsigned blocks = (unsigned)std::min<long long>(work, 1 << 20);
    hipLaunchKernelGGL((k_convert<SI, SO, C>), dim3(blocks), dim3(256), 0, st, reinterpret_cast<const SI*>(src),
                       reinterpret_cast<SO*>(dst), pm);
  } else {
    hipLaunchKernelGGL((k_permute<SI, SO, C>), dim3((unsigned)((pm.ni + 63) / 64), (unsigned)((pm.no + 63) / 64), z),
                       dim3(256), 0, st, reinterpret_cast<const SI*>(src), reinterpret_cast<SO*>(dst), pm);
  }
  return hipGetLastError();
}

hipError_t launch_permute(int sin, int sout, const void* src, void* dst, const PermMap& pm, hipStream_t st) {
  if (pm.ni <= 0 || pm.no <= 0 || pm.nbatch <= 0) return hipSuccess;
  if (sin == kF64 && sout == kF64) return permute_t<double, double, double>(src, dst, pm, st);
  const bool v4 = sin != kF64 && sout != kF64 && (sin == kF32 || sout == kF32);
  if (v4 && convert_v4_ok(pm, src, storage_bytes(sin), dst, storage_bytes(sout))) {
    const dim3 grid((unsigned)std::min<long long>(pm.nbatch, 1 << 20));
#define LFX_V4(SI_, SO_)                                                                                    \
  hipLaunchKernelGGL((k_convert_v4<SI_, SO_>), grid, dim3(256), 0, st, reinterpret_cast<const SI_*>(src), \
                     reinterpret_cast<SO_*>(dst), pm);                                                     \
  return hipGetLastError();
    if (sin == kF32 && sout == kF32) { LFX_V4(float, float) }
    if (sin == kBF16) { LFX_V4(bf16_t, float) }
    if (sin == kF16) { LFX_V4(f16_t, float) }
    if (sout == kBF16) { LFX_V4(float, bf16_t) }
    if (sout == kF16) { LFX_V4(float, f16_t) }
#undef LFX_V4
  }
  if (v4 && permute_v4_ok(pm, src, storage_bytes(sin), dst, storage_bytes(sout))) {
    const dim3 grid((unsigned)(pm.ni / 64), (unsigned)(pm.no / 64), (unsigned)std::min<long long>(pm.nbatch, 65535));
#define LFX_V4(SI_, SO_)                                                                                    \
  hipLaunchKernelGGL((k_permute_v4<SI_, SO_>), grid, dim3(256), 0, st, reinterpret_cast<const SI_*>(src), \
                     reinterpret_cast<SO_*>(dst), pm);                                                     \
  return hipGetLastError();
    if (sin == kF32 && sout == kF32) { LFX_V4(float, float) }
    if (sin == kBF16) { LFX_V4(bf16_t, float) }
    if (sin == kF16) { LFX_V4(f16_t, float) }
    if (sout == kBF16) { LFX_V4(float, bf16_t) }
    if (sout == kF16) { LFX_V4(float, f16_t) }
#undef LFX_V4
  }
  if (sin == kF32 && sout == kF32) return permute_t<float, float, float>(src, dst, pm, st);
  if (sin == kBF16 && sout == kF32) return permute_t<bf16_t, float, float>(src, dst, pm, st);
  if (sin == kF32 && sout == kBF16) return permute_t<float, bf16_t, float>(src, dst, pm, st);
  if (sin == kF16 && sout == kF32) return permute_t<f16_t, float, float>(src, dst, pm, st);
  if (sin == kF32 && sout == kF16) return permute_t<float, f16_t, float>(src, dst, pm, st);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// launchers

int block_threads() {
  static const int t = [] {
    const char* e = std::getenv("LFX_THREADS");
    const int v = e ? std::atoi(e) : 256;
    return (v == 64 || v == 128 || v == 256 || v == 512 || v == 1024) ? v : 256;
  }();
  return t;
}

template <class K>
static hipError_t ensure_lds(K kern) {
  // raise the dynamic-LDS limit once per kernel (and device)
  static std::mutex mu;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const void* key = reinterpret_cast<const void*>(kern);
  std::lock_guard<std::mutex> lk(mu);
  if (done.count({dev, key})) return hipSuccess;
  e = hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
  if (e == hipSuccess) done.insert({dev, key});
  return e;
}

template <class F>
static hipError_t grid_loop(long long nblocks, F&& launch) {
  const long long kMaxGrid = 1ll << 30;
  for (long long b0 = 0; b0 < nblocks; b0 += kMaxGrid) {
    const long long cnt = nblocks - b0 < kMaxGrid ? nblocks - b0 : kMaxGrid;
    launch(b0, (unsigned)cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// storage enum -> element type of the storage arrays
template <int S> struct StoreT;
template <> struct StoreT<kF32> { using T = float; using C = float; };
template <> struct StoreT<kF64> { using T = double; using C = double; };
template <> struct StoreT<kBF16> { using T = bf16_t; using C = float; };
template <> struct StoreT<kF16> { using T = f16_t; using C = float; };

// Array kinds of a launch: storage-typed (user arrays) or compute-typed (scratch).
template <int S, bool SCRATCH>
using ArrT = typename std::conditional<SCRATCH, typename StoreT<S>::C, typename StoreT<S>::T>::type;

template <int S, bool I_SCR, bool O_SCR>
static hipError_t line_fwd_t(const LineLaunch& ll, hipStream_t st) {
  using C = typename StoreT<S>::C;
  using SI = ArrT<S, I_SCR>;
  using SO = ArrT<S, O_SCR>;
  auto kf = k_line_fwd<C, SI, SO>;
  hipError_t e = ensure_lds(kf);
  if (e != hipSuccess) return e;
  const long long cap0 = ll.n;
  const size_t lds = line_lds_bytes(ll.n, sizeof(C));
  return grid_loop(ll.nlines, [&](long long b0, unsigned cnt) {
    hipLaunchKernelGGL(kf, dim3(cnt), dim3(block_threads()), lds, st, reinterpret_cast<const DProgram<C>*>(ll.prog), ll.bc,
                       ll.n, ll.L, reinterpret_cast<const SI*>(ll.a0), reinterpret_cast<SO*>(ll.a1), ll.lm, cap0, b0);
  });
}

template <int S, bool I_SCR, bool S_SCR, bool O_SCR>
static hipError_t line_rev_t(const LineLaunch& ll, hipStream_t st) {
  using C = typename StoreT<S>::C;
  using SI = ArrT<S, I_SCR>;
  using SS = ArrT<S, S_SCR>;
  using SO = ArrT<S, O_SCR>;
  auto kf = k_line_rev<C, SI, SS, SO>;
  hipError_t e = ensure_lds(kf);
  if (e != hipSuccess) return e;
  const long long cap0 = ll.n;
  const size_t lds = line_lds_bytes(ll.n, sizeof(C));
  return grid_loop(ll.nlines, [&](long long b0, unsigned cnt) {
    hipLaunchKernelGGL(kf, dim3(cnt), dim3(block_threads()), lds, st, reinterpret_cast<const DProgram<C>*>(ll.prog), ll.bc,
                       ll.n, ll.L, reinterpret_cast<const SI*>(ll.a0), reinterpret_cast<const SS*>(ll.a2),
                       reinterpret_cast<SO*>(ll.a1), ll.lm, cap0, b0);
  });
}

template <int S, bool I_SCR, bool O_SCR, bool K_SCR>
static hipError_t tile_fwd_t(const TileLaunch& tl, hipStream_t st) {
  using C = typename StoreT<S>::C;
  using SI = ArrT<S, I_SCR>;
  using SO = ArrT<S, O_SCR>;
  using SK = ArrT<S, K_SCR>;
  auto kf = k_tile_fwd<C, SI, SO, SK>;
  hipError_t e = ensure_lds(kf);
  if (e != hipSuccess) return e;
  const long long cap0 = tile_fwd_window(tl.ta);
  const size_t lds = tile_lds_bytes(tl.ta, true, sizeof(C));
  return grid_loop(tl.nlines * tl.ta.ntiles, [&](long long b0, unsigned cnt) {
    hipLaunchKernelGGL(kf, dim3(cnt), dim3(block_threads()), lds, st, reinterpret_cast<const DProgram<C>*>(tl.prog), tl.bc,
                       tl.ta, reinterpret_cast<const SI*>(tl.a0), reinterpret_cast<SO*>(tl.a1),
                       reinterpret_cast<SK*>(tl.a2), tl.lm, cap0, b0);
  });
}

template <int S, bool I_SCR, bool K_SCR, bool O_SCR>
static hipError_t tile_rev_t(const TileLaunch& tl, hipStream_t st) {
  using C = typename StoreT<S>::C;
  using SI = ArrT<S, I_SCR>;
  using SK = ArrT<S, K_SCR>;
  using SO = ArrT<S, O_SCR>;
  auto kf = k_tile_rev<C, SI, SK, SO>;
  hipError_t e = ensure_lds(kf);
  if (e != hipSuccess) return e;
  const long long cap0 = tile_rev_window(tl.ta);
  const size_t lds = tile_lds_bytes(tl.ta, false, sizeof(C));
  return grid_loop(tl.nlines * tl.ta.ntiles, [&](long long b0, unsigned cnt) {
    hipLaunchKernelGGL(kf, dim3(cnt), dim3(block_threads()), lds, st, reinterpret_cast<const DProgram<C>*>(tl.prog), tl.bc,
                       tl.ta, reinterpret_cast<const SI*>(tl.a0), reinterpret_cast<const SK*>(tl.a2),
                       reinterpret_cast<SO*>(tl.a1), tl.lm, cap0, b0);
  });
}

template <int S>
static hipError_t gline_t(const LineLaunch& ll, void* gscr, long long lines_per_launch, hipStream_t st) {
  using C = typename StoreT<S>::C;
  using T = typename StoreT<S>::T;
  const long long gstride = gline_elems(ll.n);
  for (long long b0 = 0; b0 < ll.nlines; b0 += lines_per_launch) {
    const unsigned cnt = (unsigned)std::min(lines_per_launch, ll.nlines - b0);
    if (ll.fwd)
      hipLaunchKernelGGL((k_gline_fwd<C, T>), dim3(cnt), dim3(kThreads), 0, st,
                         reinterpret_cast<const DProgram<C>*>(ll.prog), ll.bc, ll.n, ll.L,
                         reinterpret_cast<const T*>(ll.a0), reinterpret_cast<T*>(ll.a1), ll.lm,
                         reinterpret_cast<C*>(gscr), gstride, b0);
    else
      hipLaunchKernelGGL((k_gline_rev<C, T>), dim3(cnt), dim3(kThreads), 0, st,
                         reinterpret_cast<const DProgram<C>*>(ll.prog), ll.bc, ll.n, ll.L,
                         reinterpret_cast<const T*>(ll.a0), reinterpret_cast<T*>(ll.a1), ll.lm,
                         reinterpret_cast<C*>(gscr), gstride, b0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

#define LFX_DISPATCH_STORAGE(S, CALL) \
  switch (S) {                        \
    case kF32: return CALL(kF32);     \
    case kF64: return CALL(kF64);     \
    case kBF16: return CALL(kBF16);   \
    case kF16: return CALL(kF16);     \
  }                                   \
  return hipErrorInvalidValue;

hipError_t launch_line(const LineLaunch& ll, hipStream_t st) {
  const int code = (ll.scr[0] ? 1 : 0) | (ll.scr[1] ? 2 : 0) | (ll.scr[2] ? 4 : 0);
  if (ll.fwd) {
#define CALL_F(S)                                          \
  (code & 3) == 0 ? line_fwd_t<S, false, false>(ll, st)    \
  : (code & 3) == 1 ? line_fwd_t<S, true, false>(ll, st)   \
  : (code & 3) == 2 ? line_fwd_t<S, false, true>(ll, st)   \
                    : line_fwd_t<S, true, true>(ll, st)
    LFX_DISPATCH_STORAGE(ll.storage, CALL_F)
#undef CALL_F
  }
#define CALL_R(S)                                                   \
  code == 0 ? line_rev_t<S, false, false, false>(ll, st)            \
  : code == 4 ? line_rev_t<S, false, true, false>(ll, st)           \
  : code == 2 ? line_rev_t<S, false, false, true>(ll, st)           \
  : code == 6 ? line_rev_t<S, false, true, true>(ll, st)            \
              : hipErrorInvalidValue
  LFX_DISPATCH_STORAGE(ll.storage, CALL_R)
#undef CALL_R
}

hipError_t launch_gline(const LineLaunch& ll, void* gscr, long long lines_per_launch, hipStream_t st) {
  if (ll.nlines <= 0) return hipSuccess;
  if (lines_per_launch <= 0 || ll.scr[0] || ll.scr[1] || ll.scr[2]) return hipErrorInvalidValue;
#define CALL_G(S) gline_t<S>(ll, gscr, lines_per_launch, st)
  LFX_DISPATCH_STORAGE(ll.storage, CALL_G)
#undef CALL_G
}

hipError_t launch_tile(const TileLaunch& tl, hipStream_t st) {
  const int code = (tl.scr[0] ? 1 : 0) | (tl.scr[1] ? 2 : 0) | (tl.scr[2] ? 4 : 0);
  if (tl.fwd) {
    // input: user or scratch; output: user; s_k: user or scratch
#define CALL_F(S)                                                    \
  code == 0 ? tile_fwd_t<S, false, false, false>(tl, st)             \
  : code == 1 ? tile_fwd_t<S, true, false, false>(tl, st)            \
  : code == 4 ? tile_fwd_t<S, false, false, true>(tl, st)            \
  : code == 5 ? tile_fwd_t<S, true, false, true>(tl, st)             \
              : hipErrorInvalidValue
    LFX_DISPATCH_STORAGE(tl.storage, CALL_F)
#undef CALL_F
  }
  // input: user; output band: user or scratch; s_k in: user or scratch
#define CALL_R(S)                                                    \
  code == 0 ? tile_rev_t<S, false, false, false>(tl, st)             \
  : code == 2 ? tile_rev_t<S, false, false, true>(tl, st)            \
  : code == 4 ? tile_rev_t<S, false, true, false>(tl, st)            \
  : code == 6 ? tile_rev_t<S, false, true, true>(tl, st)             \
              : hipErrorInvalidValue
  LFX_DISPATCH_STORAGE(tl.storage, CALL_R)
#undef CALL_R
}

// ---------------------------------------------------------------------------
// batched stages: one workgroup per item; items of one stage may belong to
// different signals, wavelets and level counts (programs are per item).

template <class C, class S>
struct ArrSel {
  const S* in;
  S* out;
  C* scr;
};

// Dispatch the item's array kinds to the body with matching element types.
template <class C, class S, bool FWD>
__device__ __forceinline__ void batch_item(const BatchItem& it, int bc, const ArrSel<C, S>& a, C* P) {
  const DProgram<C>* prog = stage_program<C>(reinterpret_cast<const DProgram<C>*>(it.prog));
  const bool s0 = it.arr[0] == kArrScr, s1 = it.arr[1] == kArrScr, s2 = it.arr[2] == kArrScr;
  auto user_in = [&](int i) -> const S* { return (it.arr[i] == kArrIn ? a.in : a.out) + it.off[i]; };
  auto user_out = [&](int i) -> S* { return a.out + it.off[i]; };
  auto scr = [&](int i) -> C* { return a.scr + it.off[i]; };
  if (it.kind == kItemGLine) {
    // bands in the plan's scratch (a separate branch, so that the LDS items
    // keep LDS addressing)
    C* G = a.scr + it.a;
    C* GQ = G + it.m0;
    if (FWD) {
      if (s0) line_fwd_body<C, C, S>(prog, bc, it.m0, it.k, scr(0), 1, user_out(1), 1, G, GQ);
      else line_fwd_body<C, S, S>(prog, bc, it.m0, it.k, user_in(0), 1, user_out(1), 1, G, GQ);
    } else {
      if (s1) line_rev_body<C, S, S, C>(prog, bc, it.m0, it.k, user_in(0), 1, user_in(2), 1, scr(1), 1, G, GQ);
      else line_rev_body<C, S, S, S>(prog, bc, it.m0, it.k, user_in(0), 1, user_in(2), 1, user_out(1), 1, G, GQ);
    }
    return;
  }
  if (it.kind == kItemLine) {
    C* Q = P + it.m0;
    if (FWD) {
      if (s0) line_fwd_body<C, C, S>(prog, bc, it.m0, it.k, scr(0), 1, user_out(1), 1, P, Q);
      else line_fwd_body<C, S, S>(prog, bc, it.m0, it.k, user_in(0), 1, user_out(1), 1, P, Q);
    } else {
      if (s1) line_rev_body<C, S, S, C>(prog, bc, it.m0, it.k, user_in(0), 1, user_in(2), 1, scr(1), 1, P, Q);
      else line_rev_body<C, S, S, S>(prog, bc, it.m0, it.k, user_in(0), 1, user_in(2), 1, user_out(1), 1, P, Q);
    }
    return;
  }
  const TileArgs ta{it.m0, it.b - it.a, it.H, 0, it.k, 0};
  if (FWD) {
    C* Q = P + tile_fwd_window(ta);
    if (s0) {
      if (s2) tile_fwd_body<C, C, S, C>(prog, bc, it.m0, it.k, it.a, it.b, it.H, scr(0), 1, user_out(1), 1, scr(2), 1, P, Q);
      else tile_fwd_body<C, C, S, S>(prog, bc, it.m0, it.k, it.a, it.b, it.H, scr(0), 1, user_out(1), 1, user_out(2), 1, P, Q);
    } else {
      if (s2) tile_fwd_body<C, S, S, C>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, user_out(1), 1, scr(2), 1, P, Q);
      else tile_fwd_body<C, S, S, S>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, user_out(1), 1, user_out(2), 1, P, Q);
    }
  } else {
    C* Q = P + tile_rev_window(ta);
    if (s2) {
      if (s1) tile_rev_body<C, S, C, C>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, scr(2), 1, scr(1), 1, P, Q);
      else tile_rev_body<C, S, C, S>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, scr(2), 1, user_out(1), 1, P, Q);
    } else {
      if (s1) tile_rev_body<C, S, S, C>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, user_in(2), 1, scr(1), 1, P, Q);
      else tile_rev_body<C, S, S, S>(prog, bc, it.m0, it.k, it.a, it.b, it.H, user_in(0), 1, user_in(2), 1, user_out(1), 1, P, Q);
    }
  }
}

template <class C, class S, bool FWD>
__global__ void __launch_bounds__(kThreads) k_batch(const BatchItem* items, long long nitems, int bc, const S* in,
                                                    S* out, C* scr, long long item0) {
  const long long i = item0 + blockIdx.x;
  if (i >= nitems) return;
  const BatchItem it = items[i];
  const ArrSel<C, S> a{in, out, scr};
  batch_item<C, S, FWD>(it, bc, a, lds_base<C>());
}

template <int SI, bool FWD>
static hipError_t batch_t(int bc, const BatchItem* items, long long nitems, size_t lds, int threads, const void* in,
                          void* out, void* scr, hipStream_t st) {
  using C = typename StoreT<SI>::C;
  using S = typename StoreT<SI>::T;
  auto kf = k_batch<C, S, FWD>;
  hipError_t e = ensure_lds(kf);
  if (e != hipSuccess) return e;
  return grid_loop(nitems, [&](long long b0, unsigned cnt) {
    hipLaunchKernelGGL(kf, dim3(cnt), dim3(threads > 0 ? threads : block_threads()), lds, st, items, nitems, bc,
                       reinterpret_cast<const S*>(in), reinterpret_cast<S*>(out), reinterpret_cast<C*>(scr), b0);
  });
}

hipError_t launch_batch(int storage, bool fwd, int bc, const BatchItem* items, long long nitems, size_t lds_bytes,
                        int threads, const void* in, void* out, void* scr, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
#define CALL_B(S) fwd ? batch_t<S, true>(bc, items, nitems, lds_bytes, threads, in, out, scr, st) \
                      : batch_t<S, false>(bc, items, nitems, lds_bytes, threads, in, out, scr, st)
  LFX_DISPATCH_STORAGE(storage, CALL_B)
#undef CALL_B
}

// Streaming-copy references for the benchmark (no transform): one float4 per
// thread, and one 2048-float window per wavefront read as 32 coalesced rows of
// 64 floats and written back the same way, workgroups numbered contiguously
// per XCD (the register tiles' access pattern without the arithmetic).
__global__ void __launch_bounds__(256) k_copy_f4(const float4* __restrict__ a, float4* __restrict__ b, long long n4) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n4) b[i] = a[i];
}
__global__ void __launch_bounds__(64 * kRegWaves) k_copy_win(const float* __restrict__ a, float* __restrict__ b,
                                                             long long nwin) {
  const long long w = (long long)xcd_block(blockIdx.x, gridDim.x) * kRegWaves + (threadIdx.x >> 6);
  const int l = (int)(threadIdx.x & 63);
  if (w >= nwin) return;
  const float* p = a + w * 2048;
  float t[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) t[j] = p[j * 64 + l];
  float* q = b + w * 2048;
#pragma unroll
  for (int j = 0; j < 32; ++j) q[j * 64 + l] = t[j];
}

hipError_t launch_stream_copy(const void* src, void* dst, long long bytes, int pattern, hipStream_t st) {
  if (pattern == 0) {
    const long long n4 = bytes / 16;
    if (n4 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_f4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, (const float4*)src,
                       (float4*)dst, n4);
  } else {
    const long long nwin = bytes / (2048 * 4);
    if (nwin <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_win, dim3((unsigned)((nwin + kRegWaves - 1) / kRegWaves)), dim3(64 * kRegWaves), 0, st,
                       (const float*)src, (float*)dst, nwin);
  }
  return hipGetLastError();
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
