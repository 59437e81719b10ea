// Device-side lifting steps for gfx950, applied to bands resident in LDS.
//
// Semantics follow src/lifted-steps-inl.hpp of the reference, restated in a
// per-element *gather* form so every output element is produced by exactly one
// lane with no atomics:
//
//  normal (Forward / Inverse), element i of band y:
//     y[i] += sum_k c[k] * x[BC(i + o + k, nx)]        (terms with Zero-BC skipped)
//   The reference splits i into a front loop [0, n1), an interior [n1, n2) and a
//   back loop [n2, ny) (n1 = min(n_front, ny), n2 = nd - n_back; steps:409-410).
//   When n1 > n2 (bands shorter than the stencil) elements in [n2, n1) are
//   updated by both the front and the back loop; the gather form reproduces that.
//
//  adjoint (ForwardAdjoint / InverseAdjoint), element j of band y (the forward
//  step's x band):
//     y[j] += front folds, then the zero-BC transposed stencil
//             sum_k c_r[k] * x[j + o_r + k], then back folds,
//   where a fold is the transposed contribution of a virtual index m outside the
//   band that the BC maps onto j (steps:416-432, 466-484). Folds are applied in
//   increasing m, exactly in the reference's order.
//
// Accumulation: y = fma(c[k], x[k], y) for k = 0..N-1 (steps:97-102); Repeat
// sums its taps first and applies one fma (steps:116-122); Unit is y +/- x.
#pragma once
#include <cstdint>
#ifdef LFX_CPU_SIM
#include "lfx_cpu_sim.hpp"
#else
#include <hip/hip_runtime.h>
#define LFX_TID ((int)threadIdx.x)
#define LFX_NTHR ((int)blockDim.x)
#define LFX_SYNC __syncthreads
#endif

#include "lfx_program.hpp"

namespace lfx {
namespace dev {

template <class C>
__device__ __forceinline__ C fmac(C a, C b, C c);
template <>
__device__ __forceinline__ float fmac<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fmac<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Boundary map of the reference (src/lifted-common.hpp:719-779) on one band of
// length n: returns the mapped index, or -1 when the term is skipped (Zero BC).
// Periodic keeps the reference's unsigned modulo for negative indices.
//
// The reference iterates its Symmetric / Reflect mirrors until the index lands
// in the band. The iteration is the even extension of the band, so it has a
// closed form: Symmetric (half-sample mirrors at -1/2 and n - 1/2) has period
// 2n, Reflect (whole-sample mirrors at 0 and n - 1) period 2n - 2. The closed
// form has no loop, so a degenerate band (n = 0, or n = 1 under Reflect, where
// the reference never terminates and the host refuses the call) cannot spin a
// wave: it yields -1 (skip) or 0 instead.
__device__ __forceinline__ long long bc_map(int bc, long long m, long long n) {
  if (m >= 0 && m < n) return m;
  switch (bc) {
    case kZero:
      return -1;
    case kConstant:
      return m < 0 ? 0 : n - 1;
    case kPeriodic:
      return n > 0 ? (long long)((unsigned long long)m % (unsigned long long)n) : -1;
    case kSymmetric: {
      if (n <= 0) return -1;
      const long long p = 2 * n;
      long long r = m % p;
      if (r < 0) r += p;
      return r < n ? r : p - 1 - r;
    }
    default: {
      if (n <= 0) return -1;
      if (n == 1) return 0;
      const long long p = 2 * n - 2;
      long long r = m % p;
      if (r < 0) r += p;
      return r < n ? r : p - r;
    }
  }
}

// A band held in LDS: local index i <-> global band index g0 + i.
template <class C>
struct Band {
  C* p;
  long long g0;
  int len;
};

// Sum of the taps a Generic step reads at local x index xi (all in the window).
template <class C, int NT>
__device__ __forceinline__ C gen_interior(const C* __restrict__ x, int xi, const C* __restrict__ c, C acc) {
#pragma unroll
  for (int k = 0; k < NT; ++k) acc = fmac<C>(c[k], x[xi + k], acc);
  return acc;
}
template <class C, int NT>
__device__ __forceinline__ C rep_interior(const C* __restrict__ x, int xi, C v, C acc) {
  C t = x[xi];
#pragma unroll
  for (int k = 1; k < NT; ++k) t += x[xi + k];
  return fmac<C>(v, t, acc);
}

template <class C>
__device__ __forceinline__ C interior(const DStep& st, const C* __restrict__ cf, const C* __restrict__ x, int xi,
                                      C acc) {
  if (st.kind == kUnit) return st.sub ? acc - x[xi] : acc + x[xi];
  if (st.kind == kRepeat) {
    const C v = cf[0];
    switch (st.N) {
      case 1: return rep_interior<C, 1>(x, xi, v, acc);
      case 2: return rep_interior<C, 2>(x, xi, v, acc);
      case 3: return rep_interior<C, 3>(x, xi, v, acc);
      case 4: return rep_interior<C, 4>(x, xi, v, acc);
      default: {
        C t = x[xi];
        for (int k = 1; k < st.N; ++k) t += x[xi + k];
        return fmac<C>(v, t, acc);
      }
    }
  }
  switch (st.N) {
    case 1: return gen_interior<C, 1>(x, xi, cf, acc);
    case 2: return gen_interior<C, 2>(x, xi, cf, acc);
    case 3: return gen_interior<C, 3>(x, xi, cf, acc);
    case 4: return gen_interior<C, 4>(x, xi, cf, acc);
    case 5: return gen_interior<C, 5>(x, xi, cf, acc);
    case 6: return gen_interior<C, 6>(x, xi, cf, acc);
    case 7: return gen_interior<C, 7>(x, xi, cf, acc);
    case 8: return gen_interior<C, 8>(x, xi, cf, acc);
    case 9: return gen_interior<C, 9>(x, xi, cf, acc);
    default:
      for (int k = 0; k < st.N; ++k) acc = fmac<C>(cf[k], x[xi + k], acc);
      return acc;
  }
}

// Stencil at global x index base+k with the boundary map `bc` (normal steps) or
// only the in-range terms (bc = kZero: adjoint main loop and folds).
template <class C>
__device__ __forceinline__ C bc_stencil(const DStep& st, const C* __restrict__ cf, int bc, const Band<C>& x,
                                        long long nx, long long base, C acc) {
  if (st.kind == kUnit) {
    const long long mi = bc_map(bc, base, nx);
    if (mi >= 0) {
      const C v = x.p[mi - x.g0];
      acc = st.sub ? acc - v : acc + v;
    }
    return acc;
  }
  if (st.kind == kRepeat) {
    long long mi = bc_map(bc, base, nx);
    C t = mi >= 0 ? x.p[mi - x.g0] : C(0);
    for (int k = 1; k < st.N; ++k) {
      mi = bc_map(bc, base + k, nx);
      if (mi >= 0) t += x.p[mi - x.g0];
    }
    return fmac<C>(cf[0], t, acc);
  }
  for (int k = 0; k < st.N; ++k) {
    const long long mi = bc_map(bc, base + k, nx);
    if (mi >= 0) acc = fmac<C>(cf[k], x.p[mi - x.g0], acc);
  }
  return acc;
}

// Uniform scan of the fold sources of an adjoint step: virtual indices
// m in [m0, m1) whose BC image lies inside the band and differs from m.
// Returns the [lo, hi] range of their images (lo > hi when there are none).
__device__ __forceinline__ void fold_range(int bc, long long m0, long long m1, long long ny, long long& lo,
                                           long long& hi) {
  lo = 1;
  hi = 0;
  if (bc == kZero || ny <= 0) return;
  for (long long m = m0; m < m1; ++m) {
    const long long io = bc_map(bc, m, ny);
    if (io >= 0 && io != m) {
      if (hi < lo) { lo = io; hi = io; }
      else { lo = io < lo ? io : lo; hi = io > hi ? io : hi; }
    }
  }
}

template <class C>
__device__ __forceinline__ C apply_folds(const DStep& st, const C* __restrict__ cf, int bc, const Band<C>& x,
                                         long long nx, long long ny, long long m0, long long m1, long long j,
                                         C acc) {
  for (long long m = m0; m < m1; ++m) {
    const long long io = bc_map(bc, m, ny);
    if (io == j && io != m) acc = bc_stencil<C>(st, cf, kZero, x, nx, m + st.off, acc);
  }
  return acc;
}

// Interior elements [i0, i1) (local indices): every tap is inside the band and
// inside the window, no boundary map, no folds. One specialised loop per step
// shape; coefficients are uniform and stay in scalar registers.
template <class C, int KIND, int N, bool SUB>
__device__ __forceinline__ void interior_loop(C* __restrict__ y, const C* __restrict__ x, int i0, int i1, int xs,
                                              const C* __restrict__ cf, int tid, int nthr) {
  C c[N];
#pragma unroll
  for (int k = 0; k < N; ++k) c[k] = cf[KIND == kGeneric ? k : 0];
#pragma unroll 4
  for (int i = i0 + tid; i < i1; i += nthr) {
    const C* xp = x + (i + xs);
    C acc = y[i];
    if (KIND == kGeneric) {
#pragma unroll
      for (int k = 0; k < N; ++k) acc = fmac<C>(c[k], xp[k], acc);
    } else if (KIND == kRepeat) {
      C t = xp[0];
#pragma unroll
      for (int k = 1; k < N; ++k) t += xp[k];
      acc = fmac<C>(c[0], t, acc);
    } else {
      acc = SUB ? acc - xp[0] : acc + xp[0];
    }
    y[i] = acc;
  }
}

template <class C>
__device__ __forceinline__ void interior_dispatch(const DStep& st, C* __restrict__ y, const C* __restrict__ x, int i0, int i1,
                                  int xs, const C* __restrict__ cf, int tid, int nthr) {
  if (i1 <= i0) return;
  if (st.kind == kUnit) {
    if (st.sub) interior_loop<C, kUnit, 1, true>(y, x, i0, i1, xs, cf, tid, nthr);
    else interior_loop<C, kUnit, 1, false>(y, x, i0, i1, xs, cf, tid, nthr);
    return;
  }
  if (st.kind == kRepeat) {
    switch (st.N) {
      case 1: interior_loop<C, kRepeat, 1, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
      case 2: interior_loop<C, kRepeat, 2, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
      case 3: interior_loop<C, kRepeat, 3, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
      default:
        for (int i = i0 + tid; i < i1; i += nthr) y[i] = interior<C>(st, cf, x, i + xs, y[i]);
        return;
    }
  }
  switch (st.N) {
    case 1: interior_loop<C, kGeneric, 1, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 2: interior_loop<C, kGeneric, 2, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 3: interior_loop<C, kGeneric, 3, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 4: interior_loop<C, kGeneric, 4, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 5: interior_loop<C, kGeneric, 5, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 6: interior_loop<C, kGeneric, 6, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 8: interior_loop<C, kGeneric, 8, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    case 9: interior_loop<C, kGeneric, 9, false>(y, x, i0, i1, xs, cf, tid, nthr); return;
    default:
      for (int i = i0 + tid; i < i1; i += nthr) y[i] = interior<C>(st, cf, x, i + xs, y[i]);
      return;
  }
}

// One element with the full semantics: boundary map (normal) or zero-BC
// stencil plus folds (adjoint), and the front/back double update when the
// band is shorter than the stencil.
#ifndef LFX_EDGE_INLINE
#define LFX_EDGE_INLINE __forceinline__
#endif
template <class C>
__device__ LFX_EDGE_INLINE C edge_element(const DStep& st, const C* __restrict__ cf, int bc, const Band<C>& x,
                                          long long nx, long long ny, long long n1, long long n2, long long g,
                                          long long f0, long long f1, long long flo, long long fhi, long long b0,
                                          long long b1, long long blo, long long bhi, C acc) {
  if (!st.adjoint) {
    if (g >= n1 && g < n2) return bc_stencil<C>(st, cf, kZero, x, nx, g + st.off, acc);
    if (g < n1) acc = bc_stencil<C>(st, cf, bc, x, nx, g + st.off, acc);
    if (g >= n2) acc = bc_stencil<C>(st, cf, bc, x, nx, g + st.off, acc);
    return acc;
  }
  if (g >= flo && g <= fhi) acc = apply_folds<C>(st, cf, bc, x, nx, ny, f0, f1, g, acc);
  if (g >= n1 && g < n2) {
    acc = bc_stencil<C>(st, cf, kZero, x, nx, g + st.off, acc);
  } else {
    if (g < n1) acc = bc_stencil<C>(st, cf, kZero, x, nx, g + st.off, acc);
    if (g >= n2) acc = bc_stencil<C>(st, cf, kZero, x, nx, g + st.off, acc);
  }
  if (g >= blo && g <= bhi) acc = apply_folds<C>(st, cf, bc, x, nx, ny, b0, b1, g, acc);
  return acc;
}

// Apply one lifting step (not Scale) to the s/d bands held in LDS. ns/nd are
// the global band lengths of this level. Only elements whose global index lies
// in [ylo, yhi) are updated (the whole window when called with the full band).
// Elements in [n1, n2) away from any fold target take the specialised interior
// loop; the few others take the full per-element path. Caller synchronizes.
template <class C>
__device__ __forceinline__ void lift_step(const DStep& st, const C* __restrict__ cf, int bc, const Band<C>& s, const Band<C>& d,
                          long long ns, long long nd, long long ylo, long long yhi, int tid, int nthr) {
  const Band<C>& y = st.y_is_d ? d : s;
  const Band<C>& x = st.y_is_d ? s : d;
  const long long ny = st.y_is_d ? nd : ns;
  const long long nx = st.y_is_d ? ns : nd;
  const long long n1 = (long long)st.nf < ny ? (long long)st.nf : ny;
  const long long n2 = nd > (long long)st.nb ? nd - (long long)st.nb : 0;
  long long f0 = 0, f1 = 0, b0 = 0, b1 = 0, flo = 1, fhi = 0, blo = 1, bhi = 0;
  if (st.adjoint) {
    // fold sources: front m in [o, 0) (unit steps also stop at m = o + nx,
    // steps:697), back m in [nd, nx - o_r)
    f0 = st.fold_off;
    f1 = 0;
    if (st.kind == kUnit && f0 + nx < f1) f1 = f0 + nx;
    b0 = nd;
    b1 = nx - st.off;
    fold_range(bc, f0, f1, ny, flo, fhi);
    fold_range(bc, b0, b1, ny, blo, bhi);
  }
  // interior range (global), excluding fold targets at the band ends
  long long A = ylo > n1 ? ylo : n1;
  long long B = yhi < n2 ? yhi : n2;
  if (st.adjoint) {
    if (fhi >= flo) {
      if (fhi + 1 > A) A = fhi + 1;
      if (flo < B && flo >= A) A = B;  // targets inside the interior (Periodic wrap): no fast path
    }
    if (bhi >= blo) {
      if (blo < B) B = blo;
      if (bhi >= A && bhi < B) A = B;
    }
  }
  if (A > yhi) A = yhi;
  if (B > yhi) B = yhi;
  if (B < A) B = A;
  const int xs = (int)(y.g0 - x.g0 + st.off);
  interior_dispatch<C>(st, y.p, x.p, (int)(A - y.g0), (int)(B - y.g0), xs, cf, tid, nthr);
  // edges: [ylo, A) and [B, yhi)
  const long long e0 = A - ylo, e1 = yhi - B;
  for (long long e = tid; e < e0 + e1; e += nthr) {
    const long long g = e < e0 ? ylo + e : B + (e - e0);
    const long long i = g - y.g0;
    y.p[i] = edge_element<C>(st, cf, bc, x, nx, ny, n1, n2, g, f0, f1, flo, fhi, b0, b1, blo, bhi, y.p[i]);
  }
}

template <class C>
__device__ __forceinline__ void scale_step(const C* __restrict__ cf, const Band<C>& s, const Band<C>& d, int tid,
                                           int nthr) {
  const C sm = cf[0], dm = cf[1];
  for (int i = tid; i < s.len; i += nthr) s.p[i] = s.p[i] * sm;
  for (int i = tid; i < d.len; i += nthr) d.p[i] = d.p[i] * dm;
}

// All steps of one level on full bands held in LDS.
template <class C>
__device__ __forceinline__ void level_steps(const DProgram<C>* __restrict__ prog, int bc, const Band<C>& s, const Band<C>& d,
                            long long ns, long long nd, int tid, int nthr) {
  const int nsteps = prog->nsteps;
  for (int k = 0; k < nsteps; ++k) {
    const DStep st = prog->steps[k];
    const C* cf = prog->coef + st.coef;
    if (st.kind == kScale) {
      scale_step<C>(cf, s, d, tid, nthr);
    } else {
      const long long ylo = st.y_is_d ? d.g0 : s.g0;
      const long long yhi = ylo + (st.y_is_d ? d.len : s.len);
      lift_step<C>(st, cf, bc, s, d, ns, nd, ylo, yhi, tid, nthr);
    }
    LFX_SYNC();
  }
}

// ---- storage conversions (bf16 / fp16 storage, fp32 arithmetic) ----
struct bf16_t { uint16_t v; };
struct f16_t { uint16_t v; };

template <class S, class C>
__device__ __forceinline__ C load_as(const S* p, long long i) { return (C)p[i]; }
template <>
__device__ __forceinline__ float load_as<bf16_t, float>(const bf16_t* p, long long i) {
  return __uint_as_float(((uint32_t)p[i].v) << 16);
}
#ifndef LFX_CPU_SIM
template <>
__device__ __forceinline__ float load_as<f16_t, float>(const f16_t* p, long long i) {
  _Float16 h;
  __builtin_memcpy(&h, &p[i].v, 2);
  return (float)h;
}
#endif

template <class S, class C>
__device__ __forceinline__ void store_as(S* p, long long i, C v) { p[i] = (S)v; }
// A float32 value rounded to bfloat16: to nearest, ties to even; a NaN keeps
// its sign and leading payload bits and becomes quiet. On gfx950 this is one
// v_cvt_pk_bf16_f32, which gives the same 16 bits as the integer form below for
// every float32 pattern (checked on the device over all 2^32 of them).
__device__ __forceinline__ uint16_t to_bf16_bits(float v) {
#ifndef LFX_CPU_SIM
  __asm__("" : "+v"(v));
  return __builtin_bit_cast(uint16_t, (__bf16)v);
#else
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
#endif
}
template <>
__device__ __forceinline__ void store_as<bf16_t, float>(bf16_t* p, long long i, float v) {
  p[i].v = to_bf16_bits(v);
}
#ifndef LFX_CPU_SIM
// A float32 value rounded to float16. The value is first a float32 result:
// the empty asm keeps the compiler from folding the operation that produced
// it into the conversion (a multiply and a conversion become one v_fma_mix,
// which rounds once, to 16 bits, where the arithmetic defines two roundings).
__device__ __forceinline__ _Float16 to_f16(float v) {
  __asm__("" : "+v"(v));
  return (_Float16)v;
}
template <>
__device__ __forceinline__ void store_as<f16_t, float>(f16_t* p, long long i, float v) {
  const _Float16 h = to_f16(v);
  __builtin_memcpy(&p[i].v, &h, 2);
}
#endif

}  // namespace dev
}  // namespace lfx
