// One level of a level-by-level transform over two axes in one pass: the
// reference runs, per level, a single-level transform along each axis of the
// sub-box in turn (src/lifted-drivers-inl.hpp:488-610), reading and writing
// the whole sub-box each time. Here one wavefront owns a tile of the sub-box:
// its 64 lanes run along the contiguous axis, each lane holding kLaneW
// consecutive positions of it (so every load and store is coalesced, kLaneW
// elements per lane in one vector access where aligned), and kRows
// consecutive positions of the other axis in registers. The register axis is
// transformed within each lane (reg::level over its rows, one column at a
// time), the lane axis across lanes (reg::level over the lane chunks: each
// lane holds kLaneE s and kLaneE d values of a row, stencil neighbours in
// other lanes are lane shifts), in the order the axes were given. A tile's
// window includes a halo on each axis (sized on the host by the lfx_tiles.hpp
// rules, so host and device agree) and stores only the positions it owns, all
// exact; under the Zero boundary, windows over a band end run masked (values
// outside the band are zero and re-zeroed after every step, which is what
// skipping out-of-range terms computes).
//
// Forward-stepping levels (Forward, InverseAdjoint) read the sub-box and write
// the three detail quadrants to their places in the output and the ss quadrant
// (the next level's input) to a separate array; reverse-stepping levels read
// the ss quadrant from a separate array and the details from the input, and
// write the interleaved sub-box. So no level runs in place and every level
// reads its input once and writes its output once.
#pragma once
#include "lfx_lvl2d_types.hpp"
#include "lfx_reg.hpp"
#include "lfx_tiles.hpp"

namespace lfx {
namespace l2d {

__device__ __forceinline__ float load_any(const void* p, int st, long long i) {
  if (st == kBF16) return dev::load_as<dev::bf16_t, float>(static_cast<const dev::bf16_t*>(p), i);
#ifndef LFX_CPU_SIM
  if (st == kF16) return dev::load_as<dev::f16_t, float>(static_cast<const dev::f16_t*>(p), i);
#endif
  return static_cast<const float*>(p)[i];
}
__device__ __forceinline__ void store_any(void* p, int st, long long i, float v) {
  if (st == kBF16) {
    dev::store_as<dev::bf16_t, float>(static_cast<dev::bf16_t*>(p), i, v);
    return;
  }
#ifndef LFX_CPU_SIM
  if (st == kF16) {
    dev::store_as<dev::f16_t, float>(static_cast<dev::f16_t*>(p), i, v);
    return;
  }
#endif
  static_cast<float*>(p)[i] = v;
}

#ifndef LFX_CPU_SIM
// kLaneW (= 2) consecutive elements at i (even, array aligned) as one access
__device__ __forceinline__ void load_pair(const void* p, int st, long long i, float& a, float& b) {
  if (st == kF32) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = *reinterpret_cast<const f2*>(static_cast<const float*>(p) + i);
    a = v.x;
    b = v.y;
  } else if (st == kBF16) {
    const uint32_t u = *reinterpret_cast<const uint32_t*>(static_cast<const uint16_t*>(p) + i);
    a = __uint_as_float(u << 16);
    b = __uint_as_float(u & 0xffff0000u);
  } else {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 v = *reinterpret_cast<const h2*>(static_cast<const uint16_t*>(p) + i);
    a = (float)v.x;
    b = (float)v.y;
  }
}
__device__ __forceinline__ void store_pair(void* p, int st, long long i, float a, float b) {
  if (st == kF32) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<f2*>(static_cast<float*>(p) + i) = v;
  } else if (st == kBF16) {
    dev::bf16_t h[2];
    dev::store_as<dev::bf16_t, float>(h, 0, a);
    dev::store_as<dev::bf16_t, float>(h, 1, b);
    *reinterpret_cast<uint32_t*>(static_cast<uint16_t*>(p) + i) = (uint32_t)h[0].v | ((uint32_t)h[1].v << 16);
  } else {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    h2 v;
    v.x = (_Float16)a;
    v.y = (_Float16)b;
    *reinterpret_cast<h2*>(static_cast<uint16_t*>(p) + i) = v;
  }
}
#endif

// x[k] of lane l = fine position w0 + kLaneW l + k of the row at element
// offset b of A (0 when !row_in or outside [0, m)); `inside`: the whole window
// lies in [0, m) (wave-uniform).
template <class P>
__device__ __forceinline__ void ld_fine(const Arr& A, long long b, long long w0, long long m, bool row_in, bool inside,
                                        typename P::V (&x)[kLaneW]) {
#ifndef LFX_CPU_SIM
  static_assert(kLaneW == 2, "vector accesses take pairs");
  if (row_in && inside && A.v) {
    const long long i = b + w0 + kLaneW * (long long)(threadIdx.x & 63);
    load_pair(A.p, A.st, i, x[0], x[1]);
    return;
  }
#endif
  reg::sfor<kLaneW>([&](auto K) {
    constexpr int k = decltype(K)::value;
    x[k] = P::gen([&](int l) {
      const long long pc = w0 + (long long)kLaneW * l + k;
      return (row_in && pc >= 0 && pc < m) ? load_any(A.p, A.st, b + pc) : 0.0f;
    });
  });
}

// A.p[b + pc] = x[k] for the fine positions pc = w0 + kLaneW l + k in [lo, hi)
// (a lane whose pair lies wholly in the range stores it as one access).
template <class P>
__device__ __forceinline__ void st_fine(const Arr& A, long long b, long long w0, long long lo, long long hi,
                                        const typename P::V (&x)[kLaneW]) {
#ifndef LFX_CPU_SIM
  if (A.v) {
    const long long pc = w0 + kLaneW * (long long)(threadIdx.x & 63);
    if (pc >= lo && pc + 1 < hi) {
      store_pair(A.p, A.st, b + pc, x[0], x[1]);
    } else {
      if (pc >= lo && pc < hi) store_any(A.p, A.st, b + pc, x[0]);
      if (pc + 1 >= lo && pc + 1 < hi) store_any(A.p, A.st, b + pc + 1, x[1]);
    }
    return;
  }
#endif
  reg::sfor<kLaneW>([&](auto K) {
    constexpr int k = decltype(K)::value;
    P::each(x[k], [&](int l, float val) {
      const long long pc = w0 + (long long)kLaneW * l + k;
      if (pc >= lo && pc < hi) store_any(A.p, A.st, b + pc, val);
    });
  });
}

// x[k] of lane l = band index j0 + kLaneE l + k of the band at element offset
// b of A (0 when !ok or outside [0, n)).
template <class P>
__device__ __forceinline__ void ld_band(const Arr& A, long long b, long long j0, long long n, bool ok,
                                        typename P::V (&x)[kLaneE]) {
  const bool all = ok && j0 >= 0 && j0 + 64 * kLaneE <= n;
  reg::sfor<kLaneE>([&](auto K) {
    constexpr int k = decltype(K)::value;
    if (all) {
      x[k] = P::gen([&](int l) { return load_any(A.p, A.st, b + j0 + (long long)kLaneE * l + k); });
    } else {
      x[k] = P::gen([&](int l) {
        const long long j = j0 + (long long)kLaneE * l + k;
        return (ok && j >= 0 && j < n) ? load_any(A.p, A.st, b + j) : 0.0f;
      });
    }
  });
}

// A.p[b + j] = x[k] for the band indices j = j0 + kLaneE l + k in [lo, hi).
template <class P>
__device__ __forceinline__ void st_band(const Arr& A, long long b, long long j0, long long lo, long long hi,
                                        const typename P::V (&x)[kLaneE]) {
  const bool all = j0 >= lo && j0 + 64 * kLaneE <= hi;
  reg::sfor<kLaneE>([&](auto K) {
    constexpr int k = decltype(K)::value;
    P::each(x[k], [&](int l, float val) {
      const long long j = j0 + (long long)kLaneE * l + k;
      if (all || (j >= lo && j < hi)) store_any(A.p, A.st, b + j, val);
    });
  });
}

// One level of the wavelet's steps (transform TF) along the register axis of
// every column the lane holds: rows 2e / 2e + 1 of X are s[e] / d[e].
template <class P, class W, int TF>
__device__ __forceinline__ void rows_level(typename P::V (&X)[kRows][kLaneW], const float* cf, bool masked,
                                           const typename P::M& ms, const typename P::M& md) {
  using V = typename P::V;
  constexpr int E = kRows / 2;
  reg::sfor<kLaneW>([&](auto K) {
    constexpr int k = decltype(K)::value;
    V s[E], d[E];
    reg::sfor<E>([&](auto I) {
      constexpr int e = decltype(I)::value;
      s[e] = X[2 * e][k];
      d[e] = X[2 * e + 1][k];
    });
    reg::level<P, W, TF, E>(s, d, cf, masked, ms, md);
    reg::sfor<E>([&](auto I) {
      constexpr int e = decltype(I)::value;
      X[2 * e][k] = s[e];
      X[2 * e + 1][k] = d[e];
    });
  });
}

// Forward-stepping level of one tile (register axis first, then lane axis).
// in: a[0] at (row, col) offsets (ir, ic) of the sub-box; details to a[1]
// (rows s | d at [0, ns_r) / [ns_r, m_r), same for columns), the ss quadrant
// to a[2] at its (s row, s col).
template <class P, class W, int TF>
__device__ __forceinline__ void fwd_tile(const DProgram<float>* prog, const Span& r, const Span& c, const Arr& A0,
                                         long long o0, const Arr& A1, long long o1, const Arr& A2, long long o2) {
  using V = typename P::V;
  constexpr int E = kRows / 2;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  // X[q][k]: row r.w0 + q, fine column c.w0 + kLaneW lane + k
  V X[kRows][kLaneW];
  const bool cin = c.w0 >= 0 && c.w0 + kCols <= c.m;
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    const long long pr = r.w0 + q;
    ld_fine<P>(A0, o0 + pr * A0.sr, c.w0, c.m, pr >= 0 && pr < r.m, cin, X[q]);
  });
  const bool mr = r.w0 < 0 || r.w0 + kRows > r.m;
  typename P::M ms{}, md{};
  if (mr) {
    ms = P::umask(E, r.w0 / 2, r.ns());
    md = P::umask(E, r.w0 / 2, r.nd());
  }
  rows_level<P, W, TF>(X, cf, mr, ms, md);
  // lane axis, row by row (a row is stored, and dead, once transformed): lane
  // l holds column band indices j0 + kLaneE l + k of the s and d halves
  const bool mc = !cin;
  typename P::M msc{}, mdc{};
  const long long j0 = c.w0 / 2, rq0 = r.w0 / 2;
  if (mc) {
    msc = P::band_mask(kLaneE, j0, c.ns());
    mdc = P::band_mask(kLaneE, j0, c.nd());
  }
  const long long nsr = r.ns(), nsc = c.ns();
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    V ls[kLaneE], ld[kLaneE];
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      ls[k] = X[q][2 * k];
      ld[k] = X[q][2 * k + 1];
    });
    reg::level<P, W, TF, kLaneE>(ls, ld, cf, mc, msc, mdc);
    const long long rq = rq0 + q / 2;
    if constexpr (q % 2 == 0) {
      if (rq >= r.lo() && rq < r.hi_s()) {
        st_band<P>(A2, o2 + rq * A2.sr, j0, c.lo(), c.hi_s(), ls);
        st_band<P>(A1, o1 + rq * A1.sr + nsc, j0, c.lo(), c.hi_d(), ld);
      }
    } else {
      if (rq >= r.lo() && rq < r.hi_d()) {
        st_band<P>(A1, o1 + (nsr + rq) * A1.sr, j0, c.lo(), c.hi_s(), ls);
        st_band<P>(A1, o1 + (nsr + rq) * A1.sr + nsc, j0, c.lo(), c.hi_d(), ld);
      }
    }
  });
}

// Reverse-stepping level of one tile (register axis first, then lane axis):
// the ss quadrant from a[2], the other quadrants from a[0] (the coefficient
// layout of the sub-box), the interleaved sub-box to a[1].
template <class P, class W, int TF>
__device__ __forceinline__ void rev_tile(const DProgram<float>* prog, const Span& r, const Span& c, const Arr& A0,
                                         long long o0, const Arr& A1, long long o1, const Arr& A2, long long o2) {
  using V = typename P::V;
  constexpr int E = kRows / 2;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  const long long rq0 = r.w0 / 2, j0 = c.w0 / 2;
  const long long nsr = r.ns(), ndr = r.nd(), nsc = c.ns(), ndc = c.nd();
  // X[2e][2k] / X[2e][2k + 1]: s row rq0 + e, s / d column band index
  // j0 + kLaneE lane + k; X[2e + 1][...]: the d row
  V X[kRows][kLaneW];
  reg::sfor<E>([&](auto I) {
    constexpr int e = decltype(I)::value;
    const long long q = rq0 + e;
    const bool s_in = q >= 0 && q < nsr, d_in = q >= 0 && q < ndr;
    V a[kLaneE], b[kLaneE];
    ld_band<P>(A2, o2 + q * A2.sr, j0, nsc, s_in, a);
    ld_band<P>(A0, o0 + q * A0.sr + nsc, j0, ndc, s_in, b);
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      X[2 * e][2 * k] = a[k];
      X[2 * e][2 * k + 1] = b[k];
    });
    ld_band<P>(A0, o0 + (nsr + q) * A0.sr, j0, nsc, d_in, a);
    ld_band<P>(A0, o0 + (nsr + q) * A0.sr + nsc, j0, ndc, d_in, b);
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      X[2 * e + 1][2 * k] = a[k];
      X[2 * e + 1][2 * k + 1] = b[k];
    });
  });
  const bool mr = r.w0 < 0 || r.w0 + kRows > r.m;
  typename P::M ms{}, md{};
  if (mr) {
    ms = P::umask(E, rq0, nsr);
    md = P::umask(E, rq0, ndr);
  }
  rows_level<P, W, TF>(X, cf, mr, ms, md);
  // X[q] now holds fine row r.w0 + q
  const bool mc = c.w0 < 0 || c.w0 + kCols > c.m;
  typename P::M msc{}, mdc{};
  if (mc) {
    msc = P::band_mask(kLaneE, j0, nsc);
    mdc = P::band_mask(kLaneE, j0, ndc);
  }
  const long long cb = c.b >= c.m ? c.m : c.b;
  const long long rb = r.b >= r.m ? r.m : r.b;
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    V ls[kLaneE], ld[kLaneE];
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      ls[k] = X[q][2 * k];
      ld[k] = X[q][2 * k + 1];
    });
    reg::level<P, W, TF, kLaneE>(ls, ld, cf, mc, msc, mdc);
    const long long pr = r.w0 + q;
    if (pr >= r.a && pr < rb) {
      V y[kLaneW];
      reg::sfor<kLaneE>([&](auto K) {
        constexpr int k = decltype(K)::value;
        y[2 * k] = ls[k];
        y[2 * k + 1] = ld[k];
      });
      st_fine<P>(A1, o1 + pr * A1.sr, c.w0, c.a, cb, y);
    }
  });
}

// Offsets of batch slice b in the three arrays (the loop over dims is
// unrolled so that the launch arguments are only read at constant offsets and
// stay out of private memory).
__device__ __forceinline__ void slice_offsets(const Level& L, long long b, long long* o) {
  o[0] = o[1] = o[2] = 0;
#pragma unroll
  for (int kk = 0; kk < kMaxDims; ++kk) {
    const int k = kMaxDims - 1 - kk;
    if (k < L.nb) {
      const long long q = b % L.bshape[k];
      b /= L.bshape[k];
      o[0] += q * L.a[0].bs[k];
      o[1] += q * L.a[1].bs[k];
      o[2] += q * L.a[2].bs[k];
    }
  }
}

// Tile w of a level (slice-major, then register-axis tiles, then lane-axis tiles).
template <class P, class W, int TF>
__device__ __forceinline__ void level_tile(const Level& L, long long w) {
  const long long per = L.Tr * L.Tc;
  const long long b = w / per, t = w % per;
  const Span r = span_of(L.mr, L.Hr, L.Or, t / L.Tc);
  const Span c = span_of(L.mc, L.Hc, L.Oc, t % L.Tc);
  long long o[3];
  slice_offsets(L, b, o);
  const DProgram<float>* prog = static_cast<const DProgram<float>*>(L.prog);
  if constexpr (TF == kForward || TF == kInverseAdjoint)
    fwd_tile<P, W, TF>(prog, r, c, L.a[0], o[0], L.a[1], o[1], L.a[2], o[2]);
  else
    rev_tile<P, W, TF>(prog, r, c, L.a[0], o[0], L.a[1], o[1], L.a[2], o[2]);
}

}  // namespace l2d
}  // namespace lfx
