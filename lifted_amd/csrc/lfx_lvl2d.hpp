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

// Rows of one array as a tile sees them: the byte address of the tile
// window's origin (row 0, column 0 of the window; it may lie before the
// array, whose elements outside the band are never accessed) and the row
// pitch in bytes. Every access is that wave-uniform origin plus a 32-bit byte
// offset (the pitch times a compile-time row, plus the lane's column), so the
// 64-bit index arithmetic runs once per tile and not once per access (the
// host keeps a slice's byte offsets within 31 bits). The storage type ST is a
// compile-time parameter: a run-time type switch around every access puts a
// branch between consecutive loads, and the compiler then waits for each load
// before issuing the next.
template <int ST>
struct Io {
  static constexpr int es = ST == kF32 ? 4 : 2;  // element bytes
  char* org;
  int rowb;  // bytes between rows
  int v;     // pairs are aligned (Arr::v)
};
template <int ST>
__device__ __forceinline__ Io<ST> make_io(const Arr& A, long long o, long long row0, long long col0) {
  Io<ST> x;
  x.v = A.v;
  x.org = static_cast<char*>(A.p) + (o + row0 * A.sr + col0) * (long long)Io<ST>::es;
  x.rowb = (int)(A.sr * Io<ST>::es);
  return x;
}
template <int ST>
__device__ __forceinline__ float ldx(const Io<ST>& a, unsigned off) {
  const char* q = a.org + off;
  if constexpr (ST == kF32) {
    return *reinterpret_cast<const float*>(q);
  } else if constexpr (ST == kBF16) {
    return __uint_as_float((uint32_t)*reinterpret_cast<const uint16_t*>(q) << 16);
  } else {
#ifndef LFX_CPU_SIM
    const uint16_t h = *reinterpret_cast<const uint16_t*>(q);
    _Float16 f;
    __builtin_memcpy(&f, &h, 2);
    return (float)f;
#else
    return 0.0f;
#endif
  }
}
template <int ST>
__device__ __forceinline__ void stx(const Io<ST>& a, unsigned off, float v) {
  char* q = a.org + off;
  if constexpr (ST == kF32) {
    *reinterpret_cast<float*>(q) = v;
  } else if constexpr (ST == kBF16) {
    dev::store_as<dev::bf16_t, float>(reinterpret_cast<dev::bf16_t*>(q), 0, v);
  } else {
#ifndef LFX_CPU_SIM
    dev::store_as<dev::f16_t, float>(reinterpret_cast<dev::f16_t*>(q), 0, v);
#endif
  }
}
#ifndef LFX_CPU_SIM
// two consecutive elements (pair-aligned) as one access
template <int ST>
__device__ __forceinline__ void ld2x(const Io<ST>& a, unsigned off, float& x0, float& x1) {
  const char* q = a.org + off;
  if constexpr (ST == kF32) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = *reinterpret_cast<const f2*>(q);
    x0 = v.x;
    x1 = v.y;
  } else if constexpr (ST == kBF16) {
    const uint32_t u = *reinterpret_cast<const uint32_t*>(q);
    x0 = __uint_as_float(u << 16);
    x1 = __uint_as_float(u & 0xffff0000u);
  } else {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 v = *reinterpret_cast<const h2*>(q);
    x0 = (float)v.x;
    x1 = (float)v.y;
  }
}
template <int ST>
__device__ __forceinline__ void st2x(const Io<ST>& a, unsigned off, float x0, float x1) {
  char* q = a.org + off;
  if constexpr (ST == kF32) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 v;
    v.x = x0;
    v.y = x1;
    *reinterpret_cast<f2*>(q) = v;
  } else if constexpr (ST == kBF16) {
    *reinterpret_cast<uint32_t*>(q) = (uint32_t)dev::to_bf16_bits(x0) | ((uint32_t)dev::to_bf16_bits(x1) << 16);
  } else {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    h2 v;
    v.x = dev::to_f16(x0);
    v.y = dev::to_f16(x1);
    *reinterpret_cast<h2*>(q) = v;
  }
}
#endif

// v clamped to [0, cap] (window-relative bounds of a range)
LFX_HD int wclamp(long long v, int cap) { return v < 0 ? 0 : (v > cap ? cap : (int)v); }

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
// in: a[0] (the sub-box); details to a[1] (rows s | d at [0, ns_r) /
// [ns_r, m_r), same for columns), the ss quadrant to a[2] at its (s row,
// s col). o0..o2: the slice's element offsets.
template <class P, class W, int TF, int ST>
__device__ __forceinline__ void fwd_tile(const DProgram<float>* prog, const Span& r, const Span& c, const Arr& A0,
                                         long long o0, const Arr& A1, long long o1, const Arr& A2, long long o2) {
  using V = typename P::V;
  constexpr int E = kRows / 2;
  constexpr int NJ = 64 * kLaneE;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  // X[q][k]: row r.w0 + q, fine column c.w0 + kLaneW lane + k; in-band rows
  // [rl, rh) and columns [cl, ch) of the window
  const int rl = wclamp(-r.w0, kRows), rh = wclamp(r.m - r.w0, kRows);
  const int cl = wclamp(-c.w0, kCols), ch = wclamp(c.m - c.w0, kCols);
  const bool cin = cl == 0 && ch == kCols;
  const Io<ST> in = make_io<ST>(A0, o0, r.w0, c.w0);
  V X[kRows][kLaneW];
  if (rl == 0 && rh == kRows && cin && in.v) {
    // the whole window in the band: every row one unconditional pair load,
    // issued back to back
    reg::sfor<kRows>([&](auto R) {
      constexpr int q = decltype(R)::value;
#ifndef LFX_CPU_SIM
      ld2x(in, (unsigned)(q * in.rowb + kLaneW * Io<ST>::es * (int)(threadIdx.x & 63)), X[q][0], X[q][1]);
#else
      reg::sfor<kLaneW>([&](auto K) {
        constexpr int k = decltype(K)::value;
        X[q][k] = P::gen([&](int l) { return ldx(in, (unsigned)(q * in.rowb + (kLaneW * l + k) * Io<ST>::es)); });
      });
#endif
    });
  } else {
    reg::sfor<kRows>([&](auto R) {
      constexpr int q = decltype(R)::value;
      const bool row_in = q >= rl && q < rh;
      reg::sfor<kLaneW>([&](auto K) {
        constexpr int k = decltype(K)::value;
        X[q][k] = P::gen([&](int l) {
          const int pc = kLaneW * l + k;
          return (row_in && pc >= cl && pc < ch) ? ldx(in, (unsigned)(q * in.rowb + pc * Io<ST>::es)) : 0.0f;
        });
      });
    });
  }
  const bool mr = rl > 0 || rh < kRows;
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
  // owned window-relative ranges: band rows [el, esh) / [el, edh), band
  // columns [jl, jsh) / [jl, jdh)
  const int el = wclamp(r.lo() - rq0, E), esh = wclamp(r.hi_s() - rq0, E), edh = wclamp(r.hi_d() - rq0, E);
  const int jl = wclamp(c.lo() - j0, NJ), jsh = wclamp(c.hi_s() - j0, NJ), jdh = wclamp(c.hi_d() - j0, NJ);
  const Io<ST> ss = make_io<ST>(A2, o2, rq0, j0), sd = make_io<ST>(A1, o1, rq0, nsc + j0);
  const Io<ST> ds = make_io<ST>(A1, o1, nsr + rq0, j0), dd = make_io<ST>(A1, o1, nsr + rq0, nsc + j0);
  auto put = [&](const Io<ST>& a, int e, int hi, const V (&x)[kLaneE]) {
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      P::each(x[k], [&](int l, float val) {
        const int jj = kLaneE * l + k;
        if (jj >= jl && jj < hi) stx(a, (unsigned)(e * a.rowb + jj * Io<ST>::es), val);
      });
    });
  };
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    constexpr int e = q / 2;
    V ls[kLaneE], ld[kLaneE];
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      ls[k] = X[q][2 * k];
      ld[k] = X[q][2 * k + 1];
    });
    reg::level<P, W, TF, kLaneE>(ls, ld, cf, mc, msc, mdc);
    if constexpr (q % 2 == 0) {
      if (e >= el && e < esh) {
        put(ss, e, jsh, ls);
        put(sd, e, jdh, ld);
      }
    } else {
      if (e >= el && e < edh) {
        put(ds, e, jsh, ls);
        put(dd, e, jdh, ld);
      }
    }
  });
}

// Reverse-stepping level of one tile (register axis first, then lane axis):
// the ss quadrant from a[2], the other quadrants from a[0] (the coefficient
// layout of the sub-box), the interleaved sub-box to a[1].
template <class P, class W, int TF, int ST>
__device__ __forceinline__ void rev_tile(const DProgram<float>* prog, const Span& r, const Span& c, const Arr& A0,
                                         long long o0, const Arr& A1, long long o1, const Arr& A2, long long o2) {
  using V = typename P::V;
  constexpr int E = kRows / 2;
  constexpr int NJ = 64 * kLaneE;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  const long long rq0 = r.w0 / 2, j0 = c.w0 / 2;
  const long long nsr = r.ns(), ndr = r.nd(), nsc = c.ns(), ndc = c.nd();
  // in-band window-relative ranges: s rows [el, esh), d rows [el, edh); s
  // columns [jl, jsh), d columns [jl, jdh)
  const int el = wclamp(-rq0, E), esh = wclamp(nsr - rq0, E), edh = wclamp(ndr - rq0, E);
  const int jl = wclamp(-j0, NJ), jsh = wclamp(nsc - j0, NJ), jdh = wclamp(ndc - j0, NJ);
  const Io<ST> ss = make_io<ST>(A2, o2, rq0, j0), sd = make_io<ST>(A0, o0, rq0, nsc + j0);
  const Io<ST> ds = make_io<ST>(A0, o0, nsr + rq0, j0), dd = make_io<ST>(A0, o0, nsr + rq0, nsc + j0);
  auto get = [&](const Io<ST>& a, int e, bool row_ok, int hi, V (&x)[kLaneE]) {
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      x[k] = P::gen([&](int l) {
        const int jj = kLaneE * l + k;
        return (row_ok && jj >= jl && jj < hi) ? ldx(a, (unsigned)(e * a.rowb + jj * Io<ST>::es)) : 0.0f;
      });
    });
  };
  // X[2e][2k] / X[2e][2k + 1]: s row rq0 + e, s / d column band index
  // j0 + kLaneE lane + k; X[2e + 1][...]: the d row
  V X[kRows][kLaneW];
  if (el == 0 && esh == E && edh == E && jl == 0 && jsh == NJ && jdh == NJ) {
    // the whole window in the bands: unconditional loads, issued back to back
    auto all = [&](const Io<ST>& a, int e, int k) {
      return P::gen([&](int l) { return ldx(a, (unsigned)(e * a.rowb + (kLaneE * l + k) * Io<ST>::es)); });
    };
    reg::sfor<E>([&](auto I) {
      constexpr int e = decltype(I)::value;
      reg::sfor<kLaneE>([&](auto K) {
        constexpr int k = decltype(K)::value;
        X[2 * e][2 * k] = all(ss, e, k);
        X[2 * e][2 * k + 1] = all(sd, e, k);
        X[2 * e + 1][2 * k] = all(ds, e, k);
        X[2 * e + 1][2 * k + 1] = all(dd, e, k);
      });
    });
  } else {
    reg::sfor<E>([&](auto I) {
      constexpr int e = decltype(I)::value;
      const bool s_in = e >= el && e < esh, d_in = e >= el && e < edh;
      V a[kLaneE], b[kLaneE];
      get(ss, e, s_in, jsh, a);
      get(sd, e, s_in, jdh, b);
      reg::sfor<kLaneE>([&](auto K) {
        constexpr int k = decltype(K)::value;
        X[2 * e][2 * k] = a[k];
        X[2 * e][2 * k + 1] = b[k];
      });
      get(ds, e, d_in, jsh, a);
      get(dd, e, d_in, jdh, b);
      reg::sfor<kLaneE>([&](auto K) {
        constexpr int k = decltype(K)::value;
        X[2 * e + 1][2 * k] = a[k];
        X[2 * e + 1][2 * k + 1] = b[k];
      });
    });
  }
  const bool mr = el > 0 || esh < E || edh < E;
  typename P::M ms{}, md{};
  if (mr) {
    ms = P::umask(E, rq0, nsr);
    md = P::umask(E, rq0, ndr);
  }
  rows_level<P, W, TF>(X, cf, mr, ms, md);
  // X[q] now holds fine row r.w0 + q
  const bool mc = jl > 0 || jsh < NJ || jdh < NJ;
  typename P::M msc{}, mdc{};
  if (mc) {
    msc = P::band_mask(kLaneE, j0, nsc);
    mdc = P::band_mask(kLaneE, j0, ndc);
  }
  // owned window-relative fine rows [ra, rb) and columns [ca, cb)
  const int ra = wclamp(r.a - r.w0, kRows), rb = wclamp((r.b >= r.m ? r.m : r.b) - r.w0, kRows);
  const int ca = wclamp(c.a - c.w0, kCols), cb = wclamp((c.b >= c.m ? c.m : c.b) - c.w0, kCols);
  const Io<ST> out = make_io<ST>(A1, o1, r.w0, c.w0);
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    V ls[kLaneE], ld[kLaneE];
    reg::sfor<kLaneE>([&](auto K) {
      constexpr int k = decltype(K)::value;
      ls[k] = X[q][2 * k];
      ld[k] = X[q][2 * k + 1];
    });
    reg::level<P, W, TF, kLaneE>(ls, ld, cf, mc, msc, mdc);
    if (q >= ra && q < rb) {
      V y[kLaneW];
      reg::sfor<kLaneE>([&](auto K) {
        constexpr int k = decltype(K)::value;
        y[2 * k] = ls[k];
        y[2 * k + 1] = ld[k];
      });
#ifndef LFX_CPU_SIM
      const int pc = kLaneW * (int)(threadIdx.x & 63);
      if (out.v && pc >= ca && pc + 1 < cb) {
        st2x(out, (unsigned)(q * out.rowb + pc * Io<ST>::es), y[0], y[1]);
        return;
      }
#endif
      reg::sfor<kLaneW>([&](auto K) {
        constexpr int k = decltype(K)::value;
        P::each(y[k], [&](int l, float val) {
          const int pc2 = kLaneW * l + k;
          if (pc2 >= ca && pc2 < cb) stx(out, (unsigned)(q * out.rowb + pc2 * Io<ST>::es), val);
        });
      });
    }
  });
}

// Tile w of a level (slice-major, then register-axis tiles, then lane-axis
// tiles), its slice's element offsets in the three arrays from the batch
// dims (fast 32-bit divisions: the launch has fewer than 2^31 tiles).
template <class P, class W, int TF, int ST>
__device__ __forceinline__ void level_tile(const Level& L, long long w) {
  const uint32_t wi = (uint32_t)w;
  uint32_t b = fdiv(L.f_per, wi);
  const uint32_t t = wi - b * (uint32_t)L.f_per.d;
  const uint32_t tr = fdiv(L.f_tc, t), tc = t - tr * (uint32_t)L.f_tc.d;
  const Span r = span_of(L.mr, L.Hr, L.Or, tr);
  const Span c = span_of(L.mc, L.Hc, L.Oc, tc);
  long long o[3] = {0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < kMaxDims; ++kk) {
    const int k = kMaxDims - 1 - kk;
    if (k < L.nb) {
      const uint32_t nb = fdiv(L.f_b[k], b);
      const long long q = (long long)(b - nb * (uint32_t)L.f_b[k].d);
      b = nb;
      o[0] += q * L.a[0].bs[k];
      o[1] += q * L.a[1].bs[k];
      o[2] += q * L.a[2].bs[k];
    }
  }
  const DProgram<float>* prog = static_cast<const DProgram<float>*>(L.prog);
  if constexpr (TF == kForward || TF == kInverseAdjoint)
    fwd_tile<P, W, TF, ST>(prog, r, c, L.a[0], o[0], L.a[1], o[1], L.a[2], o[2]);
  else
    rev_tile<P, W, TF, ST>(prog, r, c, L.a[0], o[0], L.a[1], o[1], L.a[2], o[2]);
}

}  // namespace l2d
}  // namespace lfx
