// One level of a level-by-level transform over two axes in one pass: the
// reference runs, per level, a single-level transform along each axis of the
// sub-box in turn (src/lifted-drivers-inl.hpp:488-610), reading and writing
// the whole sub-box each time. Here one wavefront owns a tile of the sub-box:
// its 64 lanes are 64 adjacent positions along the contiguous axis (so every
// load and store is coalesced), and each lane holds R consecutive positions
// of the other axis in registers. The register axis is transformed within
// each lane (reg::level, as the 1-D register tiles do), the lane axis across
// lanes (even lanes hold s, odd lanes d; stencil neighbours are lane shifts),
// in the order the axes were given. A tile's window includes a halo on each
// axis (sized on the host by the lfx_tiles.hpp rules, so host and device
// agree) and stores only the positions it owns, all exact; under the Zero
// boundary, windows over a band end run masked (values outside the band are
// zero and re-zeroed after every step, which is what skipping out-of-range
// terms computes).
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

// The lane axis of one register row: lane l holds fine position w0 + l (w0
// even), an s value on even lanes and a d value on odd lanes; one level of
// the wavelet's steps (transform TF) across lanes. The x value of band index
// i + off + k for the lane of band index i sits 2 (off + k) + 1 - 2 par lanes
// away (par: parity of the updated lanes). Masked: positions outside the
// band are re-zeroed after each step.
template <class P, class W, int TF>
__device__ __forceinline__ void lane_level(typename P::V& v, const float* cf_all, bool masked, long long w0,
                                           long long ns, long long nd) {
  using V = typename P::V;
  using L = typename W::list;
  reg::sfor<L::size>([&](auto I) {
    constexpr int i = decltype(I)::value;
    using ST = wt::Exec<L, TF, i>;
    const float* cf = cf_all + wt::CoefOff<L, TF, i>::value;
    if constexpr (ST::kind == kScale) {
      const float sm = cf[0], dm = cf[1];
      v = P::where([](int l) { return (l & 1) == 0; }, P::mul(v, sm), P::mul(v, dm));
    } else {
      constexpr int par = ST::y_is_d ? 1 : 0;
      V acc;
      if constexpr (ST::kind == kGeneric) {
        acc = v;
        reg::sfor<ST::n>([&](auto K) {
          constexpr int k = decltype(K)::value;
          acc = P::fma(cf[k], P::template from<2 * (ST::off + k) + 1 - 2 * par>(v), acc);
        });
      } else if constexpr (ST::kind == kRepeat) {
        V t = P::template from<2 * ST::off + 1 - 2 * par>(v);
        reg::sfor<ST::n - 1>([&](auto K) {
          constexpr int k = decltype(K)::value + 1;
          t = P::add(t, P::template from<2 * (ST::off + k) + 1 - 2 * par>(v));
        });
        acc = P::fma(cf[0], t, v);
      } else {
        const V x = P::template from<2 * ST::off + 1 - 2 * par>(v);
        acc = ST::sub ? P::sub(v, x) : P::add(v, x);
      }
      if (masked) {
        // updated lanes keep acc inside their band, 0 outside; others keep v
        v = P::where([&](int l) { return (l & 1) == par; },
                     P::where([&](int l) {
                       const long long b = (w0 + l) >> 1;
                       return b >= 0 && b < (par ? nd : ns);
                     }, acc, P::gen([](int) { return 0.0f; })),
                     v);
      } else {
        v = P::where([&](int l) { return (l & 1) == par; }, acc, v);
      }
    }
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
  V s[E], d[E];
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    const long long pr = r.w0 + q;
    const bool row_in = pr >= 0 && pr < r.m;
    const V v = P::gen([&](int l) {
      const long long pc = c.w0 + l;
      return (row_in && pc >= 0 && pc < c.m) ? load_any(A0.p, A0.st, o0 + pr * A0.sr + pc) : 0.0f;
    });
    if constexpr (q % 2 == 0) s[q / 2] = v;
    else d[q / 2] = v;
  });
  const bool mr = r.w0 < 0 || r.w0 + kRows > r.m;
  typename P::M ms{}, md{};
  if (mr) {
    ms = P::umask(E, r.w0 / 2, r.ns());
    md = P::umask(E, r.w0 / 2, r.nd());
  }
  reg::level<P, W, TF, E>(s, d, cf, mr, ms, md);
  const bool mc = c.w0 < 0 || c.w0 + 64 > c.m;
  // stores: row band index q = w0/2 + e. The s values (even lanes) and d
  // values (odd lanes) of a row are first gathered to lanes 0..31 and 32..63,
  // so each half of the wave stores one contiguous run: lane t holds column
  // band index j = w0/2 + t % 32 of the s (t < 32) or d (t >= 32) half.
  const long long rq0 = r.w0 / 2, cj0 = c.w0 / 2;
  const long long nsr = r.ns(), nsc = c.ns();
  auto store_row = [&](const V& v, long long q, bool srow, bool row_ok) {
    if (!row_ok) return;
    const V u = P::permute(v, [](int t) { return t < 32 ? 2 * t : 2 * (t - 32) + 1; });
    P::each(u, [&](int t, float val) {
      const long long j = cj0 + (t & 31);
      const bool scol = t < 32;
      if (j < c.lo() || j >= (scol ? c.hi_s() : c.hi_d())) return;
      if (srow && scol) {
        store_any(A2.p, A2.st, o2 + q * A2.sr + j, val);
      } else {
        const long long row = srow ? q : nsr + q, col = scol ? j : nsc + j;
        store_any(A1.p, A1.st, o1 + row * A1.sr + col, val);
      }
    });
  };
  // each row: its lane-axis level, then its stores (the row is then dead)
  reg::sfor<E>([&](auto K) {
    constexpr int e = decltype(K)::value;
    const long long q = rq0 + e;
    lane_level<P, W, TF>(s[e], cf, mc, c.w0, c.ns(), c.nd());
    store_row(s[e], q, true, q >= r.lo() && q < r.hi_s());
    lane_level<P, W, TF>(d[e], cf, mc, c.w0, c.ns(), c.nd());
    store_row(d[e], q, false, q >= r.lo() && q < r.hi_d());
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
  const long long rq0 = r.w0 / 2, cj0 = c.w0 / 2;
  const long long nsr = r.ns(), ndr = r.nd(), nsc = c.ns(), ndc = c.nd();
  // loads: lanes 0..31 read one contiguous run of the s half (column band
  // index j = w0/2 + t), lanes 32..63 one of the d half; one gather then puts
  // s_j on lane 2j - w0 and d_j on the lane after it
  auto interleave = [](int l) { return (l & 1) ? 32 + (l >> 1) : (l >> 1); };
  V s[E], d[E];
  reg::sfor<E>([&](auto K) {
    constexpr int e = decltype(K)::value;
    const long long q = rq0 + e;
    const bool s_in = q >= 0 && q < nsr, d_in = q >= 0 && q < ndr;
    s[e] = P::permute(P::gen([&](int t) {
      const long long j = cj0 + (t & 31);
      if (t < 32) return (s_in && j >= 0 && j < nsc) ? load_any(A2.p, A2.st, o2 + q * A2.sr + j) : 0.0f;
      return (s_in && j >= 0 && j < ndc) ? load_any(A0.p, A0.st, o0 + q * A0.sr + (nsc + j)) : 0.0f;
    }), interleave);
    d[e] = P::permute(P::gen([&](int t) {
      const long long j = cj0 + (t & 31);
      const bool ok = d_in && j >= 0 && j < (t < 32 ? nsc : ndc);
      const long long col = t < 32 ? j : nsc + j;
      return ok ? load_any(A0.p, A0.st, o0 + (nsr + q) * A0.sr + col) : 0.0f;
    }), interleave);
  });
  const bool mr = r.w0 < 0 || r.w0 + kRows > r.m;
  typename P::M ms{}, md{};
  if (mr) {
    ms = P::umask(E, rq0, nsr);
    md = P::umask(E, rq0, ndr);
  }
  reg::level<P, W, TF, E>(s, d, cf, mr, ms, md);
  const bool mc = c.w0 < 0 || c.w0 + 64 > c.m;
  const long long cb = c.b >= c.m ? c.m : c.b;
  const long long rb = r.b >= r.m ? r.m : r.b;
  reg::sfor<kRows>([&](auto R) {
    constexpr int q = decltype(R)::value;
    V v = (q % 2 == 0) ? s[q / 2] : d[q / 2];
    lane_level<P, W, TF>(v, cf, mc, c.w0, nsc, ndc);
    const long long pr = r.w0 + q;
    if (pr >= r.a && pr < rb)
      P::each(v, [&](int l, float val) {
        const long long pc = c.w0 + l;
        if (pc >= c.a && pc < cb) store_any(A1.p, A1.st, o1 + pr * A1.sr + pc, val);
      });
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
