// Kernel bodies of the lifting DWT: whole-line and tiled multi-level
// transforms of one line (or one tile of one line) by one workgroup.
//
//  * whole-line (line_fwd_body / line_rev_body): the workgroup owns a whole
//    1-D line and runs all its levels in LDS (the reference's per-line
//    multi-level drivers, src/lifted-drivers-inl.hpp:75-274, single level when
//    L == 1, :144-179 / :276-311). The line is read once and written once.
//      forward-stepping (Forward, InverseAdjoint): deinterleave on load, steps,
//      write d_l to out[ns_l, ns_l + nd_l), deinterleave s_l in LDS for the next
//      level, finally s_L -> out[0, ns_L):  [s_L | d_L | ... | d_1].
//      reverse-stepping (Inverse, ForwardAdjoint): read s_L, per level read d_l,
//      steps, interleave into the next s; the finest level is interleaved
//      straight into the output.
//
//  * tiled (tile_fwd_body / tile_rev_body): lines longer than one workgroup's
//    LDS are cut into tiles of T samples; a tile loads its window plus halo and
//    runs k levels in LDS, tracking which elements are final (lfx_tiles.hpp).
//    Forward tiles write d_1..d_k of their own range and s_k to a scratch band;
//    reverse tiles read s_k and d_k..d_1 and write their part of the finer band.
//
// The same source also compiles for the host (LFX_CPU_SIM) with one "lane",
// so the tiling logic can be checked against the oracle without a GPU.
#pragma once
#include "lfx_device.hpp"
#include "lfx_tiles.hpp"

namespace lfx {

using dev::Band;
using dev::load_as;
using dev::store_as;

// Global <-> LDS loops over a line with element stride `ax`: a unit-stride
// path with 32-bit indices, a strided path otherwise. f(i) gives the LDS
// slot of sample i (loads) or the value to store (stores).
template <class S, class C, class F>
__device__ __forceinline__ void gload(const S* in, long long ax, int lo, int hi, int tid, int nthr, F&& f) {
  if (ax == 1) {
    for (int i = lo + tid; i < hi; i += nthr) f(i, load_as<S, C>(in, i));
  } else {
    for (int i = lo + tid; i < hi; i += nthr) f(i, load_as<S, C>(in, (long long)i * ax));
  }
}
template <class S, class C, class F>
__device__ __forceinline__ void gstore(S* out, long long ax, int lo, int hi, int tid, int nthr, F&& f) {
  if (ax == 1) {
    for (int i = lo + tid; i < hi; i += nthr) store_as<S, C>(out, i, f(i));
  } else {
    for (int i = lo + tid; i < hi; i += nthr) store_as<S, C>(out, (long long)i * ax, f(i));
  }
}

// ---------------------------------------------------------------------------
// whole-line bodies

template <class C, class SI, class SO>
__device__ __forceinline__ void line_fwd_body(const DProgram<C>* prog, int bc, long long n, int L, const SI* in, long long i_ax,
                              SO* out, long long o_ax, C* P, C* Q) {
  const int tid = LFX_TID, nthr = LFX_NTHR;
  const int ns0 = (int)(n - n / 2);
  gload<SI, C>(in, i_ax, 0, (int)n, tid, nthr, [&](int i, C v) { P[(i & 1) ? ns0 + (i >> 1) : (i >> 1)] = v; });
  LFX_SYNC();
  C* cur = P;
  C* oth = Q;
  int m = (int)n;
  for (int l = 0; l < L; ++l) {
    const int nd = m / 2, ns = m - nd;
    const Band<C> s{cur, 0, ns}, d{cur + ns, 0, nd};
    dev::level_steps<C>(prog, bc, s, d, ns, nd, tid, nthr);
    gstore<SO, C>(out, o_ax, ns, m, tid, nthr, [&](int i) { return cur[i]; });
    if (l + 1 < L) {
      const int ns2 = ns - ns / 2;
      for (int i = tid; i < ns; i += nthr) oth[(i & 1) ? ns2 + (i >> 1) : (i >> 1)] = cur[i];
      LFX_SYNC();
      C* t = cur;
      cur = oth;
      oth = t;
    }
    m = ns;
  }
  gstore<SO, C>(out, o_ax, 0, m, tid, nthr, [&](int i) { return cur[i]; });
}

// `in` holds [s_L | d_L | ... | d_1] of a band of length n; s_L may come from a
// separate array `sin` (scratch of a coarser pass) -- pass sin == in otherwise.
template <class C, class SI, class SS, class SO>
__device__ __forceinline__ void line_rev_body(const DProgram<C>* prog, int bc, long long n, int L, const SI* in, long long i_ax,
                              const SS* sin, long long s_ax, SO* out, long long o_ax, C* P, C* Q) {
  const int tid = LFX_TID, nthr = LFX_NTHR;
  // selects, not an array of pointers: the compiler then keeps the LDS
  // address space (an indexed pointer array turns every access into a flat one)
  auto buf = [&](int r) { return (r & 1) ? Q : P; };
  C* cur = buf(L - 1);
  {
    const int mL = (int)band_len(n, L - 1);
    const int nsL = mL - mL / 2;
    gload<SS, C>(sin, s_ax, 0, nsL, tid, nthr, [&](int i, C v) { cur[i] = v; });
  }
  for (int l = L - 1; l >= 0; --l) {
    const int m = (int)band_len(n, l);
    const int nd = m / 2, ns = m - nd;
    gload<SI, C>(in, i_ax, ns, m, tid, nthr, [&](int i, C v) { cur[i] = v; });
    LFX_SYNC();
    const Band<C> s{cur, 0, ns}, d{cur + ns, 0, nd};
    dev::level_steps<C>(prog, bc, s, d, ns, nd, tid, nthr);
    if (l > 0) {
      C* nxt = buf(l - 1);
      for (int j = tid; j < m; j += nthr) nxt[j] = (j & 1) ? cur[ns + (j >> 1)] : cur[j >> 1];
      LFX_SYNC();
      cur = nxt;
    } else {
      gstore<SO, C>(out, o_ax, 0, m, tid, nthr, [&](int j) { return (j & 1) ? cur[ns + (j >> 1)] : cur[j >> 1]; });
    }
  }
}

// ---------------------------------------------------------------------------
// tiled bodies

// All steps of one level on tile windows, each step updating only the
// elements whose inputs are final; `r` tracks the valid ranges (global).
template <class C>
__device__ __forceinline__ void tile_level_steps(const DProgram<C>* __restrict__ prog, int bc, const Band<C>& s, const Band<C>& d,
                                 long long ns, long long nd, Ranges& r, int tid, int nthr) {
  const int nsteps = prog->nsteps;
  for (int k = 0; k < nsteps; ++k) {
    const DStep st = prog->steps[k];
    const C* cf = prog->coef + st.coef;
    if (st.kind == kScale) {
      dev::scale_step<C>(cf, s, d, tid, nthr);
    } else {
      step_shrink(st, ns, nd, r, bc);
      const long long ylo = st.y_is_d ? r.dlo : r.slo, yhi = st.y_is_d ? r.dhi : r.shi;
      dev::lift_step<C>(st, cf, bc, s, d, ns, nd, ylo, yhi, tid, nthr);
    }
    LFX_SYNC();
  }
}

// Forward tile owning [a, b) of the input band `in` (length m0; a and, unless
// b == m0, b multiples of 2^k) -> d_r to out[(ns_r + i) * o_ax] for r = 1..k,
// s_k to sout[i * so_ax].
template <class C, class SI, class SO, class SK>
__device__ __forceinline__ void tile_fwd_body(const DProgram<C>* prog, int bc, long long m0, int k, long long a,
                                              long long b, long long H, const SI* in, long long i_ax, SO* out,
                                              long long o_ax, SK* sout, long long so_ax, C* P, C* Q) {
  const int tid = LFX_TID, nthr = LFX_NTHR;
  const long long wa = lmax(0, a - H), wb = lmin(m0, b + H);
  long long g0 = wa / 2;
  int sw = (int)((wb + 1) / 2 - g0), dw = (int)(wb / 2 - g0);
  // wa is even: local sample i of the window is an even sample iff i is even
  gload<SI, C>(in + wa * i_ax, i_ax, 0, (int)(wb - wa), tid, nthr,
               [&](int i, C v) { P[(i & 1) ? sw + (i >> 1) : (i >> 1)] = v; });
  LFX_SYNC();
  Ranges r{g0, g0 + sw, g0, g0 + dw};
  C* cur = P;
  C* oth = Q;
  long long m = m0;
  for (int lv = 1; lv <= k; ++lv) {
    const long long nd = m / 2, ns = m - nd;
    const Band<C> s{cur, g0, sw}, d{cur + sw, g0, dw};
    tile_level_steps<C>(prog, bc, s, d, ns, nd, r, tid, nthr);
    const long long ad = a >> lv, bd = (b >= m0) ? nd : (b >> lv);
    {
      const C* dsrc = cur + sw + (ad - g0);
      gstore<SO, C>(out + (ns + ad) * o_ax, o_ax, 0, (int)(bd - ad), tid, nthr, [&](int i) { return dsrc[i]; });
    }
    if (lv < k) {
      const long long ng0 = g0 / 2;
      const int nsw = (int)((g0 + sw + 1) / 2 - ng0), ndw = (int)((g0 + sw) / 2 - ng0);
      // g0 is even here (window starts are multiples of 2^k)
      for (int i = tid; i < sw; i += nthr) oth[(i & 1) ? nsw + (i >> 1) : (i >> 1)] = cur[i];
      r = Ranges{(r.slo + 1) / 2, (r.shi + 1) / 2, r.slo / 2, r.shi / 2};
      LFX_SYNC();
      C* t = cur;
      cur = oth;
      oth = t;
      g0 = ng0;
      sw = nsw;
      dw = ndw;
    } else {
      const long long bs = (b >= m0) ? ns : (b >> lv);
      const C* ssrc = cur + (ad - g0);
      gstore<SK, C>(sout + ad * so_ax, so_ax, 0, (int)(bs - ad), tid, nthr, [&](int i) { return ssrc[i]; });
    }
    m = ns;
  }
}

// Reverse tile: s_k from sin[i * si_ax], d_r from in[(ns_r + i) * i_ax]
// -> owned range [a, b) of the output band (length m0) to out[q * o_ax].
template <class C, class SI, class SK, class SO>
__device__ __forceinline__ void tile_rev_body(const DProgram<C>* prog, int bc, long long m0, int k, long long a,
                                              long long b, long long Hk, const SI* in, long long i_ax, const SK* sin,
                                              long long si_ax, SO* out, long long o_ax, C* P, C* Q) {
  const int tid = LFX_TID, nthr = LFX_NTHR;
  // level r lives in buf(r): level 1 (largest window) in P
  auto buf = [&](int r) { return (r & 1) ? P : Q; };
  long long m = band_len(m0, k - 1);
  long long ns = m - m / 2, nd = m / 2;
  long long gg = lmax(0, (a >> k) - Hk);
  const long long ghi = ((b + (1ll << k) - 1) >> k) + Hk;
  int sw = (int)lmax(0, lmin(ghi, ns) - gg), dw = (int)lmax(0, lmin(ghi, nd) - gg);
  C* cur = buf(k);
  gload<SK, C>(sin + gg * si_ax, si_ax, 0, sw, tid, nthr, [&](int i, C v) { cur[i] = v; });
  {
    C* dd = cur + sw;
    gload<SI, C>(in + (ns + gg) * i_ax, i_ax, 0, dw, tid, nthr, [&](int i, C v) { dd[i] = v; });
  }
  Ranges r{gg, gg + sw, gg, gg + dw};
  for (int lv = k; lv >= 1; --lv) {
    LFX_SYNC();
    const Band<C> s{cur, gg, sw}, d{cur + sw, gg, dw};
    tile_level_steps<C>(prog, bc, s, d, ns, nd, r, tid, nthr);
    const long long vlo = lmax(2 * r.slo, 2 * r.dlo + 1) - 1;
    const long long vhi = lmin(2 * r.shi, 2 * r.dhi + 1);
    const long long wlen = lmin(2 * (long long)sw, 2 * (long long)dw + 1);
    const long long lo = lmax(vlo, 2 * gg), hi = lmin(vhi, 2 * gg + wlen);
    if (lv == 1) {
      // band index q = 2 gg + t: even t <- s, odd t <- d
      const int t0 = (int)(a - 2 * gg);
      gstore<SO, C>(out + a * o_ax, o_ax, 0, (int)(b - a), tid, nthr, [&](int i) {
        const int t = t0 + i;
        return (t & 1) ? cur[sw + (t >> 1)] : cur[t >> 1];
      });
      break;
    }
    C* nxt = buf(lv - 1);
    for (int q = tid; q < (int)wlen; q += nthr) nxt[q] = (q & 1) ? cur[sw + (q >> 1)] : cur[q >> 1];
    m = band_len(m0, lv - 2);
    ns = m - m / 2;
    nd = m / 2;
    gg = 2 * gg;
    sw = (int)wlen;
    dw = (int)lmax(0, lmin(gg + sw, nd) - gg);
    {
      C* dd = nxt + sw;
      gload<SI, C>(in + (ns + gg) * i_ax, i_ax, 0, dw, tid, nthr, [&](int i, C v) { dd[i] = v; });
    }
    r = Ranges{lo, hi, gg, gg + dw};
    cur = nxt;
  }
}

}  // namespace lfx
