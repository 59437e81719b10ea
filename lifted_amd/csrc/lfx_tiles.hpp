// Tiling of long lines: a tile holds a window of a band plus halo and runs k
// levels of the transform in LDS. Every lifting step can only update the
// elements whose stencil lies inside the part of the other band that is
// already final ("valid"), so the valid range shrinks step by step. A window
// that touches a band end keeps that side as far as the boundary map resolves
// reads past the end to valid elements: Zero skips them, Constant reads the
// end element, Symmetric / Reflect read mirrored elements, which a short
// window may not hold (Periodic is never tiled).
//
// The same rules run on the host (to size halos) and on the device (to know
// which elements a step may update), so both agree by construction.
#pragma once
#include <cstdint>

#include "lfx_program.hpp"

#if defined(__HIPCC__) && !defined(LFX_CPU_SIM)
#define LFX_HD __host__ __device__ __forceinline__
#else
#define LFX_HD inline
#endif

namespace lfx {

LFX_HD long long lmin(long long a, long long b) { return a < b ? a : b; }
LFX_HD long long lmax(long long a, long long b) { return a > b ? a : b; }
// ceil(n / 2^l) = length of the band after l levels
LFX_HD long long band_len(long long n, int l) { return (n + (1ll << l) - 1) >> l; }

// New valid range [ylo, yhi) of the updated band after a step whose stencil
// reads x[g + off .. g + off + N - 1]; x is valid on [xlo, xhi) of a band of
// length nx.
LFX_HD void shrink(long long& ylo, long long& yhi, long long xlo, long long xhi, long long nx, int off, int N) {
  const long long lo = (xlo <= 0) ? ylo : lmax(ylo, xlo - off);
  const long long hi = (xhi >= nx) ? yhi : lmin(yhi, xhi - off - (N - 1));
  ylo = lo;
  yhi = hi > lo ? hi : lo;
}

// Valid ranges of the s and d bands of one level.
struct Ranges {
  long long slo, shi, dlo, dhi;
};

// Shrink rule of one lifting step under boundary condition bc.
// Normal steps: with Symmetric (i < 0 -> -1 - i, i >= n -> 2n - 1 - i) or
// Reflect (-i, 2n - 2 - i), a window touching a band end reads mirrored
// elements up to N - 1 + |off| samples away from that end, with Constant the
// end element itself; elements whose mapped reads leave the valid range are
// not updated.
// Adjoint steps: zero-boundary stencil plus boundary folds. A front fold
// (virtual index m in [o, 0)) lands on an element j <= -o and reads x[0, -o);
// a back fold (m in [nd, nx - o_r)) lands on j >= 2 ny - nx + o_r - 1 and reads
// x[nd + o_r, nx). A window that does not hold those reads must not update
// those elements.
LFX_HD void step_shrink(const DStep& st, long long ns, long long nd, Ranges& r, int bc) {
  long long& ylo = st.y_is_d ? r.dlo : r.slo;
  long long& yhi = st.y_is_d ? r.dhi : r.shi;
  const long long xlo = st.y_is_d ? r.slo : r.dlo, xhi = st.y_is_d ? r.shi : r.dhi;
  const long long nx = st.y_is_d ? ns : nd, ny = st.y_is_d ? nd : ns;
  shrink(ylo, yhi, xlo, xhi, nx, st.off, st.N);
  if (!st.adjoint) {
    if (bc == kSymmetric || bc == kReflect) {
      const long long e = bc == kReflect ? 1 : 0;
      if (xlo <= 0) ylo = lmax(ylo, e - xhi - st.off);
      if (xhi >= nx) yhi = lmin(yhi, 2 * nx - st.off - st.N - xlo + 1 - e);
    } else if (bc == kConstant) {
      // reads past an end take the end element, which the window may lack
      if (xlo <= 0 && xhi < 1) ylo = lmax(ylo, -(long long)st.off);
      if (xhi >= nx && xlo > nx - 1) yhi = lmin(yhi, nx - st.off - st.N + 1);
    }
  } else if (bc != kZero) {  // a Zero boundary folds nothing
    if (st.fold_off < 0 && (xlo > 0 || xhi < -(long long)st.fold_off)) ylo = lmax(ylo, 1 - (long long)st.fold_off);
    if (xhi < nx || xlo > lmax(0, nd + st.off)) yhi = lmin(yhi, 2 * ny - nx + st.off - 1);
  }
  if (yhi < ylo) yhi = ylo;
}

// Apply the shrink rule of every step of one level.
template <class C>
LFX_HD void level_ranges(const DProgram<C>& p, long long ns, long long nd, Ranges& r, int bc) {
  for (int k = 0; k < p.nsteps; ++k) {
    const DStep& st = p.steps[k];
    if (st.kind == kScale) continue;
    step_shrink(st, ns, nd, r, bc);
  }
}

// ---- forward tile (fine -> coarse) ----
// Tile [a, b) of the input band (length m0), window [wa, wb) = [a-H, b+H)
// clipped. After level r the tile must own d_r on [a>>r, b>>r) (to the band
// end for the last tile) and, after k levels, s_k on [a>>k, b>>k).
template <class C>
LFX_HD bool fwd_tile_ok(const DProgram<C>& p, long long m0, int k, long long a, long long b, long long H, int bc) {
  const long long wa = lmax(0, a - H), wb = lmin(m0, b + H);
  Ranges r{wa / 2, (wb + 1) / 2, wa / 2, wb / 2};
  long long m = m0;
  for (int lv = 1; lv <= k; ++lv) {
    const long long nd = m / 2, ns = m - nd;
    level_ranges(p, ns, nd, r, bc);
    const long long ad = a >> lv, bd = (b >= m0) ? nd : (b >> lv);
    if (ad < r.dlo || bd > r.dhi) return false;
    if (lv == k) {
      const long long bs = (b >= m0) ? ns : (b >> lv);
      if (ad < r.slo || bs > r.shi) return false;
    } else {
      // next band = s window; deinterleave its valid range
      const long long vlo = r.slo, vhi = r.shi;
      r = Ranges{(vlo + 1) / 2, (vhi + 1) / 2, vlo / 2, vhi / 2};
    }
    m = ns;
  }
  return true;
}

// ---- reverse tile (coarse -> fine) ----
// Tile [a, b) of the output band (length m0); at the coarsest level k the
// windows of s_k and d_k are [(a>>k) - Hk, ceil(b/2^k) + Hk) clipped; each
// interleave doubles the window.
template <class C>
LFX_HD bool rev_tile_ok(const DProgram<C>& p, long long m0, int k, long long a, long long b, long long Hk,
                        int bc) {
  long long m = band_len(m0, k - 1);  // band split by level k
  long long ns = m - m / 2, nd = m / 2;
  const long long g0 = lmax(0, (a >> k) - Hk);
  const long long ghi = ((b + (1ll << k) - 1) >> k) + Hk;
  long long sw = lmax(0, lmin(ghi, ns) - g0), dw = lmax(0, lmin(ghi, nd) - g0);
  long long gg = g0;
  Ranges r{g0, g0 + sw, g0, g0 + dw};
  for (int lv = k; lv >= 1; --lv) {
    level_ranges(p, ns, nd, r, bc);
    // interleave: band index q even <- s[q/2], odd <- d[(q-1)/2]
    const long long vlo = lmax(2 * r.slo, 2 * r.dlo + 1) - 1;
    const long long vhi = lmin(2 * r.shi, 2 * r.dhi + 1);
    const long long wlen = lmin(2 * sw, 2 * dw + 1);
    const long long lo = lmax(vlo, 2 * gg), hi = lmin(vhi, 2 * gg + wlen);
    if (lv == 1) return a >= lo && b <= hi;
    // next level: s window = interleaved band window, d window loaded alongside
    m = band_len(m0, lv - 2);
    ns = m - m / 2;
    nd = m / 2;
    gg = 2 * gg;
    sw = wlen;
    dw = lmax(0, lmin(gg + sw, nd) - gg);
    r = Ranges{lo, hi, gg, gg + dw};
  }
  return false;
}

}  // namespace lfx
