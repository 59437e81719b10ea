// Schedule of a 1-D multi-level transform of long lines: tiled passes of k
// levels (halo found by simulating the valid-range rules of lfx_tiles.hpp)
// until the band fits one workgroup, then one whole-line pass.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "lfx_common.hpp"
#include "lfx_error.hpp"
#include "lfx_program.hpp"
#include "lfx_tiles.hpp"

namespace lfx {

inline long long line_limit(int storage) {
  // longest line the whole-line kernels take: LDS = 1.5 n elements
  const long long cb = (long long)compute_bytes(storage);
  return (long long)kMaxLdsBytes * 2 / (3 * cb) - 2;
}

// ---------------------------------------------------------------------------
// long-line schedule: tiled passes of k levels until the band fits one
// workgroup, then one whole-line launch for the remaining levels.

struct PassPlan {
  long long m0;
  int k;
  long long T, H, ntiles;
};
struct LineSchedule {
  std::vector<PassPlan> passes;
  long long m_line;  // band length of the whole-line part
  long long L_line;  // its levels (0: none)
  // the whole-line part keeps its bands in global scratch instead of LDS
  // (Periodic lines longer than the LDS limit: the wrap reads the far end of
  // every band, which no tile holds)
  bool global_line = false;
};

inline std::vector<long long> probe_tiles(long long ntiles) {
  std::vector<long long> t;
  if (ntiles <= 64) {
    for (long long i = 0; i < ntiles; ++i) t.push_back(i);
  } else {
    for (long long i = 0; i < 4; ++i) t.push_back(i);
    t.push_back(ntiles / 2);
    for (long long i = ntiles - 4; i < ntiles; ++i) t.push_back(i);
  }
  return t;
}

template <class C>
long long find_halo(const DProgram<C>& hp, bool fwd, long long m0, int k, long long T, int bc) {
  const long long ntiles = (m0 + T - 1) / T;
  const auto tiles = probe_tiles(ntiles);
  const long long hmax = fwd ? (T >> k) : T;
  for (long long h = 0; h <= hmax; ++h) {
    const long long H = fwd ? (h << k) : h;
    bool ok = true;
    for (long long t : tiles) {
      const long long a = t * T, b = std::min(a + T, m0);
      ok = fwd ? fwd_tile_ok(hp, m0, k, a, b, H, bc) : rev_tile_ok(hp, m0, k, a, b, H, bc);
      if (!ok) break;
    }
    if (ok) return H;
  }
  return -1;
}

// Default LDS tile length (samples at a pass's finest level) by compute type.
inline long long default_tile(int storage) { return compute_bytes(storage) == 8 ? 4096 : 8192; }

// The default tile, halved (down to 512) while a pass over `nlines` lines of
// n samples would launch fewer than 512 tiles: a few long lines then still
// spread over the chip's 256 CUs. A caller's tile (tile > 0) is kept.
inline long long tile_for_lines(long long tile, int storage, long long nlines, long long n) {
  if (tile > 0) return tile;
  long long T = default_tile(storage);
  while (T > 512 && nlines * ((n + T - 1) / T) < 512) T /= 2;
  return T;
}

template <class C>
LineSchedule schedule(const DProgram<C>& hp, bool fwd, long long n, long long L, int storage, int bc,
                      long long tile_hint, int k_hint, long long limit_override = 0) {
  LineSchedule s{{}, n, L};
  const long long limit = limit_override > 0 ? std::min(limit_override, line_limit(storage)) : line_limit(storage);
  if (n <= limit || L == 0) return s;
  if (bc == kPeriodic) {
    s.global_line = n > line_limit(storage);
    return s;
  }
  long long m = n, lv = 0;
  const long long T0 = tile_hint > 0 ? tile_hint : default_tile(storage);
  while (lv < L && m > limit) {
    const int k0 = (int)std::min<long long>(L - lv, k_hint > 0 ? k_hint : 6);
    int k = k0;
    long long T = T0, Tk = T0, H = -1;
    for (;;) {
      T = ((Tk >> k) << k);
      H = find_halo(hp, fwd, m, k, T, bc);
      const long long halo_fine = fwd ? H : (H << k);
      const TileArgs ta{m, T, H, (m + T - 1) / T, k, 0};
      const size_t lds = tile_lds_bytes(ta, fwd, compute_bytes(storage));
      if (H >= 0 && halo_fine * 4 <= T && lds <= kMaxLdsBytes) break;
      if (k > 1) { --k; continue; }
      // a tile too short for the halo of even one level: longer tiles
      if (lds <= kMaxLdsBytes / 2 && Tk < m) {
        Tk *= 2;
        k = k0;
        continue;
      }
      fail(LFX_EUNSUPPORTED, "no tiling found for this wavelet");
    }
    s.passes.push_back(PassPlan{m, k, T, H, (m + T - 1) / T});
    m = band_len(m, k);
    lv += k;
  }
  s.m_line = m;
  s.L_line = L - lv;
  return s;
}

// ---------------------------------------------------------------------------
// Work items of one group of levels of one signal.
//
// A tiled group of one to four levels of a float32 table wavelet runs its
// interior as register tiles (one wavefront per window of `win` samples,
// kRegWindow by default, lfx_reg.hpp) and the parts next to the band ends as
// LDS tiles; other groups are LDS tiles of the schedule's tile length, or one
// whole-line item.

enum ItemKind { kSpecLine = 0, kSpecTile = 1, kSpecReg = 2 };
struct ItemSpec {
  int kind;
  long long m0;  // band length entering the group
  int k;         // levels
  long long a, b, H;
};

// Halo of an interior register tile of K levels in a window of `win` samples
// (units = win / 16): forward H (finest-level samples, multiple of 16, owned
// win - 2H); reverse Hk (level-4 units whatever K is, owned 16 (units - 2 Hk)).
// -1 when the stencil is too wide for the register window to pay off.
template <class C>
long long reg_halo(const DProgram<C>& hp, bool fwd, int K, int bc, long long win = kRegWindow) {
  const long long m0 = 1ll << 40, a0 = 1ll << 36, units = win >> kRegLevels;
  if (fwd) {
    for (long long H = 0; 2 * H <= win - 256; H += 16)
      if (fwd_tile_ok(hp, m0, K, a0, a0 + win - 2 * H, H, bc)) return H;
  } else {
    for (long long Hk = 0; 16 * (units - 2 * Hk) >= 256; ++Hk)
      if (rev_tile_ok(hp, m0, K, a0, a0 + 16 * (units - 2 * Hk), Hk << (kRegLevels - K), bc)) return Hk;
  }
  return -1;
}

// Register tiles may cover a band's ends (Zero boundary, masked windows) when
// every band of the pass is long enough that no step's stencil spans the band:
// the reference updates some elements twice there (n_front > nd - n_back), which
// zero padding would not reproduce.
template <class C>
bool zero_cover_ok(const DProgram<C>& hp, long long m0, int k) {
  long long reach = 0;
  for (int i = 0; i < hp.nsteps; ++i)
    if (hp.steps[i].kind != kScale) reach = std::max<long long>(reach, hp.steps[i].nf + hp.steps[i].nb);
  long long m = m0;
  for (int l = 1; l <= k; ++l) {
    if (m / 2 <= reach + 1) return false;
    m = m - m / 2;
  }
  return true;
}

// Smallest LDS-tile halo for the owned range [a, b) (-1: none below the cap).
template <class C>
long long lds_halo(const DProgram<C>& hp, bool fwd, long long m0, int k, long long a, long long b, long long cap,
                   int bc) {
  const long long step = fwd ? (1ll << k) : 1;
  for (long long H = 0; H <= cap; H += step) {
    const bool ok = fwd ? fwd_tile_ok(hp, m0, k, a, b, H, bc) : rev_tile_ok(hp, m0, k, a, b, H, bc);
    if (ok) return H;
  }
  return -1;
}

// Items of a tiled pass (band m0, k levels). `reg` allows register tiles, of
// windows of `win` samples.
template <class C>
std::vector<ItemSpec> pass_items(const DProgram<C>& hp, bool fwd, const PassPlan& pp, bool reg, int storage,
                                 int bc, long long win = kRegWindow) {
  std::vector<ItemSpec> out;
  const long long m0 = pp.m0, units = win >> kRegLevels;
  long long r0 = 0, r1 = 0, Tr = 0, H = -1;
  if (reg && pp.k >= 1 && pp.k <= kRegLevels) H = reg_halo(hp, fwd, pp.k, bc, win);
  if (H >= 0 && bc == kZero && zero_cover_ok(hp, m0, pp.k)) {
    // Zero boundary: register windows may overlap the band ends (they run
    // masked, reading zeros there), so register tiles cover the whole band
    const long long Tr = fwd ? win - 2 * H : 16 * (units - 2 * H);
    const long long nt = (m0 + Tr - 1) / Tr;
    auto ok = [&](long long t) {
      const long long a = t * Tr, b = std::min(a + Tr, m0);
      return fwd ? fwd_tile_ok(hp, m0, pp.k, a, b, H, bc)
                 : rev_tile_ok(hp, m0, pp.k, a, b, H << (kRegLevels - pp.k), bc);
    };
    bool all = true;
    for (long long t : probe_tiles(nt)) all = all && ok(t);
    if (all) {
      for (long long t = 0; t < nt; ++t)
        out.push_back(ItemSpec{kSpecReg, m0, pp.k, t * Tr, std::min(t * Tr + Tr, m0), H});
      (void)storage;
      return out;
    }
  }
  if (H >= 0) {
    // Register windows stay strictly inside the band at every level (they
    // apply no boundary map), with a margin of 64 samples at the finest level.
    if (fwd) {
      Tr = win - 2 * H;
      r0 = H + 64;
      const long long last = m0 - win + H - 64;  // window [a - H, a - H + win)
      if (last >= r0) r1 = r0 + ((last - r0) / Tr + 1) * Tr;
    } else {
      Tr = 16 * (units - 2 * H);
      r0 = 16 * (H + 4);
      // every level's window [2^(4-l) g0, 2^(4-l) (g0 + units)) inside its d band, and s_k's inside ns_k
      long long gmax = -1;
      long long m = m0;
      bool ok = true;
      for (int l = 1; l <= pp.k; ++l) {
        const long long nd = m / 2, ns = m - nd;
        const long long span = 1ll << (kRegLevels - l);
        const long long g = nd / span - (units + 4);  // g0 + units < nd / span, with a margin
        gmax = (gmax < 0 || g < gmax) ? g : gmax;
        if (l == pp.k) gmax = std::min(gmax, ns / span - (units + 4));
        m = ns;
        if (g < 0) ok = false;
      }
      if (ok && gmax >= 0) {
        const long long last = 16 * (gmax + H);  // g0 = a / 16 - H <= gmax
        if (last >= r0) r1 = r0 + ((last - r0) / Tr + 1) * Tr;
      }
    }
  }
  if (r1 > r0) {
    // The halo was sized on a tile far from the band ends; tiles near an end
    // also lose the elements that boundary folds and maps reach, so check the
    // first and last tiles with the exact rule and move them inward until they
    // hold (interior tiles are translates of a checked one).
    auto ok = [&](long long a) {
      return fwd ? fwd_tile_ok(hp, m0, pp.k, a, a + Tr, H, bc)
                 : rev_tile_ok(hp, m0, pp.k, a, a + Tr, H << (kRegLevels - pp.k), bc);
    };
    const long long ntl = (r1 - r0) / Tr;
    long long first = 0, last = ntl - 1;
    while (first <= last && !(ok(r0 + first * Tr) && ok(r0 + std::min(first + 1, last) * Tr) &&
                              ok(r0 + std::min(first + 2, last) * Tr)))
      ++first;
    while (last >= first && !(ok(r0 + last * Tr) && ok(r0 + std::max(last - 1, first) * Tr) &&
                              ok(r0 + std::max(last - 2, first) * Tr)))
      --last;
    if (last >= first) {
      r1 = r0 + (last + 1) * Tr;
      r0 = r0 + first * Tr;
    } else {
      r1 = r0;
    }
  }
  if (r1 <= r0) {
    for (long long t = 0; t < pp.ntiles; ++t) {
      const long long a = t * pp.T, b = std::min(a + pp.T, m0);
      out.push_back(ItemSpec{kSpecTile, m0, pp.k, a, b, pp.H});
    }
    return out;
  }
  const long long cap = std::max<long long>(4 * pp.T, 4096);
  auto lds_range = [&](long long a, long long b) {
    // cut [a, b) into pieces of at most pp.T, each with its own halo
    for (long long x = a; x < b; x += pp.T) {
      const long long y = std::min(b, x + pp.T);
      const long long h = lds_halo(hp, fwd, m0, pp.k, x, y, cap, bc);
      if (h < 0) fail(LFX_EUNSUPPORTED, "no halo found for a band-end tile");
      out.push_back(ItemSpec{kSpecTile, m0, pp.k, x, y, h});
    }
  };
  if (r0 > 0) lds_range(0, r0);
  for (long long a = r0; a < r1; a += Tr) out.push_back(ItemSpec{kSpecReg, m0, pp.k, a, a + Tr, H});
  if (r1 < m0) lds_range(r1, m0);
  (void)storage;
  return out;
}


// ---------------------------------------------------------------------------
// Two-axis level tiles (lfx_lvl2d.hpp): halo (fine positions, even) of a
// window of W positions along one axis of band length m, for one level of
// program hp under the Zero boundary, so that every tile owns an exact range;
// -1 when no halo up to W / 4 works, or the band is too short for masked
// windows (zero_cover_ok).
template <class C>
long long lvl2d_axis_halo(const DProgram<C>& hp, bool fwd, long long m, long long W) {
  if (!zero_cover_ok(hp, m, 1)) return -1;
  const long long M = 1ll << 40, a0 = 1ll << 36;
  long long H = -1;
  for (long long h = 0; 4 * h <= W; h += 2) {
    const bool ok = fwd ? fwd_tile_ok(hp, M, 1, a0, a0 + W - 2 * h, h, kZero)
                        : rev_tile_ok(hp, M, 1, a0, a0 + W - 2 * h, h / 2, kZero);
    if (ok) {
      H = h;
      break;
    }
  }
  if (H < 0) return -1;
  // tiles next to the band ends lose what the boundary reaches: check them
  // with the exact rule (interior tiles are translates of the checked one)
  const long long O = W - 2 * H, T = (m + O - 1) / O;
  for (long long t : probe_tiles(T)) {
    const long long a = t * O, b = std::min(a + O, m);
    const bool ok = fwd ? fwd_tile_ok(hp, m, 1, a, b, H, kZero) : rev_tile_ok(hp, m, 1, a, b, H / 2, kZero);
    if (!ok) return -1;
  }
  return H;
}

}  // namespace lfx
