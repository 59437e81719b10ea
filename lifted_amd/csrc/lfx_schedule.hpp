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
long long find_halo(const DProgram<C>& hp, bool fwd, long long m0, int k, long long T) {
  const long long ntiles = (m0 + T - 1) / T;
  const auto tiles = probe_tiles(ntiles);
  const long long hmax = fwd ? (T >> k) : T;
  for (long long h = 0; h <= hmax; ++h) {
    const long long H = fwd ? (h << k) : h;
    bool ok = true;
    for (long long t : tiles) {
      const long long a = t * T, b = std::min(a + T, m0);
      ok = fwd ? fwd_tile_ok(hp, m0, k, a, b, H) : rev_tile_ok(hp, m0, k, a, b, H);
      if (!ok) break;
    }
    if (ok) return H;
  }
  return -1;
}

template <class C>
LineSchedule schedule(const DProgram<C>& hp, bool fwd, long long n, long long L, int storage, int bc,
                      long long tile_hint, int k_hint, long long limit_override = 0) {
  LineSchedule s{{}, n, L};
  const long long limit = limit_override > 0 ? limit_override : line_limit(storage);
  if (n <= limit || L == 0) return s;
  if (bc == kPeriodic)
    fail(LFX_EUNSUPPORTED, "Periodic boundary on lines longer than " + std::to_string(limit) +
                               " samples is not supported (its wrap reads the far end of the band)");
  long long m = n, lv = 0;
  const long long T0 = tile_hint > 0 ? tile_hint : (compute_bytes(storage) == 8 ? 4096 : 8192);
  while (lv < L && m > limit) {
    int k = (int)std::min<long long>(L - lv, k_hint > 0 ? k_hint : 6);
    long long T = T0, H = -1;
    for (;;) {
      T = ((T >> k) << k);
      H = find_halo(hp, fwd, m, k, T);
      const long long halo_fine = fwd ? H : (H << k);
      const TileArgs ta{m, T, H, (m + T - 1) / T, k, 0};
      if (H >= 0 && halo_fine * 4 <= T && tile_lds_bytes(ta, fwd, compute_bytes(storage)) <= kMaxLdsBytes) break;
      if (k > 1) { --k; continue; }
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

}  // namespace lfx
