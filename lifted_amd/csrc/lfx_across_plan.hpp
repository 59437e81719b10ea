// Host planning of the across-line kernel (lfx_across.hpp), shared by the
// library and the CPU simulation.
#pragma once
#include <algorithm>

#include "lfx_across_types.hpp"
#include "lfx_schedule.hpp"

namespace lfx {

// Passes of at most kRegLevels levels of a line of n samples (fine ->
// coarse), each admitting the row tiles (Zero boundary): false when some pass
// does not, or when the line is too long for the slab.
template <class C>
bool across_passes(const DProgram<C>& hp, bool fwd, long long n, long long L, acr::Args& A) {
  if (n < 2 || n > acr::kMaxN || L < 1) return false;
  A.npass = 0;
  long long m = n, left = L;
  while (left > 0) {
    const int K = (int)std::min<long long>(kRegLevels, left);
    if (A.npass >= acr::kMaxPasses || reg_halo(hp, fwd, K, kZero) < 0 || !zero_cover_ok(hp, m, K)) return false;
    A.pm[A.npass] = m;
    A.pk[A.npass] = K;
    ++A.npass;
    m = band_len(m, K);
    left -= K;
  }
  A.fwd = fwd ? 1 : 0;
  return true;
}

// Blocks of the lines of axis ax: dim c (unit stride on both sides) gives a
// block's acr::kLines lines, the other dims are the batch.
inline void across_geometry(int ndim, const long long* shp, int ax, int c, const long long* sst,
                            const long long* dstr, acr::Args& A) {
  A.n = shp[ax];
  A.nc = shp[c];
  A.ncb = (A.nc + acr::kLines - 1) / acr::kLines;
  A.sax_in = sst[ax];
  A.sax_out = dstr[ax];
  A.nb = 0;
  long long nbatch = 1;
  for (int k = 0; k < ndim; ++k) {
    if (k == ax || k == c) continue;
    A.bshape[A.nb] = shp[k];
    A.bs_in[A.nb] = sst[k];
    A.bs_out[A.nb] = dstr[k];
    ++A.nb;
    nbatch *= shp[k];
  }
  A.nblocks = nbatch * A.ncb;
}

}  // namespace lfx
