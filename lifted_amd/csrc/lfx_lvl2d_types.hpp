// Launch arguments and tile geometry of the two-axis level tiles
// (lfx_lvl2d.hpp); no device code, so host planning can include it.
#pragma once
#include <cstdint>

#include "lfx_common.hpp"
#include "lfx_tiles.hpp"

namespace lfx {
namespace l2d {

constexpr int kRows = 32;  // register-axis window (positions per lane)
// Lane axis: each lane holds kLaneE values of the s band and kLaneE of the d
// band, i.e. kLaneW consecutive fine positions; the wavefront's window is
// kCols fine positions.
constexpr int kLaneE = 1;
constexpr int kLaneW = 2 * kLaneE;
constexpr int kCols = 64 * kLaneW;

// An array of one tile: base pointer and storage type, element strides of the
// register axis (sr) and lane axis (sc; the kernels require 1: lanes are
// contiguous), and of the batch dims. v: the base pointer and every stride
// other than sc are multiples of kLaneW elements, so a lane's kLaneW
// consecutive fine positions (starting at an even position) are one aligned
// vector access.
struct Arr {
  void* p;
  int32_t st;
  int32_t v;
  long long sr, sc;
  long long bs[kMaxDims];
};

// Division of 31-bit unsigned values by an invariant d >= 1 as a multiply-
// high and a shift (round-up method): q = (umulhi(n, m) + n) >> l with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1; exact for n < 2^31.
struct FDiv {
  uint32_t d, m, l, pad;
};
inline FDiv make_fdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  return FDiv{d, (uint32_t)m, l, 0};
}
LFX_HD uint32_t fdiv(const FDiv& f, uint32_t n) {
  const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
  return (hi + n) >> f.l;
}

// One level's launch: arrays 0 (input: the sub-box, or the details for a
// reverse level), 1 (output: the details, or the sub-box) and 2 (the ss
// quadrant: written by a forward level, read by a reverse one).
struct Level {
  Arr a[3];
  int32_t nb;    // batch dims
  int32_t pad;
  long long bshape[kMaxDims];
  long long nbatch;
  long long mr, mc;  // band lengths entering the level (register axis, lane axis)
  long long Hr, Hc;  // halos (fine positions, even)
  long long Or, Oc;  // owned positions per tile (even)
  long long Tr, Tc;  // tiles per axis
  long long ntiles;  // nbatch * Tr * Tc
  const void* prog;  // DProgram<float>
  uint32_t wavelet;
  int32_t pad2;
  // tile index decomposition (ntiles < 2^31): w = (b Tr + tr) Tc + tc
  FDiv f_tc, f_per, f_b[kMaxDims];
};

// Geometry of one tile along one axis: window start w0 (even, may be
// negative), owned fine range [a, b) (b == m for the last tile), band length m.
struct Span {
  long long m, w0, a, b;
  LFX_HD long long ns() const { return m - m / 2; }
  LFX_HD long long nd() const { return m / 2; }
  // owned band indices [lo, hi) of the s / d band after the level
  LFX_HD long long lo() const { return a >> 1; }
  LFX_HD long long hi_s() const { return b >= m ? ns() : (b >> 1); }
  LFX_HD long long hi_d() const { return b >= m ? nd() : (b >> 1); }
};

LFX_HD Span span_of(long long m, long long H, long long O, long long t) {
  const long long a = t * O;
  const long long b = a + O < m ? a + O : m;
  return Span{m, a - H, a, b};
}

}  // namespace l2d
}  // namespace lfx
