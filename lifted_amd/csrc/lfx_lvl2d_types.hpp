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
