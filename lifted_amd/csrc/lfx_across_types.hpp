// Launch arguments of the across-line kernel (lfx_across.hpp); no device
// code, so host planning can include it.
#pragma once
#include <cstdint>

#include "lfx_common.hpp"

namespace lfx {
namespace acr {

constexpr int kLines = 32;       // lines per workgroup
constexpr int kMinLines = 16;    // fewest lines along the contiguous dim worth a launch
constexpr int kMaxN = 512;       // samples per line (row tiles: 16 lanes x 32)
constexpr int kGroups = kLines / 4;  // groups of four lines (one row tile each)
constexpr int kWaves = 4;            // wavefronts per workgroup, kGroups / kWaves groups each
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxPasses = 8;    // passes of at most four levels
constexpr int kPer = kLines * kMaxN / kThreads;  // slab values per thread
constexpr int kPosStep = kThreads / kLines;      // positions between a thread's values
static_assert(kGroups % kWaves == 0 && 32 % kPosStep == 0 && kMaxN == 512,
              "the slab index of a thread's values (lfx_across.hpp) assumes these");
LFX_HDI int phys(int a) { return a + (a >> 5) + (a >> 9); }
constexpr int kSlabFloats = kLines * kMaxN + kLines * kMaxN / 32 + kLines * kMaxN / 512;

// One launch: every line of the transform axis of a box. The contiguous dim
// (unit stride on both sides) supplies the kLines lines of a workgroup; the
// other dims are the batch. Passes run fine -> coarse for forward-stepping
// transforms, coarse -> fine otherwise.
struct Args {
  const void* in;
  void* out;
  int32_t st_in, st_out;  // storage types (float32 or 16-bit)
  long long n;            // line length
  long long nc;           // lines along the contiguous dim
  long long ncb;          // blocks of kLines of them
  long long sax_in, sax_out;
  int32_t nb;
  int32_t npass;
  long long bshape[kMaxDims], bs_in[kMaxDims], bs_out[kMaxDims];
  long long nblocks;  // batch slices x ncb
  long long pm[kMaxPasses];  // band length entering each pass
  int32_t pk[kMaxPasses];    // levels of each pass
  const void* prog;          // DProgram<float>
  uint32_t wavelet;
  int32_t fwd;
};

}  // namespace acr
}  // namespace lfx
