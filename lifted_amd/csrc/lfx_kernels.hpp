// Host-visible declarations of the gfx950 kernels' launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace lfx {

enum Storage : int32_t { kF32 = 0, kF64 = 1, kBF16 = 2, kF16 = 3 };

inline size_t storage_bytes(int s) { return s == kF64 ? 8 : (s == kF32 ? 4 : 2); }
inline size_t compute_bytes(int s) { return s == kF64 ? 8 : 4; }

constexpr int kLineThreads = 256;
constexpr int kMaxDims = 8;
// LDS a single workgroup may use (gfx950 has 160 KiB per CU).
constexpr size_t kMaxLdsBytes = 160 * 1024;

// Maps a line index to element offsets of source and destination: the line
// index is decomposed over the non-transformed dims (last one fastest).
struct LineMap {
  int32_t nod;
  int32_t pad;
  long long shape[kMaxDims];
  long long sst[kMaxDims];
  long long dstr[kMaxDims];
  long long s_ax, d_ax;  // element strides along the transform axis
};

struct CopyMap {
  int32_t nd;
  int32_t pad;
  long long shape[kMaxDims];
  long long sst[kMaxDims];
  long long dstr[kMaxDims];
};

// LDS bytes the whole-line kernels need for a line of n samples.
inline size_t line_lds_bytes(long long n, size_t csize) { return (size_t)(n + (n - n / 2)) * csize; }

hipError_t launch_line(int storage, bool fwd, const void* prog, int bc, long long n, int L, const void* src,
                       void* dst, const LineMap& lm, long long nlines, hipStream_t st);
hipError_t launch_copy(int storage, const void* src, void* dst, const CopyMap& cm, long long total,
                       hipStream_t st);

}  // namespace lfx
