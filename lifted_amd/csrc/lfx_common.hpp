// Types shared by the host planner, the kernels and the CPU simulation build
// (no HIP dependency).
#pragma once
#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__) && !defined(LFX_CPU_SIM)
#define LFX_HDI __host__ __device__ inline
#else
#define LFX_HDI inline
#endif

namespace lfx {

enum Storage : int32_t { kF32 = 0, kF64 = 1, kBF16 = 2, kF16 = 3 };

inline size_t storage_bytes(int s) { return s == kF64 ? 8 : (s == kF32 ? 4 : 2); }
inline size_t compute_bytes(int s) { return s == kF64 ? 8 : 4; }

constexpr int kThreads = 1024;  // launch bound; the block size is chosen at launch
int block_threads();             // threads per workgroup used by the launchers
constexpr int kMaxDims = 8;
// LDS of one workgroup (gfx950: 160 KiB per CU): the LDS kernels keep a copy
// of the step program (kProgLdsBytes, static) next to their dynamic LDS of at
// most kMaxLdsBytes
constexpr size_t kProgLdsBytes = 3072;
constexpr size_t kMaxLdsBytes = 160 * 1024 - kProgLdsBytes;
// Register tiles: one wavefront holds a window of kRegWindow samples,
// kRegPerLane consecutive ones per lane, and runs up to kRegLevels levels;
// reverse tiles count their window in level-kRegLevels units (kRegUnits).
// Plans may take windows of kRegPerLaneSmall samples per lane instead (the
// `strip` tuning knob). Row mode holds one band of at most kRowMax samples per
// 16-lane row.
constexpr int kRegLevels = 4;
constexpr int kRegPerLane = 32;
constexpr int kRegPerLaneSmall = 16;
constexpr long long kRegWindow = 64ll * kRegPerLane;
constexpr long long kRegUnits = kRegWindow >> kRegLevels;
constexpr long long kRowMax = 16ll * kRegPerLane;
// Wavefronts per workgroup of the register-tile launches (each wave runs its
// own item; workgroups only group waves for the dispatcher).
constexpr int kRegWaves = 1;

// Base offsets of a line in three arrays: the line index is decomposed over
// the non-transformed dims (last fastest); str[a][k] is array a's stride of
// dim k, ax[a] its element stride along the transform axis.
struct LineMap3 {
  int32_t nod;
  int32_t pad;
  long long shape[kMaxDims];
  long long str[3][kMaxDims];
  long long ax[3];
};

struct CopyMap {
  int32_t nd;
  int32_t pad;
  long long shape[kMaxDims];
  long long sst[kMaxDims];
  long long dstr[kMaxDims];
};

// Strided copy with type conversion between two boxes of the same shape where
// a different dim is contiguous on each side: dim fi is the source's fast dim,
// dim fo the destination's (fi == fo: a plain strided copy along fi). The other
// dims are the batch (bshape, with their source / destination strides).
struct PermMap {
  int32_t nb;
  int32_t same;  // fi == fo
  long long ni, no;
  long long si_i, si_o, so_i, so_o;
  long long bshape[kMaxDims], bss[kMaxDims], bds[kMaxDims];
  long long nbatch;
};

// LDS bytes of the whole-line kernels for a line of n samples
inline size_t line_lds_bytes(long long n, size_t csize) { return (size_t)(n + (n - n / 2)) * csize; }
// Elements of global scratch per line of the global whole-line kernels (the
// same two buffers, rounded up to 64 elements)
inline long long gline_elems(long long n) { return (n + (n - n / 2) + 63) / 64 * 64; }

// Whole-line launch: arrays 0 (input), 1 (output), 2 (reverse: s_L source).
struct LineLaunch {
  int storage;
  bool fwd;
  const void* prog;
  int bc;
  long long n;  // band length the launch starts from
  int L;        // levels
  const void* a0;
  void* a1;
  const void* a2;
  bool scr[3];  // array holds compute-type scratch instead of storage type
  LineMap3 lm;
  long long nlines;
};

struct TileArgs {
  long long m0;      // band length at the pass's finest level
  long long T;       // tile length at the finest level (multiple of 2^k)
  long long H;       // forward: halo (finest level units); reverse: halo at level k
  long long ntiles;  // tiles per line
  int k;             // levels in the pass
  int pad;
};

// Reverse tile: LDS window (elements) of the finest level.
LFX_HDI long long tile_rev_window(const TileArgs& t) { return (1ll << t.k) * ((t.T >> t.k) + 2 + 2 * t.H); }
LFX_HDI long long tile_fwd_window(const TileArgs& t) { return t.T + 2 * t.H + 2; }
inline size_t tile_lds_bytes(const TileArgs& t, bool fwd, size_t csize) {
  const long long W = fwd ? tile_fwd_window(t) : tile_rev_window(t);
  return (size_t)(W + W / 2 + 2) * csize;
}

// Tiled launch. Forward: arrays 0 input band, 1 output (d regions), 2 s_k out.
// Reverse: 0 input (d regions), 1 output band, 2 s_k in.
struct TileLaunch {
  int storage;
  bool fwd;
  const void* prog;
  int bc;
  TileArgs ta;
  const void* a0;
  void* a1;
  void* a2;
  bool scr[3];
  LineMap3 lm;
  long long nlines;
};

// One work item of a batched stage (see lfx_batch_*). Each of the item's
// three logical arrays (as in LineLaunch / TileLaunch) lives in one of the
// user input, the user output or the plan's scratch, at an element offset.
enum BatchArr : int32_t { kArrIn = 0, kArrOut = 1, kArrScr = 2 };
// kItemGLine: a whole line whose bands live in the plan's scratch (global
// memory) instead of LDS, at element offset BatchItem::a (Periodic lines
// longer than the LDS limit).
enum BatchKind : int32_t { kItemLine = 0, kItemTile = 1, kItemReg = 2, kItemGLine = 3 };
struct BatchItem {
  const void* prog;  // device program
  long long m0;      // band length
  long long a, b;    // owned range of the band (tile items)
  long long H;       // halo (tile items)
  long long off[3];  // element offsets of the three logical arrays
  int32_t kind;      // BatchKind
  int32_t k;         // levels of this item
  int32_t arr[3];    // BatchArr of the three logical arrays
  int32_t pidx;      // register items: index into the plan's program table
  uint32_t wavelet;  // register items: wavelet id (selects the compiled step shapes)
  int32_t rows;      // row items: lines in the item (1..4), one per 16-lane row; 0 otherwise
  long long rstride[3];  // row items: element stride between consecutive lines of each array
};

}  // namespace lfx
