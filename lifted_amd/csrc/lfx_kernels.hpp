// Host-visible declarations of the gfx950 kernels' launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "lfx_common.hpp"

namespace lfx {

hipError_t launch_line(const LineLaunch& ll, hipStream_t st);
hipError_t launch_tile(const TileLaunch& tl, hipStream_t st);
// Whole lines with their bands in global scratch (gscr: gline_elems(n)
// compute-type elements per line for lines_per_launch lines; launches run in
// turn over chunks of lines_per_launch lines). User arrays only (no scr flags).
hipError_t launch_gline(const LineLaunch& ll, void* gscr, long long lines_per_launch, hipStream_t st);
// One batched stage: `nitems` items (sorted by decreasing cost), arrays
// `in` / `out` in storage type, `scr` in compute type.
// threads: workgroup size (0: the default of block_threads())
hipError_t launch_batch(int storage, bool fwd, int bc, const BatchItem* items, long long nitems, size_t lds_bytes,
                        int threads, const void* in, void* out, void* scr, hipStream_t st);
// Register-tile items (float32) of one transform type, any table wavelet, in
// windows of pl (16 or 32) samples per lane: launch_reg_<transform>_<pl>;
// progs: the plan's device table of DProgram<float>, indexed by BatchItem::pidx.
#define LFX_REG_DECL(tf, pl)                                                                               \
  hipError_t launch_reg_##tf##_##pl(const void* progs, const BatchItem* items, long long nitems, const void* in, \
                                    void* out, void* scr, hipStream_t st);
LFX_REG_DECL(0, 16)
LFX_REG_DECL(1, 16)
LFX_REG_DECL(2, 16)
LFX_REG_DECL(3, 16)
LFX_REG_DECL(0, 32)
LFX_REG_DECL(1, 32)
LFX_REG_DECL(2, 32)
LFX_REG_DECL(3, 32)
#undef LFX_REG_DECL
inline hipError_t launch_reg(int tf, int pl, const void* progs, const BatchItem* items, long long nitems,
                             const void* in, void* out, void* scr, hipStream_t st) {
  switch (tf * 100 + pl) {
    case 16: return launch_reg_0_16(progs, items, nitems, in, out, scr, st);
    case 116: return launch_reg_1_16(progs, items, nitems, in, out, scr, st);
    case 216: return launch_reg_2_16(progs, items, nitems, in, out, scr, st);
    case 316: return launch_reg_3_16(progs, items, nitems, in, out, scr, st);
    case 32: return launch_reg_0_32(progs, items, nitems, in, out, scr, st);
    case 132: return launch_reg_1_32(progs, items, nitems, in, out, scr, st);
    case 232: return launch_reg_2_32(progs, items, nitems, in, out, scr, st);
    case 332: return launch_reg_3_32(progs, items, nitems, in, out, scr, st);
  }
  return hipErrorInvalidValue;
}
// One level of a two-axis level-by-level transform (l2d::Level, lfx_lvl2d.hpp).
namespace l2d {
struct Level;
}
// launch_lvl2d_<transform>_<storage>: one per transform type and storage
// type (float32 0, bfloat16 2, float16 3)
#define LFX_L2D_DECL(tf)                                              \
  hipError_t launch_lvl2d_##tf##_0(const l2d::Level& L, hipStream_t st); \
  hipError_t launch_lvl2d_##tf##_2(const l2d::Level& L, hipStream_t st); \
  hipError_t launch_lvl2d_##tf##_3(const l2d::Level& L, hipStream_t st);
LFX_L2D_DECL(0)
LFX_L2D_DECL(1)
LFX_L2D_DECL(2)
LFX_L2D_DECL(3)
#undef LFX_L2D_DECL
// every array of a level has the storage type of its input (the kernels are
// compiled per storage type; a launch whose arrays differ is refused)
inline hipError_t launch_lvl2d(int tf, int storage, const l2d::Level& L, hipStream_t st) {
  switch (tf * 4 + storage) {
#define LFX_L2D_CASE(tf, s) \
  case tf * 4 + s:          \
    return launch_lvl2d_##tf##_##s(L, st);
#define LFX_L2D_CASES(tf) LFX_L2D_CASE(tf, 0) LFX_L2D_CASE(tf, 2) LFX_L2D_CASE(tf, 3)
    LFX_L2D_CASES(0)
    LFX_L2D_CASES(1)
    LFX_L2D_CASES(2)
    LFX_L2D_CASES(3)
#undef LFX_L2D_CASES
#undef LFX_L2D_CASE
  }
  return hipErrorInvalidValue;
}
// Every line of a strided axis in blocks of lines staged in LDS
// (acr::Args, lfx_across.hpp).
namespace acr {
struct Args;
}
hipError_t launch_across_0(const acr::Args& A, hipStream_t st);
hipError_t launch_across_1(const acr::Args& A, hipStream_t st);
hipError_t launch_across_2(const acr::Args& A, hipStream_t st);
hipError_t launch_across_3(const acr::Args& A, hipStream_t st);
inline hipError_t launch_across(int tf, const acr::Args& A, hipStream_t st) {
  switch (tf) {
    case 0: return launch_across_0(A, st);
    case 1: return launch_across_1(A, st);
    case 2: return launch_across_2(A, st);
    case 3: return launch_across_3(A, st);
  }
  return hipErrorInvalidValue;
}
// Row items (float32, Zero boundary): four equal short lines per wavefront.
hipError_t launch_rows_0(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st);
hipError_t launch_rows_1(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st);
hipError_t launch_rows_2(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st);
hipError_t launch_rows_3(const void* progs, const BatchItem* items, long long nitems, const void* in, void* out,
                         void* scr, hipStream_t st);
inline hipError_t launch_rows(int tf, const void* progs, const BatchItem* items, long long nitems, const void* in,
                              void* out, void* scr, hipStream_t st) {
  switch (tf) {
    case 0: return launch_rows_0(progs, items, nitems, in, out, scr, st);
    case 1: return launch_rows_1(progs, items, nitems, in, out, scr, st);
    case 2: return launch_rows_2(progs, items, nitems, in, out, scr, st);
    case 3: return launch_rows_3(progs, items, nitems, in, out, scr, st);
  }
  return hipErrorInvalidValue;
}
// Streaming-copy references (pattern 0: float4 per thread; 1: 2048-float
// window per wavefront); copies bytes rounded down to the pattern's unit.
hipError_t launch_stream_copy(const void* src, void* dst, long long bytes, int pattern, hipStream_t st);
hipError_t launch_copy(int storage, const void* src, void* dst, const CopyMap& cm, long long total,
                       hipStream_t st);
// Box copy between storage types (kF32 <-> kBF16 / kF16, or same type f32 /
// f64) where each side is contiguous along its own dim (PermMap).
hipError_t launch_permute(int storage_in, int storage_out, const void* src, void* dst, const PermMap& pm,
                          hipStream_t st);

}  // namespace lfx
