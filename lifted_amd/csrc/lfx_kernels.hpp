// Host-visible declarations of the gfx950 kernels' launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "lfx_common.hpp"

namespace lfx {

hipError_t launch_line(const LineLaunch& ll, hipStream_t st);
hipError_t launch_tile(const TileLaunch& tl, hipStream_t st);
// One batched stage: `nitems` items (sorted by decreasing cost), arrays
// `in` / `out` in storage type, `scr` in compute type.
hipError_t launch_batch(int storage, bool fwd, int bc, const BatchItem* items, long long nitems, size_t lds_bytes,
                        const void* in, void* out, void* scr, hipStream_t st);
hipError_t launch_copy(int storage, const void* src, void* dst, const CopyMap& cm, long long total,
                       hipStream_t st);

}  // namespace lfx
