// Fixed device memory of the no-allocation entry point (lfx_lwt_noalloc).
#pragma once
#include <cstddef>

namespace lfx {

// Largest scratch a no-allocation call may use: the band scratch of tiled
// passes (two bands per line) plus, in place, a copy of the input.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
// Program slots: every table wavelet x 4 transforms, per compute type.
constexpr int kNoAllocWavelets = 40;

// Device addresses of the current device's fixed arrays (found with
// hipGetSymbolAddress at set-up, once per device) and its program tables,
// filled at set-up; slot 4 * table_index(wavelet) + transform.
struct NoAllocDevice {
  void* scratch = nullptr;
  const void* progs_f32 = nullptr;  // DProgram<float>[4 * kNoAllocWavelets]
  const void* progs_f64 = nullptr;  // DProgram<double>[4 * kNoAllocWavelets]
};

// Set-up of the current device on first use (thread-safe); throws Error on
// a HIP failure. Allocates no device memory.
const NoAllocDevice& noalloc_device();

}  // namespace lfx
