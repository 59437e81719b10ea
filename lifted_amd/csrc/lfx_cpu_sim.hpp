// Host build of the kernel bodies (LFX_CPU_SIM): one lane per "workgroup",
// barriers are no-ops, device qualifiers vanish. Used only by the CPU
// simulation that checks the tiling logic against the oracle.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __forceinline__ inline
#define LFX_TID 0
#define LFX_NTHR 1
#define LFX_SYNC() ((void)0)

static inline float __uint_as_float(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline uint32_t __float_as_uint(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
