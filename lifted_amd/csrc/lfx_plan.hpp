// A validated transform request (the lfx_desc of the C ABI): the reference's
// lwt arguments (src/lifted.hpp:13-39) with axes normalised. Host-only; shared
// by the GPU driver (lfx_api.cpp) and the CPU driver (lfx_cpu.cpp).
#pragma once
#include <string>
#include <vector>

#include "../../include/lifted_hip.h"
#include "lfx_common.hpp"
#include "lfx_error.hpp"

namespace lfx {

struct Plan {
  int tf, bc, storage;
  uint32_t wavelet;
  int ndim;
  long long shape[kMaxDims], sin[kMaxDims], sout[kMaxDims];
  std::vector<int> axes;
  bool separable;
  std::vector<long long> levels;
  uint32_t flags;
};

inline Plan make_plan(const lfx_desc* d) {
  if (!d) fail(LFX_EINVAL, "null descriptor");
  Plan p;
  if (d->transform > 3) fail(LFX_EINVAL, "unknown transform " + std::to_string(d->transform));
  if (d->bc > 4) fail(LFX_EINVAL, "Unknown Boundary Condition");
  if (d->storage < 0 || d->storage > 3) fail(LFX_EINVAL, "unknown storage type");
  if (d->ndim < 1 || d->ndim > kMaxDims) fail(LFX_EINVAL, "ndim must be in 1..8");
  if (d->naxes < 1) fail(LFX_EINVAL, "at least one axis is required");
  p.tf = (int)d->transform;
  p.bc = (int)d->bc;
  p.storage = d->storage;
  p.wavelet = d->wavelet;
  p.ndim = d->ndim;
  p.flags = d->flags;
  for (int i = 0; i < d->ndim; ++i) {
    if (d->shape[i] < 0) fail(LFX_EINVAL, "negative shape");
    p.shape[i] = d->shape[i];
    p.sin[i] = d->stride_in[i];
    p.sout[i] = d->stride_out[i];
  }
  for (int i = 0; i < d->naxes; ++i) {
    long long a = d->axes[i];
    if (a < 0) a += d->ndim;
    if (a < 0 || a >= d->ndim) fail(LFX_EINVAL, "axis " + std::to_string(d->axes[i]) + " out of range");
    p.axes.push_back((int)a);
  }
  p.separable = d->separable != 0;
  const int nl = p.separable ? d->naxes : 1;
  for (int i = 0; i < nl; ++i) {
    if (d->levels[i] < 0) fail(LFX_EINVAL, "negative level");
    p.levels.push_back(d->levels[i]);
  }
  return p;
}

}  // namespace lfx
