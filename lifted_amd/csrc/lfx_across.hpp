// Lines along a strided axis ("across" lines, the reference's Across step
// variants: src/lifted-steps-inl.hpp:502, :756, :1013, :1180, chosen at
// src/lifted-drivers-inl.hpp:475-482) without a transpose pass: a workgroup
// takes kLines lines that are neighbours along the array's contiguous dim,
// reads them into an LDS slab position by position (kLines consecutive
// elements per position: coalesced), runs every level of every line from the
// slab with the row-mode register tiles (lfx_reg.hpp: four lines per
// wavefront, 16 lanes of 32 samples per line, DPP row shifts, Zero boundary),
// in place, and writes the slab back the same way. Each sample is read once
// and written once per axis; the slab never leaves the CU.
//
// Slab layout: line c, sample p at virtual float index c * kMaxN + p, stored
// at phys(index) = index + index / 32 + index / 512. The 32-sample chunks of
// a lane then start in distinct banks, and the lines sit 529 floats apart (an
// odd number of banks), so both the row tiles' chunk accesses and the
// position-by-position loads and stores meet at most two-way bank conflicts.
#pragma once
#include "lfx_across_types.hpp"
#include "lfx_reg.hpp"

namespace lfx {
namespace acr {

// First line of block `blk` along the contiguous dim and the element offsets
// of its slice in the input and the output.
LFX_HDI void block_origin(const Args& A, long long blk, long long& c0, long long& oi, long long& oo) {
  long long b = blk / A.ncb;
  c0 = (blk % A.ncb) * kLines;
  oi = c0;
  oo = c0;
  for (int kk = 0; kk < kMaxDims; ++kk) {
    const int k = kMaxDims - 1 - kk;
    if (k < A.nb) {
      const long long q = b % A.bshape[k];
      b /= A.bshape[k];
      oi += q * A.bs_in[k];
      oo += q * A.bs_out[k];
    }
  }
}

// Thread tid of kThreads moves line tid % kLines of a block at positions
// tid / kLines + j kThreads / kLines (j < kPer): a wavefront's access covers
// kLines consecutive elements (coalesced) at 64 / kLines positions.
//
// fetch: v <- the block's lines (zeros past the dim's end or the line's end);
// the loads are only issued here, so a caller can overlap them with work on
// the previous block. Every load is issued unconditionally, from a position
// clamped into the block (column 0 of a block and position 0 of a line always
// exist), and the values outside are zeroed afterwards: a branch around each
// load would make the compiler wait for every load before issuing the next.
// put: slab <- v.
template <class S>
__device__ __forceinline__ void fetch_t(const Args& A, long long c0, long long oi, int tid, float (&v)[kPer]) {
  const int c = tid % kLines, p0 = tid / kLines, n = (int)A.n;
  const bool cin = c < A.nc - c0;
  const S* in = static_cast<const S*>(A.in) + oi + (cin ? c : 0);
  if (n == kMaxN) {
    // every position of the block in the line: one stride between loads
    const S* q = in + (long long)p0 * A.sax_in;
    const long long step = (long long)kPosStep * A.sax_in;
#pragma unroll
    for (int j = 0; j < kPer; ++j) v[j] = dev::load_as<S, float>(q, j * step);
    if (!cin) {
#pragma unroll
      for (int j = 0; j < kPer; ++j) v[j] = 0.0f;
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int p = p0 + j * kPosStep;
    v[j] = dev::load_as<S, float>(in, (long long)(p < n ? p : n - 1) * A.sax_in);
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int p = p0 + j * kPosStep;
    if (!(cin && p < n)) v[j] = 0.0f;
  }
}
__device__ __forceinline__ void fetch(const Args& A, long long c0, long long oi, int tid, float (&v)[kPer]) {
  if (A.st_in == kBF16)
    fetch_t<dev::bf16_t>(A, c0, oi, tid, v);
#ifndef LFX_CPU_SIM
  else if (A.st_in == kF16)
    fetch_t<dev::f16_t>(A, c0, oi, tid, v);
#endif
  else
    fetch_t<float>(A, c0, oi, tid, v);
}
// Position p0 + kPosStep j of line c lies at physical slab index
// phys(c kMaxN + p0 + kPosStep j) = 529 c + p0 + kPosStep j + kPosStep j / 32
// (p0 < kPosStep, which divides 32): one base address per thread and
// compile-time offsets.
LFX_HDI int slab_line(int c, int p0) { return c * (kMaxN + kMaxN / 32 + 1) + p0; }
LFX_HDI int slab_off(int j) { return kPosStep * j + kPosStep * j / 32; }
__device__ __forceinline__ void put(const Args& A, float* slab, int tid, const float (&v)[kPer]) {
  const int c = tid % kLines, p0 = tid / kLines, n = (int)A.n;
  float* b = slab + slab_line(c, p0);
  if (n == kMaxN) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) b[slab_off(j)] = v[j];
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int p = p0 + j * kPosStep;
    if (p < n) b[slab_off(j)] = v[j];
  }
}

// The block's lines <- slab (same thread mapping).
template <class S>
__device__ __forceinline__ void store_block_t(const Args& A, const float* slab, long long c0, long long oo, int tid) {
  const int c = tid % kLines, p0 = tid / kLines, n = (int)A.n;
  if (c >= A.nc - c0) return;
  S* out = static_cast<S*>(A.out) + oo + c;
  const float* b = slab + slab_line(c, p0);
  if (n == kMaxN) {
    S* q = out + (long long)p0 * A.sax_out;
    const long long step = (long long)kPosStep * A.sax_out;
#pragma unroll
    for (int j = 0; j < kPer; ++j) dev::store_as<S, float>(q, j * step, b[slab_off(j)]);
    return;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int p = p0 + j * kPosStep;
    if (p < n) dev::store_as<S, float>(out, (long long)p * A.sax_out, b[slab_off(j)]);
  }
}
__device__ __forceinline__ void store_block(const Args& A, const float* slab, long long c0, long long oo, int tid) {
  if (A.st_out == kBF16)
    store_block_t<dev::bf16_t>(A, slab, c0, oo, tid);
#ifndef LFX_CPU_SIM
  else if (A.st_out == kF16)
    store_block_t<dev::f16_t>(A, slab, c0, oo, tid);
#endif
  else
    store_block_t<float>(A, slab, c0, oo, tid);
}

// Every level of the four lines of slab rows [4 w, 4 w + 4) (virtual base
// vb = slab-virtual index 4 w kMaxN, passed as a pointer the policy maps).
template <class P, class W, int TF>
__device__ __forceinline__ void run_lines(const Args& A, float* vb) {
  const DProgram<float>* prog = static_cast<const DProgram<float>*>(A.prog);
  const reg::RowArgs ra{{kMaxN, kMaxN, kMaxN}, 4};
  if constexpr (TF == kForward || TF == kInverseAdjoint) {
    for (int i = 0; i < A.npass; ++i)
      reg::fwd_tile<P, W, TF>(prog, A.pm[i], A.pk[i], 0, A.pm[i], 0, (const float*)vb, vb, vb, nullptr, ra);
  } else {
    for (int i = A.npass - 1; i >= 0; --i)
      reg::rev_tile<P, W, TF>(prog, A.pm[i], A.pk[i], 0, A.pm[i], 0, (const float*)vb, (const float*)vb, vb, nullptr,
                              ra);
  }
}

}  // namespace acr
}  // namespace lfx
