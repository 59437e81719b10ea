// Host planning of a two-axis level-by-level transform as one launch per level
// of lfx_lvl2d.hpp's tiles (shared by the library and the CPU simulation).
//
// Level l works on the sub-box of the two axes entering it. A forward-stepping
// transform (Forward, InverseAdjoint) walks the levels fine -> coarse: level 1
// reads the input, each level writes its detail quadrants into the output and
// its ss quadrant into workspace A or B in turn (the last level into the
// output), which the next level reads. A reverse-stepping one (Inverse,
// ForwardAdjoint) walks coarse -> fine: the coarsest level reads its ss
// quadrant from the input, each level reads the details from the input and
// writes its interleaved sub-box to a workspace in turn (the finest level into
// the output), which the next finer level reads as its ss quadrant. The
// workspaces are dense [batch dims..., rows, cols] in the storage type, so a
// 16-bit transform rounds every level's result once, like the other
// level-by-level paths.
#pragma once
#include <cstdint>
#include <vector>

#include "lfx_lvl2d_types.hpp"
#include "lfx_schedule.hpp"

namespace lfx {

// Band lengths entering each level along an axis of length n.
inline std::vector<long long> level_lengths(long long n, long long L) {
  std::vector<long long> v;
  for (long long l = 0; l < L; ++l) {
    v.push_back(n);
    n = n - n / 2;
  }
  return v;
}

// Elements of one workspace: the largest ss quadrant of a forward level or
// interleaved sub-box of a reverse one (level 2's box).
inline long long lvl2d_ws_elems(int ndim, const long long* shape, int ra, int ca) {
  long long e = 1;
  for (int k = 0; k < ndim; ++k) e *= (k == ra || k == ca) ? shape[k] - shape[k] / 2 : shape[k];
  return e;
}

// Levels of the transform (empty when some level's bands admit no tile halo:
// the caller then takes another path). ra: register axis (transformed first),
// ca: lane axis (unit stride expected on both sides for coalescing).
// prog: device program of (wavelet, transform) for float32 arithmetic; hp: the
// same on the host.
inline std::vector<l2d::Level> lvl2d_plan(const DProgram<float>& hp, bool fwd, int ndim, const long long* shape,
                                          int ra, int ca, long long L, const void* in, const long long* sin,
                                          void* out, const long long* sout, int storage, void* ws0, void* ws1,
                                          const void* prog, uint32_t wavelet) {
  std::vector<l2d::Level> lv;
  if (sin[ca] != 1 || sout[ca] != 1) return lv;  // lanes run along a unit-stride axis
  // tiles address their rows as 32-bit byte offsets from a window origin:
  // positive row strides, and a window's span within 31 bits
  if (sin[ra] <= 0 || sout[ra] <= 0) return lv;
  {
    const long long es_ = storage == kF32 ? 4 : 2;
    for (const long long sr : {sin[ra], sout[ra], shape[ca]})
      if ((2 * l2d::kRows * sr + 2 * l2d::kCols) * es_ >= (1ll << 31)) return lv;
  }
  const std::vector<long long> mr = level_lengths(shape[ra], L), mc = level_lengths(shape[ca], L);
  std::vector<long long> Hr(L), Hc(L);
  for (long long l = 0; l < L; ++l) {
    Hr[l] = lvl2d_axis_halo(hp, fwd, mr[l], l2d::kRows);
    Hc[l] = lvl2d_axis_halo(hp, fwd, mc[l], l2d::kCols);
    if (Hr[l] < 0 || Hc[l] < 0) return lv;
  }
  // batch dims: every other dim
  l2d::Level base{};
  base.nbatch = 1;
  std::vector<int> bdims;
  for (int k = 0; k < ndim; ++k)
    if (k != ra && k != ca) {
      base.bshape[base.nb++] = shape[k];
      base.nbatch *= shape[k];
      bdims.push_back(k);
    }
  const long long es = storage == kF32 ? 4 : 2;
  // vector accesses: the pointer and every stride but the lane axis's are
  // multiples of kLaneW elements
  auto vec_ok = [&](const void* p, const l2d::Arr& a) {
    if (reinterpret_cast<uintptr_t>(p) % (uintptr_t)(l2d::kLaneW * es) != 0 || a.sr % l2d::kLaneW != 0) return 0;
    for (int i = 0; i < base.nb; ++i)
      if (a.bs[i] % l2d::kLaneW != 0) return 0;
    return 1;
  };
  auto user = [&](const void* p, const long long* str) {
    l2d::Arr a{};
    a.p = const_cast<void*>(p);
    a.st = storage;
    a.sr = str[ra];
    a.sc = str[ca];
    for (size_t i = 0; i < bdims.size(); ++i) a.bs[i] = str[bdims[i]];
    a.v = vec_ok(p, a);
    return a;
  };
  auto dense = [&](void* p, long long rows, long long cols) {
    l2d::Arr a{};
    a.p = p;
    a.st = storage;
    a.sr = cols;
    a.sc = 1;
    long long acc = rows * cols;
    for (int i = (int)bdims.size() - 1; i >= 0; --i) {
      a.bs[i] = acc;
      acc *= shape[bdims[i]];
    }
    a.v = vec_ok(p, a);
    return a;
  };
  void* ws[2] = {ws0, ws1};
  for (long long il = 0; il < L; ++il) {
    const long long l = fwd ? il : L - 1 - il;
    l2d::Level v = base;
    v.mr = mr[l];
    v.mc = mc[l];
    v.Hr = Hr[l];
    v.Hc = Hc[l];
    v.Or = l2d::kRows - 2 * Hr[l];
    v.Oc = l2d::kCols - 2 * Hc[l];
    v.Tr = (v.mr + v.Or - 1) / v.Or;
    v.Tc = (v.mc + v.Oc - 1) / v.Oc;
    v.ntiles = v.nbatch * v.Tr * v.Tc;
    if (v.ntiles >= (1ll << 31)) return std::vector<l2d::Level>();
    v.f_tc = l2d::make_fdiv((uint32_t)v.Tc);
    v.f_per = l2d::make_fdiv((uint32_t)(v.Tr * v.Tc));
    for (int k = 0; k < v.nb; ++k) v.f_b[k] = l2d::make_fdiv((uint32_t)v.bshape[k]);
    v.prog = prog;
    v.wavelet = wavelet;
    const long long nsr = v.mr - v.mr / 2, nsc = v.mc - v.mc / 2;
    if (fwd) {
      v.a[0] = il == 0 ? user(in, sin) : dense(ws[(il - 1) & 1], v.mr, v.mc);
      v.a[1] = user(out, sout);
      v.a[2] = il == L - 1 ? user(out, sout) : dense(ws[il & 1], nsr, nsc);
    } else {
      v.a[0] = user(in, sin);
      v.a[2] = il == 0 ? user(in, sin) : dense(ws[(il - 1) & 1], nsr, nsc);
      v.a[1] = il == L - 1 ? user(out, sout) : dense(ws[il & 1], v.mr, v.mc);
    }
    lv.push_back(v);
  }
  return lv;
}

}  // namespace lfx
