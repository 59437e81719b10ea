// Host side of lifted-hip: wavelet table, step-program resolution, the ND
// drivers (which lines, which levels, in which order) and the C ABI.
//
// ND semantics follow src/lifted-drivers-inl.hpp:417-610:
//  * separable (levels per axis): axes in the given order, each a full
//    multi-level transform of every line along it; the first axis reads the
//    input, later axes read the output (:459-485);
//  * level by level (one level count): for each level, each axis in the given
//    order transforms one level on the shrinking sub-box [0, ns) of the
//    transformed axes; Inverse / ForwardAdjoint copy in -> out first and walk the
//    levels coarse -> fine (:530-609).
// The reference sorts dims by output stride only to order its memory walk
// (:432-454); results do not depend on that order, so the GPU plan enumerates
// lines directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lifted_hip.h"
#include "lfx_kernels.hpp"
#include "lfx_program.hpp"

namespace lfx {

static thread_local std::string g_err;

struct Error {
  int code;
  std::string msg;
};
[[noreturn]] static void fail(int code, const std::string& m) { throw Error{code, m}; }
static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(LFX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// wavelet table (generated from the reference's src/lifted-wavelets.hpp)

static std::vector<WaveletDefL> build_table() {
  std::vector<WaveletDefL> v;
  WaveletDefL* cur = nullptr;
  auto add = [&](int kind, int target, int off, int n, int sign, const long double* t, int nt) {
    StepDefL& s = cur->steps[cur->nsteps++];
    s = StepDefL{};
    s.kind = kind;
    s.target = target;
    s.offset = off;
    s.ntaps = n;
    s.sign = sign;
    for (int k = 0; k < nt; ++k) s.taps[k] = t[k];
  };
  const int S = 0, D = 1;
#define LFX_WAVELET(nm, idv, w)  \
  v.emplace_back();              \
  cur = &v.back();               \
  *cur = WaveletDefL{};          \
  cur->id = idv;                 \
  cur->name = #nm;               \
  cur->width = w;
#define LFX_END(nm)
#define LFX_GENERIC(b, off, ...)                                         \
  {                                                                      \
    const long double t_[] = {__VA_ARGS__};                              \
    const int n_ = int(sizeof(t_) / sizeof(t_[0]));                      \
    add(kGeneric, b, off, n_, 1, t_, n_);                                \
  }
#define LFX_UNIT(b, off, sg) add(kUnit, b, off, 1, sg, nullptr, 0);
#define LFX_REPEAT(b, off, n, val)   \
  {                                  \
    const long double t_[] = {val};  \
    add(kRepeat, b, off, n, 1, t_, 1); \
  }
#define LFX_SCALE(a, b)                 \
  {                                     \
    const long double t_[] = {a, b};    \
    add(kScale, 0, 0, 0, 1, t_, 2);     \
  }
#include "wavelet_table.inc"
#undef LFX_WAVELET
#undef LFX_END
#undef LFX_GENERIC
#undef LFX_UNIT
#undef LFX_REPEAT
#undef LFX_SCALE
  return v;
}

static std::mutex g_mu;
static std::vector<WaveletDefL>& table() {
  static std::vector<WaveletDefL> t = build_table();
  return t;
}
static std::vector<WaveletDefL> g_custom;  // ids 0x80000000 + index
constexpr uint32_t kCustomBase = 0x80000000u;

static WaveletDefL find_wavelet(uint32_t id) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (id >= kCustomBase) {
    if (id - kCustomBase < g_custom.size()) return g_custom[id - kCustomBase];
  } else {
    for (auto& w : table())
      if (w.id == id) return w;
  }
  fail(LFX_EINVAL, "unknown wavelet id " + std::to_string(id));
}

// Resolve a wavelet for one transform type (see lfx_program.hpp).
template <class C>
static DProgram<C> resolve(const WaveletDefL& w, int tf) {
  DProgram<C> p{};
  p.width = (int32_t)w.width;
  const bool fwd_stepping = tf == kForward || tf == kInverseAdjoint;
  const bool normal = tf == kForward || tf == kInverse;
  const bool fwd_tf = tf == kForward || tf == kForwardAdjoint;
  int nc = 0;
  for (int q = 0; q < w.nsteps; ++q) {
    const StepDefL& sd = w.steps[fwd_stepping ? q : w.nsteps - 1 - q];
    DStep& st = p.steps[p.nsteps++];
    st = DStep{};
    st.kind = sd.kind;
    st.coef = nc;
    if (sd.kind == kScale) {
      // ScaleStep (lifted-common.hpp:936-964): Forward/ForwardAdjoint multiply by
      // (s_mul, d_mul = 1/d_div), Inverse/InverseAdjoint by (s_div = 1/s_mul, d_div);
      // reciprocals in long double, then cast.
      p.coef[nc++] = fwd_tf ? (C)sd.taps[0] : (C)(1.0L / sd.taps[0]);
      p.coef[nc++] = fwd_tf ? (C)(1.0L / sd.taps[1]) : (C)sd.taps[1];
      continue;
    }
    const int N = sd.ntaps, o = sd.offset;
    const int maxo = N - 1 + o;
    const int n_front = o < 0 ? -o : 0, n_back = maxo < 0 ? 0 : maxo;
    const int o_r = -maxo;
    const int maxo_r = N - 1 + o_r;
    const int n_front_r = o_r < 0 ? -o_r : 0, n_back_r = maxo_r < 0 ? 0 : maxo_r;
    // updated band: normal transforms update the declared band, adjoints the other
    st.y_is_d = ((sd.target == 1) == normal) ? 1 : 0;
    st.adjoint = normal ? 0 : 1;
    st.off = normal ? o : o_r;
    st.N = N;
    st.nf = normal ? n_front : n_front_r;
    st.nb = normal ? n_back : n_back_r;
    st.fold_off = o;
    if (sd.kind == kUnit) {
      st.sub = ((sd.sign > 0) == fwd_tf) ? 0 : 1;
    } else {
      const bool sub = !fwd_tf;
      if (sd.kind == kGeneric) {
        for (int k = 0; k < N; ++k) {
          C c = (C)sd.taps[normal ? k : N - 1 - k];
          p.coef[nc++] = sub ? -c : c;
        }
      } else {
        C c = (C)sd.taps[0];
        p.coef[nc++] = sub ? -c : c;
      }
    }
    // reach of the stencil relative to the updated element (band index units)
    const int lo = -st.off, hi = st.off + N - 1;
    p.reach_lo = std::max(p.reach_lo, std::max(lo, 0));
    p.reach_hi = std::max(p.reach_hi, std::max(hi, 0));
  }
  return p;
}

// Device-resident programs, one per (device, wavelet, transform, compute type).
static std::map<std::tuple<int, uint32_t, int, int>, void*> g_dev_progs;

static const void* device_program(uint32_t wavelet, int tf, int storage) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  const int ctype = storage == kF64 ? 1 : 0;
  auto key = std::make_tuple(dev, wavelet, tf, ctype);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev_progs.find(key);
    if (it != g_dev_progs.end()) return it->second;
  }
  const WaveletDefL w = find_wavelet(wavelet);
  void* d = nullptr;
  if (ctype == 1) {
    DProgram<double> p = resolve<double>(w, tf);
    hip_check(hipMalloc(&d, sizeof(p)), "hipMalloc(program)");
    hip_check(hipMemcpy(d, &p, sizeof(p), hipMemcpyHostToDevice), "hipMemcpy(program)");
  } else {
    DProgram<float> p = resolve<float>(w, tf);
    hip_check(hipMalloc(&d, sizeof(p)), "hipMalloc(program)");
    hip_check(hipMemcpy(d, &p, sizeof(p), hipMemcpyHostToDevice), "hipMemcpy(program)");
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_dev_progs.find(key);
  if (it != g_dev_progs.end()) {
    (void)hipFree(d);
    return it->second;
  }
  g_dev_progs[key] = d;
  return d;
}

// max_level (lifted-common.hpp:1099-1108); -1 for width 1 (never terminates there)
static long long max_level_w(long long width, long long n) {
  if (width == 0) return 0;
  if (width == 1) return -1;
  if (n < width - 1) return 0;
  long long lvl = 0;
  while (n >= 2 * (width - 1)) {
    ++lvl;
    n = (n + 1) / 2;
  }
  return lvl;
}

static bool has_lift_steps(const WaveletDefL& w) {
  for (int i = 0; i < w.nsteps; ++i)
    if (w.steps[i].kind != kScale) return true;
  return false;
}

// Bands the reference cannot process: empty bands under a non-Zero BC divide
// by zero (Periodic) or loop forever (Symmetric, Reflect), as do bands of
// length 1 under Reflect.
static void check_bands(int bc, bool steps, long long n, long long L) {
  if (!steps || bc == kZero) return;
  long long m = n;
  for (long long l = 0; l < L; ++l) {
    const long long nd = m / 2, ns = m - nd;
    if (nd == 0 || (bc == kReflect && (nd < 2 || ns < 2)))
      fail(LFX_EINVAL, "level " + std::to_string(l + 1) + " of a length-" + std::to_string(n) +
                           " axis leaves a band too short for this boundary condition");
    m = ns;
  }
}

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

static Plan make_plan(const lfx_desc* d) {
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

static long long line_limit(int storage) {
  // longest line the whole-line kernel takes: LDS = 1.5 n elements
  const size_t cb = compute_bytes(storage);
  return (long long)(kMaxLdsBytes / (cb * 3 / 2)) - 2;
}

// Build the line map of all lines along `ax` of the box `shp`.
static LineMap line_map(int ndim, const long long* shp, const long long* sst, const long long* dst, int ax,
                        long long& nlines) {
  LineMap m{};
  nlines = 1;
  for (int k = 0; k < ndim; ++k) {
    if (k == ax) continue;
    m.shape[m.nod] = shp[k];
    m.sst[m.nod] = sst[k];
    m.dstr[m.nod] = dst[k];
    ++m.nod;
    nlines *= shp[k];
  }
  m.s_ax = sst[ax];
  m.d_ax = dst[ax];
  return m;
}

static void copy_box(int storage, int ndim, const long long* shp, const void* src, const long long* sst, void* dst,
                     const long long* dstr, hipStream_t st) {
  CopyMap cm{};
  cm.nd = ndim;
  long long total = 1;
  for (int k = 0; k < ndim; ++k) {
    cm.shape[k] = shp[k];
    cm.sst[k] = sst[k];
    cm.dstr[k] = dstr[k];
    total *= shp[k];
  }
  hip_check(launch_copy(storage, src, dst, cm, total, st), "copy kernel");
}

// One 1-D transform of L levels (L == 1 for a single level) of all lines of `shp` along `ax`.
static void transform_lines(const Plan& p, const void* prog, bool fwd_stepping, const long long* shp, int ax,
                            long long L, const void* src, const long long* sst, void* dst, const long long* dstr,
                            hipStream_t st) {
  long long nlines = 0;
  LineMap lm = line_map(p.ndim, shp, sst, dstr, ax, nlines);
  const long long n = shp[ax];
  if (nlines == 0 || n == 0) return;
  if (n > line_limit(p.storage))
    fail(LFX_EUNSUPPORTED, "line length " + std::to_string(n) + " exceeds the single-workgroup limit");
  hip_check(launch_line(p.storage, fwd_stepping, prog, p.bc, n, (int)L, src, dst, lm, nlines, st), "line kernel");
}

static void run_plan(const Plan& p, const void* in, void* out, hipStream_t st) {
  const WaveletDefL w = find_wavelet(p.wavelet);
  const bool steps = has_lift_steps(w);
  const bool fwd_stepping = p.tf == kForward || p.tf == kInverseAdjoint;
  long long total = 1;
  for (int k = 0; k < p.ndim; ++k) total *= p.shape[k];
  if (in == out) {
    for (int k = 0; k < p.ndim; ++k)
      if (p.shape[k] > 1 && p.sin[k] != p.sout[k])
        fail(LFX_EINVAL, "in-place transforms need identical input and output strides");
  }
  const void* prog = device_program(p.wavelet, p.tf, p.storage);

  if (p.separable) {
    // resolve levels and validate before launching anything
    std::vector<long long> L(p.axes.size());
    for (size_t ia = 0; ia < p.axes.size(); ++ia) {
      const long long len = p.shape[p.axes[ia]];
      L[ia] = p.levels[ia];
      if (L[ia] == 0) {
        L[ia] = max_level_w(w.width, len);
        if (L[ia] < 0) fail(LFX_EINVAL, "level 0 (max level) is undefined for a width-1 wavelet");
      }
      check_bands(p.bc, steps, len, L[ia]);
    }
    if (total == 0) return;
    for (size_t ia = 0; ia < p.axes.size(); ++ia) {
      const void* src = ia == 0 ? in : out;
      const long long* sst = ia == 0 ? p.sin : p.sout;
      if (L[ia] == 0) {
        // zero levels: identity (the reference leaves `out` unwritten here)
        if (ia == 0 && in != out) copy_box(p.storage, p.ndim, p.shape, in, p.sin, out, p.sout, st);
        continue;
      }
      if (p.flags & LFX_FLAG_PLAIN) fail(LFX_EUNSUPPORTED, "plain per-level path not built yet");
      transform_lines(p, prog, fwd_stepping, p.shape, p.axes[ia], L[ia], src, sst, out, p.sout, st);
    }
    return;
  }

  // level by level
  long long L = p.levels[0];
  if (L == 0) {
    long long mn = p.shape[p.axes[0]];
    for (int a : p.axes) mn = std::min(mn, p.shape[a]);
    L = max_level_w(w.width, mn);
    if (L < 0) fail(LFX_EINVAL, "level 0 (max level) is undefined for a width-1 wavelet");
  }
  for (int a : p.axes) check_bands(p.bc, steps, p.shape[a], L);
  if (total == 0) return;
  std::vector<std::vector<long long>> shapes(L > 0 ? L : 0);
  if (L > 0) shapes[0].assign(p.shape, p.shape + p.ndim);
  for (long long l = 1; l < L; ++l) {
    shapes[l] = shapes[l - 1];
    for (int a : p.axes) shapes[l][a] = shapes[l - 1][a] - shapes[l - 1][a] / 2;
  }
  const void* src = in;
  const long long* sst = p.sin;
  if (L == 0) {
    if (in != out) copy_box(p.storage, p.ndim, p.shape, in, p.sin, out, p.sout, st);
    return;
  }
  if (!fwd_stepping && in != out) {
    copy_box(p.storage, p.ndim, p.shape, in, p.sin, out, p.sout, st);
    src = out;
    sst = p.sout;
  }
  for (long long il = 0; il < L; ++il) {
    const std::vector<long long>& shp = fwd_stepping ? shapes[il] : shapes[L - 1 - il];
    for (int a : p.axes) {
      transform_lines(p, prog, fwd_stepping, shp.data(), a, 1, src, sst, out, p.sout, st);
      src = out;
      sst = p.sout;
    }
  }
}

// element span [lo, hi] (offsets relative to the base pointer) of a strided array
static void span_of(int ndim, const long long* shape, const long long* str, long long& lo, long long& hi) {
  lo = 0;
  hi = 0;
  for (int k = 0; k < ndim; ++k) {
    if (shape[k] == 0) { lo = 0; hi = -1; return; }
    const long long e = (shape[k] - 1) * str[k];
    if (e < 0) lo += e; else hi += e;
  }
}

static void run_host(const lfx_desc* d, const void* in, void* out) {
  Plan p = make_plan(d);
  const size_t es = storage_bytes(p.storage);
  long long ilo, ihi, olo, ohi;
  span_of(p.ndim, p.shape, p.sin, ilo, ihi);
  span_of(p.ndim, p.shape, p.sout, olo, ohi);
  if (ihi < ilo) {  // empty array: still validate levels / bands
    run_plan(p, nullptr, nullptr, nullptr);
    return;
  }
  const size_t ibytes = (size_t)(ihi - ilo + 1) * es, obytes = (size_t)(ohi - olo + 1) * es;
  hipStream_t st = nullptr;
  void *din = nullptr, *dout = nullptr;
  const bool alias = in == out;
  try {
    hip_check(hipMalloc(&din, ibytes), "hipMalloc(in)");
    hip_check(hipMemcpy(din, (const char*)in + ilo * (long long)es, ibytes, hipMemcpyHostToDevice), "copy in");
    if (alias) {
      dout = din;
    } else {
      hip_check(hipMalloc(&dout, obytes), "hipMalloc(out)");
      // keep elements of the output span that are not part of the array
      hip_check(hipMemcpy(dout, (const char*)out + olo * (long long)es, obytes, hipMemcpyHostToDevice), "copy out");
    }
    const void* din0 = (const char*)din - ilo * (long long)es;
    void* dout0 = (char*)dout - olo * (long long)es;
    run_plan(p, din0, alias ? const_cast<void*>(din0) : dout0, st);
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    hip_check(hipMemcpy((char*)out + olo * (long long)es, dout, obytes, hipMemcpyDeviceToHost), "copy back");
  } catch (...) {
    if (din) (void)hipFree(din);
    if (dout && !alias) (void)hipFree(dout);
    throw;
  }
  (void)hipFree(din);
  if (!alias) (void)hipFree(dout);
}

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return LFX_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return LFX_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return LFX_EINVAL;
  }
}

template <class T>
static int lwt_overload(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                        const ptrdiff_t* sin, const ptrdiff_t* sout, size_t naxes, const size_t* axes,
                        const size_t* levels, bool separable, const T* in, T* out) {
  return guarded([&] {
    if (ndim < 1 || ndim > (size_t)kMaxDims) fail(LFX_EINVAL, "ndim must be in 1..8");
    std::vector<int64_t> shp(ndim), si(ndim), so(ndim), ax(naxes), lv(separable ? naxes : 1);
    for (size_t i = 0; i < ndim; ++i) {
      shp[i] = (int64_t)shape[i];
      si[i] = sin[i];
      so[i] = sout[i];
    }
    for (size_t i = 0; i < naxes; ++i) ax[i] = (int64_t)axes[i];
    for (size_t i = 0; i < lv.size(); ++i) lv[i] = (int64_t)levels[i];
    lfx_desc d{};
    d.wavelet = wavelet;
    d.transform = transform;
    d.bc = bc;
    d.ndim = (int32_t)ndim;
    d.shape = shp.data();
    d.stride_in = si.data();
    d.stride_out = so.data();
    d.naxes = (int32_t)naxes;
    d.axes = ax.data();
    d.separable = separable ? 1 : 0;
    d.levels = lv.data();
    d.storage = sizeof(T) == 8 ? kF64 : kF32;
    run_host(&d, in, out);
  });
}

}  // namespace lfx

using namespace lfx;

extern "C" {

const char* lfx_last_error(void) { return g_err.c_str(); }

int lfx_lwt(const lfx_desc* desc, const void* in_dev, void* out_dev, void* stream) {
  return guarded([&] {
    Plan p = make_plan(desc);
    run_plan(p, in_dev, out_dev, (hipStream_t)stream);
  });
}

int lfx_lwt_host(const lfx_desc* desc, const void* in_host, void* out_host) {
  return guarded([&] { run_host(desc, in_host, out_host); });
}

int lfx_lwt_f32(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                size_t level, const float* data_in, float* data_out, size_t) {
  return lwt_overload<float>(wavelet, transform, bc, ndim, shape, stride_in, stride_out, naxes, axes, &level, false,
                             data_in, data_out);
}
int lfx_lwt_f64(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                size_t level, const double* data_in, double* data_out, size_t) {
  return lwt_overload<double>(wavelet, transform, bc, ndim, shape, stride_in, stride_out, naxes, axes, &level,
                              false, data_in, data_out);
}
int lfx_lwt_sep_f32(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                    const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                    const size_t* levels, const float* data_in, float* data_out, size_t) {
  return lwt_overload<float>(wavelet, transform, bc, ndim, shape, stride_in, stride_out, naxes, axes, levels, true,
                             data_in, data_out);
}
int lfx_lwt_sep_f64(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                    const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                    const size_t* levels, const double* data_in, double* data_out, size_t) {
  return lwt_overload<double>(wavelet, transform, bc, ndim, shape, stride_in, stride_out, naxes, axes, levels, true,
                              data_in, data_out);
}

int64_t lfx_max_level(uint32_t wavelet, int64_t n) {
  int64_t r = LFX_EINVAL;
  int rc = guarded([&] {
    const WaveletDefL w = find_wavelet(wavelet);
    r = max_level_w(w.width, n);
    if (r < 0) fail(LFX_EINVAL, "max_level is undefined for a width-1 wavelet");
  });
  return rc == LFX_OK ? r : rc;
}

int64_t lfx_wavelet_width(uint32_t wavelet) {
  int64_t r = 0;
  int rc = guarded([&] { r = find_wavelet(wavelet).width; });
  return rc == LFX_OK ? r : rc;
}

int lfx_program_create(const lfx_step* steps, int32_t nsteps, uint32_t width, uint32_t* wavelet_id) {
  return guarded([&] {
    if (!steps || nsteps < 0 || nsteps > kMaxSteps) fail(LFX_EINVAL, "1..16 steps are supported");
    WaveletDefL w{};
    w.width = width;
    w.name = "custom";
    for (int i = 0; i < nsteps; ++i) {
      const lfx_step& s = steps[i];
      if (s.kind < 0 || s.kind > 3) fail(LFX_EINVAL, "unknown step kind");
      if (s.kind != kScale && (s.ntaps < 1 || s.ntaps > kMaxTaps)) fail(LFX_EINVAL, "taps must be 1..16");
      if (s.kind == kUnit && s.ntaps != 1) fail(LFX_EINVAL, "unit steps have one tap");
      if (s.target != 0 && s.target != 1) fail(LFX_EINVAL, "target must be 0 (s) or 1 (d)");
      StepDefL& d = w.steps[w.nsteps++];
      d = StepDefL{};
      d.kind = s.kind;
      d.target = s.target;
      d.offset = s.offset;
      d.ntaps = s.kind == kScale ? 0 : s.ntaps;
      d.sign = s.sign >= 0 ? 1 : -1;
      for (int k = 0; k < kMaxTaps; ++k) d.taps[k] = (long double)s.taps[k];
    }
    std::lock_guard<std::mutex> lk(g_mu);
    w.id = kCustomBase + (uint32_t)g_custom.size();
    g_custom.push_back(w);
    *wavelet_id = w.id;
  });
}

int lfx_program_dump(uint32_t wavelet, uint32_t transform, int32_t storage, double* out, int32_t cap_steps) {
  int n = 0;
  int rc = guarded([&] {
    if (transform > 3) fail(LFX_EINVAL, "unknown transform");
    const WaveletDefL w = find_wavelet(wavelet);
    auto dump = [&](auto prog) {
      for (int i = 0; i < prog.nsteps && i < cap_steps; ++i) {
        const DStep& s = prog.steps[i];
        double* o = out + 25 * i;
        o[0] = s.kind; o[1] = s.y_is_d; o[2] = s.adjoint; o[3] = s.sub; o[4] = s.off; o[5] = s.N;
        o[6] = s.nf; o[7] = s.nb; o[8] = s.fold_off;
        const int nc = s.kind == kScale ? 2 : (s.kind == kGeneric ? s.N : (s.kind == kRepeat ? 1 : 0));
        for (int k = 0; k < 16; ++k) o[9 + k] = k < nc ? (double)prog.coef[s.coef + k] : 0.0;
        ++n;
      }
    };
    if (storage == kF64) dump(resolve<double>(w, (int)transform));
    else dump(resolve<float>(w, (int)transform));
  });
  return rc == LFX_OK ? n : rc;
}

}  // extern "C"
