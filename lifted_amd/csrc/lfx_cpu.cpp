// CPU product path of lifted-hip (liblifted_cpu.so): the GPU kernels' line
// bodies (lfx_bodies.hpp) compiled for the host, one line per task, the lines
// of an axis split evenly over a pool of threads, as the reference partitions
// them (src/lifted-common.hpp:446-479, 655-685). Lines are whole on the CPU
// (no LDS limit), so every line runs the whole-line bodies with its bands in
// per-thread buffers -- the same operations in the same order as the GPU, so
// the results are identical.
//
// ND drivers follow src/lifted-drivers-inl.hpp:417-610: separable (levels per
// axis): each axis in order, a full multi-level transform of every line (the
// first axis reads the input, later axes the output); level by level: for each
// level (coarse -> fine for Inverse / ForwardAdjoint, which first copy in ->
// out) each axis in order transforms one level of the sub-box [0, ns) of the
// transformed axes. Every line is gathered into a private buffer before its
// output is written, so in == out is safe.
#define LFX_CPU_SIM 1
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lifted_cpu.h"
#include "lfx_bodies.hpp"
#include "lfx_plan.hpp"
#include "lfx_table.hpp"

namespace lfx {

static thread_local std::string g_cpu_err;

// Base offsets of line `i` (lines enumerated over the dims other than ax,
// last fastest) in two arrays with strides s0 / s1.
static void line_offsets(int ndim, const long long* shp, int ax, const long long* s0, const long long* s1, long long i,
                         long long& o0, long long& o1) {
  o0 = o1 = 0;
  for (int k = ndim - 1; k >= 0; --k) {
    if (k == ax) continue;
    const long long q = i % shp[k];
    i /= shp[k];
    o0 += q * s0[k];
    o1 += q * s1[k];
  }
}

static int pool_threads(int n_threads, long long work) {
  int n = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<long long>(1, std::min<long long>(n, work));
}

// Runs f(begin, end) over [0, n) split evenly over the threads (the calling
// thread takes the first part).
template <class F>
static void parallel_lines(long long n, int threads, F&& f) {
  if (threads <= 1 || n <= 1) {
    f(0ll, n);
    return;
  }
  std::vector<std::thread> pool;
  const long long per = n / threads, extra = n % threads;
  long long b = 0;
  std::vector<std::pair<long long, long long>> parts;
  for (int t = 0; t < threads; ++t) {
    const long long e = b + per + (t < extra ? 1 : 0);
    parts.push_back({b, e});
    b = e;
  }
  for (int t = 1; t < threads; ++t) pool.emplace_back([&, t] { f(parts[t].first, parts[t].second); });
  f(parts[0].first, parts[0].second);
  for (auto& th : pool) th.join();
}

// All lines of box `shp` along `ax`: L levels (forward or reverse stepping).
template <class C>
static void lines(const DProgram<C>& prog, int bc, bool fwd, int ndim, const long long* shp, int ax, long long L,
                  const C* src, const long long* sst, C* dst, const long long* dstr, int n_threads) {
  const long long n = shp[ax];
  long long nlines = 1;
  for (int k = 0; k < ndim; ++k)
    if (k != ax) nlines *= shp[k];
  if (nlines == 0 || n == 0 || L == 0) return;
  parallel_lines(nlines, pool_threads(n_threads, nlines), [&](long long b, long long e) {
    std::vector<C> buf((size_t)n), work((size_t)(n + (n - n / 2)));
    for (long long i = b; i < e; ++i) {
      long long oi, oo;
      line_offsets(ndim, shp, ax, sst, dstr, i, oi, oo);
      const C* in = src + oi;
      for (long long j = 0; j < n; ++j) buf[(size_t)j] = in[j * sst[ax]];
      C* P = work.data();
      C* Q = P + n;
      if (fwd)
        line_fwd_body<C, C, C>(&prog, bc, n, (int)L, buf.data(), 1, dst + oo, dstr[ax], P, Q);
      else
        line_rev_body<C, C, C, C>(&prog, bc, n, (int)L, buf.data(), 1, buf.data(), 1, dst + oo, dstr[ax], P, Q);
    }
  });
}

template <class C>
static void copy_box(int ndim, const long long* shp, const C* src, const long long* sst, C* dst,
                     const long long* dstr) {
  long long total = 1;
  for (int k = 0; k < ndim; ++k) total *= shp[k];
  for (long long e = 0; e < total; ++e) {
    long long r = e, so = 0, dof = 0;
    for (int k = ndim - 1; k >= 0; --k) {
      const long long q = r % shp[k];
      r /= shp[k];
      so += q * sst[k];
      dof += q * dstr[k];
    }
    dst[dof] = src[so];
  }
}

template <class C>
static void run_cpu(const Plan& p, const C* in, C* out, int n_threads) {
  const WaveletDefL w = find_wavelet(p.wavelet);
  const bool steps = has_lift_steps(w);
  const bool fwd = p.tf == kForward || p.tf == kInverseAdjoint;
  const DProgram<C> prog = resolve<C>(w, p.tf);
  long long total = 1;
  for (int k = 0; k < p.ndim; ++k) total *= p.shape[k];
  if (in == out)
    for (int k = 0; k < p.ndim; ++k)
      if (p.shape[k] > 1 && p.sin[k] != p.sout[k])
        fail(LFX_EINVAL, "in-place transforms need identical input and output strides");
  if (p.separable) {
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
      const C* src = ia == 0 ? in : out;
      const long long* sst = ia == 0 ? p.sin : p.sout;
      if (L[ia] == 0) {
        if (ia == 0 && in != out) copy_box<C>(p.ndim, p.shape, in, p.sin, out, p.sout);
        continue;
      }
      lines<C>(prog, p.bc, fwd, p.ndim, p.shape, p.axes[ia], L[ia], src, sst, out, p.sout, n_threads);
    }
    return;
  }
  long long L = p.levels[0];
  if (L == 0) {
    long long mn = p.shape[p.axes[0]];
    for (int a : p.axes) mn = std::min(mn, p.shape[a]);
    L = max_level_w(w.width, mn);
    if (L < 0) fail(LFX_EINVAL, "level 0 (max level) is undefined for a width-1 wavelet");
  }
  for (int a : p.axes) check_bands(p.bc, steps, p.shape[a], L);
  if (total == 0) return;
  if (L == 0) {
    if (in != out) copy_box<C>(p.ndim, p.shape, in, p.sin, out, p.sout);
    return;
  }
  std::vector<std::vector<long long>> shapes((size_t)L);
  shapes[0].assign(p.shape, p.shape + p.ndim);
  for (long long l = 1; l < L; ++l) {
    shapes[l] = shapes[l - 1];
    for (int a : p.axes) shapes[l][a] = shapes[l - 1][a] - shapes[l - 1][a] / 2;
  }
  const C* src = in;
  const long long* sst = p.sin;
  if (!fwd && in != out) {
    copy_box<C>(p.ndim, p.shape, in, p.sin, out, p.sout);
    src = out;
    sst = p.sout;
  }
  for (long long il = 0; il < L; ++il) {
    const std::vector<long long>& shp = fwd ? shapes[il] : shapes[L - 1 - il];
    for (int a : p.axes) {
      lines<C>(prog, p.bc, fwd, p.ndim, shp.data(), a, 1, src, sst, out, p.sout, n_threads);
      src = out;
      sst = p.sout;
    }
  }
}

}  // namespace lfx

using namespace lfx;

extern "C" {

int lfx_cpu_lwt(const lfx_desc* desc, const void* in_host, void* out_host, int32_t n_threads) {
  try {
    const Plan p = make_plan(desc);
    if (p.storage == kF64)
      run_cpu<double>(p, static_cast<const double*>(in_host), static_cast<double*>(out_host), n_threads);
    else if (p.storage == kF32)
      run_cpu<float>(p, static_cast<const float*>(in_host), static_cast<float*>(out_host), n_threads);
    else
      fail(LFX_EUNSUPPORTED, "the CPU path computes float32 and float64 arrays");
    return LFX_OK;
  } catch (const Error& e) {
    g_cpu_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_cpu_err = "host allocation failed";
    return LFX_ENOMEM;
  } catch (const std::exception& e) {
    g_cpu_err = e.what();
    return LFX_EINVAL;
  }
}

const char* lfx_cpu_last_error(void) { return g_cpu_err.c_str(); }

int64_t lfx_cpu_max_level(uint32_t wavelet, int64_t n) {
  try {
    const long long r = max_level_w(find_wavelet(wavelet).width, n);
    if (r < 0) fail(LFX_EINVAL, "max_level is undefined for a width-1 wavelet");
    return r;
  } catch (const Error& e) {
    g_cpu_err = e.msg;
    return e.code;
  }
}

}  // extern "C"
