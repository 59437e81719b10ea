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
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lifted_hip.h"
#include "lfx_kernels.hpp"
#include "lfx_program.hpp"
#include "lfx_tiles.hpp"
#include "lfx_error.hpp"
#include "lfx_table.hpp"
#include "lfx_schedule.hpp"
#include "lfx_noalloc.hpp"
#include "lfx_plan.hpp"
#include "lfx_across_plan.hpp"
#include "lfx_lvl2d_plan.hpp"

namespace lfx {

static std::mutex g_mu;  // guards the program cache and the workspaces
static thread_local std::string g_err;

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(LFX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
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

// ---------------------------------------------------------------------------
// workspace: one grow-only device buffer per device, used by one call at a
// time. A lease holds the workspace mutex from the moment it is handed out
// until the call has enqueued all its work and recorded the workspace's event
// on its stream; the next lease makes its stream wait on that event. So two
// host threads on two streams never share the buffer, and the buffer is only
// replaced (grown) by the thread that holds it, after a device
// synchronization. Under stream capture a lease neither waits on nor records
// the event (capture cannot); a buffer that a captured graph may use is never
// freed (kept until the process ends), and it is not grown during a capture.

static std::mutex g_ws_mu;  // taken after g_line_mu when both are held

static bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
  return cs != hipStreamCaptureStatusNone;
}

struct Workspace {
  void* p = nullptr;
  size_t bytes = 0;
  hipEvent_t ev = nullptr;
  bool captured = false;           // p is referenced by a captured graph
  std::vector<void*> retired;      // replaced buffers a captured graph may still use
};
static std::map<int, Workspace> g_ws;

struct WsLease {
  void* p = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  std::unique_lock<std::mutex> lk;
  WsLease() = default;
  WsLease(const WsLease&) = delete;
  WsLease& operator=(const WsLease&) = delete;
  ~WsLease() {
    if (ev) (void)hipEventRecord(ev, st);
  }  // the mutex is released after the event is recorded (member destroyed last)
};

static void lease_workspace(size_t bytes, hipStream_t st, WsLease& lease) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  const bool capturing = stream_capturing(st);
  lease.lk = std::unique_lock<std::mutex>(g_ws_mu);
  Workspace* wp = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    wp = &g_ws[dev];
  }
  Workspace& w = *wp;
  if (w.bytes < bytes) {
    if (capturing)
      fail(LFX_EUNSUPPORTED, "the transform needs a larger workspace (" + std::to_string(bytes) +
                                 " bytes) while the stream is capturing: run it once outside capture first");
    if (w.p) {
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (w.captured) w.retired.push_back(w.p);
      else (void)hipFree(w.p);
      w.p = nullptr;
      w.bytes = 0;
      w.captured = false;
    }
    const size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&w.p, want) != hipSuccess) {
      (void)hipGetLastError();
      fail(LFX_ENOMEM, "workspace allocation of " + std::to_string(want) + " bytes failed");
    }
    w.bytes = want;
  }
  if (!w.ev) {
    if (capturing) fail(LFX_EUNSUPPORTED, "run the transform once outside stream capture first");
    hip_check(hipEventCreateWithFlags(&w.ev, hipEventDisableTiming), "hipEventCreate");
  }
  if (!capturing) {
    hip_check(hipStreamWaitEvent(st, w.ev, 0), "hipStreamWaitEvent");
    lease.ev = w.ev;
  } else {
    w.captured = true;
  }
  lease.p = w.p;
  lease.st = st;
}

// LineMap3 over the non-axis dims of box `shp`, with per-array strides.
static LineMap3 line_map3(int ndim, const long long* shp, int ax, const long long* s0, long long a0,
                          const long long* s1, long long a1, const long long* s2, long long a2, long long& nlines) {
  LineMap3 m{};
  nlines = 1;
  for (int k = 0; k < ndim; ++k) {
    if (k == ax) continue;
    m.shape[m.nod] = shp[k];
    m.str[0][m.nod] = s0 ? s0[k] : 0;
    m.str[1][m.nod] = s1 ? s1[k] : 0;
    m.str[2][m.nod] = s2 ? s2[k] : 0;
    ++m.nod;
    nlines *= shp[k];
  }
  m.ax[0] = a0;
  m.ax[1] = a1;
  m.ax[2] = a2;
  return m;
}

// Dense strides of a scratch array holding, per line, `len` contiguous samples.
static std::vector<long long> dense_strides(int ndim, const long long* shp, int ax, long long len) {
  std::vector<long long> s(ndim, 0);
  long long acc = len;
  for (int k = ndim - 1; k >= 0; --k) {
    if (k == ax) continue;
    s[k] = acc;
    acc *= shp[k];
  }
  return s;
}

// Identical strides on both sides over a box without gaps (any dim order,
// negative strides allowed): the box is one memory range, copied flat (the
// streaming float4 copy for its 16-byte aligned part). Returns false when the
// box does not qualify.
static bool flat_copy(size_t es, int ndim, const long long* shp, const void* src, const long long* sst, void* dst,
                      const long long* dstr, hipStream_t st) {
  std::vector<std::pair<long long, long long>> dims;  // (|stride|, size) of dims longer than 1
  long long total = 1, lo = 0;
  for (int k = 0; k < ndim; ++k) {
    if (sst[k] != dstr[k]) return false;
    total *= shp[k];
    if (shp[k] == 1) continue;
    dims.push_back({sst[k] < 0 ? -sst[k] : sst[k], shp[k]});
    if (sst[k] < 0) lo += (shp[k] - 1) * sst[k];
  }
  if (total == 0) return true;
  std::sort(dims.begin(), dims.end());
  long long expect = 1;
  for (const auto& d : dims) {
    if (d.first != expect) return false;
    expect *= d.second;
  }
  const char* s = static_cast<const char*>(src) + lo * (long long)es;
  char* t = static_cast<char*>(dst) + lo * (long long)es;
  const long long bytes = total * (long long)es;
  long long head = 0;
  if ((((uintptr_t)s | (uintptr_t)t) & 15) == 0) {
    head = bytes & ~15ll;
    hip_check(launch_stream_copy(s, t, head, 0, st), "copy kernel");
  }
  if (bytes > head) hip_check(hipMemcpyAsync(t + head, s + head, (size_t)(bytes - head), hipMemcpyDeviceToDevice, st),
                              "hipMemcpyAsync");
  return true;
}

static void copy_box(int storage, int ndim, const long long* shp, const void* src, const long long* sst, void* dst,
                     const long long* dstr, hipStream_t st) {
  if (flat_copy(storage_bytes(storage), ndim, shp, src, sst, dst, dstr, st)) return;
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

struct Tuning {
  long long tile = 0;    // 0: default tile length
  int k = 0;             // 0: default levels per pass
  long long line = 0;    // 0: default whole-line threshold of register-tiled plans
  int streams = -1;      // side streams of a batched stage (-1: default)
  int nd_batch = -1;     // ND lines through batched plans (0: off)
  long long small = -1;  // LDS bytes up to which batch items run as one-wave workgroups (-1: default)
  int across = -1;       // strided axes through the across-line kernel (0: off; -1: default, on)
  int strip = 0;         // register window of float32 plans in samples (0: default kRegWindow; 1024, 2048)
};
static Tuning env_tuning() {
  Tuning t;
  if (const char* e = std::getenv("LFX_TILE")) t.tile = std::atoll(e);
  if (const char* e = std::getenv("LFX_KPASS")) t.k = std::atoi(e);
  if (const char* e = std::getenv("LFX_LINE")) t.line = std::atoll(e);
  if (const char* e = std::getenv("LFX_STREAMS")) t.streams = std::atoi(e);
  if (const char* e = std::getenv("LFX_SMALL")) t.small = std::atoll(e);
  if (const char* e = std::getenv("LFX_ND_BATCH")) t.nd_batch = std::atoi(e);
  if (const char* e = std::getenv("LFX_ACROSS")) t.across = std::atoi(e);
  if (const char* e = std::getenv("LFX_STRIP")) t.strip = std::atoi(e);
  return t;
}
static thread_local Tuning g_tuning = env_tuning();

// Fixed memory of a no-allocation call (lfx_lwt_noalloc): its scratch and
// program tables. Such a call runs the LDS line and tile kernels only (the
// batched-plan paths need allocations) and fails when its scratch would not
// fit.
struct RunCtx {
  void* ws = nullptr;
  size_t ws_bytes = 0;
  const void* progs_f32 = nullptr;
  const void* progs_f64 = nullptr;
};

static bool lines_fast(const Plan& p, const long long* shp, int ax, long long L, const void* src, const long long* sst,
                       void* dst, const long long* dstr, hipStream_t st);
static bool sep_chain(const Plan& p, const long long* shape, const long long* sin, const long long* sout,
                      const std::vector<int>& axes, const std::vector<long long>& L, const void* in, void* out,
                      hipStream_t st);

template <class C>
static void transform_lines_t(const Plan& p, const WaveletDefL& w, const void* prog, bool fwd, const long long* shp,
                              int ax, long long L, const void* src, const long long* sst, void* dst,
                              const long long* dstr, hipStream_t st, const RunCtx* ctx) {
  const long long n = shp[ax];
  long long nlines = 1;
  for (int k = 0; k < p.ndim; ++k)
    if (k != ax) nlines *= shp[k];
  if (nlines == 0 || n == 0 || L == 0) return;
  if (!ctx && lines_fast(p, shp, ax, L, src, sst, dst, dstr, st)) return;
  const DProgram<C> hp = resolve<C>(w, p.tf);
  // LFX_FLAG_PLAIN: one tiled pass per level (one launch per level, the band
  // round-trips through memory between levels); Periodic keeps whole lines
  const bool plain = (p.flags & LFX_FLAG_PLAIN) != 0 && p.bc != kPeriodic;
  const long long tile = tile_for_lines(g_tuning.tile, p.storage, nlines, n);
  const LineSchedule sc = plain ? schedule(hp, fwd, n, L, p.storage, p.bc, tile, 1, 1)
                                : schedule(hp, fwd, n, L, p.storage, p.bc, tile, g_tuning.k);
  const bool alias = src == dst;
  if (sc.global_line) {
    // one workgroup per line with its bands in global scratch (in place is
    // safe: a line is read in full before any of it is written), lines in
    // chunks whose scratch fits the budget
    const size_t per_line = (size_t)gline_elems(n) * sizeof(C);
    const size_t budget = ctx ? ctx->ws_bytes : ((size_t)256 << 20);
    const long long lpl = std::min<long long>(nlines, std::max<long long>(1, (long long)(budget / per_line)));
    const size_t need = (size_t)lpl * per_line;
    WsLease lease;
    void* g = nullptr;
    if (ctx) {
      if (need > ctx->ws_bytes)
        fail(LFX_EINVAL, "this transform needs " + std::to_string(need) +
                             " bytes of scratch; the no-allocation entry point holds at most " +
                             std::to_string(ctx->ws_bytes));
      g = ctx->ws;
    } else {
      lease_workspace(need, st, lease);
      g = lease.p;
    }
    LineLaunch ll{};
    ll.storage = p.storage; ll.fwd = fwd; ll.prog = prog; ll.bc = p.bc; ll.n = n; ll.L = (int)L;
    ll.a0 = src; ll.a1 = dst; ll.a2 = src;
    ll.lm = line_map3(p.ndim, shp, ax, sst, sst[ax], dstr, dstr[ax], sst, sst[ax], ll.nlines);
    hip_check(launch_gline(ll, g, lpl, st), "global line kernel");
    return;
  }
  if (sc.passes.empty()) {
    LineLaunch ll{};
    ll.storage = p.storage; ll.fwd = fwd; ll.prog = prog; ll.bc = p.bc; ll.n = n; ll.L = (int)L;
    ll.a0 = src; ll.a1 = dst; ll.a2 = src;
    ll.lm = line_map3(p.ndim, shp, ax, sst, sst[ax], dstr, dstr[ax], sst, sst[ax], ll.nlines);
    hip_check(launch_line(ll, st), "line kernel");
    return;
  }
  // workspace: [copy of the input if it aliases the output][scratch A][scratch B]
  const size_t es = storage_bytes(p.storage), cs = compute_bytes(p.storage);
  long long mmax = 0;
  for (size_t i = 1; i < sc.passes.size(); ++i) mmax = std::max(mmax, sc.passes[i].m0);
  mmax = std::max(mmax, sc.m_line);
  const size_t copy_bytes = alias ? (size_t)(nlines * n) * es : 0;
  const size_t copy_pad = (copy_bytes + 255) / 256 * 256;
  const size_t scr_bytes = (size_t)(nlines * mmax) * cs;
  const size_t scr_pad = (scr_bytes + 255) / 256 * 256;
  WsLease lease;
  char* ws = nullptr;
  if (ctx) {
    if (copy_pad + 2 * scr_pad > ctx->ws_bytes)
      fail(LFX_EINVAL, "this transform needs " + std::to_string(copy_pad + 2 * scr_pad) +
                           " bytes of scratch; the no-allocation entry point holds at most " +
                           std::to_string(ctx->ws_bytes));
    ws = (char*)ctx->ws;
  } else {
    lease_workspace(copy_pad + 2 * scr_pad, st, lease);
    ws = (char*)lease.p;
  }
  void* scr[2] = {ws + copy_pad, ws + copy_pad + scr_pad};
  const void* in = src;
  std::vector<long long> in_str(sst, sst + p.ndim);
  if (alias) {
    in_str = dense_strides(p.ndim, shp, ax, n);
    std::vector<long long> s2(in_str);
    s2[ax] = 1;
    copy_box(p.storage, p.ndim, shp, src, sst, ws, s2.data(), st);
    in = ws;
  }
  const long long in_ax = alias ? 1 : sst[ax];
  const size_t P = sc.passes.size();
  if (fwd) {
    const void* cur = in;
    bool cur_scr = false;
    std::vector<long long> cur_str = in_str;
    long long cur_ax = in_ax;
    for (size_t i = 0; i < P; ++i) {
      const PassPlan& pp = sc.passes[i];
      const bool last = (i + 1 == P) && sc.L_line == 0;
      const long long mk = band_len(pp.m0, pp.k);
      std::vector<long long> ks = dense_strides(p.ndim, shp, ax, mk);
      TileLaunch tl{};
      tl.storage = p.storage; tl.fwd = true; tl.prog = prog; tl.bc = p.bc;
      tl.ta = TileArgs{pp.m0, pp.T, pp.H, pp.ntiles, pp.k, 0};
      tl.a0 = cur; tl.a1 = dst; tl.a2 = last ? dst : scr[i & 1];
      tl.scr[0] = cur_scr; tl.scr[1] = false; tl.scr[2] = !last;
      tl.lm = line_map3(p.ndim, shp, ax, cur_str.data(), cur_ax, dstr, dstr[ax], last ? dstr : ks.data(),
                        last ? dstr[ax] : 1, tl.nlines);
      hip_check(launch_tile(tl, st), "tile kernel");
      cur = scr[i & 1];
      cur_scr = true;
      cur_str = ks;
      cur_ax = 1;
    }
    if (sc.L_line > 0) {
      LineLaunch ll{};
      ll.storage = p.storage; ll.fwd = true; ll.prog = prog; ll.bc = p.bc; ll.n = sc.m_line; ll.L = (int)sc.L_line;
      ll.a0 = cur; ll.a1 = dst; ll.a2 = cur;
      ll.scr[0] = true;
      ll.lm = line_map3(p.ndim, shp, ax, cur_str.data(), 1, dstr, dstr[ax], cur_str.data(), 1, ll.nlines);
      hip_check(launch_line(ll, st), "line kernel");
    }
    return;
  }
  // reverse: whole-line part first (coarsest levels), then tiled passes coarse -> fine
  const void* s_src = in;
  bool s_scr = false;
  std::vector<long long> s_str = in_str;
  long long s_ax = in_ax;
  int buf = 0;
  if (sc.L_line > 0) {
    std::vector<long long> os = dense_strides(p.ndim, shp, ax, sc.m_line);
    LineLaunch ll{};
    ll.storage = p.storage; ll.fwd = false; ll.prog = prog; ll.bc = p.bc; ll.n = sc.m_line; ll.L = (int)sc.L_line;
    ll.a0 = in; ll.a1 = scr[buf]; ll.a2 = in;
    ll.scr[1] = true;
    ll.lm = line_map3(p.ndim, shp, ax, in_str.data(), in_ax, os.data(), 1, in_str.data(), in_ax, ll.nlines);
    hip_check(launch_line(ll, st), "line kernel");
    s_src = scr[buf];
    s_scr = true;
    s_str = os;
    s_ax = 1;
    buf ^= 1;
  }
  for (size_t j = P; j-- > 0;) {
    const PassPlan& pp = sc.passes[j];
    const bool final_pass = j == 0;
    std::vector<long long> os = dense_strides(p.ndim, shp, ax, pp.m0);
    TileLaunch tl{};
    tl.storage = p.storage; tl.fwd = false; tl.prog = prog; tl.bc = p.bc;
    tl.ta = TileArgs{pp.m0, pp.T, pp.H, pp.ntiles, pp.k, 0};
    tl.a0 = in; tl.a1 = final_pass ? dst : scr[buf]; tl.a2 = const_cast<void*>(s_src);
    tl.scr[0] = false; tl.scr[1] = !final_pass; tl.scr[2] = s_scr;
    tl.lm = line_map3(p.ndim, shp, ax, in_str.data(), in_ax, final_pass ? dstr : os.data(),
                      final_pass ? dstr[ax] : 1, s_str.data(), s_ax, tl.nlines);
    hip_check(launch_tile(tl, st), "tile kernel");
    s_src = scr[buf];
    s_scr = true;
    s_str = os;
    s_ax = 1;
    buf ^= 1;
  }
}

// One 1-D transform of L levels (L == 1 for a single level) of all lines of `shp` along `ax`.
static void transform_lines(const Plan& p, const WaveletDefL& w, const void* prog, bool fwd, const long long* shp,
                            int ax, long long L, const void* src, const long long* sst, void* dst,
                            const long long* dstr, hipStream_t st, const RunCtx* ctx) {
  if (compute_bytes(p.storage) == 8)
    transform_lines_t<double>(p, w, prog, fwd, shp, ax, L, src, sst, dst, dstr, st, ctx);
  else
    transform_lines_t<float>(p, w, prog, fwd, shp, ax, L, src, sst, dst, dstr, st, ctx);
}

// A level-by-level transform over two axes as one launch per level of
// two-axis register tiles (lfx_lvl2d.hpp): float32 arithmetic (float32 or
// 16-bit storage), table wavelets, Zero boundary, the second-named axis with
// unit stride on both sides (coalesced lanes), input and output distinct, and
// every level's bands long enough for the tiles. Returns false (nothing
// launched) when the call does not qualify.
static bool lbl2d(const Plan& p, const WaveletDefL& w, long long L, const void* in, void* out, hipStream_t st) {
  if (g_tuning.nd_batch == 0 || p.axes.size() != 2 || p.bc != kZero || p.storage == kF64 ||
      (p.flags & LFX_FLAG_PLAIN) || p.wavelet >= kCustomBase || !has_lift_steps(w) || L < 1 || in == out)
    return false;
  const int ra = p.axes[0], ca = p.axes[1];
  if (ra == ca || p.sin[ca] != 1 || p.sout[ca] != 1) return false;
  // the arrays' memory spans must not overlap (levels read the input while
  // they write the output)
  const long long es_ = (long long)storage_bytes(p.storage);
  long long ilo = 0, ihi = 0, olo = 0, ohi = 0;
  for (int k = 0; k < p.ndim; ++k) {
    if (p.shape[k] == 0) return false;
    const long long ei = (p.shape[k] - 1) * p.sin[k], eo = (p.shape[k] - 1) * p.sout[k];
    (ei < 0 ? ilo : ihi) += ei;
    (eo < 0 ? olo : ohi) += eo;
  }
  const char* ib = static_cast<const char*>(in);
  const char* ob = static_cast<const char*>(out);
  if (ib + ilo * es_ <= ob + (ohi + 1) * es_ - 1 && ob + olo * es_ <= ib + (ihi + 1) * es_ - 1) return false;
  const bool fwd = p.tf == kForward || p.tf == kInverseAdjoint;
  const DProgram<float> hp = resolve<float>(w, p.tf);
  // plan once without buffers to learn whether every level admits tiles
  if (lvl2d_plan(hp, fwd, p.ndim, p.shape, ra, ca, L, in, p.sin, out, p.sout, p.storage, nullptr, nullptr,
                 nullptr, p.wavelet).empty())
    return false;
  const void* prog = device_program(p.wavelet, p.tf, kF32);
  const size_t es = storage_bytes(p.storage);
  const size_t wsb = ((size_t)lvl2d_ws_elems(p.ndim, p.shape, ra, ca) * es + 255) / 256 * 256;
  WsLease lease;
  lease_workspace(L > 1 ? 2 * wsb : 256, st, lease);
  char* ws = (char*)lease.p;
  const auto levels = lvl2d_plan(hp, fwd, p.ndim, p.shape, ra, ca, L, in, p.sin, out, p.sout, p.storage, ws,
                                 ws + wsb, prog, p.wavelet);
  for (const auto& lv : levels) hip_check(launch_lvl2d(p.tf, p.storage, lv, st), "two-axis level kernel");
  return true;
}

// Device program of a no-allocation call: its slot in the fixed tables.
static const void* fixed_program(const RunCtx& ctx, uint32_t wavelet, int tf, int storage) {
  const int i = wavelet < kCustomBase ? table_index(wavelet) : -1;
  if (i < 0)
    fail(LFX_EUNSUPPORTED, "the no-allocation entry point takes table wavelets only (a caller-defined program "
                           "needs a device allocation)");
  if (storage == kF64) return (const DProgram<double>*)ctx.progs_f64 + 4 * i + tf;
  return (const DProgram<float>*)ctx.progs_f32 + 4 * i + tf;
}

// ctx: null for lfx_lwt / lfx_lwt_host; the fixed memory of a no-allocation call.
static void run_plan(const Plan& p, const void* in, void* out, hipStream_t st, const RunCtx* ctx = nullptr) {
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
  const void* prog = ctx ? fixed_program(*ctx, p.wavelet, p.tf, p.storage) : device_program(p.wavelet, p.tf, p.storage);

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
    if (!ctx && sep_chain(p, p.shape, p.sin, p.sout, p.axes, L, in, out, st)) return;
    for (size_t ia = 0; ia < p.axes.size(); ++ia) {
      const void* src = ia == 0 ? in : out;
      const long long* sst = ia == 0 ? p.sin : p.sout;
      if (L[ia] == 0) {
        // zero levels: identity (the reference leaves `out` unwritten here)
        if (ia == 0 && in != out) copy_box(p.storage, p.ndim, p.shape, in, p.sin, out, p.sout, st);
        continue;
      }
      transform_lines(p, w, prog, fwd_stepping, p.shape, p.axes[ia], L[ia], src, sst, out, p.sout, st, ctx);
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
  if (!ctx && lbl2d(p, w, L, in, out, st)) return;
  if (!fwd_stepping && in != out) {
    copy_box(p.storage, p.ndim, p.shape, in, p.sin, out, p.sout, st);
    src = out;
    sst = p.sout;
  }
  for (long long il = 0; il < L; ++il) {
    const std::vector<long long>& shp = fwd_stepping ? shapes[il] : shapes[L - 1 - il];
    // one level on every axis of the box, chained through workspaces (the
    // source is read in full by the first axis before the last axis writes
    // `out`, so this is safe in place)
    if (!ctx && sep_chain(p, shp.data(), sst, p.sout, p.axes, std::vector<long long>(p.axes.size(), 1), src, out, st)) {
      src = out;
      sst = p.sout;
      continue;
    }
    for (int a : p.axes) {
      transform_lines(p, w, prog, fwd_stepping, shp.data(), a, 1, src, sst, out, p.sout, st, ctx);
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

// ---------------------------------------------------------------------------
// batched plans: many 1-D signals of uneven length, wavelet and level count,
// packed back to back in one input and one output array.
//
// Each signal's levels are cut into groups (tiled passes of k levels while the
// band exceeds one workgroup, then one whole-line group; with LFX_FLAG_PLAIN
// every group is a single level). Group g of every signal runs in stage g
// (forward: fine -> coarse; reverse: coarse -> fine); a stage is one launch
// whose items are tiles or whole lines of any signal, longest first.

// LDS class marker of global-line items (they need no dynamic LDS)
static constexpr size_t kGlobalLineClass = ~(size_t)0;

struct BatchLaunch {
  long long item0 = 0, nitems = 0;
  size_t lds = 0;
  int threads = 0;   // 0: default workgroup size
  double bytes = 0;  // algorithmic bytes of its items
};
struct BatchStage {
  long long nitems = 0;
  size_t lds = 0;
  double alg_bytes = 0;  // bytes the stage must read and write (owned ranges)
  std::vector<BatchLaunch> launches;
  long long reg0 = 0, nreg = 0;  // register-tile items (one launch, all wavelets)
  double reg_bytes = 0;
  long long rows0 = 0, nrows = 0;  // row items: four equal short lines per wavefront (one launch)
  double rows_bytes = 0;
};

struct BatchPlanImpl {
  int dev = 0;
  int tf = 0, bc = 0, storage = 0;
  bool fwd = true;
  long long total = 0;  // samples
  std::vector<BatchStage> stages;
  BatchItem* d_items = nullptr;
  void* d_scr = nullptr;
  size_t scr_bytes = 0;
  void* d_progs = nullptr;  // DProgram<float> table of the register items
  int pl = kRegPerLane;      // samples per lane of its register windows
  // side streams: the launches of one stage are independent (disjoint owned
  // outputs), so they are spread over the caller's stream and these, joined
  // by events before the next stage
  std::vector<hipStream_t> side;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
};

struct Group {
  bool tile;
  long long m0;  // band length entering the group (fine side)
  int k;         // levels
  long long T, H, ntiles;
  bool global = false;  // whole-line group with its bands in the plan's scratch
};

template <class C>
static std::vector<Group> signal_groups(const WaveletDefL& w, int tf, int storage, int bc, long long n, long long L,
                                        bool plain, bool reg, bool rows = false) {
  const DProgram<C> hp = resolve<C>(w, tf);
  const bool fwd = tf == kForward || tf == kInverseAdjoint;
  std::vector<Group> g;
  const int k_hint = plain ? 1 : (g_tuning.k > 0 ? g_tuning.k : (reg ? kRegLevels : 0));
  // with register tiles, bands longer than the whole-line threshold are tiled so
  // that whole-line items stay small (several workgroups per CU); with a Zero
  // boundary register tiles also cover band ends, so only short bands are left
  // to the (latency-bound) whole-line kernels
  // (many equal lines, Zero boundary: row items take bands down to 32 samples)
  const long long zl = rows ? 32 : 256;
  const long long limit = reg ? (g_tuning.line > 0 ? g_tuning.line : (bc == kZero ? zl : 4096)) : 0;
  const LineSchedule sc = schedule(hp, fwd, n, L, storage, bc, g_tuning.tile, k_hint, bc == kPeriodic ? 0 : limit);
  for (auto& p : sc.passes) g.push_back(Group{true, p.m0, p.k, p.T, p.H, p.ntiles});
  if (sc.L_line > 0) {
    if (plain) {
      long long m = sc.m_line;
      for (long long l = 0; l < sc.L_line; ++l) {
        g.push_back(Group{false, m, 1, 0, 0, 1, sc.global_line && m > line_limit(storage)});
        m = band_len(m, 1);
      }
    } else {
      g.push_back(Group{false, sc.m_line, (int)sc.L_line, 0, 0, 1, sc.global_line});
    }
  }
  return g;  // fine -> coarse
}

static void free_plan(BatchPlanImpl* p);
static void plan_streams(BatchPlanImpl* p);
static void batch_stage(BatchPlanImpl* p, int st, const void* in, void* out, hipStream_t s,
                        const std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev);
static void batch_stage_all(BatchPlanImpl* p, const void* in, void* out, hipStream_t s) {
  for (int i = 0; i < (int)p->stages.size(); ++i) batch_stage(p, i, in, out, s, nullptr);
}

// nrep > 1: the signal list is repeated nrep times, copy r shifted by
// r * rep_stride elements (many equal lines: items are computed once).
static BatchPlanImpl* batch_create(const lfx_signal* sigs, int nsig, int tf, int bc, int storage, uint32_t flags,
                                   long long nrep = 1, long long rep_stride = 0) {
  if (tf < 0 || tf > 3) fail(LFX_EINVAL, "unknown transform");
  if (bc < 0 || bc > 4) fail(LFX_EINVAL, "Unknown Boundary Condition");
  if (storage < 0 || storage > 3) fail(LFX_EINVAL, "unknown storage type");
  if (nsig < 0 || (nsig > 0 && !sigs)) fail(LFX_EINVAL, "bad signal list");
  auto plan = std::make_unique<BatchPlanImpl>();
  hip_check(hipGetDevice(&plan->dev), "hipGetDevice");
  plan->tf = tf;
  plan->bc = bc;
  plan->storage = storage;
  plan->fwd = tf == kForward || tf == kInverseAdjoint;
  plan->pl = g_tuning.strip == 64 * kRegPerLaneSmall ? kRegPerLaneSmall : kRegPerLane;
  // plain: one level per group (one launch per level, the band round-tripping
  // through memory between levels); register tiles still run each level
  // unless LFX_FLAG_LDS_ONLY
  const bool plain = (flags & LFX_FLAG_PLAIN) != 0;
  const bool lds_only = (flags & LFX_FLAG_LDS_ONLY) != 0;
  const size_t es = storage_bytes(storage), cs = compute_bytes(storage);
  std::vector<std::vector<BatchItem>> per_stage;
  std::vector<std::vector<double>> cost;
  std::vector<size_t> lds;
  std::vector<std::vector<size_t>> item_lds;
  std::vector<double> bytes;
  std::vector<std::vector<BatchItem>> reg_stage;
  std::vector<std::vector<double>> reg_ibytes;  // algorithmic bytes of each register item
  std::vector<std::vector<double>> item_bytes;
  std::vector<double> reg_bytes;
  std::vector<DProgram<float>> reg_progs;  // program table of register items
  std::vector<std::vector<BatchItem>> rows_stage;
  std::vector<double> rows_bytes;
  std::map<uint32_t, int> reg_pidx;
  long long scr_elems = 0;
  // Level groups of every signal first. Forward plans put each signal's
  // trailing whole-line groups in the plan's last stages (aligned on the
  // deepest signal), so that the latency-bound LDS items of all signals share
  // one launch instead of trailing every stage; a signal's groups keep their
  // order, so its dependencies are unchanged.
  std::vector<std::vector<Group>> groups(nsig);
  std::vector<long long> levels(nsig, 0);
  int Gmax = 0;
  for (int i = 0; i < nsig; ++i) {
    const lfx_signal& sg = sigs[i];
    if (sg.length < 0 || sg.offset < 0 || sg.levels < 0) fail(LFX_EINVAL, "bad signal " + std::to_string(i));
    const WaveletDefL w = find_wavelet(sg.wavelet);
    long long L = sg.levels;
    if (L == 0) {
      L = max_level_w(w.width, sg.length);
      if (L < 0) fail(LFX_EINVAL, "level 0 (max level) is undefined for a width-1 wavelet");
    }
    check_bands(bc, has_lift_steps(w), sg.length, L);
    if (L == 0 && sg.length > 0) fail(LFX_EINVAL, "batched signals need at least one level");
    levels[i] = L;
    if (sg.length == 0 || L == 0) continue;
    const bool reg = storage == kF32 && !lds_only && sg.wavelet < kCustomBase;
    const bool rows = reg && !plain && nrep >= 4 && bc == kZero;
    groups[i] = cs == 8 ? signal_groups<double>(w, tf, storage, bc, sg.length, L, plain, reg)
                        : signal_groups<float>(w, tf, storage, bc, sg.length, L, plain, reg, rows);
    Gmax = std::max(Gmax, (int)groups[i].size());
  }
  for (int i = 0; i < nsig; ++i) {
    const lfx_signal& sg = sigs[i];
    const WaveletDefL w = find_wavelet(sg.wavelet);
    const long long L = levels[i];
    plan->total += sg.length;
    if (sg.length == 0 || L == 0) continue;
    const void* prog = device_program(sg.wavelet, tf, storage);
    const bool reg = storage == kF32 && !lds_only && sg.wavelet < kCustomBase;
    const std::vector<Group>& gs = groups[i];
    const DProgram<float> hpf = resolve<float>(w, tf);
    const DProgram<double> hpd = resolve<double>(w, tf);
    auto items_of = [&](const Group& gr) {
      const PassPlan pp{gr.m0, gr.k, gr.T, gr.H, gr.ntiles};
      return cs == 8 ? pass_items(hpd, plan->fwd, pp, false, storage, bc)
                     : pass_items(hpf, plan->fwd, pp, reg, storage, bc, 64ll * plan->pl);
    };
    const int G = (int)gs.size();
    // where each group's output band goes: user output or a scratch region
    // forward: group g writes d's to out, s to (last ? out : next input)
    // reverse: group g (coarse -> fine order) writes its band to
    //          (finest or next is a whole-line group) ? out : scratch
    std::vector<long long> scr_off(G, -1);
    auto add_item = [&](int stage, const BatchItem& it, double c, size_t l, double b, uint32_t wid) {
      if (it.kind == kItemReg) {
        if ((int)reg_stage.size() <= stage) reg_stage.resize(stage + 1);
        auto f = reg_pidx.find(wid);
        if (f == reg_pidx.end()) {
          f = reg_pidx.emplace(wid, (int)reg_progs.size()).first;
          reg_progs.push_back(resolve<float>(find_wavelet(wid), tf));
        }
        BatchItem ri = it;
        ri.wavelet = wid;
        ri.pidx = f->second;
        reg_stage[stage].push_back(ri);
        if ((int)reg_ibytes.size() <= stage) reg_ibytes.resize(stage + 1);
        reg_ibytes[stage].push_back(b);
        if ((int)bytes.size() <= stage) {
          per_stage.resize(stage + 1);
          cost.resize(stage + 1);
          lds.resize(stage + 1, 0);
          item_lds.resize(stage + 1);
          item_bytes.resize(stage + 1);
          bytes.resize(stage + 1, 0.0);
        }
        if ((int)reg_bytes.size() <= stage) reg_bytes.resize(stage + 1, 0.0);
        bytes[stage] += b;
        reg_bytes[stage] += b;
        return;
      }
      if ((int)per_stage.size() <= stage) {
        per_stage.resize(stage + 1);
        cost.resize(stage + 1);
        lds.resize(stage + 1, 0);
        item_lds.resize(stage + 1);
        item_bytes.resize(stage + 1);
        bytes.resize(stage + 1, 0.0);
      }
      per_stage[stage].push_back(it);
      item_bytes[stage].push_back(b);
      item_lds[stage].push_back(l);
      cost[stage].push_back(c);
      if (l != kGlobalLineClass) lds[stage] = std::max(lds[stage], l);
      bytes[stage] += b;
    };
    auto arr_bytes = [&](int arr) { return arr == kArrScr ? (double)cs : (double)es; };
    if (plan->fwd) {
      int in_arr = kArrIn;
      long long in_off = sg.offset;
      for (int g = 0; g < G; ++g) {
        const Group& gr = gs[g];
        const bool last = g + 1 == G;
        const long long mk = band_len(gr.m0, gr.k);
        int s_arr;
        long long s_off;
        if (last || !gs[g + 1].tile) {
          if (!last && gr.tile) {  // tile group followed by a whole-line group: s via scratch
            s_arr = kArrScr;
            s_off = scr_elems;
            scr_elems += mk;
          } else {
            s_arr = kArrOut;
            s_off = sg.offset;
          }
        } else {
          s_arr = kArrScr;
          s_off = scr_elems;
          scr_elems += mk;
        }
        if (gr.tile) {
          for (const ItemSpec& sp : items_of(gr)) {
            BatchItem it{};
            it.prog = prog; it.m0 = gr.m0; it.a = sp.a; it.b = sp.b; it.H = sp.H; it.k = gr.k;
            it.kind = sp.kind == kSpecReg ? kItemReg : kItemTile;
            it.arr[0] = in_arr; it.off[0] = in_off;
            it.arr[1] = kArrOut; it.off[1] = sg.offset;
            it.arr[2] = s_arr; it.off[2] = s_off;
            const TileArgs ta{gr.m0, sp.b - sp.a, sp.H, 0, gr.k, 0};
            add_item(g, it, (double)(sp.b - sp.a + 2 * sp.H),
                     sp.kind == kSpecReg ? 0 : tile_lds_bytes(ta, true, cs),
                     (sp.b - sp.a) * (arr_bytes(in_arr) + (double)es), sg.wavelet);
          }
        } else {
          BatchItem it{};
          it.prog = prog; it.kind = kItemLine; it.m0 = gr.m0; it.k = gr.k;
          it.arr[0] = in_arr; it.off[0] = in_off;
          it.arr[1] = kArrOut; it.off[1] = sg.offset;
          it.arr[2] = in_arr; it.off[2] = in_off;
          if (gr.global) {
            it.kind = kItemGLine;
            it.a = scr_elems;
            scr_elems += gline_elems(gr.m0);
          }
          add_item(g + (Gmax - G), it, (double)gr.m0, gr.global ? kGlobalLineClass : line_lds_bytes(gr.m0, cs),
                   gr.m0 * (arr_bytes(in_arr) + (double)es), sg.wavelet);
        }
        in_arr = s_arr;
        in_off = s_off;
      }
    } else {
      // coarse -> fine: groups in reverse order. A signal's whole-line group
      // (its coarsest levels) runs in stage 0 with those of all signals (one
      // latency-bound LDS launch), and its tiled groups are aligned on the
      // plan's last stage, so that the finest group of every signal -- most
      // of the plan's bytes -- shares one register launch.
      int s_arr = kArrIn;
      long long s_off = sg.offset;
      for (int gi = 0; gi < G; ++gi) {
        const int g = G - 1 - gi;
        const int stage = gs[g].tile ? gi + (Gmax - G) : gi;
        const Group& gr = gs[g];
        const bool finest = g == 0;
        int o_arr;
        long long o_off;
        if (finest || !gs[g - 1].tile) {
          o_arr = kArrOut;
          o_off = sg.offset;
        } else {
          o_arr = kArrScr;
          o_off = scr_elems;
          scr_elems += gr.m0;
        }
        if (gr.tile) {
          for (const ItemSpec& sp : items_of(gr)) {
            BatchItem it{};
            it.prog = prog; it.m0 = gr.m0; it.a = sp.a; it.b = sp.b; it.H = sp.H; it.k = gr.k;
            it.kind = sp.kind == kSpecReg ? kItemReg : kItemTile;
            it.arr[0] = kArrIn; it.off[0] = sg.offset;
            it.arr[1] = o_arr; it.off[1] = o_off;
            it.arr[2] = s_arr; it.off[2] = s_off;
            const TileArgs ta{gr.m0, sp.b - sp.a, sp.H, 0, gr.k, 0};
            add_item(stage, it, (double)tile_rev_window(ta), sp.kind == kSpecReg ? 0 : tile_lds_bytes(ta, false, cs),
                     (sp.b - sp.a) * ((double)es + arr_bytes(o_arr)), sg.wavelet);
          }
        } else {
          BatchItem it{};
          it.prog = prog; it.kind = kItemLine; it.m0 = gr.m0; it.k = gr.k;
          it.arr[0] = kArrIn; it.off[0] = sg.offset;
          it.arr[1] = o_arr; it.off[1] = o_off;
          it.arr[2] = s_arr; it.off[2] = s_off;
          if (gr.global) {
            it.kind = kItemGLine;
            it.a = scr_elems;
            scr_elems += gline_elems(gr.m0);
          }
          add_item(stage, it, (double)gr.m0, gr.global ? kGlobalLineClass : line_lds_bytes(gr.m0, cs),
                   gr.m0 * ((double)es + arr_bytes(o_arr)), sg.wavelet);
        }
        s_arr = o_arr;
        s_off = o_off;
      }
    }
  }
  if (nrep > 1) {
    // replicate: user-array offsets shift by rep_stride, scratch by scr_elems
    auto shift = [&](BatchItem it, long long r) {
      for (int a = 0; a < 3; ++a) it.off[a] += r * (it.arr[a] == kArrScr ? scr_elems : rep_stride);
      if (it.kind == kItemGLine) it.a += r * scr_elems;
      return it;
    };
    for (size_t st = 0; st < per_stage.size(); ++st) {
      const size_t n0 = per_stage[st].size();
      per_stage[st].reserve(n0 * nrep);
      for (long long r = 1; r < nrep; ++r)
        for (size_t j = 0; j < n0; ++j) {
          per_stage[st].push_back(shift(per_stage[st][j], r));
          cost[st].push_back(cost[st][j]);
          item_lds[st].push_back(item_lds[st][j]);
          item_bytes[st].push_back(item_bytes[st][j]);
        }
      bytes[st] *= (double)nrep;
    }
    // Register items that hold a whole band of at most 512 samples (Zero
    // boundary) become row items: four copies per wavefront, one per 16-lane
    // row, starting at band position 0 (no halo: reads past a row read zeros).
    rows_stage.resize(reg_stage.size());
    rows_bytes.resize(reg_stage.size(), 0.0);
    for (size_t st = 0; st < reg_stage.size(); ++st) {
      std::vector<BatchItem> keep;
      double kept = 0;
      for (size_t j = 0; j < reg_stage[st].size(); ++j) {
        const BatchItem& it = reg_stage[st][j];
        const bool row = nrep >= 4 && bc == kZero && it.a == 0 && it.b == it.m0 && it.m0 <= kRowMax;
        if (!row) {
          keep.push_back(it);
          kept += reg_ibytes[st][j];
          continue;
        }
        for (long long r = 0; r < nrep; r += 4) {
          BatchItem ri = shift(it, r);
          ri.rows = (int32_t)std::min<long long>(4, nrep - r);
          for (int a = 0; a < 3; ++a) ri.rstride[a] = ri.arr[a] == kArrScr ? scr_elems : rep_stride;
          ri.a = 0;
          ri.b = ri.m0;
          ri.H = 0;
          rows_stage[st].push_back(ri);
        }
        rows_bytes[st] += reg_ibytes[st][j] * (double)nrep;
      }
      const size_t n0 = keep.size();
      keep.reserve(n0 * nrep);
      for (long long r = 1; r < nrep; ++r)
        for (size_t j = 0; j < n0; ++j) keep.push_back(shift(keep[j], r));
      reg_stage[st] = std::move(keep);
      reg_bytes[st] = kept * (double)nrep;
    }
    scr_elems *= nrep;
    plan->total *= nrep;
  }
  // each stage: items grouped into LDS classes (one launch per class, so small
  // tiles are not launched with the LDS reservation of the longest line), each
  // class ordered by decreasing cost; upload. Items of the smallest class
  // (band-end pieces, short lines) are latency-bound: they run as one-wave
  // workgroups, so that four times as many are in flight and their barriers
  // are wavefront-local.
  std::vector<BatchItem> all;
  const size_t small = g_tuning.small >= 0 ? (size_t)g_tuning.small : 8 * 1024;
  const size_t kClasses[] = {small, 40 * 1024, 80 * 1024, kMaxLdsBytes};
  for (size_t st = 0; st < per_stage.size(); ++st) {
    BatchStage bs;
    bs.alg_bytes = bytes[st];
    // class 4: global-line items (bands in scratch, no dynamic LDS, 1024 threads)
    for (int c = 4; c >= 0; --c) {
      std::vector<size_t> idx;
      for (size_t j = 0; j < per_stage[st].size(); ++j) {
        const size_t l = item_lds[st][j];
        const int cls = l == kGlobalLineClass ? 4
                        : (l <= kClasses[0] ? 0 : (l <= kClasses[1] ? 1 : (l <= kClasses[2] ? 2 : 3)));
        if (cls == c) idx.push_back(j);
      }
      if (idx.empty()) continue;
      std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[st][a] > cost[st][b]; });
      BatchLaunch bl;
      bl.threads = c == 0 ? 64 : (c == 4 ? kThreads : 0);
      bl.item0 = (long long)all.size();
      bl.nitems = (long long)idx.size();
      for (size_t j : idx) {
        all.push_back(per_stage[st][j]);
        if (c != 4) bl.lds = std::max(bl.lds, item_lds[st][j]);
        bl.bytes += item_bytes[st][j];
      }
      bs.nitems += bl.nitems;
      bs.lds = std::max(bs.lds, bl.lds);
      bs.launches.push_back(bl);
    }
    if (st < reg_stage.size() && !reg_stage[st].empty()) {
      // grouped by wavelet, so that neighbouring waves run the same code
      auto& rs = reg_stage[st];
      std::stable_sort(rs.begin(), rs.end(), [](const BatchItem& x, const BatchItem& y) { return x.pidx < y.pidx; });
      bs.reg0 = (long long)all.size();
      bs.nreg = (long long)rs.size();
      bs.reg_bytes = reg_bytes[st];
      all.insert(all.end(), rs.begin(), rs.end());
      bs.nitems += bs.nreg;
    }
    if (st < rows_stage.size() && !rows_stage[st].empty()) {
      bs.rows0 = (long long)all.size();
      bs.nrows = (long long)rows_stage[st].size();
      bs.rows_bytes = rows_bytes[st];
      all.insert(all.end(), rows_stage[st].begin(), rows_stage[st].end());
      bs.nitems += bs.nrows;
    }
    plan->stages.push_back(bs);
  }
  if (!all.empty()) {
    hip_check(hipMalloc(&plan->d_items, all.size() * sizeof(BatchItem)), "hipMalloc(items)");
    hip_check(hipMemcpy(plan->d_items, all.data(), all.size() * sizeof(BatchItem), hipMemcpyHostToDevice),
              "hipMemcpy(items)");
  }
  if (!reg_progs.empty()) {
    const size_t pb = reg_progs.size() * sizeof(DProgram<float>);
    hip_check(hipMalloc(&plan->d_progs, pb), "hipMalloc(programs)");
    hip_check(hipMemcpy(plan->d_progs, reg_progs.data(), pb, hipMemcpyHostToDevice), "hipMemcpy(programs)");
  }
  plan->scr_bytes = (size_t)std::max<long long>(scr_elems, 1) * cs;
  if (hipMalloc(&plan->d_scr, plan->scr_bytes) != hipSuccess) {
    (void)hipGetLastError();
    plan->d_scr = nullptr;
    free_plan(plan.release());
    fail(LFX_ENOMEM, "batch scratch allocation failed");
  }
  try {
    plan_streams(plan.get());
  } catch (...) {
    free_plan(plan.release());
    throw;
  }
  return plan.release();
}

static void free_plan(BatchPlanImpl* p) {
  for (hipStream_t q : p->side) (void)hipStreamSynchronize(q);
  if (p->d_items) (void)hipFree(p->d_items);
  if (p->d_scr) (void)hipFree(p->d_scr);
  if (p->d_progs) (void)hipFree(p->d_progs);
  for (hipStream_t q : p->side) (void)hipStreamDestroy(q);
  for (hipEvent_t e : p->join) (void)hipEventDestroy(e);
  if (p->fork) (void)hipEventDestroy(p->fork);
  delete p;
}

// created with the plan, not at first execution: stream and event creation is
// not allowed while the caller's stream is being captured into a graph
static void plan_streams(BatchPlanImpl* p) {
  if (p->fork) return;
  const int n = g_tuning.streams >= 0 ? std::min(g_tuning.streams, 8) : 3;
  hip_check(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming), "hipEventCreate");
  for (int i = 0; i < n; ++i) {
    hipStream_t q = nullptr;
    hipEvent_t e = nullptr;
    hip_check(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "hipStreamCreate");
    p->side.push_back(q);
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    p->join.push_back(e);
  }
}

// ev (optional): one start/end event pair per launch, recorded on the
// launch's stream (LDS launches in order, then the register launch)
static void batch_stage(BatchPlanImpl* p, int st, const void* in, void* out, hipStream_t s,
                        const std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev) {
  if (st < 0 || st >= (int)p->stages.size()) fail(LFX_EINVAL, "stage out of range");
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != p->dev) fail(LFX_EINVAL, "plan was created on another device");
  const BatchStage& bs = p->stages[st];
  // the LDS launches (latency-bound: few, long-lived workgroups) run on the
  // side streams next to the register launch on the caller's stream
  const int nside = std::min<int>((int)p->side.size(), (int)bs.launches.size());
  std::vector<hipStream_t> lanes;
  if (nside > 0 && bs.launches.size() + (bs.nreg > 0 ? 1 : 0) + (bs.nrows > 0 ? 1 : 0) >= 2) {
    hip_check(hipEventRecord(p->fork, s), "hipEventRecord");
    for (int i = 0; i < nside; ++i) {
      hip_check(hipStreamWaitEvent(p->side[i], p->fork, 0), "hipStreamWaitEvent");
      lanes.push_back(p->side[i]);
    }
  } else {
    lanes.push_back(s);
  }
  size_t ei = 0;
  auto mark = [&](bool start, hipStream_t q) {
    if (!ev) return;
    hip_check(hipEventRecord(start ? (*ev)[ei].first : (*ev)[ei].second, q), "hipEventRecord");
    if (!start) ++ei;
  };
  for (size_t j = 0; j < bs.launches.size(); ++j) {
    const BatchLaunch& bl = bs.launches[j];
    hipStream_t q = lanes[j % lanes.size()];
    mark(true, q);
    hip_check(launch_batch(p->storage, p->fwd, p->bc, p->d_items + bl.item0, bl.nitems, bl.lds, bl.threads, in, out,
                           p->d_scr, q),
              "batch kernel");
    mark(false, q);
  }
  if (bs.nreg > 0) {
    mark(true, s);
    hip_check(launch_reg(p->tf, p->pl, p->d_progs, p->d_items + bs.reg0, bs.nreg, in, out, p->d_scr, s),
              "register kernel");
    mark(false, s);
  }
  if (bs.nrows > 0) {
    mark(true, s);
    hip_check(launch_rows(p->tf, p->d_progs, p->d_items + bs.rows0, bs.nrows, in, out, p->d_scr, s), "row kernel");
    mark(false, s);
  }
  if (lanes[0] != s) {
    for (int i = 0; i < nside; ++i) {
      hip_check(hipEventRecord(p->join[i], p->side[i]), "hipEventRecord");
      hip_check(hipStreamWaitEvent(s, p->join[i], 0), "hipStreamWaitEvent");
    }
  }
}

// ---------------------------------------------------------------------------
// Chains: a sequence of transforms of one signal list (in -> out_0 -> out_1
// ...), each a batched plan, pipelined over groups ("chunks") of signals.
//
// Each plan has one "big" stage that holds most of its bytes: stage 0 of a
// forward plan (the finest levels of every signal) and the last stage of a
// reverse plan; its other stages are short, latency-bound launches (coarse
// levels, whole short bands). Chunk c runs its plans in order on lane c (lane
// 0 is the caller's stream, the others are streams the chain owns); the big
// stages of all (chunk, plan) pairs run one after another -- big(c, t) waits
// for big(c - 1, t), big(0, t) for big(last, t - 1) -- so while one chunk's
// big stage streams HBM, the other chunks' short stages run beside it instead
// of leaving the GPU idle between stages. Every signal still sees its own
// transforms in order, so results equal plan-by-plan execution.

struct ChainImpl {
  int dev = 0;
  int nchunks = 0, ntf = 0;
  std::vector<std::vector<BatchPlanImpl*>> plans;  // [chunk][transform]
  std::vector<hipStream_t> lanes;                  // lanes 1.. (lane 0 = caller's stream)
  std::vector<std::vector<hipEvent_t>> big_done;   // [chunk][transform]
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> done;                    // per lane 1..
};

static int big_stage(const BatchPlanImpl* p) { return p->fwd ? 0 : (int)p->stages.size() - 1; }

static void free_chain(ChainImpl* c) {
  for (hipStream_t q : c->lanes) (void)hipStreamSynchronize(q);
  for (auto& v : c->plans)
    for (BatchPlanImpl* p : v)
      if (p) free_plan(p);
  for (hipStream_t q : c->lanes) (void)hipStreamDestroy(q);
  for (auto& v : c->big_done)
    for (hipEvent_t e : v)
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->done)
    if (e) (void)hipEventDestroy(e);
  if (c->fork) (void)hipEventDestroy(c->fork);
  delete c;
}

static ChainImpl* chain_create(const lfx_signal* sigs, int nsig, const uint32_t* tfs, int ntf, int bc, int storage,
                               uint32_t flags, int nchunks) {
  if (ntf < 1 || ntf > 8 || !tfs) fail(LFX_EINVAL, "a chain has 1..8 transforms");
  if (nchunks < 1 || nchunks > 8) fail(LFX_EINVAL, "a chain has 1..8 chunks");
  if (nsig < 0 || (nsig > 0 && !sigs)) fail(LFX_EINVAL, "bad signal list");
  auto c = std::unique_ptr<ChainImpl, void (*)(ChainImpl*)>(new ChainImpl, free_chain);
  hip_check(hipGetDevice(&c->dev), "hipGetDevice");
  // chunks: consecutive signals of about equal total length
  long long total = 0;
  for (int i = 0; i < nsig; ++i) total += std::max<long long>(sigs[i].length, 0);
  std::vector<int> cut{0};
  long long acc = 0;
  for (int i = 0; i < nsig && (int)cut.size() < nchunks; ++i) {
    acc += std::max<long long>(sigs[i].length, 0);
    if (acc * nchunks >= total * (long long)cut.size() && i + 1 < nsig) cut.push_back(i + 1);
  }
  cut.push_back(nsig);
  c->nchunks = (int)cut.size() - 1;
  c->ntf = ntf;
  // the chunks' plans run their short launches on the lanes unless the
  // caller's tuning names side streams
  const int saved_streams = g_tuning.streams;
  if (g_tuning.streams < 0) g_tuning.streams = 0;
  try {
    for (int k = 0; k < c->nchunks; ++k) {
      c->plans.emplace_back();
      for (int t = 0; t < ntf; ++t)
        c->plans.back().push_back(batch_create(sigs + cut[k], cut[k + 1] - cut[k], (int)tfs[t], bc, storage, flags));
    }
  } catch (...) {
    g_tuning.streams = saved_streams;
    throw;
  }
  g_tuning.streams = saved_streams;
  hip_check(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming), "hipEventCreate");
  for (int k = 0; k < c->nchunks; ++k) {
    c->big_done.emplace_back();
    for (int t = 0; t < ntf; ++t) {
      hipEvent_t e = nullptr;
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      c->big_done.back().push_back(e);
    }
    if (k > 0) {
      hipStream_t q = nullptr;
      hipEvent_t e = nullptr;
      hip_check(hipStreamCreateWithFlags(&q, hipStreamNonBlocking), "hipStreamCreate");
      c->lanes.push_back(q);
      hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
      c->done.push_back(e);
    }
  }
  return c.release();
}

// ev (optional): one start/end event pair per kernel launch, in the order
// chain_launch_info lists them.
static void chain_execute(ChainImpl* c, const void* in, void* const* outs, hipStream_t s,
                          const std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev != c->dev) fail(LFX_EINVAL, "chain was created on another device");
  for (int t = 0; t < c->ntf; ++t) {
    if (!outs[t]) fail(LFX_EINVAL, "null output array");
    if (outs[t] == (t == 0 ? in : outs[t - 1])) fail(LFX_EINVAL, "a chain step's input and output must not alias");
  }
  auto lane = [&](int k) { return k == 0 ? s : c->lanes[k - 1]; };
  if (c->nchunks > 1) {
    hip_check(hipEventRecord(c->fork, s), "hipEventRecord");
    for (hipStream_t q : c->lanes) hip_check(hipStreamWaitEvent(q, c->fork, 0), "hipStreamWaitEvent");
  }
  size_t ei = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> sub;
  for (int t = 0; t < c->ntf; ++t) {
    const void* src = t == 0 ? in : outs[t - 1];
    for (int k = 0; k < c->nchunks; ++k) {
      BatchPlanImpl* p = c->plans[k][t];
      const hipStream_t q = lane(k);
      const int big = big_stage(p);
      for (int st = 0; st < (int)p->stages.size(); ++st) {
        const bool is_big = st == big && c->nchunks > 1;
        if (is_big) {
          const hipEvent_t prev = k > 0 ? c->big_done[k - 1][t] : (t > 0 ? c->big_done[c->nchunks - 1][t - 1] : nullptr);
          if (prev) hip_check(hipStreamWaitEvent(q, prev, 0), "hipStreamWaitEvent");
        }
        const std::vector<std::pair<hipEvent_t, hipEvent_t>>* sev = nullptr;
        if (ev) {
          const BatchStage& bs = p->stages[st];
          const size_t n = bs.launches.size() + (bs.nreg > 0 ? 1 : 0) + (bs.nrows > 0 ? 1 : 0);
          if (ei + n > ev->size()) fail(LFX_EINVAL, "one start and one end event per launch are needed");
          sub.assign(ev->begin() + ei, ev->begin() + ei + n);
          ei += n;
          sev = &sub;
        }
        batch_stage(p, st, src, outs[t], q, sev);
        if (is_big) hip_check(hipEventRecord(c->big_done[k][t], q), "hipEventRecord");
      }
    }
  }
  for (int k = 1; k < c->nchunks; ++k) {
    hip_check(hipEventRecord(c->done[k - 1], c->lanes[k - 1]), "hipEventRecord");
    hip_check(hipStreamWaitEvent(s, c->done[k - 1], 0), "hipStreamWaitEvent");
  }
}

// Launch records of a chain execution, in launch order: per transform, per
// chunk, per stage, the stage's launches (lfx_batch_launch_info order).
struct ChainLaunch {
  int chunk, tf, stage;
  lfx_launch_info info;
};
static std::vector<ChainLaunch> chain_launches(const ChainImpl* c) {
  std::vector<ChainLaunch> out;
  for (int t = 0; t < c->ntf; ++t)
    for (int k = 0; k < c->nchunks; ++k) {
      const BatchPlanImpl* p = c->plans[k][t];
      for (int st = 0; st < (int)p->stages.size(); ++st) {
        const BatchStage& bs = p->stages[st];
        for (const BatchLaunch& bl : bs.launches) {
          lfx_launch_info r{};
          r.kind = 0;
          r.threads = bl.threads > 0 ? bl.threads : block_threads();
          r.nitems = bl.nitems;
          r.alg_bytes = bl.bytes;
          r.lds_bytes = (int64_t)bl.lds;
          out.push_back(ChainLaunch{k, t, st, r});
        }
        if (bs.nreg > 0) out.push_back(ChainLaunch{k, t, st, lfx_launch_info{1, 64 * kRegWaves, bs.nreg, bs.reg_bytes, 0}});
        if (bs.nrows > 0)
          out.push_back(ChainLaunch{k, t, st, lfx_launch_info{2, 64 * kRegWaves, bs.nrows, bs.rows_bytes, 0}});
      }
    }
  return out;
}

// ---------------------------------------------------------------------------
// Lines of an ND box through batched plans: the box is copied into a dense
// float workspace with the transform axis contiguous (a tiled transpose when
// another dim is contiguous in the box, with the storage conversion), the
// lines run as one batched plan (register tiles), and the result is copied
// back. Plans are cached per (device, lines, length, wavelet, levels,
// transform, boundary).

// The key holds the calling thread's planning knobs too: a plan built under
// one tuning is not reused by a call made under another.
struct LinePlanKey {
  int dev;
  long long nlines, n, L;
  uint32_t wavelet;
  int tf, bc;
  long long tile, line, small;
  int k, streams, strip;
  static LinePlanKey make(int dev, long long nlines, long long n, long long L, uint32_t wavelet, int tf, int bc,
                          const Tuning& t) {
    return LinePlanKey{dev, nlines, n, L, wavelet, tf, bc, t.tile, t.line, t.small, t.k, t.streams, t.strip};
  }
  bool operator<(const LinePlanKey& o) const {
    return std::tie(dev, nlines, n, L, wavelet, tf, bc, tile, line, small, k, streams, strip) <
           std::tie(o.dev, o.nlines, o.n, o.L, o.wavelet, o.tf, o.bc, o.tile, o.line, o.small, o.k, o.streams,
                    o.strip);
  }
};
struct CachedPlan {
  BatchPlanImpl* plan;
  unsigned long long tick;
  bool captured;  // used during a stream capture: a graph may replay it
};
static std::map<LinePlanKey, CachedPlan> g_line_plans;
static std::vector<BatchPlanImpl*> g_retired_plans;  // evicted after a capture used them: kept
static unsigned long long g_line_tick = 0;
static constexpr size_t kLinePlanCache = 32;
// Guards the line-plan cache and each cached plan's launches (a plan's fork /
// join events must not interleave between host threads). Distinct from g_mu:
// batch_create takes g_mu (device programs), so holding g_mu here would
// self-deadlock. Lock order: g_line_mu, then g_ws_mu.
static std::mutex g_line_mu;

// Caller holds g_line_mu. Creating or evicting a plan allocates, frees and
// synchronizes, none of which a capturing stream allows: under capture only
// cached plans are used.
static BatchPlanImpl* line_plan(const LinePlanKey& k, hipStream_t st) {
  const bool capturing = stream_capturing(st);
  auto it = g_line_plans.find(k);
  if (it != g_line_plans.end()) {
    it->second.tick = ++g_line_tick;
    it->second.captured = it->second.captured || capturing;
    return it->second.plan;
  }
  if (capturing)
    fail(LFX_EUNSUPPORTED, "no cached plan for this transform while the stream is capturing: run it once outside "
                           "capture first");
  if (g_line_plans.size() >= kLinePlanCache) {
    auto old = g_line_plans.begin();
    for (auto j = g_line_plans.begin(); j != g_line_plans.end(); ++j)
      if (j->second.tick < old->second.tick) old = j;
    if (old->second.captured) {
      g_retired_plans.push_back(old->second.plan);
    } else {
      (void)hipDeviceSynchronize();  // the evicted plan may still be in flight on a caller's stream
      free_plan(old->second.plan);
    }
    g_line_plans.erase(old);
  }
  lfx_signal sg{0, k.n, k.wavelet, (int32_t)k.L};
  BatchPlanImpl* bp = batch_create(&sg, 1, k.tf, k.bc, kF32, 0, k.nlines, k.n);
  g_line_plans[k] = CachedPlan{bp, ++g_line_tick, false};
  return bp;
}

static PermMap perm_map(int ndim, const long long* shp, const long long* s_in, const long long* s_out, int fi,
                        int fo) {
  PermMap pm{};
  pm.same = fi == fo;
  pm.ni = shp[fi];
  pm.si_i = s_in[fi];
  pm.so_i = s_out[fi];
  pm.no = pm.same ? 1 : shp[fo];
  pm.si_o = pm.same ? 0 : s_in[fo];
  pm.so_o = pm.same ? 0 : s_out[fo];
  pm.nbatch = 1;
  for (int k = 0; k < ndim; ++k) {
    if (k == fi || k == fo) continue;
    pm.bshape[pm.nb] = shp[k];
    pm.bss[pm.nb] = s_in[k];
    pm.bds[pm.nb] = s_out[k];
    ++pm.nb;
    pm.nbatch *= shp[k];
  }
  return pm;
}

static int fast_dim(int ndim, const long long* shp, const long long* str) {
  int f = -1;
  for (int k = 0; k < ndim; ++k) {
    if (shp[k] <= 1) continue;
    const long long a = str[k] < 0 ? -str[k] : str[k];
    if (f < 0 || a < (str[f] < 0 ? -str[f] : str[f])) f = k;
  }
  return f < 0 ? ndim - 1 : f;
}

// A dim other than ax, longer than 1, with unit stride in str (-1: none).
static int unit_dim(int ndim, const long long* shp, const long long* str, int ax) {
  for (int k = ndim - 1; k >= 0; --k)
    if (k != ax && shp[k] > 1 && str[k] == 1) return k;
  return -1;
}

// Every line of a strided axis `ax` through the across-line kernel
// (lfx_across.hpp), reading the lines where they are: float32 arithmetic
// (float32 or 16-bit storage on either side), Zero boundary, table wavelets,
// lines of at most acr::kMaxN samples whose every pass of up to four levels
// admits the row tiles, and a dim other than ax with unit stride on both
// sides (the workgroup's lines). src == dst is allowed with identical strides
// (a workgroup reads its lines whole before writing them); otherwise the two
// must not overlap. Returns false (nothing launched) when the call does not
// qualify.
static bool across_axis(const Plan& p, const long long* shp, int ax, long long L, const void* src,
                        const long long* sst, int st_in, void* dst, const long long* dstr, int st_out,
                        hipStream_t st) {
  if (g_tuning.nd_batch == 0 || g_tuning.across == 0) return false;
  if (p.bc != kZero || p.storage == kF64 || (p.flags & LFX_FLAG_PLAIN) || p.wavelet >= kCustomBase ||
      p.wavelet == 0)
    return false;
  const long long n = shp[ax];
  if (n < 2 || n > acr::kMaxN || L < 1) return false;
  const int c = unit_dim(p.ndim, shp, sst, ax);
  // the contiguous dim supplies a workgroup's lines
  if (c < 0 || dstr[c] != 1 || shp[c] < acr::kMinLines) return false;
  long long total = 1;
  for (int k = 0; k < p.ndim; ++k) total *= shp[k];
  if (total <= 0) return false;
  if (src == dst) {
    for (int k = 0; k < p.ndim; ++k)
      if (shp[k] > 1 && sst[k] != dstr[k]) return false;
  } else {
    long long ilo = 0, ihi = 0, olo = 0, ohi = 0;
    for (int k = 0; k < p.ndim; ++k) {
      const long long ei = (shp[k] - 1) * sst[k], eo = (shp[k] - 1) * dstr[k];
      (ei < 0 ? ilo : ihi) += ei;
      (eo < 0 ? olo : ohi) += eo;
    }
    const char* ib = static_cast<const char*>(src);
    const char* ob = static_cast<const char*>(dst);
    const long long bi = (long long)storage_bytes(st_in), bo = (long long)storage_bytes(st_out);
    if (ib + ilo * bi < ob + (ohi + 1) * bo && ob + olo * bo < ib + (ihi + 1) * bi) return false;
  }
  const WaveletDefL& w = find_wavelet(p.wavelet);
  if (!has_lift_steps(w)) return false;
  const bool fwd = p.tf == kForward || p.tf == kInverseAdjoint;
  const DProgram<float> hp = resolve<float>(w, p.tf);
  acr::Args A{};
  if (!across_passes(hp, fwd, n, L, A)) return false;
  across_geometry(p.ndim, shp, ax, c, sst, dstr, A);
  if (A.nblocks > 0x7fffffffll) return false;
  A.in = src;
  A.out = dst;
  A.st_in = st_in;
  A.st_out = st_out;
  A.prog = device_program(p.wavelet, p.tf, kF32);
  A.wavelet = p.wavelet;
  hip_check(launch_across(p.tf, A, st), "across-line kernel");
  return true;
}

static bool lines_fast(const Plan& p, const long long* shp, int ax, long long L, const void* src, const long long* sst,
                       void* dst, const long long* dstr, hipStream_t st) {
  if (g_tuning.nd_batch == 0) return false;
  // a strided axis: its lines are read where they are
  if (sst[ax] != 1 && across_axis(p, shp, ax, L, src, sst, p.storage, dst, dstr, p.storage, st)) return true;
  if (p.storage == kF64 || p.bc == kPeriodic || (p.flags & LFX_FLAG_PLAIN) || p.wavelet >= kCustomBase ||
      p.wavelet == 0)
    return false;
  const long long n = shp[ax];
  long long total = 1;
  for (int k = 0; k < p.ndim; ++k) total *= shp[k];
  if (n < 2 || total <= 0) return false;
  const long long nlines = total / n;
  if (nlines > (1ll << 24)) return false;
  // dense workspace layout: the other dims in order, then the axis
  std::vector<long long> ws_str(p.ndim);
  {
    long long acc = n;
    ws_str[ax] = 1;
    for (int k = p.ndim - 1; k >= 0; --k) {
      if (k == ax) continue;
      ws_str[k] = acc;
      acc *= shp[k];
    }
  }
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_line_mu);
  BatchPlanImpl* bp = line_plan(LinePlanKey::make(dev, nlines, n, L, p.wavelet, p.tf, p.bc, g_tuning), st);
  // A float32 side whose lines are already packed along the axis (its strides
  // equal the dense layout's on every dim longer than 1) is used directly:
  // the plan reads the input there, or writes the output there, saving that
  // side's transpose. The plan's input and output must not alias, so an
  // in-place call copies its input once.
  auto packed = [&](const long long* str) {
    if (p.storage != kF32) return false;
    for (int k = 0; k < p.ndim; ++k)
      if (shp[k] > 1 && str[k] != ws_str[k]) return false;
    return true;
  };
  const bool in_packed = packed(sst) && src != dst;
  const bool out_packed = packed(dstr);
  const size_t bytes = (size_t)total * sizeof(float);
  const size_t pad = (bytes + 255) / 256 * 256;
  const int nws = (in_packed ? 0 : 1) + (out_packed ? 0 : 1);
  // the lease is taken even when no workspace is needed: its event also
  // orders this call's use of the cached plan's scratch after earlier calls
  // on other streams
  WsLease lease;
  lease_workspace(std::max<size_t>((size_t)nws * pad, 256), st, lease);
  float* ws1 = in_packed ? nullptr : (float*)lease.p;
  float* ws2 = out_packed ? nullptr : (float*)((char*)lease.p + (in_packed ? 0 : pad));
  const void* pin = src;
  if (!in_packed) {
    const int fi = fast_dim(p.ndim, shp, sst);
    hip_check(launch_permute(p.storage, kF32, src, ws1, perm_map(p.ndim, shp, sst, ws_str.data(), fi, ax), st),
              "permute kernel");
    pin = ws1;
  }
  batch_stage_all(bp, pin, out_packed ? dst : (void*)ws2, st);
  if (!out_packed) {
    const int fo = fast_dim(p.ndim, shp, dstr);
    hip_check(launch_permute(kF32, p.storage, ws2, dst, perm_map(p.ndim, shp, ws_str.data(), dstr, ax, fo), st),
              "permute kernel");
  }
  return true;
}

// Separable transforms of several axes through batched plans, chained in two
// float32 workspaces: each axis transposes its input into the dense layout
// with that axis contiguous (unless it already is) and runs its plan into the
// other workspace; only the last axis writes the output (directly when its
// lines are packed). One transpose and one plan per axis, where the axis-by-
// axis path pays a transpose in and a transpose out (or an in-place copy) per
// axis. Results are the same: every axis pass computes in float32 from the
// previous axis's float32 values (16-bit storage is rounded once, at the end).
static bool sep_chain(const Plan& p, const long long* shape, const long long* sin, const long long* sout,
                      const std::vector<int>& axes, const std::vector<long long>& L, const void* in, void* out,
                      hipStream_t st) {
  if (g_tuning.nd_batch == 0 || axes.size() < 2) return false;
  if (p.storage == kF64 || p.bc == kPeriodic || (p.flags & LFX_FLAG_PLAIN) || p.wavelet >= kCustomBase ||
      p.wavelet == 0)
    return false;
  long long total = 1;
  for (int k = 0; k < p.ndim; ++k) total *= shape[k];
  if (total <= 0) return false;
  for (size_t ia = 0; ia < axes.size(); ++ia) {
    const long long n = shape[axes[ia]];
    if (L[ia] < 1 || n < 2 || total / n > (1ll << 24)) return false;
  }
  // dense layout with `ax` contiguous, the other dims in order
  auto layout = [&](int ax) {
    std::vector<long long> str(p.ndim);
    long long acc = shape[ax];
    str[ax] = 1;
    for (int k = p.ndim - 1; k >= 0; --k) {
      if (k == ax) continue;
      str[k] = acc;
      acc *= shape[k];
    }
    return str;
  };
  auto same = [&](const long long* a, const std::vector<long long>& b) {
    for (int k = 0; k < p.ndim; ++k)
      if (shape[k] > 1 && a[k] != b[k]) return false;
    return true;
  };
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_line_mu);
  const size_t pad = ((size_t)total * sizeof(float) + 255) / 256 * 256;
  WsLease lease;
  lease_workspace(2 * pad, st, lease);
  float* buf[2] = {(float*)lease.p, (float*)((char*)lease.p + pad)};
  const void* cur = in;
  int cur_buf = -1;  // -1: the user input
  int cur_fast = -1;
  std::vector<long long> cur_str(sin, sin + p.ndim);
  for (size_t ia = 0; ia < axes.size(); ++ia) {
    const int ax = axes[ia];
    const bool last = ia + 1 == axes.size();
    const long long n = shape[ax], nlines = total / n;
    const std::vector<long long> wstr = layout(ax);
    const bool direct = cur_buf < 0 ? (p.storage == kF32 && same(cur_str.data(), wstr)) : same(cur_str.data(), wstr);
    if (!direct) {
      // A strided axis whose lines have a unit-stride neighbour dim: the
      // across-line kernel reads them where they are, into a workspace of
      // the same dim order (in place once in a workspace), or straight into
      // the output for the last axis.
      const int cu = unit_dim(p.ndim, shape, cur_str.data(), ax);
      const int st_cur = cur_buf < 0 ? p.storage : kF32;
      if (cu >= 0 && last) {
        if (across_axis(p, shape, ax, L[ia], cur, cur_str.data(), st_cur, out, sout, p.storage, st)) return true;
      } else if (cu >= 0) {
        const int ob = cur_buf < 0 ? 0 : cur_buf;
        const std::vector<long long> ostr = cur_buf < 0 ? layout(cu) : cur_str;
        if (across_axis(p, shape, ax, L[ia], cur, cur_str.data(), st_cur, buf[ob], ostr.data(), kF32, st)) {
          cur = buf[ob];
          cur_buf = ob;
          cur_fast = cu;
          cur_str = ostr;
          continue;
        }
      }
    }
    const void* pin = cur;
    int in_buf = cur_buf;
    if (!direct) {
      in_buf = cur_buf < 0 ? 0 : 1 - cur_buf;
      const int fi = cur_buf < 0 ? fast_dim(p.ndim, shape, cur_str.data()) : cur_fast;
      hip_check(launch_permute(cur_buf < 0 ? p.storage : kF32, kF32, cur, buf[in_buf],
                               perm_map(p.ndim, shape, cur_str.data(), wstr.data(), fi, ax), st),
                "permute kernel");
      pin = buf[in_buf];
    }
    BatchPlanImpl* bp = line_plan(LinePlanKey::make(dev, nlines, n, L[ia], p.wavelet, p.tf, p.bc, g_tuning), st);
    const int out_buf = in_buf < 0 ? 0 : 1 - in_buf;
    if (last) {
      // pin is a workspace here (the first axis is never the last)
      if (p.storage == kF32 && same(sout, wstr)) {
        batch_stage_all(bp, pin, out, st);
      } else {
        batch_stage_all(bp, pin, buf[out_buf], st);
        const int fo = fast_dim(p.ndim, shape, sout);
        hip_check(launch_permute(kF32, p.storage, buf[out_buf], out,
                                 perm_map(p.ndim, shape, wstr.data(), sout, ax, fo), st),
                  "permute kernel");
      }
    } else {
      batch_stage_all(bp, pin, buf[out_buf], st);
      cur = buf[out_buf];
      cur_buf = out_buf;
      cur_fast = ax;
      cur_str = wstr;
    }
  }
  return true;
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

int lfx_lwt_noalloc(const lfx_desc* desc, const void* in_dev, void* out_dev, void* stream) {
  return guarded([&] {
    Plan p = make_plan(desc);
    if ((p.storage == kBF16 || p.storage == kF16) && p.axes.size() > 1)
      fail(LFX_EUNSUPPORTED, "16-bit storage over several axes runs through float32 workspaces, which the "
                             "no-allocation entry point does not hold");
    const NoAllocDevice& d = noalloc_device();
    const RunCtx ctx{d.scratch, kNoAllocScratchBytes, d.progs_f32, d.progs_f64};
    run_plan(p, in_dev, out_dev, (hipStream_t)stream, &ctx);
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
    *wavelet_id = register_custom(w);
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

int lfx_batch_plan_create(const lfx_signal* signals, int32_t nsignals, uint32_t transform, uint32_t bc,
                          int32_t storage, uint32_t flags, lfx_batch_plan* plan) {
  return guarded([&] {
    if (!plan) fail(LFX_EINVAL, "null plan pointer");
    *plan = reinterpret_cast<lfx_batch_plan>(batch_create(signals, nsignals, (int)transform, (int)bc, storage, flags));
  });
}

int lfx_batch_execute(lfx_batch_plan plan, const void* in_dev, void* out_dev, void* stream) {
  return guarded([&] {
    BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
    if (!p) fail(LFX_EINVAL, "null plan");
    if (in_dev == out_dev) fail(LFX_EINVAL, "batched transforms need distinct input and output arrays");
    for (int s = 0; s < (int)p->stages.size(); ++s) batch_stage(p, s, in_dev, out_dev, (hipStream_t)stream, nullptr);
  });
}

int lfx_batch_num_stages(lfx_batch_plan plan) {
  BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
  return p ? (int)p->stages.size() : LFX_EINVAL;
}

int lfx_batch_execute_stage(lfx_batch_plan plan, int32_t stage, const void* in_dev, void* out_dev, void* stream) {
  return guarded([&] {
    BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
    if (!p) fail(LFX_EINVAL, "null plan");
    if (in_dev == out_dev) fail(LFX_EINVAL, "batched transforms need distinct input and output arrays");
    batch_stage(p, stage, in_dev, out_dev, (hipStream_t)stream, nullptr);
  });
}

int lfx_batch_stage_info(lfx_batch_plan plan, int32_t stage, int64_t* nitems, double* alg_bytes, int64_t* lds_bytes) {
  return guarded([&] {
    BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
    if (!p) fail(LFX_EINVAL, "null plan");
    if (stage < 0 || stage >= (int)p->stages.size()) fail(LFX_EINVAL, "stage out of range");
    if (nitems) *nitems = p->stages[stage].nitems;
    if (alg_bytes) *alg_bytes = p->stages[stage].alg_bytes;
    if (lds_bytes) *lds_bytes = (int64_t)p->stages[stage].lds;
  });
}

int lfx_batch_launch_info(lfx_batch_plan plan, int32_t stage, int32_t max_launches, lfx_launch_info* info) {
  int count = 0;
  const int rc = guarded([&] {
    BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
    if (!p) fail(LFX_EINVAL, "null plan");
    if (stage < 0 || stage >= (int)p->stages.size()) fail(LFX_EINVAL, "stage out of range");
    const BatchStage& bs = p->stages[stage];
    const int nl = (int)bs.launches.size();
    const int n = nl + (bs.nreg > 0 ? 1 : 0) + (bs.nrows > 0 ? 1 : 0);
    for (int i = 0; i < n && i < max_launches && info; ++i) {
      lfx_launch_info& r = info[i];
      if (i < nl) {
        r.kind = 0;
        r.threads = bs.launches[i].threads > 0 ? bs.launches[i].threads : block_threads();
        r.nitems = bs.launches[i].nitems;
        r.alg_bytes = bs.launches[i].bytes;
        r.lds_bytes = (int64_t)bs.launches[i].lds;
      } else {
        const bool reg = i == nl && bs.nreg > 0;
        r.kind = reg ? 1 : 2;
        r.threads = 64 * kRegWaves;
        r.nitems = reg ? bs.nreg : bs.nrows;
        r.alg_bytes = reg ? bs.reg_bytes : bs.rows_bytes;
        r.lds_bytes = 0;
      }
    }
    count = n;
  });
  return rc == LFX_OK ? count : rc;
}

int lfx_batch_execute_stage_timed(lfx_batch_plan plan, int32_t stage, const void* in_dev, void* out_dev, void* stream,
                                  int32_t nevents, void* const* ev_start, void* const* ev_end) {
  return guarded([&] {
    BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
    if (!p) fail(LFX_EINVAL, "null plan");
    if (stage < 0 || stage >= (int)p->stages.size()) fail(LFX_EINVAL, "stage out of range");
    const BatchStage& bs = p->stages[stage];
    const int n = (int)bs.launches.size() + (bs.nreg > 0 ? 1 : 0) + (bs.nrows > 0 ? 1 : 0);
    if (nevents < n || !ev_start || !ev_end) fail(LFX_EINVAL, "one start and one end event per launch are needed");
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev(n);
    for (int i = 0; i < n; ++i) {
      if (!ev_start[i] || !ev_end[i]) fail(LFX_EINVAL, "null event");
      ev[i] = {(hipEvent_t)ev_start[i], (hipEvent_t)ev_end[i]};
    }
    batch_stage(p, stage, in_dev, out_dev, (hipStream_t)stream, &ev);
  });
}

void lfx_batch_plan_destroy(lfx_batch_plan plan) {
  BatchPlanImpl* p = reinterpret_cast<BatchPlanImpl*>(plan);
  if (!p) return;
  free_plan(p);
}

int lfx_batch_chain_create(const lfx_signal* signals, int32_t nsignals, const uint32_t* transforms,
                           int32_t ntransforms, uint32_t bc, int32_t storage, uint32_t flags, int32_t nchunks,
                           lfx_batch_chain* chain) {
  return guarded([&] {
    if (!chain) fail(LFX_EINVAL, "null chain pointer");
    *chain = reinterpret_cast<lfx_batch_chain>(
        chain_create(signals, nsignals, transforms, ntransforms, (int)bc, storage, flags, nchunks));
  });
}

int lfx_batch_chain_execute(lfx_batch_chain chain, const void* in_dev, void* const* out_devs, void* stream) {
  return guarded([&] {
    ChainImpl* c = reinterpret_cast<ChainImpl*>(chain);
    if (!c || !out_devs) fail(LFX_EINVAL, "null chain or output list");
    chain_execute(c, in_dev, out_devs, (hipStream_t)stream, nullptr);
  });
}

int lfx_batch_chain_launch_info(lfx_batch_chain chain, int32_t max_launches, lfx_chain_launch_info* info) {
  int count = 0;
  const int rc = guarded([&] {
    ChainImpl* c = reinterpret_cast<ChainImpl*>(chain);
    if (!c) fail(LFX_EINVAL, "null chain");
    const std::vector<ChainLaunch> v = chain_launches(c);
    for (int i = 0; i < (int)v.size() && i < max_launches && info; ++i) {
      info[i].chunk = v[i].chunk;
      info[i].transform_index = v[i].tf;
      info[i].stage = v[i].stage;
      info[i].big = v[i].stage == big_stage(c->plans[v[i].chunk][v[i].tf]) ? 1 : 0;
      info[i].launch = v[i].info;
    }
    count = (int)v.size();
  });
  return rc == LFX_OK ? count : rc;
}

int lfx_batch_chain_execute_timed(lfx_batch_chain chain, const void* in_dev, void* const* out_devs, void* stream,
                                  int32_t nevents, void* const* ev_start, void* const* ev_end) {
  return guarded([&] {
    ChainImpl* c = reinterpret_cast<ChainImpl*>(chain);
    if (!c || !out_devs) fail(LFX_EINVAL, "null chain or output list");
    const size_t n = chain_launches(c).size();
    if (nevents < (int32_t)n || !ev_start || !ev_end) fail(LFX_EINVAL, "one start and one end event per launch are needed");
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev(n);
    for (size_t i = 0; i < n; ++i) {
      if (!ev_start[i] || !ev_end[i]) fail(LFX_EINVAL, "null event");
      ev[i] = {(hipEvent_t)ev_start[i], (hipEvent_t)ev_end[i]};
    }
    chain_execute(c, in_dev, out_devs, (hipStream_t)stream, &ev);
  });
}

int lfx_batch_chain_num_chunks(lfx_batch_chain chain) {
  ChainImpl* c = reinterpret_cast<ChainImpl*>(chain);
  return c ? c->nchunks : LFX_EINVAL;
}

void lfx_batch_chain_destroy(lfx_batch_chain chain) {
  ChainImpl* c = reinterpret_cast<ChainImpl*>(chain);
  if (c) free_chain(c);
}

static void check_tuning(long long tile, int levels_per_pass) {
  if (tile < 0 || levels_per_pass < 0 || levels_per_pass > 12) fail(LFX_EINVAL, "bad tuning values");
  if (tile > 0 && (tile < 64 || tile > (1 << 16))) fail(LFX_EINVAL, "tile must be in [64, 65536]");
}

int lfx_stream_copy(const void* src_dev, void* dst_dev, int64_t bytes, int32_t pattern, void* stream) {
  return guarded([&] {
    if (pattern < 0 || pattern > 1) fail(LFX_EINVAL, "pattern must be 0 or 1");
    if (bytes < 0 || (bytes > 0 && (!src_dev || !dst_dev))) fail(LFX_EINVAL, "bad copy arguments");
    if (((uintptr_t)src_dev | (uintptr_t)dst_dev) & 15) fail(LFX_EINVAL, "copy buffers must be 16-byte aligned");
    hip_check(launch_stream_copy(src_dev, dst_dev, bytes, pattern, (hipStream_t)stream), "copy kernel");
  });
}

int lfx_set_tuning(int64_t tile, int32_t levels_per_pass) {
  return guarded([&] {
    check_tuning(tile, levels_per_pass);
    g_tuning.tile = tile;
    g_tuning.k = levels_per_pass;
  });
}

int lfx_set_tuning_ex(const lfx_tuning* t) {
  return guarded([&] {
    if (!t) fail(LFX_EINVAL, "null tuning");
    check_tuning(t->tile, t->levels_per_pass);
    if (t->side_streams < -1 || t->side_streams > 8) fail(LFX_EINVAL, "side_streams must be in [-1, 8]");
    if (t->line_limit < 0 || (t->line_limit > 0 && t->line_limit < 16)) fail(LFX_EINVAL, "line_limit must be 0 or >= 16");
    if (t->across < -1 || t->across > 1) fail(LFX_EINVAL, "across must be -1, 0 or 1");
    if (t->strip != 0 && t->strip != 64 * kRegPerLaneSmall && t->strip != kRegWindow)
      fail(LFX_EINVAL, "strip must be 0, 1024 or 2048");
    g_tuning.tile = t->tile;
    g_tuning.k = t->levels_per_pass;
    g_tuning.streams = t->side_streams;
    g_tuning.line = t->line_limit;
    g_tuning.across = t->across;
    g_tuning.strip = t->strip;
  });
}

int lfx_get_tuning(lfx_tuning* t) {
  return guarded([&] {
    if (!t) fail(LFX_EINVAL, "null tuning");
    t->tile = g_tuning.tile;
    t->levels_per_pass = g_tuning.k;
    t->side_streams = g_tuning.streams;
    t->line_limit = g_tuning.line;
    t->across = g_tuning.across;
    t->strip = g_tuning.strip;
  });
}

}  // extern "C"
