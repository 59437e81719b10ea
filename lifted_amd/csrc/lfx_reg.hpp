// Register-tiled lifting: one wavefront owns one tile and runs up to four
// levels in registers; LDS serves only to transpose global loads and stores
// into coalesced ones (no workgroup barriers).
//
// Each of the 64 lanes holds PL consecutive samples of the tile's window (the
// tile functions' template parameter: 32 by default, a 2048-sample window; 16
// for 1024-sample windows). Level 1 splits them into PL/2 s and PL/2 d values
// per lane (s[e] = x[2e], d[e] = x[2e+1]); level l + 1 splits s_l the same
// way, so after four levels each lane holds PL/16 s_4 and PL/16 d_4. A
// lifting step
//     y[e] += sum_k c[k] * x[e + off + k]
// reads x values past the lane's own chunk from neighbouring lanes (DPP shifts
// by whole chunks). The step shapes (kind, band, offset, taps) are compile-time
// (lfx_wavelet_types.hpp); the coefficients come from the runtime program.
// The window's outer lanes compute values that depend on samples outside the
// window; the halo is sized on the host (lfx_tiles.hpp rules) so that only
// values owned by the tile -- all exact -- are stored. A window inside a band
// applies no boundary map or adjoint fold; under the Zero boundary a window
// may overlap a band end and runs masked (the band is zeroed outside its range
// after every step, which is what skipping out-of-range terms computes).
// Other boundaries keep the parts of a band next to its ends in the LDS tile
// bodies.
//
// Row mode (lane policies with ROWS = 4) holds four equal lines of at most 512
// samples, one per 16-lane row, each from band position 0; reads past a row's
// ends yield zeros, so whole bands need no halo (Zero boundary).
//
// The lane policy P abstracts the wavefront: on the GPU a value is one float
// per lane and cross-lane reads are lane shifts; in the CPU simulation a value is
// an array of 64 floats.
#pragma once
#include <cmath>
#include <cstdint>
#include <utility>

#include "lfx_common.hpp"
#include "lfx_device.hpp"
#include "lfx_wavelet_types.hpp"

namespace lfx {
namespace reg {

constexpr int kLanes = 64;
constexpr int kLevels = kRegLevels;           // levels per register tile
constexpr int kPerLane = kRegPerLane;         // default samples per lane at the finest level
static_assert(kLevels == 4 && kRegWindow == kLanes * kPerLane, "register tiles run four levels");

// Lane policies. ROWS = 1: the wavefront holds one window of a band. ROWS = 4
// ("row mode"): it holds four equal lines of at most 512 samples, one per row
// of 16 lanes, each row starting at band position 0; a lane's band position is
// counted within its row, a read past a row's ends yields 0 (the Zero
// boundary), and memory helpers address row r of an array at r * rs. The
// trailing (rs, nr) arguments of the memory helpers (row stride, rows in use)
// are ignored when ROWS = 1.
#ifdef LFX_CPU_SIM
// CPU simulation: a value is the 64 lanes of the wavefront. The elementwise
// operations stay out of line: inlining a 64-lane loop into every unrolled
// step makes the simulation too large to compile in reasonable time.
#define LFX_SIM_NOINLINE __attribute__((noinline))
template <int ROWS>
struct LaneCpuT {
  static constexpr int kRows = ROWS;
  static constexpr int kRowLanes = kLanes / ROWS;
  struct V {
    float v[kLanes];
  };
  // value of x held by lane (lane + Q): one window -- the source lane clamped
  // to the wavefront (as the GPU's chained one-lane shifts give); rows -- 0
  // when the source lies outside the lane's row
  template <int Q>
  LFX_SIM_NOINLINE static V from(const V& x) {
    V r;
    for (int l = 0; l < kLanes; ++l) {
      const int src = l + Q;
      if constexpr (ROWS == 1) {
        r.v[l] = x.v[src < 0 ? 0 : (src >= kLanes ? kLanes - 1 : src)];
      } else {
        const bool in_row = src >= 0 && src < kLanes && src / kRowLanes == l / kRowLanes;
        r.v[l] = in_row ? x.v[src] : 0.0f;
      }
    }
    return r;
  }
  LFX_SIM_NOINLINE static V fma(float c, V x, V acc) {
    for (int l = 0; l < kLanes; ++l) acc.v[l] = std::fma(c, x.v[l], acc.v[l]);
    return acc;
  }
  LFX_SIM_NOINLINE static V add(V a, const V& b) {
    for (int l = 0; l < kLanes; ++l) a.v[l] = a.v[l] + b.v[l];
    return a;
  }
  LFX_SIM_NOINLINE static V sub(V a, const V& b) {
    for (int l = 0; l < kLanes; ++l) a.v[l] = a.v[l] - b.v[l];
    return a;
  }
  LFX_SIM_NOINLINE static V mul(V a, float c) {
    for (int l = 0; l < kLanes; ++l) a.v[l] = a.v[l] * c;
    return a;
  }
  template <class S, class F>
  static V ld(const S* p, F idx) {
    V r;
    for (int l = 0; l < kLanes; ++l) r.v[l] = dev::load_as<S, float>(p, idx(l));
    return r;
  }
  template <class S, class F, class G>
  static void st(S* p, const V& v, F idx, G pred) {
    for (int l = 0; l < kLanes; ++l)
      if (pred(l)) dev::store_as<S, float>(p, idx(l), v.v[l]);
  }
  // bit e of lane l: band position p0 + E l' + e lies in [0, n) (l' = lane within its row)
  struct M {
    uint32_t m[kLanes];
  };
  static M band_mask(int E, long long p0, long long n) {
    M r;
    for (int l = 0; l < kLanes; ++l) {
      r.m[l] = 0;
      for (int e = 0; e < E; ++e) {
        const long long g = p0 + (long long)E * (l % kRowLanes) + e;
        if (g >= 0 && g < n) r.m[l] |= 1u << e;
      }
    }
    return r;
  }
  LFX_SIM_NOINLINE static V keep(const M& m, int e, V v) {
    for (int l = 0; l < kLanes; ++l)
      if (!((m.m[l] >> e) & 1u)) v.v[l] = 0.0f;
    return v;
  }
  template <class S, class F, class G>
  static V ld(const S* p, F idx, G pred) {
    V r;
    for (int l = 0; l < kLanes; ++l) r.v[l] = pred(l) ? dev::load_as<S, float>(p, idx(l)) : 0.0f;
    return r;
  }
  // x[e] of lane l = p[pos0 + E l' + e] of its row where that position lies in
  // [0, n), else 0. The model has no coalescing order: ld_raw yields the chunk
  // values and xpose_in copies them.
  template <int E, class S>
  static void ld_raw(const S* p, long long pos0, long long n, V (&t)[E], long long rs = 0, int nr = 1) {
    for (int e = 0; e < E; ++e)
      for (int l = 0; l < kLanes; ++l) {
        const int r = l / kRowLanes;
        const long long g = pos0 + (long long)E * (l % kRowLanes) + e;
        t[e].v[l] = (r < nr && g >= 0 && g < n) ? dev::load_as<S, float>(p + r * rs, g) : 0.0f;
      }
  }
  template <int E>
  static void xpose_in(const V (&t)[E], V (&x)[E], float*) {
    for (int e = 0; e < E; ++e) x[e] = t[e];
  }
  template <int E, class S>
  static void ld_chunks(const S* p, long long pos0, long long n, V (&x)[E], float* lds, long long rs = 0,
                        int nr = 1) {
    V t[E];
    ld_raw<E>(p, pos0, n, t, rs, nr);
    xpose_in<E>(t, x, lds);
  }
  // the N step coefficients of a program, read once per tile
  template <int N>
  static void coefs(const float* p, float* c) {
    for (int i = 0; i < N; ++i) c[i] = p[i];
  }
  template <int E, class S>
  static void st_chunks(S* p, long long pos0, const V (&v)[E], long long g0, long long lo, long long hi, float*,
                        long long rs = 0, int nr = 1) {
    for (int e = 0; e < E; ++e)
      for (int l = 0; l < kLanes; ++l) {
        const int r = l / kRowLanes;
        const long long g = g0 + (long long)E * (l % kRowLanes) + e;
        if (r < nr && g >= lo && g < hi) dev::store_as<S, float>(p + r * rs, pos0 + g, v[e].v[l]);
      }
  }
  // per-lane values, selects and visits (lane-indexed functions)
  template <class F>
  static V gen(F f) {
    V r;
    for (int l = 0; l < kLanes; ++l) r.v[l] = f(l);
    return r;
  }
  template <class F>
  static V where(F pred, const V& a, const V& b) {
    V r;
    for (int l = 0; l < kLanes; ++l) r.v[l] = pred(l) ? a.v[l] : b.v[l];
    return r;
  }
  template <class F>
  static void each(const V& v, F f) {
    for (int l = 0; l < kLanes; ++l) f(l, v.v[l]);
  }
  // value of v held by lane src(l), for every lane l
  template <class F>
  static V permute(const V& v, F src) {
    V r;
    for (int l = 0; l < kLanes; ++l) r.v[l] = v.v[src(l)];
    return r;
  }
  // the same mask on every lane: bit e iff p0 + e lies in [0, n)
  static M umask(int E, long long p0, long long n) {
    M r;
    for (int l = 0; l < kLanes; ++l) {
      r.m[l] = 0;
      for (int e = 0; e < E; ++e)
        if (p0 + e >= 0 && p0 + e < n) r.m[l] |= 1u << e;
    }
    return r;
  }
};
using LaneCpu = LaneCpuT<1>;
using RowCpu = LaneCpuT<4>;
#else
template <int ROWS>
struct LaneGpuT {
  static constexpr int kRows = ROWS;
  static constexpr int kRowLanes = kLanes / ROWS;
  using V = float;
  static __device__ __forceinline__ int lane() { return (int)(threadIdx.x & (kLanes - 1)); }
  // One-lane shifts by DPP. One window: wave_shl:1 / wave_shr:1, lane l gets
  // the value of lane l + 1 (l - 1) and the lane with no source keeps its own
  // value. Rows: row_shl:1 / row_shr:1 within each 16-lane row, the lane with
  // no source in its row gets 0.
  static __device__ __forceinline__ V shl1(V v) {
    const int b = __builtin_bit_cast(int, v);
    if constexpr (ROWS == 1)
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, 0x130, 0xF, 0xF, false));
    else
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, b, 0x101, 0xF, 0xF, false));
  }
  static __device__ __forceinline__ V shr1(V v) {
    const int b = __builtin_bit_cast(int, v);
    if constexpr (ROWS == 1)
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, 0x138, 0xF, 0xF, false));
    else
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, b, 0x111, 0xF, 0xF, false));
  }
  // value of v held by lane (lane + Q): chains of one-lane DPP shifts for
  // |Q| <= 3 (a few cycles each, no LDS round trip), a ds_bpermute beyond; one
  // window clamps the source lane to the wavefront, rows give 0 outside the row
  template <int Q>
  static __device__ __forceinline__ V from(V v) {
    if constexpr (Q == 0) {
      return v;
    } else if constexpr (Q > 0 && Q <= 3) {
      return from<Q - 1>(shl1(v));
    } else if constexpr (Q < 0 && Q >= -3) {
      return from<Q + 1>(shr1(v));
    } else {
      const int l = lane();
      int src = l + Q;
      const bool in_row = src >= 0 && src < kLanes && src / kRowLanes == l / kRowLanes;
      src = src < 0 ? 0 : (src >= kLanes ? kLanes - 1 : src);
      const float r = __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src << 2, __builtin_bit_cast(int, v)));
      if constexpr (ROWS == 1) return r;
      else return in_row ? r : 0.0f;
    }
  }
  static __device__ __forceinline__ V fma(float c, V x, V acc) { return __builtin_fmaf(c, x, acc); }
  static __device__ __forceinline__ V add(V a, V b) { return a + b; }
  static __device__ __forceinline__ V sub(V a, V b) { return a - b; }
  static __device__ __forceinline__ V mul(V a, float c) { return a * c; }
  template <class S, class F>
  static __device__ __forceinline__ V ld(const S* p, F idx) {
    return dev::load_as<S, float>(p, idx(lane()));
  }
  template <class S, class F, class G>
  static __device__ __forceinline__ void st(S* p, V v, F idx, G pred) {
    const int l = lane();
    if (pred(l)) dev::store_as<S, float>(p, idx(l), v);
  }
  static __device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // bit e: band position p0 + E l' + e lies in [0, n) (l' = lane within its row)
  using M = uint32_t;
  static __device__ __forceinline__ M band_mask(int E, long long p0, long long n) {
    const long long g = p0 + (long long)E * (lane() % kRowLanes);
    M r = 0;
    for (int e = 0; e < E; ++e) r |= (g + e >= 0 && g + e < n) ? (1u << e) : 0u;
    return r;
  }
  static __device__ __forceinline__ V keep(M m, int e, V v) { return ((m >> e) & 1u) ? v : 0.0f; }
  template <class S, class F, class G>
  static __device__ __forceinline__ V ld(const S* p, F idx, G pred) {
    const int l = lane();
    return pred(l) ? dev::load_as<S, float>(p, idx(l)) : 0.0f;
  }
  // x[e] = p[pos0 + E * lane + e] (0 outside [0, n)): coalesced global loads
  // (lane l reads pos0 + 64 j + l), transposed through the wave's LDS buffer
  // (row stride E + 1 keeps the chunk reads conflict-free). The bounds checks
  // run only for windows that overlap a band end (a wave-uniform test).
  // Split in two so that a tile can issue all its global loads first and
  // transpose each band when a level needs it: ld_raw leaves t[j] = element
  // u = 64 j + lane of the window, xpose_in moves it to the chunk layout. In
  // row mode element u is position u % (16 E) of row u / (16 E).
  template <int E, class S>
  static __device__ __forceinline__ void ld_raw(const S* p, long long pos0, long long n, V (&t)[E], long long rs = 0,
                                                int nr = 1) {
    const int l = lane();
    if constexpr (ROWS == 1) {
      if (pos0 >= 0 && pos0 + E * kLanes <= n) {
        const S* q = p + pos0;
#pragma unroll
        for (int j = 0; j < E; ++j) t[j] = dev::load_as<S, float>(q, j * kLanes + l);  // all loads in flight
      } else if (n > 0) {
        // a window over a band end: every load issued unconditionally, from a
        // position clamped into the band, and the values outside zeroed
        // afterwards (a branch around each load would make the compiler wait
        // for it before issuing the next)
        // byte offsets from the window start: the clamped position lies
        // within the window (it overlaps the band), so they are 32-bit
        const char* q = reinterpret_cast<const char*>(p + pos0);
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const long long g = pos0 + j * kLanes + l;
          const unsigned off = (unsigned)((g < 0 ? 0 : (g < n ? g : n - 1)) - pos0) * (unsigned)sizeof(S);
          t[j] = dev::load_as<S, float>(reinterpret_cast<const S*>(q + off), 0);
        }
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const long long g = pos0 + j * kLanes + l;
          if (!(g >= 0 && g < n)) t[j] = 0.0f;
        }
      } else {
#pragma unroll
        for (int j = 0; j < E; ++j) t[j] = 0.0f;
      }
    } else {
      constexpr int SEG = E * kRowLanes;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int u = j * kLanes + l, r = u / SEG;
        const long long g = pos0 + u % SEG;
        t[j] = (r < nr && g >= 0 && g < n) ? dev::load_as<S, float>(p + r * rs, g) : 0.0f;
      }
    }
  }
  // Up to 16 values per lane go through the buffer at once (64 rows of E + 1
  // floats). E = 32 or 64 per lane go in E / 16 rounds of 1024 values: round
  // h fills the chunks of lanes [h P, h P + P), P = 1024 / E (P rows of E + 1
  // floats), so the buffer never exceeds 64 x 17 floats.
  template <int E>
  static __device__ __forceinline__ void xpose_in(const V (&t)[E], V (&x)[E], float* lds) {
    const int l = lane();
    if constexpr (E <= 16) {
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int u = j * kLanes + l;
        lds[(u / E) * (E + 1) + (u % E)] = t[j];
      }
      wave_sync();
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = lds[l * (E + 1) + e];
      wave_sync();
    } else {
      static_assert(E == 32 || E == 64, "register chunks hold at most 64 samples per lane");
      constexpr int P = 1024 / E;
#pragma unroll
      for (int h = 0; h < E / 16; ++h) {
        // value j of round h: u = (16 h + j) 64 + l -> lane u / E (row u / E - h P), slot u % E
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int u = (16 * h + j) * kLanes + l;
          lds[(u / E - h * P) * (E + 1) + (u % E)] = t[16 * h + j];
        }
        wave_sync();
        if (l / P == h) {
#pragma unroll
          for (int e = 0; e < E; ++e) x[e] = lds[(l - h * P) * (E + 1) + e];
        }
        wave_sync();
      }
    }
  }
  template <int E, class S>
  static __device__ __forceinline__ void ld_chunks(const S* p, long long pos0, long long n, V (&x)[E], float* lds,
                                                   long long rs = 0, int nr = 1) {
    V t[E];
    ld_raw<E>(p, pos0, n, t, rs, nr);
    xpose_in<E>(t, x, lds);
  }
  // The N step coefficients of a program, read once per tile through the
  // constant address space: the address is wave-uniform, so these become
  // scalar loads into SGPRs, issued before any global store (a vector load
  // after a store would wait for the store on the shared memory counter).
  template <int N>
  static __device__ __forceinline__ void coefs(const float* p, float* c) {
    typedef const __attribute__((address_space(4))) float* cptr;
    const cptr q = (cptr)p;
#pragma unroll
    for (int i = 0; i < N; ++i) c[i] = q[i];
  }
  // Store value u (coalesced order) of the window if its band position lies
  // in [lo, hi). One window: position g0 + u, owned iff u in [t0, t1) (window
  // offsets, 32-bit); rows: position g0 + u % (16 E) of row u / (16 E).
  template <int E, class S>
  static __device__ __forceinline__ void st_one(S* p, long long pos0, long long g0, long long lo, long long hi, int t0,
                                                int t1, int u, float val, long long rs, int nr) {
    if constexpr (ROWS == 1) {
      if (u >= t0 && u < t1) dev::store_as<S, float>(p + pos0 + g0, u, val);
    } else {
      constexpr int SEG = E * kRowLanes;
      const int r = u / SEG;
      const long long g = g0 + u % SEG;
      if (r < nr && g >= lo && g < hi) dev::store_as<S, float>(p + r * rs, pos0 + g, val);
    }
  }
  // p[pos0 + g] = v for the owned band positions g = g0 + E * lane + e in
  // [lo, hi), transposed through LDS so that the global stores are coalesced.
  template <int E, class S>
  static __device__ __forceinline__ void st_chunks(S* p, long long pos0, const V (&v)[E], long long g0, long long lo,
                                                   long long hi, float* lds, long long rs = 0, int nr = 1) {
    const int l = lane();
    // owned window-relative range [t0, t1) of the E * 64 values (one window)
    const int t0 = (int)(lo - g0 < 0 ? 0 : (lo - g0 > E * kLanes ? E * kLanes : lo - g0));
    const int t1 = (int)(hi - g0 > E * kLanes ? E * kLanes : (hi - g0 < 0 ? 0 : hi - g0));
    // Every value is read back from the buffer before the first (predicated)
    // store: a read inside each store's branch would wait for the LDS before
    // every store instead of once.
    if constexpr (E <= 16) {
#pragma unroll
      for (int e = 0; e < E; ++e) lds[l * (E + 1) + e] = v[e];
      wave_sync();
      float t[E];
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int u = j * kLanes + l;
        t[j] = lds[(u / E) * (E + 1) + (u % E)];
      }
#pragma unroll
      for (int j = 0; j < E; ++j) st_one<E>(p, pos0, g0, lo, hi, t0, t1, j * kLanes + l, t[j], rs, nr);
      wave_sync();
    } else {
      static_assert(E == 32 || E == 64, "register chunks hold at most 64 samples per lane");
      // E / 16 rounds of 1024 values (see xpose_in)
      constexpr int P = 1024 / E;
#pragma unroll
      for (int h = 0; h < E / 16; ++h) {
        if (l / P == h) {
#pragma unroll
          for (int e = 0; e < E; ++e) lds[(l - h * P) * (E + 1) + e] = v[e];
        }
        wave_sync();
        float t[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int u = (16 * h + j) * kLanes + l;
          t[j] = lds[(u / E - h * P) * (E + 1) + (u % E)];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) st_one<E>(p, pos0, g0, lo, hi, t0, t1, (16 * h + j) * kLanes + l, t[j], rs, nr);
        wave_sync();
      }
    }
  }
  template <class F>
  static __device__ __forceinline__ V gen(F f) {
    return f(lane());
  }
  template <class F>
  static __device__ __forceinline__ V where(F pred, V a, V b) {
    return pred(lane()) ? a : b;
  }
  template <class F>
  static __device__ __forceinline__ void each(V v, F f) {
    f(lane(), v);
  }
  // value of v held by lane src(lane) (one ds_bpermute)
  template <class F>
  static __device__ __forceinline__ V permute(V v, F src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(src(lane()) << 2, __builtin_bit_cast(int, v)));
  }
  static __device__ __forceinline__ M umask(int E, long long p0, long long n) {
    M r = 0;
    for (int e = 0; e < E; ++e) r |= (p0 + e >= 0 && p0 + e < n) ? (1u << e) : 0u;
    return r;
  }
};
using LaneGpu = LaneGpuT<1>;
using RowGpu = LaneGpuT<4>;
#endif

template <int A, int B>
constexpr int fdiv() {
  return (A >= 0) ? A / B : -((-A + B - 1) / B);
}

template <int... I, class F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// x value at chunk-relative index J (may lie in another lane's chunk)
template <class P, int E, int J>
__device__ __forceinline__ typename P::V xat(const typename P::V (&x)[E]) {
  constexpr int q = fdiv<J, E>();
  constexpr int r = J - q * E;
  if constexpr (q == 0) return x[r];
  else return P::template from<q>(x[r]);
}

// One lifting step on E elements per lane: y[e] (+)= stencil of x.
template <class P, int E, int KIND, int OFF, int N, bool SUB>
__device__ __forceinline__ void rstep(typename P::V (&y)[E], const typename P::V (&x)[E], const float* cf) {
  using V = typename P::V;
  constexpr int NH = E + N - 1;  // x indices OFF .. OFF + E + N - 2
  V hx[NH];
  sfor<NH>([&](auto j) { hx[j] = xat<P, E, OFF + decltype(j)::value>(x); });
  if constexpr (KIND == kGeneric) {
    float c[N];
    sfor<N>([&](auto k) { c[k] = cf[k]; });
    sfor<E>([&](auto e) {
      V acc = y[e];
      sfor<N>([&](auto k) { acc = P::fma(c[k], hx[decltype(e)::value + decltype(k)::value], acc); });
      y[e] = acc;
    });
  } else if constexpr (KIND == kRepeat) {
    const float c0 = cf[0];
    sfor<E>([&](auto e) {
      V t = hx[e];
      sfor<N - 1>([&](auto k) { t = P::add(t, hx[decltype(e)::value + decltype(k)::value + 1]); });
      y[e] = P::fma(c0, t, y[e]);
    });
  } else {
    sfor<E>([&](auto e) { y[e] = SUB ? P::sub(y[e], hx[e]) : P::add(y[e], hx[e]); });
  }
}

// All steps of one level (wavelet W, transform TF) on E elements per lane.
// With masks (windows overlapping a band end, Zero boundary), the updated band
// is zeroed outside its range after every step: the next steps then read exact
// zeros there, which is what skipping out-of-range terms computes.
// cf: the program's coefficients (wt::NCoef of them), held in registers.
template <class P, class W, int TF, int E>
__device__ __forceinline__ void level(typename P::V (&s)[E], typename P::V (&d)[E], const float* cf_all,
                                      bool masked, const typename P::M& ms, const typename P::M& md) {
  using L = typename W::list;
  sfor<L::size>([&](auto I) {
    constexpr int i = decltype(I)::value;
    using ST = wt::Exec<L, TF, i>;
    const float* cf = cf_all + wt::CoefOff<L, TF, i>::value;
    if constexpr (ST::kind == kScale) {
      const float sm = cf[0], dm = cf[1];
      sfor<E>([&](auto e) {
        s[e] = P::mul(s[e], sm);
        d[e] = P::mul(d[e], dm);
      });
    } else if constexpr (ST::y_is_d) {
      rstep<P, E, ST::kind, ST::off, ST::n, ST::sub>(d, s, cf);
      if (masked) sfor<E>([&](auto e) { d[e] = P::keep(md, decltype(e)::value, d[e]); });
    } else {
      rstep<P, E, ST::kind, ST::off, ST::n, ST::sub>(s, d, cf);
      if (masked) sfor<E>([&](auto e) { s[e] = P::keep(ms, decltype(e)::value, s[e]); });
    }
  });
}

template <class P, int E>
__device__ __forceinline__ void split(const typename P::V (&x)[2 * E], typename P::V (&s)[E],
                                      typename P::V (&d)[E]) {
  sfor<E>([&](auto e) {
    s[e] = x[2 * decltype(e)::value];
    d[e] = x[2 * decltype(e)::value + 1];
  });
}
template <class P, int E>
__device__ __forceinline__ void merge(const typename P::V (&s)[E], const typename P::V (&d)[E],
                                      typename P::V (&x)[2 * E]) {
  sfor<E>([&](auto e) {
    x[2 * decltype(e)::value] = s[e];
    x[2 * decltype(e)::value + 1] = d[e];
  });
}

// store the owned part [lo, hi) of a band held E per lane from global index g0
template <class P, int E, class S>
__device__ __forceinline__ void store_band(S* p, long long pos0, const typename P::V (&v)[E], long long g0,
                                           long long lo, long long hi) {
  sfor<E>([&](auto e) {
    constexpr int ee = decltype(e)::value;
    P::st(p, v[ee], [&](int l) { return pos0 + g0 + (long long)l * E + ee; },
          [&](int l) {
            const long long g = g0 + (long long)l * E + ee;
            return g >= lo && g < hi;
          });
  });
}

// Lines of a row-mode tile: element stride between consecutive lines in each
// of the tile's three arrays (input, output, s band) and the number of lines
// in use (1..4). Ignored by the one-window policies.
struct RowArgs {
  long long rs[3];
  int n;
};

// Forward tile (Forward / InverseAdjoint) of K (1..4) levels: owned [a, b)
// of the input band of length m0 (b == m0 for the last tile), window
// [w0, w0 + 64 PL) with w0 = a - H (a multiple of 16). Lane l holds the
// PL samples w0 + PL l + e; after level l it holds PL / 2^l
// values of s_l and of d_l. Writes d_l of the owned range to out[ns_l + i]
// (l = 1..K) and s_K to sout[i]. A window inside the band reads no boundary;
// one that overlaps a band end (Zero boundary only) runs masked. K is uniform
// over the wavefront.
template <class P, class W, int TF, int PL = kPerLane, class SI, class SO, class SK>
__device__ __forceinline__ void fwd_tile(const DProgram<float>* prog, long long m0, int K, long long a, long long b,
                                         long long H, const SI* in, SO* out, SK* sout, float* lds,
                                         const RowArgs& ra = RowArgs{{0, 0, 0}, 1}) {
  using V = typename P::V;
  using M = typename P::M;
  static_assert(PL == 16 || PL == 32 || PL == 64, "register windows of 16, 32 or 64 samples per lane");
  constexpr int E1 = PL / 2, E2 = PL / 4, E3 = PL / 8, E4 = PL / 16;
  static_assert(P::kRows == 1 || E4 > 1, "row mode stores level-4 bands through the chunk transposes");
  const long long w0 = a - H;
  const bool masked = P::kRows > 1 || w0 < 0 || w0 + kLanes * PL > m0;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  V x[PL];
  P::template ld_chunks<PL>(in, w0, m0, x, lds, ra.rs[0], ra.n);
  V s1[E1], d1[E1], s2[E2], d2[E2], s3[E3], d3[E3], s4[E4], d4[E4];
  M ms{}, md{};
  long long m = m0, ns = m - m / 2, nd = m / 2;
  // owned end of a band after l levels: the band end for the last tile
  auto hi_d = [&](int l) { return b >= m0 ? nd : (b >> l); };
  auto hi_s = [&](int l) { return b >= m0 ? ns : (b >> l); };
  split<P, E1>(x, s1, d1);
  if (masked) {
    ms = P::band_mask(E1, w0 / 2, ns);
    md = P::band_mask(E1, w0 / 2, nd);
  }
  level<P, W, TF, E1>(s1, d1, cf, masked, ms, md);
  P::template st_chunks<E1>(out, ns, d1, w0 / 2, a >> 1, hi_d(1), lds, ra.rs[1], ra.n);
  if (K == 1) {
    P::template st_chunks<E1>(sout, 0, s1, w0 / 2, a >> 1, hi_s(1), lds, ra.rs[2], ra.n);
    return;
  }
  m = ns;
  ns = m - m / 2;
  nd = m / 2;
  split<P, E2>(s1, s2, d2);
  if (masked) {
    ms = P::band_mask(E2, w0 / 4, ns);
    md = P::band_mask(E2, w0 / 4, nd);
  }
  level<P, W, TF, E2>(s2, d2, cf, masked, ms, md);
  P::template st_chunks<E2>(out, ns, d2, w0 / 4, a >> 2, hi_d(2), lds, ra.rs[1], ra.n);
  if (K == 2) {
    P::template st_chunks<E2>(sout, 0, s2, w0 / 4, a >> 2, hi_s(2), lds, ra.rs[2], ra.n);
    return;
  }
  m = ns;
  ns = m - m / 2;
  nd = m / 2;
  split<P, E3>(s2, s3, d3);
  if (masked) {
    ms = P::band_mask(E3, w0 / 8, ns);
    md = P::band_mask(E3, w0 / 8, nd);
  }
  level<P, W, TF, E3>(s3, d3, cf, masked, ms, md);
  P::template st_chunks<E3>(out, ns, d3, w0 / 8, a >> 3, hi_d(3), lds, ra.rs[1], ra.n);
  if (K == 3) {
    P::template st_chunks<E3>(sout, 0, s3, w0 / 8, a >> 3, hi_s(3), lds, ra.rs[2], ra.n);
    return;
  }
  m = ns;
  ns = m - m / 2;
  nd = m / 2;
  split<P, E4>(s3, s4, d4);
  if (masked) {
    ms = P::band_mask(E4, w0 / 16, ns);
    md = P::band_mask(E4, w0 / 16, nd);
  }
  level<P, W, TF, E4>(s4, d4, cf, masked, ms, md);
  if constexpr (E4 == 1) {
    store_band<P, 1>(out, ns, d4, w0 / 16, a >> 4, hi_d(4));
    store_band<P, 1>(sout, 0, s4, w0 / 16, a >> 4, hi_s(4));
  } else {
    P::template st_chunks<E4>(out, ns, d4, w0 / 16, a >> 4, hi_d(4), lds, ra.rs[1], ra.n);
    P::template st_chunks<E4>(sout, 0, s4, w0 / 16, a >> 4, hi_s(4), lds, ra.rs[2], ra.n);
  }
}

// Reverse tile (Inverse / ForwardAdjoint) of K (1..4) levels: owned [a, b) of
// the output band of length m0 (b == m0 for the last tile). Windows are
// counted in level-4 units whatever K is: with g0 = (a >> 4) - Hk, the level-l
// windows are [2^(4-l) g0, 2^(4-l) (g0 + 4 PL)), PL / 2^l values per
// lane. Reads s_K from sin[i] and d_l from in[ns_l + i]; writes the owned band
// values to out[q]. Windows that overlap a band end (Zero boundary only) run
// masked. All global loads of the tile are issued first (the coarse levels
// then compute while the fine bands are in flight); each band is transposed
// to the chunk layout when its level runs.
template <class P, class W, int TF, int K, int PL, class SI, class SK, class SO>
__device__ __forceinline__ void rev_tile_k(const DProgram<float>* prog, long long m0, long long a, long long b,
                                           long long Hk, const SI* in, const SK* sin, SO* out, float* lds,
                                           const RowArgs& ra) {
  using V = typename P::V;
  using M = typename P::M;
  static_assert(PL == 16 || PL == 32 || PL == 64, "register windows of 16, 32 or 64 samples per lane");
  constexpr int E1 = PL / 2, E2 = PL / 4, E3 = PL / 8, E4 = PL / 16;
  constexpr int kU = (kLanes * PL) >> kLevels;  // the window in level-4 units
  static_assert(P::kRows == 1 || E4 > 1, "row mode loads level-4 bands through the chunk transposes");
  const long long g0 = (a >> 4) - Hk;
  const bool masked = P::kRows > 1 || g0 < 0 || 16 * (g0 + kU) > m0;
  const long long m1 = m0, m2 = m1 - m1 / 2, m3 = m2 - m2 / 2, m4 = m3 - m3 / 2;  // band split by level l
  const long long ns1 = m1 - m1 / 2, ns2 = m2 - m2 / 2, ns3 = m3 - m3 / 2, ns4 = m4 - m4 / 2;
  constexpr int NC = wt::NCoef<typename W::list, TF>::value;
  float cf[NC > 0 ? NC : 1];
  P::template coefs<NC>(prog->coef, cf);
  // raw (coalesced-order) loads: d_1 .. d_K and s_K
  constexpr int EK = PL >> K;  // values per lane of the level-K bands
  V rs[EK], r1[E1], r2[E2], r3[E3], r4[E4];
  if constexpr (K == 4) {
    if constexpr (E4 == 1) {
      rs[0] = P::ld(sin, [&](int l) { return g0 + l; }, [&](int l) { return g0 + l >= 0 && g0 + l < ns4; });
      r4[0] = P::ld(in + ns4, [&](int l) { return g0 + l; }, [&](int l) { return g0 + l >= 0 && g0 + l < m4 / 2; });
    } else {
      P::template ld_raw<E4>(sin, g0, ns4, rs, ra.rs[2], ra.n);
      P::template ld_raw<E4>(in + ns4, g0, m4 / 2, r4, ra.rs[0], ra.n);
    }
  } else if constexpr (K == 3) {
    P::template ld_raw<E3>(sin, 2 * g0, ns3, rs, ra.rs[2], ra.n);
  } else if constexpr (K == 2) {
    P::template ld_raw<E2>(sin, 4 * g0, ns2, rs, ra.rs[2], ra.n);
  } else {
    P::template ld_raw<E1>(sin, 8 * g0, ns1, rs, ra.rs[2], ra.n);
  }
  if constexpr (K >= 3) P::template ld_raw<E3>(in + ns3, 2 * g0, m3 / 2, r3, ra.rs[0], ra.n);
  if constexpr (K >= 2) P::template ld_raw<E2>(in + ns2, 4 * g0, m2 / 2, r2, ra.rs[0], ra.n);
  P::template ld_raw<E1>(in + ns1, 8 * g0, m1 / 2, r1, ra.rs[0], ra.n);
  V s3[E3], d3[E3], s2[E2], d2[E2], s1[E1], d1[E1], x[PL];
  M ms{}, md{};
  if constexpr (K == 4) {
    V s4[E4], d4[E4];
    if constexpr (E4 == 1) {
      s4[0] = rs[0];
      d4[0] = r4[0];
    } else {
      P::template xpose_in<E4>(rs, s4, lds);
      P::template xpose_in<E4>(r4, d4, lds);
    }
    if (masked) {
      ms = P::band_mask(E4, g0, ns4);
      md = P::band_mask(E4, g0, m4 / 2);
    }
    level<P, W, TF, E4>(s4, d4, cf, masked, ms, md);
    merge<P, E4>(s4, d4, s3);
  } else if constexpr (K == 3) {
    P::template xpose_in<E3>(rs, s3, lds);
  }
  if constexpr (K >= 3) {
    P::template xpose_in<E3>(r3, d3, lds);
    if (masked) {
      ms = P::band_mask(E3, 2 * g0, ns3);
      md = P::band_mask(E3, 2 * g0, m3 / 2);
    }
    level<P, W, TF, E3>(s3, d3, cf, masked, ms, md);
    merge<P, E3>(s3, d3, s2);
  } else if constexpr (K == 2) {
    P::template xpose_in<E2>(rs, s2, lds);
  }
  if constexpr (K >= 2) {
    P::template xpose_in<E2>(r2, d2, lds);
    if (masked) {
      ms = P::band_mask(E2, 4 * g0, ns2);
      md = P::band_mask(E2, 4 * g0, m2 / 2);
    }
    level<P, W, TF, E2>(s2, d2, cf, masked, ms, md);
    merge<P, E2>(s2, d2, s1);
  } else {
    P::template xpose_in<E1>(rs, s1, lds);
  }
  P::template xpose_in<E1>(r1, d1, lds);
  if (masked) {
    ms = P::band_mask(E1, 8 * g0, ns1);
    md = P::band_mask(E1, 8 * g0, m1 / 2);
  }
  level<P, W, TF, E1>(s1, d1, cf, masked, ms, md);
  merge<P, E1>(s1, d1, x);
  P::template st_chunks<PL>(out, 0, x, 16 * g0, a, b >= m0 ? m0 : b, lds, ra.rs[1], ra.n);
}

template <class P, class W, int TF, int PL = kPerLane, class SI, class SK, class SO>
__device__ __forceinline__ void rev_tile(const DProgram<float>* prog, long long m0, int K, long long a, long long b,
                                         long long Hk, const SI* in, const SK* sin, SO* out, float* lds,
                                         const RowArgs& ra = RowArgs{{0, 0, 0}, 1}) {
  switch (K) {
    case 4: rev_tile_k<P, W, TF, 4, PL>(prog, m0, a, b, Hk, in, sin, out, lds, ra); return;
    case 3: rev_tile_k<P, W, TF, 3, PL>(prog, m0, a, b, Hk, in, sin, out, lds, ra); return;
    case 2: rev_tile_k<P, W, TF, 2, PL>(prog, m0, a, b, Hk, in, sin, out, lds, ra); return;
    default: rev_tile_k<P, W, TF, 1, PL>(prog, m0, a, b, Hk, in, sin, out, lds, ra); return;
  }
}

}  // namespace reg
}  // namespace lfx
