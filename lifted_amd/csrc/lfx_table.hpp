// Wavelet definitions and their resolution into device step programs
// (host side; shared by the library and the CPU simulation build).
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>

#include "lfx_error.hpp"
#include "lfx_program.hpp"

namespace lfx {

constexpr uint32_t kCustomBase = 0x80000000u;

// Definition of a table wavelet (reference enum value) or a registered custom
// program (ids from kCustomBase); throws Error for unknown ids.
WaveletDefL find_wavelet(uint32_t id);
uint32_t register_custom(WaveletDefL w);
// Position of a table wavelet in the table (-1: not a table wavelet), the
// table's size, and its i-th entry.
int table_index(uint32_t id);
int table_size();
WaveletDefL table_entry(int i);

// Resolve a wavelet for one transform type (see lfx_program.hpp).
template <class C>
inline DProgram<C> resolve(const WaveletDefL& w, int tf) {
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

// max_level (lifted-common.hpp:1099-1108); -1 for width 1 (never terminates there)
inline long long max_level_w(long long width, long long n) {
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

inline bool has_lift_steps(const WaveletDefL& w) {
  for (int i = 0; i < w.nsteps; ++i)
    if (w.steps[i].kind != kScale) return true;
  return false;
}

// Bands the reference cannot process: empty bands under a non-Zero BC divide
// by zero (Periodic) or loop forever (Symmetric, Reflect), as do bands of
// length 1 under Reflect.
inline void check_bands(int bc, bool steps, long long n, long long L) {
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

}  // namespace lfx
