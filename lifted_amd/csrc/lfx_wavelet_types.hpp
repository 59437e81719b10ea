// Compile-time step shapes of the predefined wavelets.
//
// The reference defines each wavelet as a compile-time tuple of steps
// (src/lifted-wavelets.hpp); the register-tiled kernels need the same: the
// kind, updated band, offset and tap count of every step must be constants so
// that stencils over register arrays unroll with constant indices. This header
// expands the same generated table (wavelet_table.inc) into one type per
// wavelet. Coefficients are not part of the type: kernels read them, already
// cast and signed for the transform, from the runtime program.
#pragma once
#include <cstdint>

#include "lfx_program.hpp"

namespace lfx {
namespace wt {

// One step as declared: kind, target (0 s, 1 d), offset, taps, unit sign.
template <int KIND, int TGT, int OFF, int N, int SIGN>
struct SDef {
  static constexpr int kind = KIND, tgt = TGT, off = OFF, n = N, sign = SIGN;
};
struct SEnd {};

template <class... S>
struct SList;
template <class... S>
struct SList<SEnd, S...> {  // the table's lists end with SEnd (trailing comma of the expansion)
  static constexpr int size = 0;
};
template <class H, class... S>
struct SList<H, S...> {
  static constexpr int size = 1 + SList<S...>::size;
};
template <>
struct SList<> {
  static constexpr int size = 0;
};

template <int I, class L>
struct At;
template <class H, class... S>
struct At<0, SList<H, S...>> {
  using type = H;
};
template <int I, class H, class... S>
struct At<I, SList<H, S...>> {
  using type = typename At<I - 1, SList<S...>>::type;
};

// Step I (execution order) of a wavelet for transform TF, resolved like the
// runtime program (lfx_api / lfx_table resolve()).
template <class L, int TF, int I>
struct Exec {
  static constexpr bool fwd_stepping = TF == kForward || TF == kInverseAdjoint;
  static constexpr bool normal = TF == kForward || TF == kInverse;
  static constexpr bool fwd_tf = TF == kForward || TF == kForwardAdjoint;
  using D = typename At<fwd_stepping ? I : L::size - 1 - I, L>::type;
  static constexpr int kind = D::kind;
  static constexpr int y_is_d = ((D::tgt == 1) == normal) ? 1 : 0;
  static constexpr int off = normal ? D::off : -(D::n - 1 + D::off);
  static constexpr int n = D::n;
  static constexpr bool sub = D::kind == kUnit ? ((D::sign > 0) != fwd_tf) : false;
};

// Index of step I's first coefficient in the resolved program (the runtime
// resolution packs them in execution order: scale 2, generic N, repeat 1,
// unit none), and the program's coefficient count.
template <class L, int TF, int I>
struct CoefOff {
  using P = Exec<L, TF, I - 1>;
  static constexpr int value =
      CoefOff<L, TF, I - 1>::value + (P::kind == kScale ? 2 : (P::kind == kGeneric ? P::n : (P::kind == kRepeat ? 1 : 0)));
};
template <class L, int TF>
struct CoefOff<L, TF, 0> {
  static constexpr int value = 0;
};
template <class L, int TF>
struct NCoef {
  static constexpr int value = CoefOff<L, TF, L::size>::value;
};

#define LFX_NARGS_(a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15, a16, n, ...) n
#define LFX_NARGS(...) LFX_NARGS_(__VA_ARGS__, 16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0)

constexpr int S = 0, D = 1;
#define LFX_WAVELET(nm, idv, w)                     \
  struct WL_##nm {                                  \
    static constexpr uint32_t id = idv;             \
    static constexpr int width = w;                 \
    using list = SList<
#define LFX_GENERIC(b, off, ...) SDef<kGeneric, b, off, LFX_NARGS(__VA_ARGS__), 1>,
#define LFX_UNIT(b, off, sg) SDef<kUnit, b, off, 1, sg>,
#define LFX_REPEAT(b, off, n, val) SDef<kRepeat, b, off, n, 1>,
#define LFX_SCALE(a, b) SDef<kScale, 0, 0, 0, 1>,
#define LFX_END(nm) \
  SEnd > ;          \
  }                 \
  ;
#include "wavelet_table.inc"
#undef LFX_WAVELET
#undef LFX_GENERIC
#undef LFX_UNIT
#undef LFX_REPEAT
#undef LFX_SCALE
#undef LFX_END

}  // namespace wt
}  // namespace lfx
