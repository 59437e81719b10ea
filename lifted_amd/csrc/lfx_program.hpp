// Step programs: the reference's compile-time wavelet step tuples
// (src/lifted-wavelets.hpp, src/lifted-common.hpp:798-964) resolved at run time
// into a flat, device-resident form for one (wavelet, transform, compute type).
//
// A resolved program lists its steps in *execution order* for that transform
// (declared order for Forward / InverseAdjoint, reversed for Inverse /
// ForwardAdjoint -- src/lifted-drivers-inl.hpp:51-63) and carries coefficients
// already cast to the compute type and already signed for the transform
// (y -= c*x becomes y += (-c)*x; negation is exact, so results are unchanged).
#pragma once
#include <cstdint>

namespace lfx {

constexpr int kMaxSteps = 16;  // Daubechies10 has 11 lifting steps + scale
constexpr int kMaxTaps = 16;   // predefined wavelets use at most 9 taps

enum StepKind : int32_t { kGeneric = 0, kUnit = 1, kRepeat = 2, kScale = 3 };
enum Bc : int32_t { kZero = 0, kConstant = 1, kPeriodic = 2, kSymmetric = 3, kReflect = 4 };
enum Tf : int32_t { kForward = 0, kInverse = 1, kForwardAdjoint = 2, kInverseAdjoint = 3 };

// One resolved step (device POD, 48 bytes).
struct DStep {
  int32_t kind;      // StepKind
  int32_t y_is_d;    // updated band is d (odd samples) -- else s (even samples)
  int32_t adjoint;   // transposed form: zero-BC stencil plus boundary folds
  int32_t sub;       // unit step: subtract (generic/repeat carry the sign in coef)
  int32_t off;       // stencil offset: o (normal) or o_r = -(N-1+o) (adjoint)
  int32_t N;         // taps (generic), summed taps (repeat), 1 (unit)
  int32_t nf, nb;    // n_front / n_back (reversed for adjoint): interior split
  int32_t fold_off;  // forward offset o: first virtual index of the front folds
  int32_t coef;      // index of the first coefficient in DProgram::coef
  int32_t pad0, pad1;
};

template <class C>
struct DProgram {
  int32_t nsteps;
  int32_t width;     // filter width (for max_level)
  int32_t reach_lo;  // max left reach of any step (for tile halos), >= 0
  int32_t reach_hi;  // max right reach of any step, >= 0
  DStep steps[kMaxSteps];
  C coef[kMaxSteps * kMaxTaps];
};

// Host-side definition of a step with long double coefficients, as the
// reference writes them (static_cast<T>(long double literal)).
struct StepDefL {
  int32_t kind;    // StepKind
  int32_t target;  // 0: update evens (s), 1: update odds (d)
  int32_t offset;
  int32_t ntaps;   // generic: taps; repeat: N; unit: 1
  int32_t sign;    // unit: +1 / -1
  long double taps[kMaxTaps];  // generic taps; repeat: taps[0]; scale: {s_mul, d_div}
};

struct WaveletDefL {
  uint32_t id;
  const char* name;
  uint32_t width;
  int32_t nsteps;
  StepDefL steps[kMaxSteps];
};

}  // namespace lfx
