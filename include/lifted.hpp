// Drop-in C++ interface of jcapriot/lifted (src/lifted.hpp:11-39) on top of
// the lifted-hip C ABI (lifted_hip.h): the same namespace, enum types and
// values, vector aliases and the four lwt overloads, so that a caller of the
// reference compiles unchanged against this header and links liblifted_hip.so.
//
// The work runs on the current HIP device: host arrays are copied in, the
// transform runs, the result is copied back, and the call returns when done
// (the reference's blocking semantics). n_threads is accepted and ignored.
// Errors throw: std::invalid_argument for invalid arguments (the reference
// throws it for an unknown boundary condition, src/lifted-steps-inl.hpp:428),
// std::runtime_error for device failures. Header-only; C++17 or later.
#ifndef LIFTED_HIP_LIFTED_HPP_
#define LIFTED_HIP_LIFTED_HPP_

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lifted_hip.h"

namespace lifted {

using std::ptrdiff_t;
using std::size_t;

using size_v = std::vector<size_t>;
using stride_v = std::vector<ptrdiff_t>;
using stride_t = stride_v::value_type;

// Values are the reference's ABI (src/lifted-common.hpp:33-98).
enum class Wavelet : std::uint32_t {
  Lazy = 0,
  Daubechies1 = 1, Daubechies2 = 2, Daubechies3 = 3, Daubechies4 = 4, Daubechies5 = 5,
  Daubechies6 = 6, Daubechies7 = 7, Daubechies8 = 8, Daubechies9 = 9, Daubechies10 = 10,
  Symlet4 = 104, Symlet5 = 105, Symlet6 = 106,
  Coiflet2 = 202, Coiflet3 = 203,
  Bior1_3 = 10103, Bior1_5 = 10105,
  Bior2_2 = 10202, Bior2_4 = 10204, Bior2_6 = 10206, Bior2_8 = 10208,
  Bior3_1 = 10301, Bior3_3 = 10303, Bior3_5 = 10305, Bior3_7 = 10307, Bior3_9 = 10309,
  Bior4_2 = 10402, Bior4_4 = 10404, Bior4_6 = 10406,
  Bior5_5 = 10505, Bior6_8 = 10608,
  CDF5_3 = 100503, CDF9_7 = 100907,
};

enum class BoundaryCondition : std::uint32_t { Zero = 0, Constant = 1, Periodic = 2, Symmetric = 3, Reflect = 4 };

enum class Transform : std::uint32_t { Forward = 0, Inverse = 1, ForwardAdjoint = 2, InverseAdjoint = 3 };

namespace hip_detail {

inline void check(int rc) {
  if (rc >= 0) return;
  const std::string msg = std::string("lifted-hip: ") + lfx_last_error();
  if (rc == LFX_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline void check_sizes(const size_v& shape, const stride_v& si, const stride_v& so) {
  if (si.size() != shape.size() || so.size() != shape.size())
    throw std::invalid_argument("lifted-hip: shape and strides differ in length");
}

}  // namespace hip_detail

// Scalar `levels`: the level-by-level path (src/lifted.hpp:13-25).
inline void lwt(const Wavelet wvlt, const Transform op, const BoundaryCondition bc, const size_v& shape,
                const stride_v& stride_in, const stride_v& stride_out, const size_v& axes, const size_t levels,
                const float* data_in, float* data_out, const size_t n_threads = 1) {
  hip_detail::check_sizes(shape, stride_in, stride_out);
  hip_detail::check(lfx_lwt_f32((uint32_t)wvlt, (uint32_t)op, (uint32_t)bc, shape.size(), shape.data(),
                                stride_in.data(), stride_out.data(), axes.size(), axes.data(), levels, data_in,
                                data_out, n_threads));
}

inline void lwt(const Wavelet wvlt, const Transform op, const BoundaryCondition bc, const size_v& shape,
                const stride_v& stride_in, const stride_v& stride_out, const size_v& axes, const size_t levels,
                const double* data_in, double* data_out, const size_t n_threads = 1) {
  hip_detail::check_sizes(shape, stride_in, stride_out);
  hip_detail::check(lfx_lwt_f64((uint32_t)wvlt, (uint32_t)op, (uint32_t)bc, shape.size(), shape.data(),
                                stride_in.data(), stride_out.data(), axes.size(), axes.data(), levels, data_in,
                                data_out, n_threads));
}

// Per-axis `levels`: the separable path (src/lifted.hpp:27-39).
inline void lwt(const Wavelet wvlt, const Transform op, const BoundaryCondition bc, const size_v& shape,
                const stride_v& stride_in, const stride_v& stride_out, const size_v& axes, const size_v& levels,
                const float* data_in, float* data_out, const size_t n_threads = 1) {
  hip_detail::check_sizes(shape, stride_in, stride_out);
  if (levels.size() != axes.size()) throw std::invalid_argument("lifted-hip: levels and axes differ in length");
  hip_detail::check(lfx_lwt_sep_f32((uint32_t)wvlt, (uint32_t)op, (uint32_t)bc, shape.size(), shape.data(),
                                    stride_in.data(), stride_out.data(), axes.size(), axes.data(), levels.data(),
                                    data_in, data_out, n_threads));
}

inline void lwt(const Wavelet wvlt, const Transform op, const BoundaryCondition bc, const size_v& shape,
                const stride_v& stride_in, const stride_v& stride_out, const size_v& axes, const size_v& levels,
                const double* data_in, double* data_out, const size_t n_threads = 1) {
  hip_detail::check_sizes(shape, stride_in, stride_out);
  if (levels.size() != axes.size()) throw std::invalid_argument("lifted-hip: levels and axes differ in length");
  hip_detail::check(lfx_lwt_sep_f64((uint32_t)wvlt, (uint32_t)op, (uint32_t)bc, shape.size(), shape.data(),
                                    stride_in.data(), stride_out.data(), axes.size(), axes.data(), levels.data(),
                                    data_in, data_out, n_threads));
}

// max_level(width(wavelet), n) (src/lifted-common.hpp:1099-1108); throws for
// the width-1 Lazy wavelet, where the reference's loop does not terminate.
inline size_t max_level(const Wavelet wvlt, const size_t n) {
  const int64_t r = lfx_max_level((uint32_t)wvlt, (int64_t)n);
  hip_detail::check((int)(r < 0 ? r : 0));
  return (size_t)r;
}

}  // namespace lifted

#endif  // LIFTED_HIP_LIFTED_HPP_
