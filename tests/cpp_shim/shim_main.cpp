// Uses include/lifted.hpp exactly as a caller of the reference's lifted.hpp
// would (src/lifted.hpp:13-39; tests/test_dwt.cpp builds its calls the same
// way): the four lwt overloads on host arrays, then checks
//   - Inverse(Forward(x)) == x  (both paths, float and double)
//   - <Forward x, y> == <x, ForwardAdjoint y>
//   - that invalid arguments throw std::invalid_argument.
// Mode "errors" runs only the argument checks that fail before any device
// work; mode "all" runs everything and prints one line of results.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "lifted.hpp"

using lifted::BoundaryCondition;
using lifted::size_v;
using lifted::stride_v;
using lifted::Transform;
using lifted::Wavelet;

template <class T>
static double roundtrip(bool sep, Wavelet w, BoundaryCondition bc) {
  const size_v shape{33, 40, 27};
  const stride_v st{40 * 27, 27, 1};
  const size_v axes{0, 2, 1};
  std::mt19937_64 g(5);
  std::normal_distribution<double> nd;
  std::vector<T> x(33 * 40 * 27), y(x.size()), z(x.size());
  for (auto& v : x) v = (T)nd(g);
  if (sep) {
    const size_v lv{2, 3, 1};
    lifted::lwt(w, Transform::Forward, bc, shape, st, st, axes, lv, x.data(), y.data());
    lifted::lwt(w, Transform::Inverse, bc, shape, st, st, axes, lv, y.data(), z.data());
  } else {
    lifted::lwt(w, Transform::Forward, bc, shape, st, st, axes, (size_t)2, x.data(), y.data());
    lifted::lwt(w, Transform::Inverse, bc, shape, st, st, axes, (size_t)2, y.data(), z.data());
  }
  double e = 0, m = 0;
  for (size_t i = 0; i < x.size(); ++i) {
    e = std::fmax(e, std::fabs((double)z[i] - (double)x[i]));
    m = std::fmax(m, std::fabs((double)x[i]));
  }
  return e / m;
}

static double dot_test() {
  const size_v shape{64, 50};
  const stride_v st{50, 1};
  const size_v axes{0, 1};
  const size_v lv{3, 2};
  std::mt19937_64 g(9);
  std::normal_distribution<double> nd;
  std::vector<double> x(64 * 50), y(x.size()), wx(x.size()), wty(x.size());
  for (auto& v : x) v = nd(g);
  for (auto& v : y) v = nd(g);
  lifted::lwt(Wavelet::CDF9_7, Transform::Forward, BoundaryCondition::Symmetric, shape, st, st, axes, lv, x.data(),
              wx.data());
  lifted::lwt(Wavelet::CDF9_7, Transform::ForwardAdjoint, BoundaryCondition::Symmetric, shape, st, st, axes, lv,
              y.data(), wty.data());
  double a = 0, b = 0;
  for (size_t i = 0; i < x.size(); ++i) {
    a += wx[i] * y[i];
    b += x[i] * wty[i];
  }
  return std::fabs(a - b) / std::fabs(a);
}

static int errors() {
  int ok = 0;
  const size_v shape{16};
  const stride_v st{1};
  std::vector<float> x(16), y(16);
  try {  // unknown boundary condition
    lifted::lwt(Wavelet::CDF9_7, Transform::Forward, (BoundaryCondition)7, shape, st, st, size_v{0}, (size_t)1,
                x.data(), y.data());
  } catch (const std::invalid_argument&) {
    ++ok;
  }
  try {  // strides of the wrong length
    lifted::lwt(Wavelet::CDF9_7, Transform::Forward, BoundaryCondition::Zero, shape, stride_v{1, 1}, st, size_v{0},
                (size_t)1, x.data(), y.data());
  } catch (const std::invalid_argument&) {
    ++ok;
  }
  try {  // max_level of the width-1 Lazy wavelet
    (void)lifted::max_level(Wavelet::Lazy, 64);
  } catch (const std::invalid_argument&) {
    ++ok;
  }
  return ok;
}

int main(int argc, char** argv) {
  const bool all = argc > 1 && std::strcmp(argv[1], "all") == 0;
  const int ok = errors();
  std::printf("errors_thrown %d/3 max_level_db2_65536 %zu\n", ok, lifted::max_level(Wavelet::Daubechies2, 65536));
  if (!all) return ok == 3 ? 0 : 1;
  double worst = 0;
  worst = std::fmax(worst, roundtrip<float>(true, Wavelet::CDF9_7, BoundaryCondition::Symmetric) / 1e-5);
  worst = std::fmax(worst, roundtrip<float>(false, Wavelet::Daubechies4, BoundaryCondition::Zero) / 1e-5);
  worst = std::fmax(worst, roundtrip<double>(true, Wavelet::Bior4_4, BoundaryCondition::Constant) / 1e-12);
  worst = std::fmax(worst, roundtrip<double>(false, Wavelet::Symlet5, BoundaryCondition::Reflect) / 1e-8);  // Symlet5 lifting is ill-conditioned
  const double d = dot_test();
  std::printf("roundtrip_worst_over_tol %.3g dot_rel %.3g\n", worst, d);
  return (ok == 3 && worst <= 1.0 && d <= 1e-12) ? 0 : 1;
}
