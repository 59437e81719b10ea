// Uses include/lifted.hpp exactly as a caller of the reference's lifted.hpp
// would (src/lifted.hpp:13-39; tests/test_dwt.cpp builds its calls the same
// way): the four lwt overloads on host arrays, then checks
//   - Inverse(Forward(x)) == x  (both paths, float and double)
//   - <Forward x, y> == <x, ForwardAdjoint y>
//   - that invalid arguments throw std::invalid_argument.
// Mode "errors" runs only the argument checks that fail before any device
// work; mode "all" runs everything and prints one line of results; mode
// "dump DIR" runs one call of each overload on seeded inputs and writes the
// inputs and outputs as raw little-endian arrays (x_<case>.bin,
// y_<case>.bin) for the test suite to compare with the oracle.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
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

template <class T>
static void write_raw(const std::string& path, const std::vector<T>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) throw std::runtime_error("cannot write " + path);
  std::fclose(f);
}

template <class T>
static std::vector<T> seeded(size_t n, unsigned seed) {
  std::mt19937_64 g(seed);
  std::normal_distribution<double> nd;
  std::vector<T> x(n);
  for (auto& v : x) v = (T)nd(g);
  return x;
}

// One call of each lifted::lwt overload (src/lifted.hpp:13-39) on a 33 x 40 x 27
// box, axes {0, 2, 1}: level-by-level float / double (scalar level 2) and
// separable float / double (levels {2, 3, 1}); the double separable call reads
// a Fortran-ordered input and writes a C-ordered output.
static int dump(const std::string& dir) {
  const size_v shape{33, 40, 27};
  const stride_v c_st{40 * 27, 27, 1};
  const stride_v f_st{1, 33, 33 * 40};
  const size_v axes{0, 2, 1};
  const size_v lv{2, 3, 1};
  const size_t n = 33 * 40 * 27;
  {
    auto x = seeded<float>(n, 11);
    std::vector<float> y(n);
    lifted::lwt(Wavelet::CDF9_7, Transform::Forward, BoundaryCondition::Symmetric, shape, c_st, c_st, axes, (size_t)2,
                x.data(), y.data());
    write_raw(dir + "/x_lbl_f32.bin", x);
    write_raw(dir + "/y_lbl_f32.bin", y);
  }
  {
    auto x = seeded<double>(n, 12);
    std::vector<double> y(n);
    lifted::lwt(Wavelet::Daubechies4, Transform::Inverse, BoundaryCondition::Zero, shape, c_st, c_st, axes, (size_t)2,
                x.data(), y.data());
    write_raw(dir + "/x_lbl_f64.bin", x);
    write_raw(dir + "/y_lbl_f64.bin", y);
  }
  {
    auto x = seeded<float>(n, 13);
    std::vector<float> y(n);
    lifted::lwt(Wavelet::Bior4_4, Transform::ForwardAdjoint, BoundaryCondition::Constant, shape, c_st, c_st, axes, lv,
                x.data(), y.data());
    write_raw(dir + "/x_sep_f32.bin", x);
    write_raw(dir + "/y_sep_f32.bin", y);
  }
  {
    auto x = seeded<double>(n, 14);  // element (i, j, k) at i + 33 j + 1320 k
    std::vector<double> y(n);
    lifted::lwt(Wavelet::Symlet5, Transform::InverseAdjoint, BoundaryCondition::Reflect, shape, f_st, c_st, axes, lv,
                x.data(), y.data());
    write_raw(dir + "/x_sep_f64.bin", x);
    write_raw(dir + "/y_sep_f64.bin", y);
  }
  std::printf("dumped 4 cases to %s\n", dir.c_str());
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 2 && std::strcmp(argv[1], "dump") == 0) return dump(argv[2]);
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
