// Checks the loop-free boundary maps of the kernels (lfx::dev::bc_map, host
// build) against the reference's iterated mirrors (src/lifted-common.hpp:
// 750-777, restated here as loops): Symmetric for every band length n >= 1 and
// Reflect for n >= 2, at every index |i| <= 4n + 3; Constant and Periodic at
// the same indices; and that degenerate bands (n = 0; n = 1 under Reflect)
// return without looping. Prints "ok" or the first mismatch.
#define LFX_CPU_SIM 1
#include <cstdio>
#include <initializer_list>

#include "../../lifted_amd/csrc/lfx_device.hpp"

static long long sym_iter(long long i, long long n) {
  long long io = i;
  while (io >= n || io < 0) io = io < 0 ? -(io + 1) : 2 * (n - 1) - (io - 1);
  return io;
}
static long long refl_iter(long long i, long long n) {
  long long io = i;
  while (io >= n || io < 0) io = io < 0 ? -io : 2 * (n - 1) - io;
  return io;
}

int main() {
  using lfx::dev::bc_map;
  long long checked = 0;
  for (long long n = 1; n <= 300; ++n) {
    for (long long i = -4 * n - 3; i <= 4 * n + 3; ++i) {
      const long long s = bc_map(lfx::kSymmetric, i, n), se = sym_iter(i, n);
      if (s != se) {
        std::printf("Symmetric n=%lld i=%lld: %lld, iterated %lld\n", n, i, s, se);
        return 1;
      }
      if (n >= 2) {
        const long long r = bc_map(lfx::kReflect, i, n), re = refl_iter(i, n);
        if (r != re) {
          std::printf("Reflect n=%lld i=%lld: %lld, iterated %lld\n", n, i, r, re);
          return 1;
        }
      }
      const long long c = bc_map(lfx::kConstant, i, n);
      if (c != (i < 0 ? 0 : (i >= n ? n - 1 : i))) {
        std::printf("Constant n=%lld i=%lld: %lld\n", n, i, c);
        return 1;
      }
      const long long p = bc_map(lfx::kPeriodic, i, n);
      if (p != (long long)((unsigned long long)i % (unsigned long long)n)) {
        std::printf("Periodic n=%lld i=%lld: %lld\n", n, i, p);
        return 1;
      }
      checked += 4;
    }
  }
  // large offsets (many periods away) still land in the band
  for (long long n = 2; n <= 40; ++n)
    for (long long i : {-1000003ll, 999983ll, -4000037ll, 8000017ll}) {
      if (bc_map(lfx::kSymmetric, i, n) != sym_iter(i, n) || bc_map(lfx::kReflect, i, n) != refl_iter(i, n)) {
        std::printf("far index n=%lld i=%lld\n", n, i);
        return 1;
      }
    }
  // degenerate bands: no loop, a skip (-1) or the only element
  if (bc_map(lfx::kSymmetric, -1, 0) != -1 || bc_map(lfx::kReflect, 5, 0) != -1 || bc_map(lfx::kReflect, -3, 1) != 0 ||
      bc_map(lfx::kReflect, 7, 1) != 0 || bc_map(lfx::kPeriodic, 3, 0) != -1) {
    std::printf("degenerate band\n");
    return 1;
  }
  std::printf("ok %lld\n", checked);
  return 0;
}
