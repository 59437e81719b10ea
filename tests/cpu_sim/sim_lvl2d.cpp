// CPU simulation of the two-axis level tiles (lfx_lvl2d.hpp) on the 64-lane
// model, driven by the library's own plan (lfx_lvl2d_plan.hpp): every tile of
// every level, in launch order (test infrastructure).
//
// usage: sim_lvl2d <wavelet> <tf> <storage 0=f32 2=bf16> <batch> <rows> <cols> <levels>
//   reads batch x rows x cols values of the storage type on stdin (C order;
//   axes (1, 2), rows transformed first), writes the result on stdout; prints
//   "no plan" on stderr and exits 3 when the plan refuses the shape.
#define LFX_CPU_SIM 1
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../lifted_amd/csrc/lfx_lvl2d.hpp"
#include "../../lifted_amd/csrc/lfx_lvl2d_plan.hpp"
#include "../../lifted_amd/csrc/lfx_table.hpp"

using namespace lfx;

template <int TF, int ST>
static bool run_tile_st(uint32_t wid, const l2d::Level& L, long long w) {
  switch (wid) {
#define LFX_SIM_W(nm)                                          \
  case wt::WL_##nm::id:                                        \
    l2d::level_tile<reg::LaneCpu, wt::WL_##nm, TF, ST>(L, w);  \
    return true;
    LFX_SIM_W(Daubechies1)
    LFX_SIM_W(Daubechies2)
    LFX_SIM_W(Daubechies3)
    LFX_SIM_W(Daubechies10)
    LFX_SIM_W(Symlet5)
    LFX_SIM_W(Bior2_2)
    LFX_SIM_W(Bior3_1)
    LFX_SIM_W(Bior6_8)
    LFX_SIM_W(CDF5_3)
    LFX_SIM_W(CDF9_7)
#undef LFX_SIM_W
  }
  return false;
}

// the kernels are compiled per storage type (float32 or bfloat16 here)
template <int TF>
static bool run_tile(uint32_t wid, const l2d::Level& L, long long w) {
  return L.a[0].st == kBF16 ? run_tile_st<TF, kBF16>(wid, L, w) : run_tile_st<TF, kF32>(wid, L, w);
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: sim_lvl2d wavelet tf storage batch rows cols levels\n");
    return 2;
  }
  const uint32_t wid = (uint32_t)std::atol(argv[1]);
  const int tf = std::atoi(argv[2]), st = std::atoi(argv[3]);
  const long long B = std::atoll(argv[4]), R = std::atoll(argv[5]), C = std::atoll(argv[6]), L = std::atoll(argv[7]);
  const size_t es = st == kF32 ? 4 : 2, n = (size_t)(B * R * C);
  std::vector<unsigned char> in(n * es), out(n * es, 0);
  if (std::fread(in.data(), 1, in.size(), stdin) != in.size()) return 2;
  try {
    const WaveletDefL w = find_wavelet(wid);
    const DProgram<float> hp = resolve<float>(w, tf);
    const long long shape[3] = {B, R, C}, str[3] = {R * C, C, 1};
    const long long wse = lvl2d_ws_elems(3, shape, 1, 2);
    std::vector<unsigned char> ws0((size_t)wse * es), ws1((size_t)wse * es);
    const bool fwd = tf == kForward || tf == kInverseAdjoint;
    const auto levels = lvl2d_plan(hp, fwd, 3, shape, 1, 2, L, in.data(), str, out.data(), str, st, ws0.data(),
                                   ws1.data(), &hp, wid);
    if (levels.empty()) {
      std::fprintf(stderr, "no plan\n");
      return 3;
    }
    for (const auto& lv : levels)
      for (long long t = 0; t < lv.ntiles; ++t) {
        const bool ok = tf == 0 ? run_tile<0>(wid, lv, t)
                        : tf == 1 ? run_tile<1>(wid, lv, t)
                        : tf == 2 ? run_tile<2>(wid, lv, t)
                                  : run_tile<3>(wid, lv, t);
        if (!ok) {
          std::fprintf(stderr, "wavelet not in the simulation\n");
          return 2;
        }
      }
    std::fprintf(stderr, "levels %zu tiles0 %lld H %lld/%lld\n", levels.size(), levels[0].ntiles, levels[0].Hr,
                 levels[0].Hc);
  } catch (const Error& e) {
    std::fprintf(stderr, "error: %s\n", e.msg.c_str());
    return 1;
  }
  std::fwrite(out.data(), 1, out.size(), stdout);
  return 0;
}
