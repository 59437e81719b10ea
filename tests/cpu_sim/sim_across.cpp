// CPU simulation of the across-line kernel (lfx_across.hpp) on the 64-lane
// model, driven by the library's own planning (lfx_across_plan.hpp): every
// workgroup in turn loads its block into a model of the LDS slab (the same
// phys() layout), runs its eight wavefronts one after another, and stores
// the block (test infrastructure).
//
// usage: sim_across <wavelet> <tf> <storage 0=f32 2=bf16> <batch> <n> <lines> <levels>
//   reads batch x n x lines values of the storage type on stdin (C order;
//   axis 1 transformed, axis 2 contiguous), writes the result on stdout;
//   prints "no plan" on stderr and exits 3 when the planner refuses.
#define LFX_CPU_SIM 1
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../lifted_amd/csrc/lfx_across.hpp"
#include "../../lifted_amd/csrc/lfx_across_plan.hpp"
#include "../../lifted_amd/csrc/lfx_table.hpp"

using namespace lfx;

static float g_slab[acr::kSlabFloats];

// The row-mode lane model whose memory is the slab model (virtual addresses
// mapped through acr::phys), as the GPU's slab policy.
struct RowLdsCpu : reg::LaneCpuT<4> {
  static float& at(const float* p, long long off) {
    const long long a = (p - g_slab) + off;
    if (a < 0 || a >= (long long)acr::kLines * acr::kMaxN) {
      std::fprintf(stderr, "slab access out of range: %lld\n", a);
      std::abort();
    }
    return g_slab[acr::phys((int)a)];
  }
  template <int E, class S>
  static void ld_raw(const S* p, long long pos0, long long n, V (&t)[E], long long rs = 0, int nr = 1) {
    for (int e = 0; e < E; ++e)
      for (int l = 0; l < reg::kLanes; ++l) {
        const int r = l / 16;
        const long long g = pos0 + (long long)E * (l % 16) + e;
        t[e].v[l] = (r < nr && g >= 0 && g < n) ? at(reinterpret_cast<const float*>(p), r * rs + g) : 0.0f;
      }
  }
  template <int E>
  static void xpose_in(const V (&t)[E], V (&x)[E], float*) {
    for (int e = 0; e < E; ++e) x[e] = t[e];
  }
  template <int E, class S>
  static void ld_chunks(const S* p, long long pos0, long long n, V (&x)[E], float*, long long rs = 0, int nr = 1) {
    ld_raw<E, S>(p, pos0, n, x, rs, nr);
  }
  template <int E, class S>
  static void st_chunks(S* p, long long pos0, const V (&v)[E], long long g0, long long lo, long long hi, float*,
                        long long rs = 0, int nr = 1) {
    for (int e = 0; e < E; ++e)
      for (int l = 0; l < reg::kLanes; ++l) {
        const int r = l / 16;
        const long long g = g0 + (long long)E * (l % 16) + e;
        if (r < nr && g >= lo && g < hi) at(reinterpret_cast<const float*>(p), r * rs + pos0 + g) = v[e].v[l];
      }
  }
};

template <int TF>
static bool run_wave(uint32_t wid, const acr::Args& A, float* vb) {
  switch (wid) {
#define LFX_SIM_W(nm)                                       \
  case wt::WL_##nm::id:                                     \
    acr::run_lines<RowLdsCpu, wt::WL_##nm, TF>(A, vb);      \
    return true;
    LFX_SIM_W(Daubechies1)
    LFX_SIM_W(Daubechies2)
    LFX_SIM_W(Daubechies10)
    LFX_SIM_W(Symlet5)
    LFX_SIM_W(Bior3_1)
    LFX_SIM_W(Bior6_8)
    LFX_SIM_W(CDF5_3)
    LFX_SIM_W(CDF9_7)
#undef LFX_SIM_W
  }
  return false;
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: sim_across wavelet tf storage batch n lines levels\n");
    return 2;
  }
  const uint32_t wid = (uint32_t)std::atol(argv[1]);
  const int tf = std::atoi(argv[2]), st = std::atoi(argv[3]);
  const long long B = std::atoll(argv[4]), N = std::atoll(argv[5]), C = std::atoll(argv[6]), L = std::atoll(argv[7]);
  const size_t es = st == kF32 ? 4 : 2, tot = (size_t)(B * N * C);
  std::vector<unsigned char> in(tot * es), out(tot * es, 0);
  if (std::fread(in.data(), 1, in.size(), stdin) != in.size()) return 2;
  try {
    const WaveletDefL w = find_wavelet(wid);
    const DProgram<float> hp = resolve<float>(w, tf);
    const bool fwd = tf == kForward || tf == kInverseAdjoint;
    acr::Args A{};
    if (!has_lift_steps(w) || !across_passes(hp, fwd, N, L, A)) {
      std::fprintf(stderr, "no plan\n");
      return 3;
    }
    const long long shape[3] = {B, N, C}, str[3] = {N * C, C, 1};
    across_geometry(3, shape, 1, 2, str, str, A);
    A.in = in.data();
    A.out = out.data();
    A.st_in = st;
    A.st_out = st;
    A.prog = &hp;
    A.wavelet = wid;
    for (long long blk = 0; blk < A.nblocks; ++blk) {
      long long c0, oi, oo;
      acr::block_origin(A, blk, c0, oi, oo);
      for (int i = 0; i < acr::kSlabFloats; ++i) g_slab[i] = -12345.0f;  // poison: every read slot is written first
      for (int t = 0; t < acr::kThreads; ++t) {
        float v[acr::kPer];
        acr::fetch(A, c0, oi, t, v);
        acr::put(A, g_slab, t, v);
      }
      for (int wv = 0; wv < acr::kGroups; ++wv) {
        float* vb = g_slab + wv * 4 * acr::kMaxN;
        const bool ok = tf == 0 ? run_wave<0>(wid, A, vb)
                        : tf == 1 ? run_wave<1>(wid, A, vb)
                        : tf == 2 ? run_wave<2>(wid, A, vb)
                                  : run_wave<3>(wid, A, vb);
        if (!ok) {
          std::fprintf(stderr, "wavelet not in the simulation\n");
          return 2;
        }
      }
      for (int t = 0; t < acr::kThreads; ++t) acr::store_block(A, g_slab, c0, oo, t);
    }
    std::fprintf(stderr, "blocks %lld passes %d\n", A.nblocks, A.npass);
  } catch (const Error& e) {
    std::fprintf(stderr, "error: %s\n", e.msg.c_str());
    return 1;
  }
  std::fwrite(out.data(), 1, out.size(), stdout);
  return 0;
}
