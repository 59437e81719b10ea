// CPU simulation of the GPU kernel bodies (test infrastructure).
//
// Compiles lifted_amd/csrc/lfx_bodies.hpp for the host with one lane per
// workgroup and runs the same schedule the library would launch for a single
// line (tiled passes + whole-line pass), with "LDS" buffers of exactly the size
// the launcher requests, so AddressSanitizer flags any out-of-window access.
//
// usage: lfx_sim <wavelet|custom> <tf> <bc> <dtype 0=f32 1=f64> <n> <levels>
//                <line_limit> <tile> <k> [steps]
//   reads n samples (dtype) on stdin, writes n samples to stdout.
//   steps (when wavelet == custom): "kind,target,offset,sign,ntaps,t0,t1,...;..."
#define LFX_CPU_SIM 1
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "../../lifted_amd/csrc/lfx_bodies.hpp"
#include "../../lifted_amd/csrc/lfx_schedule.hpp"
#include "../../lifted_amd/csrc/lfx_table.hpp"

using namespace lfx;

// register tiles of every table wavelet on the 64-lane CPU model: one
// translation unit per transform type (sim_reg_tf.inc), compiled in parallel
namespace lfx {
bool sim_reg_tf0(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg_tf1(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg_tf2(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg_tf3(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
}  // namespace lfx

template <class C>
static void reg_dispatch(int tf, uint32_t wid, const DProgram<C>& prog, const ItemSpec& it, const C* in, C* out, C* sk) {
  if constexpr (sizeof(C) == 4) {
    bool ok = false;
    switch (tf) {
      case 0: ok = sim_reg_tf0(wid, prog, it, in, out, sk); break;
      case 1: ok = sim_reg_tf1(wid, prog, it, in, out, sk); break;
      case 2: ok = sim_reg_tf2(wid, prog, it, in, out, sk); break;
      case 3: ok = sim_reg_tf3(wid, prog, it, in, out, sk); break;
    }
    if (!ok) fail(LFX_EINVAL, "no register kernel for this wavelet");
  } else {
    fail(LFX_EINVAL, "register tiles are float32 only");
  }
}

static WaveletDefL parse_custom(const std::string& spec) {
  WaveletDefL w{};
  w.width = 2;
  std::stringstream ss(spec);
  std::string item;
  while (std::getline(ss, item, ';')) {
    if (item.empty()) continue;
    std::vector<double> v;
    std::stringstream is(item);
    std::string tok;
    while (std::getline(is, tok, ',')) v.push_back(std::strtod(tok.c_str(), nullptr));
    StepDefL& s = w.steps[w.nsteps++];
    s = StepDefL{};
    s.kind = (int)v[0];
    s.target = (int)v[1];
    s.offset = (int)v[2];
    s.sign = (int)v[3];
    s.ntaps = (int)v[4];
    for (size_t k = 5; k < v.size(); ++k) s.taps[k - 5] = (long double)v[k];
  }
  return w;
}

template <class C>
static void run(const WaveletDefL& w, bool use_reg, int tf, int bc, long long n, long long L, long long limit,
                long long tile, int k, const std::vector<C>& in, std::vector<C>& out) {
  const DProgram<C> prog = resolve<C>(w, tf);
  const bool fwd = tf == kForward || tf == kInverseAdjoint;
  const int storage = sizeof(C) == 8 ? kF64 : kF32;
  const LineSchedule sc = schedule(prog, fwd, n, L, storage, bc, tile, k, limit);
  std::fprintf(stderr, "passes=%zu m_line=%lld L_line=%lld\n", sc.passes.size(), sc.m_line, sc.L_line);
  for (auto& p : sc.passes)
    std::fprintf(stderr, "  pass m0=%lld k=%d T=%lld H=%lld ntiles=%lld\n", p.m0, p.k, p.T, p.H, p.ntiles);
  auto line_run = [&](bool f, long long m, long long Ll, const C* src, const C* ssrc, C* dst) {
    std::vector<C> lds(line_lds_bytes(m, sizeof(C)) / sizeof(C));
    if (f) line_fwd_body<C, C, C>(&prog, bc, m, (int)Ll, src, 1, dst, 1, lds.data(), lds.data() + m);
    else line_rev_body<C, C, C, C>(&prog, bc, m, (int)Ll, src, 1, ssrc, 1, dst, 1, lds.data(), lds.data() + m);
  };
  if (sc.passes.empty()) {
    line_run(fwd, n, L, in.data(), in.data(), out.data());
    return;
  }
  if (fwd) {
    std::vector<C> cur(in);
    for (size_t i = 0; i < sc.passes.size(); ++i) {
      const PassPlan& pp = sc.passes[i];
      const bool last = (i + 1 == sc.passes.size()) && sc.L_line == 0;
      std::vector<C> sk(band_len(pp.m0, pp.k));
      for (const ItemSpec& it : pass_items(prog, true, pp, use_reg, storage, bc)) {
        C* skp = last ? out.data() : sk.data();
        if (it.kind == kSpecReg && !std::getenv("LFX_SIM_REG_AS_LDS")) {
          std::fprintf(stderr, "  reg [%lld, %lld) H=%lld\n", it.a, it.b, it.H);
          reg_dispatch<C>(tf, w.id, prog, it, cur.data(), out.data(), skp);
          continue;
        }
        const TileArgs ta{pp.m0, it.b - it.a, it.H, 0, pp.k, 0};
        std::vector<C> lds(tile_lds_bytes(ta, true, sizeof(C)) / sizeof(C));
        tile_fwd_body<C, C, C, C>(&prog, bc, pp.m0, pp.k, it.a, it.b, it.H, cur.data(), 1, out.data(), 1, skp, 1,
                                  lds.data(), lds.data() + tile_fwd_window(ta));
      }
      cur = sk;
    }
    if (sc.L_line > 0) line_run(true, sc.m_line, sc.L_line, cur.data(), cur.data(), out.data());
    return;
  }
  std::vector<C> s;
  bool s_user = true;
  if (sc.L_line > 0) {
    s.resize(sc.m_line);
    line_run(false, sc.m_line, sc.L_line, in.data(), in.data(), s.data());
    s_user = false;
  }
  for (size_t j = sc.passes.size(); j-- > 0;) {
    const PassPlan& pp = sc.passes[j];
    std::vector<C> band(j == 0 ? 0 : pp.m0);
    C* dst = j == 0 ? out.data() : band.data();
    for (const ItemSpec& it : pass_items(prog, false, pp, use_reg, storage, bc)) {
      const C* sp = s_user ? in.data() : s.data();
      if (it.kind == kSpecReg && !std::getenv("LFX_SIM_REG_AS_LDS")) {
        std::fprintf(stderr, "  reg [%lld, %lld) Hk=%lld\n", it.a, it.b, it.H);
        reg_dispatch<C>(tf, w.id, prog, it, in.data(), dst, const_cast<C*>(sp));
        continue;
      }
      const TileArgs ta{pp.m0, it.b - it.a, it.H, 0, pp.k, 0};
      std::vector<C> lds(tile_lds_bytes(ta, false, sizeof(C)) / sizeof(C));
      tile_rev_body<C, C, C, C>(&prog, bc, pp.m0, pp.k, it.a, it.b, it.H, in.data(), 1, sp, 1, dst, 1, lds.data(),
                                lds.data() + tile_rev_window(ta));
    }
    if (j > 0) {
      s = band;
      s_user = false;
    }
  }
}

int main(int argc, char** argv) {
  if (argc < 10) {
    std::fprintf(stderr, "usage: lfx_sim wavelet tf bc dtype n L limit tile k [steps]\n");
    return 2;
  }
  try {
    const std::string wname = argv[1];
    const int tf = std::atoi(argv[2]), bc = std::atoi(argv[3]), dt = std::atoi(argv[4]);
    const long long n = std::atoll(argv[5]), L = std::atoll(argv[6]), limit = std::atoll(argv[7]);
    const long long tile = std::atoll(argv[8]);
    const int k = std::atoi(argv[9]);
    // register tiles for table wavelets unless LFX_SIM_NOREG is set
    const bool use_reg = wname != "custom" && !std::getenv("LFX_SIM_NOREG");
    const WaveletDefL w = wname == "custom" ? parse_custom(argc > 10 ? argv[10] : "")
                                            : find_wavelet((uint32_t)std::strtoul(wname.c_str(), nullptr, 10));
    if (dt == 1) {
      std::vector<double> in(n), out(n, 0.0);
      if (std::fread(in.data(), 8, n, stdin) != (size_t)n) return 3;
      run<double>(w, false, tf, bc, n, L, limit, tile, k, in, out);
      std::fwrite(out.data(), 8, n, stdout);
    } else {
      std::vector<float> in(n), out(n, 0.0f);
      if (std::fread(in.data(), 4, n, stdin) != (size_t)n) return 3;
      run<float>(w, use_reg, tf, bc, n, L, limit, tile, k, in, out);
      std::fwrite(out.data(), 4, n, stdout);
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.msg.c_str());
    return 4;
  }
  return 0;
}
