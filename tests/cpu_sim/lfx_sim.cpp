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
#include "../../lifted_amd/csrc/lfx_reg.hpp"
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
bool sim_reg16_tf0(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg16_tf1(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg16_tf2(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_reg16_tf3(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*);
bool sim_rows_tf0(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*, const reg::RowArgs&);
bool sim_rows_tf1(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*, const reg::RowArgs&);
bool sim_rows_tf2(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*, const reg::RowArgs&);
bool sim_rows_tf3(uint32_t, const DProgram<float>&, const ItemSpec&, const float*, float*, float*, const reg::RowArgs&);
}  // namespace lfx

// register window of the simulated plans: LFX_SIM_STRIP=1024 for 16 samples
// per lane (a subset of the table), else the default 2048
static long long sim_window() {
  const char* e = std::getenv("LFX_SIM_STRIP");
  return e && std::atoll(e) == 1024 ? 1024 : kRegWindow;
}

template <class C>
static void reg_dispatch(int tf, uint32_t wid, const DProgram<C>& prog, const ItemSpec& it, const C* in, C* out, C* sk) {
  if constexpr (sizeof(C) == 4) {
    bool ok = false;
    const bool small = sim_window() == 1024;
    switch (tf) {
      case 0: ok = small ? sim_reg16_tf0(wid, prog, it, in, out, sk) : sim_reg_tf0(wid, prog, it, in, out, sk); break;
      case 1: ok = small ? sim_reg16_tf1(wid, prog, it, in, out, sk) : sim_reg_tf1(wid, prog, it, in, out, sk); break;
      case 2: ok = small ? sim_reg16_tf2(wid, prog, it, in, out, sk) : sim_reg_tf2(wid, prog, it, in, out, sk); break;
      case 3: ok = small ? sim_reg16_tf3(wid, prog, it, in, out, sk) : sim_reg_tf3(wid, prog, it, in, out, sk); break;
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
      for (const ItemSpec& it : pass_items(prog, true, pp, use_reg, storage, bc, sim_window())) {
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
    for (const ItemSpec& it : pass_items(prog, false, pp, use_reg, storage, bc, sim_window())) {
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

// R equal lines (float32) through the same schedule; register items that hold
// a whole band of at most 512 samples under the Zero boundary run as row items
// (up to four lines per wavefront), like a plan of many equal lines.
static void run_rows(const WaveletDefL& w, int tf, int bc, long long n, long long L, long long limit, long long tile,
                     int k, int R, const std::vector<float>& in, std::vector<float>& out) {
  const DProgram<float> prog = resolve<float>(w, tf);
  const bool fwd = tf == kForward || tf == kInverseAdjoint;
  const LineSchedule sc = schedule(prog, fwd, n, L, kF32, bc, tile, k, limit);
  std::fprintf(stderr, "rows=%d passes=%zu m_line=%lld L_line=%lld\n", R, sc.passes.size(), sc.m_line, sc.L_line);
  auto rows_call = [&](const ItemSpec& it, const float* a0, float* a1, float* a2, long long rs0, long long rs1,
                       long long rs2, int nr) {
    const reg::RowArgs ra{{rs0, rs1, rs2}, nr};
    bool ok = false;
    switch (tf) {
      case 0: ok = sim_rows_tf0(w.id, prog, it, a0, a1, a2, ra); break;
      case 1: ok = sim_rows_tf1(w.id, prog, it, a0, a1, a2, ra); break;
      case 2: ok = sim_rows_tf2(w.id, prog, it, a0, a1, a2, ra); break;
      case 3: ok = sim_rows_tf3(w.id, prog, it, a0, a1, a2, ra); break;
    }
    if (!ok) fail(LFX_EINVAL, "no row kernel for this wavelet");
  };
  auto row_ok = [&](const ItemSpec& it) {
    return it.kind == kSpecReg && bc == kZero && it.a == 0 && it.b == it.m0 && it.m0 <= kRowMax;
  };
  auto line_run = [&](bool f, long long m, long long Ll, const float* src, const float* ssrc, float* dst) {
    std::vector<float> lds(line_lds_bytes(m, 4) / 4);
    if (f) line_fwd_body<float, float, float>(&prog, bc, m, (int)Ll, src, 1, dst, 1, lds.data(), lds.data() + m);
    else line_rev_body<float, float, float, float>(&prog, bc, m, (int)Ll, src, 1, ssrc, 1, dst, 1, lds.data(),
                                                   lds.data() + m);
  };
  if (fwd) {
    std::vector<float> cur(in);
    long long cs = n;  // row stride of cur
    for (size_t i = 0; i < sc.passes.size(); ++i) {
      const PassPlan& pp = sc.passes[i];
      const bool last = (i + 1 == sc.passes.size()) && sc.L_line == 0;
      const long long mk = band_len(pp.m0, pp.k);
      std::vector<float> sk((size_t)R * mk);
      const long long ss = last ? n : mk;
      for (const ItemSpec& it : pass_items(prog, true, pp, true, kF32, bc)) {
        if (row_ok(it)) {
          std::fprintf(stderr, "  rows [0, %lld) k=%d\n", it.m0, it.k);
          for (int r0 = 0; r0 < R; r0 += 4)
            rows_call(it, cur.data() + r0 * cs, out.data() + r0 * n, (last ? out.data() + r0 * n : sk.data() + r0 * mk),
                      cs, n, ss, std::min(4, R - r0));
          continue;
        }
        for (int r = 0; r < R; ++r) {
          float* skp = last ? out.data() + (size_t)r * n : sk.data() + (size_t)r * mk;
          const float* cp = cur.data() + (size_t)r * cs;
          float* op = out.data() + (size_t)r * n;
          if (it.kind == kSpecReg) {
            reg_dispatch<float>(tf, w.id, prog, it, cp, op, skp);
            continue;
          }
          const TileArgs ta{pp.m0, it.b - it.a, it.H, 0, pp.k, 0};
          std::vector<float> lds(tile_lds_bytes(ta, true, 4) / 4);
          tile_fwd_body<float, float, float, float>(&prog, bc, pp.m0, pp.k, it.a, it.b, it.H, cp, 1, op, 1, skp, 1,
                                                    lds.data(), lds.data() + tile_fwd_window(ta));
        }
      }
      cur = sk;
      cs = mk;
    }
    if (sc.L_line > 0)
      for (int r = 0; r < R; ++r)
        line_run(true, sc.m_line, sc.L_line, cur.data() + (size_t)r * cs, cur.data() + (size_t)r * cs,
                 out.data() + (size_t)r * n);
    return;
  }
  // reverse: s of the coarsest part, then passes coarse -> fine
  std::vector<float> s;
  long long sst = n;  // row stride of s
  bool s_user = true;
  if (sc.L_line > 0) {
    s.resize((size_t)R * sc.m_line);
    for (int r = 0; r < R; ++r)
      line_run(false, sc.m_line, sc.L_line, in.data() + (size_t)r * n, in.data() + (size_t)r * n,
               s.data() + (size_t)r * sc.m_line);
    s_user = false;
    sst = sc.m_line;
  }
  for (size_t j = sc.passes.size(); j-- > 0;) {
    const PassPlan& pp = sc.passes[j];
    std::vector<float> band(j == 0 ? 0 : (size_t)R * pp.m0);
    float* dst = j == 0 ? out.data() : band.data();
    const long long ds = j == 0 ? n : pp.m0;
    const float* sp = s_user ? in.data() : s.data();
    for (const ItemSpec& it : pass_items(prog, false, pp, true, kF32, bc)) {
      if (row_ok(it)) {
        std::fprintf(stderr, "  rows [0, %lld) k=%d\n", it.m0, it.k);
        for (int r0 = 0; r0 < R; r0 += 4)
          rows_call(it, in.data() + r0 * n, dst + r0 * ds, const_cast<float*>(sp) + r0 * sst, n, ds, sst,
                    std::min(4, R - r0));
        continue;
      }
      for (int r = 0; r < R; ++r) {
        const float* ip = in.data() + (size_t)r * n;
        const float* spr = sp + (size_t)r * sst;
        float* dp = dst + (size_t)r * ds;
        if (it.kind == kSpecReg) {
          reg_dispatch<float>(tf, w.id, prog, it, ip, dp, const_cast<float*>(spr));
          continue;
        }
        const TileArgs ta{pp.m0, it.b - it.a, it.H, 0, pp.k, 0};
        std::vector<float> lds(tile_lds_bytes(ta, false, 4) / 4);
        tile_rev_body<float, float, float, float>(&prog, bc, pp.m0, pp.k, it.a, it.b, it.H, ip, 1, spr, 1, dp, 1,
                                                  lds.data(), lds.data() + tile_rev_window(ta));
      }
    }
    if (j > 0) {
      s = band;
      s_user = false;
      sst = pp.m0;
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
    } else if (const char* rv = std::getenv("LFX_SIM_ROWS")) {
      // R equal lines back to back on stdin (float32, table wavelets)
      const int R = std::atoi(rv);
      std::vector<float> in((size_t)R * n), out((size_t)R * n, 0.0f);
      if (std::fread(in.data(), 4, in.size(), stdin) != in.size()) return 3;
      run_rows(w, tf, bc, n, L, limit, tile, k, R, in, out);
      std::fwrite(out.data(), 4, out.size(), stdout);
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
