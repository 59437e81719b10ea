// The predefined wavelet table (generated from the reference's
// src/lifted-wavelets.hpp into wavelet_table.inc) and the registry of
// caller-defined step programs.
#include <mutex>
#include <string>
#include <vector>

#include "lfx_error.hpp"
#include "lfx_table.hpp"

namespace lfx {

static std::vector<WaveletDefL> build_table() {
  std::vector<WaveletDefL> v;
  WaveletDefL* cur = nullptr;
  auto add = [&](int kind, int target, int off, int n, int sign, const long double* t, int nt) {
    StepDefL& s = cur->steps[cur->nsteps++];
    s = StepDefL{};
    s.kind = kind;
    s.target = target;
    s.offset = off;
    s.ntaps = n;
    s.sign = sign;
    for (int k = 0; k < nt; ++k) s.taps[k] = t[k];
  };
  const int S = 0, D = 1;
#define LFX_WAVELET(nm, idv, w)  \
  v.emplace_back();              \
  cur = &v.back();               \
  *cur = WaveletDefL{};          \
  cur->id = idv;                 \
  cur->name = #nm;               \
  cur->width = w;
#define LFX_END(nm)
#define LFX_GENERIC(b, off, ...)                                         \
  {                                                                      \
    const long double t_[] = {__VA_ARGS__};                              \
    const int n_ = int(sizeof(t_) / sizeof(t_[0]));                      \
    add(kGeneric, b, off, n_, 1, t_, n_);                                \
  }
#define LFX_UNIT(b, off, sg) add(kUnit, b, off, 1, sg, nullptr, 0);
#define LFX_REPEAT(b, off, n, val)   \
  {                                  \
    const long double t_[] = {val};  \
    add(kRepeat, b, off, n, 1, t_, 1); \
  }
#define LFX_SCALE(a, b)                 \
  {                                     \
    const long double t_[] = {a, b};    \
    add(kScale, 0, 0, 0, 1, t_, 2);     \
  }
#include "wavelet_table.inc"
#undef LFX_WAVELET
#undef LFX_END
#undef LFX_GENERIC
#undef LFX_UNIT
#undef LFX_REPEAT
#undef LFX_SCALE
  return v;
}

static std::mutex g_mu;
static std::vector<WaveletDefL>& table() {
  static std::vector<WaveletDefL> t = build_table();
  return t;
}
static std::vector<WaveletDefL> g_custom;  // ids 0x80000000 + index

WaveletDefL find_wavelet(uint32_t id) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (id >= kCustomBase) {
    if (id - kCustomBase < g_custom.size()) return g_custom[id - kCustomBase];
  } else {
    for (auto& w : table())
      if (w.id == id) return w;
  }
  fail(LFX_EINVAL, "unknown wavelet id " + std::to_string(id));
}

int table_index(uint32_t id) {
  std::lock_guard<std::mutex> lk(g_mu);
  const auto& t = table();
  for (size_t i = 0; i < t.size(); ++i)
    if (t[i].id == id) return (int)i;
  return -1;
}

int table_size() {
  std::lock_guard<std::mutex> lk(g_mu);
  return (int)table().size();
}

WaveletDefL table_entry(int i) {
  std::lock_guard<std::mutex> lk(g_mu);
  const auto& t = table();
  if (i < 0 || i >= (int)t.size()) fail(LFX_EINVAL, "wavelet table index out of range");
  return t[i];
}

uint32_t register_custom(WaveletDefL w) {
  std::lock_guard<std::mutex> lk(g_mu);
  w.id = kCustomBase + (uint32_t)g_custom.size();
  g_custom.push_back(w);
  return w.id;
}

}  // namespace lfx
