// Fixed device memory of the no-allocation entry point (lfx_lwt_noalloc):
// a scratch array (band scratch of tiled passes, the input copy of in-place
// calls) and the step programs of every table wavelet and transform in both
// compute types. They are file-scope __device__ arrays: each device holds
// them from the time the library's code object is loaded until the process
// ends, and a call through that entry point never allocates.
//
// In host code the arrays' names are host-side stand-ins, not device
// addresses: set-up looks up each array's address on the current device
// with hipGetSymbolAddress (each device holds its own copies at its own
// addresses), and only those addresses are ever handed to kernels, copies or
// fills.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "lfx_error.hpp"
#include "lfx_noalloc.hpp"
#include "lfx_program.hpp"
#include "lfx_table.hpp"

namespace lfx {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static __device__ __align__(256) DProgram<float> noalloc_progs_f32[4 * kNoAllocWavelets];
static __device__ __align__(256) DProgram<double> noalloc_progs_f64[4 * kNoAllocWavelets];

static std::mutex g_noalloc_mu;
static std::map<int, NoAllocDevice> g_noalloc;

static void na_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(LFX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

const NoAllocDevice& noalloc_device() {
  int dev = 0;
  na_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc.find(dev);
  if (it != g_noalloc.end()) return it->second;
  const int nw = table_size();
  if (nw > kNoAllocWavelets) fail(LFX_EUNSUPPORTED, "wavelet table larger than the fixed program slots");
  NoAllocDevice d;
  void* p = nullptr;
  na_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  d.scratch = p;
  na_check(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_progs_f32)), "hipGetSymbolAddress(programs)");
  d.progs_f32 = p;
  na_check(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_progs_f64)), "hipGetSymbolAddress(programs)");
  d.progs_f64 = p;
  // fill the program slots from the host (plain copies into the fixed arrays)
  for (int i = 0; i < nw; ++i) {
    const WaveletDefL w = table_entry(i);
    for (int tf = 0; tf < 4; ++tf) {
      const DProgram<float> pf = resolve<float>(w, tf);
      const DProgram<double> pd = resolve<double>(w, tf);
      na_check(hipMemcpy((DProgram<float>*)const_cast<void*>(d.progs_f32) + 4 * i + tf, &pf, sizeof(pf),
                         hipMemcpyHostToDevice),
               "hipMemcpy(program)");
      na_check(hipMemcpy((DProgram<double>*)const_cast<void*>(d.progs_f64) + 4 * i + tf, &pd, sizeof(pd),
                         hipMemcpyHostToDevice),
               "hipMemcpy(program)");
    }
  }
  return g_noalloc.emplace(dev, d).first->second;
}

}  // namespace lfx
