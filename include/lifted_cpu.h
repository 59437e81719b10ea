/* C ABI of the CPU product path of lifted-hip (liblifted_cpu.so): the same
 * transforms as lfx_lwt_host (include/lifted_hip.h) computed on the host by
 * the GPU kernels' line bodies compiled for the CPU, the lines of each axis
 * split evenly over n_threads threads (0: all cores), as the reference's
 * thread pool does (src/lifted-common.hpp:143-480; lifted::lwt's n_threads
 * argument, src/lifted.hpp:13-39). float32 and float64 storage; table
 * wavelets. Results are bit-identical to the GPU path. Returns 0 or a
 * negative LFX_E* code; lfx_cpu_last_error() then holds a message. */
#ifndef LIFTED_CPU_H_
#define LIFTED_CPU_H_

#include "lifted_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

int lfx_cpu_lwt(const lfx_desc* desc, const void* in_host, void* out_host, int32_t n_threads);
const char* lfx_cpu_last_error(void);
/* max_level of src/lifted-common.hpp:1099-1108 (negative code for width 1). */
int64_t lfx_cpu_max_level(uint32_t wavelet, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* LIFTED_CPU_H_ */
