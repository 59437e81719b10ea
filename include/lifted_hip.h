/* C ABI of lifted-hip: the lifting-scheme DWT of jcapriot/lifted on gfx950.
 *
 * The reference's boundary is four C++ overloads of lifted::lwt
 * (src/lifted.hpp:13-39; defined in src/lifted-{float,double,sep-float,
 * sep-double}.cpp:20-31) taking std::vector shapes/strides. A C ABI cannot pass
 * std::vector, so each overload becomes one function taking (pointer, count)
 * pairs with the same meaning; include/lifted.hpp rebuilds the exact C++
 * overloads on top of these for C++ callers.
 *
 * Enum values are the reference's ABI (src/lifted-common.hpp:33-98):
 *   wavelets 0..100907, BoundaryCondition 0..4, Transform 0..3.
 * Strides are in elements and may be negative. in == out is allowed when both
 * use the same strides. All functions return 0 on success and a negative
 * LFX_E* code on error; lfx_last_error() then holds a message (per thread).
 */
#ifndef LIFTED_HIP_H_
#define LIFTED_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LFX_OK 0
#define LFX_EINVAL (-1)       /* invalid argument (shape, axes, enum, level) */
#define LFX_EUNSUPPORTED (-2) /* valid but not supported by this build */
#define LFX_EDEVICE (-3)      /* HIP runtime error */
#define LFX_ENOMEM (-4)       /* device allocation failed */

/* storage types of lfx_desc.storage; bf16/fp16 are computed in fp32 */
#define LFX_F32 0
#define LFX_F64 1
#define LFX_BF16 2
#define LFX_F16 3

/* lfx_desc.flags and lfx_batch_plan_create flags */
#define LFX_FLAG_PLAIN 1u    /* one launch per level (no multi-level fusion); float32 batched
                                plans still run each level as register tiles */
#define LFX_FLAG_LDS_ONLY 2u /* batched plans: LDS kernels only (no register tiles) */

typedef struct {
  uint32_t wavelet;         /* lifted::Wavelet value, or an id from lfx_program_create */
  uint32_t transform;       /* lifted::Transform value */
  uint32_t bc;              /* lifted::BoundaryCondition value */
  int32_t ndim;             /* 1..8 */
  const int64_t* shape;     /* ndim sizes */
  const int64_t* stride_in; /* ndim element strides of the input */
  const int64_t* stride_out;/* ndim element strides of the output */
  int32_t naxes;            /* number of transformed axes (1..ndim) */
  const int64_t* axes;      /* axes in application order; negative counts from the end */
  int32_t separable;        /* 1: levels[naxes] (separable), 0: levels[0] (level by level) */
  const int64_t* levels;    /* 0 = max_level */
  int32_t storage;          /* LFX_F32 / LFX_F64 / LFX_BF16 / LFX_F16 */
  uint32_t flags;           /* LFX_FLAG_* */
} lfx_desc;

/* Transform device-resident arrays, stream-ordered and asynchronous.
 * stream is a hipStream_t (NULL = default stream). Replaces the work done by
 * lifting_transform / seperable_lifting_transform (src/lifted-drivers-inl.hpp:630-656). */
int lfx_lwt(const lfx_desc* desc, const void* in_dev, void* out_dev, void* stream);

/* lfx_lwt without any device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same descriptor, arrays and
 * results as lfx_lwt, restricted to table wavelets (and, for 16-bit storage,
 * to one transformed axis); it runs the LDS line and
 * tile kernels with their band scratch in a fixed device array of at most
 * 64 MiB (the library holds one per GPU until the process ends; the step
 * programs of every table wavelet sit in a fixed table beside it). A call
 * whose scratch would exceed that -- long lines (tiled passes), or an
 * in-place call of them -- fails with LFX_EINVAL. Nothing in the scratch
 * needs to start at zero. Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (two streams, or two host threads) race and
 * must be serialized by the caller. */
int lfx_lwt_noalloc(const lfx_desc* desc, const void* in_dev, void* out_dev, void* stream);

/* Same on host arrays: copies to the current device, transforms, copies back,
 * blocking -- the calling convention of lifted::lwt. */
int lfx_lwt_host(const lfx_desc* desc, const void* in_host, void* out_host);

/* The four lifted::lwt overloads (src/lifted.hpp:13-39), host pointers.
 * level: scalar level (level-by-level path, src/lifted.hpp:13-25);
 * levels[naxes]: separable path (src/lifted.hpp:27-39). n_threads is accepted
 * for signature parity and ignored (the work runs on the GPU). */
int lfx_lwt_f32(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                size_t level, const float* data_in, float* data_out, size_t n_threads);
int lfx_lwt_f64(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                size_t level, const double* data_in, double* data_out, size_t n_threads);
int lfx_lwt_sep_f32(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                    const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                    const size_t* levels, const float* data_in, float* data_out, size_t n_threads);
int lfx_lwt_sep_f64(uint32_t wavelet, uint32_t transform, uint32_t bc, size_t ndim, const size_t* shape,
                    const ptrdiff_t* stride_in, const ptrdiff_t* stride_out, size_t naxes, const size_t* axes,
                    const size_t* levels, const double* data_in, double* data_out, size_t n_threads);

/* max_level(width(wavelet), n) of src/lifted-common.hpp:1099-1108.
 * Returns a negative code for an unknown wavelet or a width-1 wavelet (the
 * reference's loop never terminates there). */
int64_t lfx_max_level(uint32_t wavelet, int64_t n);

/* Filter width of a wavelet (the `width` member, src/lifted-wavelets.hpp). */
int64_t lfx_wavelet_width(uint32_t wavelet);

/* Caller-defined lifting scheme (the reference only has compile-time ones,
 * src/lifted-common.hpp:966-1032). Coefficients are doubles, cast to the
 * transform type like the reference casts its literals. */
typedef struct {
  int32_t kind;    /* 0 generic (update_s/update_d), 1 unit, 2 repeat, 3 scale */
  int32_t target;  /* 0 update evens (s), 1 update odds (d) */
  int32_t offset;
  int32_t ntaps;   /* generic: taps; repeat: summed taps; unit: 1 */
  int32_t sign;    /* unit: +1 add / -1 subtract */
  int32_t pad;
  double taps[16]; /* generic: taps; repeat: taps[0]; scale: taps[0]=s_mul, taps[1]=d_div */
} lfx_step;

/* Registers a step program; *wavelet_id receives an id usable as lfx_desc.wavelet. */
int lfx_program_create(const lfx_step* steps, int32_t nsteps, uint32_t width, uint32_t* wavelet_id);

/* Coefficients the GPU uses for (wavelet, transform, storage), in execution
 * order, as doubles: per step [kind, y_is_d, adjoint, sub, off, N, nf, nb,
 * fold_off, c0..c15] (25 values). Returns the step count. */
int lfx_program_dump(uint32_t wavelet, uint32_t transform, int32_t storage, double* out, int32_t cap_steps);

/* ---- batched transforms of many 1-D signals ----
 * Signals are packed back to back (any order) in one input and one output
 * array of the plan's storage type; each has its own length, wavelet and
 * level count. A plan fixes the transform type, boundary condition and the
 * signal list, owns its device scratch, and runs as a few stage launches that
 * can be captured in a hipGraph. This is the reference's per-line driver
 * (src/lifted-drivers-inl.hpp:75-274) applied to many lines at once; the
 * reference has no batched entry point. */
typedef struct {
  int64_t offset;   /* element offset of the signal in the packed arrays */
  int64_t length;   /* samples */
  uint32_t wavelet; /* lifted::Wavelet value or custom program id */
  int32_t levels;   /* 0 = max_level */
} lfx_signal;

typedef struct lfx_batch_plan_s* lfx_batch_plan;

int lfx_batch_plan_create(const lfx_signal* signals, int32_t nsignals, uint32_t transform, uint32_t bc,
                          int32_t storage, uint32_t flags, lfx_batch_plan* plan);
/* All stages, stream-ordered; in and out must not alias. The launches of a
 * stage are spread over side streams the plan owns, forked from and joined
 * back to `stream` with events (graph capture of `stream` captures them too);
 * one plan must not be executed from two host threads at once. */
int lfx_batch_execute(lfx_batch_plan plan, const void* in_dev, void* out_dev, void* stream);
int lfx_batch_num_stages(lfx_batch_plan plan);
int lfx_batch_execute_stage(lfx_batch_plan plan, int32_t stage, const void* in_dev, void* out_dev, void* stream);
/* Work items, algorithmic bytes (owned samples read + written) and LDS bytes of a stage. */
int lfx_batch_stage_info(lfx_batch_plan plan, int32_t stage, int64_t* nitems, double* alg_bytes, int64_t* lds_bytes);
void lfx_batch_plan_destroy(lfx_batch_plan plan);

/* One kernel launch of a batched stage. */
typedef struct {
  int32_t kind;      /* 0: LDS items (tiles, whole lines), 1: register tiles, 2: row items (four equal short lines per wavefront) */
  int32_t threads;   /* workgroup size */
  int64_t nitems;    /* work items (workgroups for kind 0, wavefronts for kind 1) */
  double alg_bytes;  /* algorithmic bytes: owned samples read + written */
  int64_t lds_bytes; /* dynamic LDS per workgroup (kind 0) */
} lfx_launch_info;

/* Launches of a stage (LDS launches first, then the register launch, then the row launch): fills
 * info[0..min(n, max_launches)) and returns n, or an error code. */
int lfx_batch_launch_info(lfx_batch_plan plan, int32_t stage, int32_t max_launches, lfx_launch_info* info);

/* lfx_batch_execute_stage that also records caller-created hipEvent_t pairs
 * around each launch, on the stream the launch goes to (launch order of
 * lfx_batch_launch_info). Nothing is synchronized: read the events later. */
int lfx_batch_execute_stage_timed(lfx_batch_plan plan, int32_t stage, const void* in_dev, void* out_dev, void* stream,
                                  int32_t nevents, void* const* ev_start, void* const* ev_end);

/* ---- chains: a sequence of batched transforms pipelined over signal chunks ----
 * A chain runs transforms[0..n) in sequence on one signal list (in -> outs[0]
 * -> outs[1] ...), each as a batched plan, with the signals split into
 * nchunks groups of about equal length. Each plan's "big" stage (the finest
 * levels of every signal: stage 0 of a forward-stepping plan, the last stage
 * of a reverse one) runs one chunk after another, and each chunk's short,
 * latency-bound stages run beside the next chunk's big stage on a stream the
 * chain owns, forked from and joined back to `stream` (graph capture of
 * `stream` captures them too). Results equal running the plans in turn. One
 * chain must not be executed from two host threads at once. No counterpart
 * in the reference. */
typedef struct lfx_batch_chain_s* lfx_batch_chain;
int lfx_batch_chain_create(const lfx_signal* signals, int32_t nsignals, const uint32_t* transforms,
                           int32_t ntransforms, uint32_t bc, int32_t storage, uint32_t flags, int32_t nchunks,
                           lfx_batch_chain* chain);
/* outs[i]: output of transform i (and input of transform i + 1); a step's
 * input and output must not alias. */
int lfx_batch_chain_execute(lfx_batch_chain chain, const void* in_dev, void* const* out_devs, void* stream);
int lfx_batch_chain_num_chunks(lfx_batch_chain chain);
/* One kernel launch of a chain execution. */
typedef struct {
  int32_t chunk;           /* signal chunk */
  int32_t transform_index; /* position in the chain's transform list */
  int32_t stage;           /* stage of that chunk's plan */
  int32_t big;             /* 1: the plan's big stage */
  lfx_launch_info launch;
} lfx_chain_launch_info;
/* Launches of one execution in launch order (per transform, per chunk, per
 * stage, the stage's launches); fills info[0..min(n, max)) and returns n. */
int lfx_batch_chain_launch_info(lfx_batch_chain chain, int32_t max_launches, lfx_chain_launch_info* info);
/* lfx_batch_chain_execute with caller-created hipEvent_t pairs recorded
 * around each launch (lfx_batch_chain_launch_info order) on its stream. */
int lfx_batch_chain_execute_timed(lfx_batch_chain chain, const void* in_dev, void* const* out_devs, void* stream,
                                  int32_t nevents, void* const* ev_start, void* const* ev_end);
void lfx_batch_chain_destroy(lfx_batch_chain chain);

/* Tile length (samples at a pass's finest level, 0 = default) and levels per
 * tiled pass (0 = default) used by plans and transforms created afterwards on
 * this thread. */
int lfx_set_tuning(int64_t tile, int32_t levels_per_pass);

/* All planning knobs of the calling thread, for plans and transforms created
 * afterwards; lifted_amd.tune searches them by timing candidate plans. */
typedef struct {
  int64_t tile;            /* LDS tile length (samples at a pass's finest level), 0 = default */
  int32_t levels_per_pass; /* levels per tiled pass (1..12), 0 = default (4 for register tiles) */
  int32_t side_streams;    /* side streams of a batched plan (0..8), -1 = default (3) */
  int64_t line_limit;      /* register-tiled plans: longest band handed to the whole-line kernels, 0 = default */
  int32_t across;          /* ND calls: strided float32 axes of at most 512 samples (Zero boundary) read in
                              place by the across-line kernel instead of transposed; 0 = off, 1 = on,
                              -1 = default (on) */
  int32_t strip;           /* register window (strip) of float32 plans in samples: 1024 or 2048,
                              0 = default (2048) */
} lfx_tuning;
int lfx_set_tuning_ex(const lfx_tuning* t);
int lfx_get_tuning(lfx_tuning* t);

/* Streaming-copy reference for measurements (no counterpart in the
 * reference): copies `bytes` (rounded down to the pattern's unit) from src to
 * dst, stream-ordered. pattern 0: one float4 per thread; pattern 1: one
 * 2048-float window per wavefront, read and written as 32 coalesced rows of
 * 64 floats, each XCD taking one contiguous range of windows (the register
 * tiles' access pattern without the arithmetic).
 * Buffers must be 16-byte aligned and must not overlap. */
int lfx_stream_copy(const void* src_dev, void* dst_dev, int64_t bytes, int32_t pattern, void* stream);

/* Message of the last error on this thread. */
const char* lfx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LIFTED_HIP_H_ */
