/*
 * mcfft.h -- C ABI of the MI355X batched complex-to-complex FFT library.
 *
 * Plain pointers and sizes only: device buffers are HIP device pointers passed
 * as void*, the stream is a hipStream_t passed as void* (NULL = default stream).
 *
 * Data layout: interleaved complex, transforms stored back to back, row-major
 * dimensions with the last dimension contiguous ([b][x][y] for 2D, as in the
 * reference's source/mcfft_2d_utils.cpp:49-50).  Input and output are in natural
 * order; in-place (in == out) and out-of-place execution are both supported.
 *
 * Element formats:
 *   MCFFT_FMT_FP16        4 B: binary16 re, binary16 im
 *   MCFFT_FMT_BF16        4 B: bfloat16 re, bfloat16 im
 *   MCFFT_FMT_FP16_SPLIT  8 B: binary16 (re_hi, im_hi, re_lo, im_lo) with
 *                         value = hi + lo * 2^-11 (about 22 significant bits)
 * Supported (input, output) pairs: FP16->FP16, BF16->BF16, FP16->SPLIT,
 * SPLIT->SPLIT.  FP16/BF16 compute on MFMA with FP32 accumulation; SPLIT forms
 * the compensated products hi*hi + 2^-11 (hi*lo + lo*hi) on FP16 MFMA.
 *
 * Scaling (block floating point): the stored result of transform t is
 * true_result * 2^-E[t].  AUTO (the default of the mcfft_plan_1d/2d/3d
 * shorthands) follows the data: the first pass scales each tile of its input
 * so that its largest |component| lies just below 2^13, every radix-R stage
 * divides by R, and E[t] is found on the GPU per transform (read it with
 * mcfft_get_transform_exponents or mcfft_get_exponent_buffer).  No finite input
 * overflows, and small inputs are scaled up rather than lost to underflow.  The
 * static modes use one exponent for every transform (mcfft_get_output_exponent):
 * NONE keeps E = 0 (unnormalised, FFTW convention, may overflow FP16),
 * BALANCED keeps the RMS of random data roughly constant (E ~ log2(N)/2), FULL
 * divides by N (E = log2 N; never overflows, but small inputs underflow).  A
 * statically scaled exec that stores an Inf/NaN, or an AUTO exec whose input
 * holds one, sets MCFFT_STATUS_NONFINITE (mcfft_get_status_flags); not watched:
 * the stores of statically scaled 1D plans of exactly 65536 points (their
 * single-pass kernel has no registers to spare).
 *
 * Sizes: every length a power of two; 1D 16 <= N <= 2^27 (one pass up to 2^14,
 * 2^16 for FP16/BF16 output; two passes up to 2^23; three passes from 2^24,
 * one transform at a time, with a workspace of one transform); 2D/3D outer sides
 * 16..4096 and innermost side 16..2^23 (one single-pass row kernel up to 2^16
 * for FP16/BF16 output and 2^14 for split output; longer rows run a two-pass
 * four-step over every row, with a workspace of one batch chunk).  The number of
 * columns an outer axis has (the product of the sides inside it) must be a
 * multiple of the width of a compiled column tile for that axis length (4..64;
 * 16 or more for every axis but the 16-point one, which needs 32: 16 x 16 is
 * MCFFT_UNSUPPORTED).
 *
 * Reference interface replaced: mcfftCreate / mcfftExec of
 * include/mcfft_1d_device.h:37-38, include/mcfft_2d_device.h:37-38 and
 * include/mcfft_3d_device.h:37-38 (void functions, digit-reversed in-place
 * input, no status, null stream).  The reference's exact signatures are kept by
 * the compatibility headers mcfft_{1,2,3}d_device.h / mcfft_{1,2,3}d_utils.h.
 */
#ifndef MCFFT_H
#define MCFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mcfftStatus_t {
  MCFFT_SUCCESS = 0,
  MCFFT_INVALID_VALUE = 1,   /* bad size, batch, rank, pointer or enum */
  MCFFT_UNSUPPORTED = 2,     /* valid request this build does not implement */
  MCFFT_ALLOC_FAILED = 3,
  MCFFT_HIP_ERROR = 4,       /* a HIP runtime call failed */
  MCFFT_MISALIGNED = 5,      /* buffer not aligned to its element size */
  MCFFT_INVALID_PLAN = 6,
  MCFFT_CACHE_ERROR = 7,     /* tuning cache unreadable or of another version */
  MCFFT_NONFINITE = 8        /* a result holds Inf/NaN (FP16 overflow under static scaling; the
                                reference-compatible layer's mcfftGetLastError) */
} mcfftStatus;

typedef enum mcfftFormat_t {
  MCFFT_FMT_FP16 = 0,
  MCFFT_FMT_BF16 = 1,
  MCFFT_FMT_FP16_SPLIT = 2
} mcfftFormat;

typedef enum mcfftDirection_t { MCFFT_FORWARD = -1, MCFFT_INVERSE = 1 } mcfftDirection;

typedef enum mcfftScaling_t {
  MCFFT_SCALE_NONE = 0,
  MCFFT_SCALE_BALANCED = 1,
  MCFFT_SCALE_FULL = 2,
  MCFFT_SCALE_AUTO = 3   /* per-transform exponents chosen from the data (not for ragged plans) */
} mcfftScaling;

/* bits of mcfft_get_status_flags */
#define MCFFT_STATUS_WAIT_BOUND 1u  /* ragged single launch: a dependency wait hit its bound */
#define MCFFT_STATUS_NONFINITE 2u   /* an Inf/NaN was stored (static scaling) or read (AUTO) */

typedef struct mcfftPlan_st* mcfftPlan;

/* Create a plan for `batch` transforms of shape dims[0..rank-1] (rank 1..3).
 * Allocates the plan's tables (DFT matrices and, for split precision, the
 * twiddles; generated on the GPU in FP64) and, for 1D lengths above the
 * single-pass limit (2^16 for FP16/BF16 output, 2^14 for split), a workspace; both live on the
 * current device.  FP16/BF16 passes evaluate their twiddles in the kernel.
 * Replaces mcfftCreate (source/mcfft_1d_device.cpp:680-790). */
mcfftStatus mcfft_plan_many(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                            mcfftFormat out_fmt, mcfftScaling scaling);
/* Shorthands with MCFFT_SCALE_AUTO. */
mcfftStatus mcfft_plan_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt, mcfftFormat out_fmt);
mcfftStatus mcfft_plan_2d(mcfftPlan* plan, long long nx, long long ny, long long batch, mcfftFormat in_fmt,
                          mcfftFormat out_fmt);
mcfftStatus mcfft_plan_3d(mcfftPlan* plan, long long nx, long long ny, long long nz, long long batch,
                          mcfftFormat in_fmt, mcfftFormat out_fmt);

/* Execute on `stream` (hipStream_t or NULL); asynchronous, no host sync.
 * direction: MCFFT_FORWARD (exp(-2 pi i kn/N)) or MCFFT_INVERSE (exp(+...),
 * unnormalised apart from the plan's scaling).  Capturable into a hipGraph.
 * A plan's workspace, real-transform buffer, layout staging buffers and ragged
 * queue counters belong to the plan: execs of one plan on different streams
 * must not overlap in time (use one plan per stream).
 * Replaces mcfftExec (source/mcfft_1d_device.cpp:596-678). */
mcfftStatus mcfft_exec(mcfftPlan plan, const void* in, void* out, int direction, void* stream);

/* mcfft_exec whose input transforms carry their own exponents: in_exp is a
 * device array of one int per transform (true input = stored * 2^in_exp[t]),
 * or NULL for all 0; the output exponents then include them.  Chaining plans
 * this way keeps every result in true units, e.g. an inverse over a forward's
 * output with in_exp = the forward plan's mcfft_get_exponent_buffer (which may
 * be the same plan's own buffer).  No reference counterpart. */
mcfftStatus mcfft_exec_exp(mcfftPlan plan, const void* in, const int* in_exp, void* out, int direction, void* stream);

/* Real-input forward (R2C) and real-output inverse (C2R) transforms of rank 1
 * to 3; no reference counterpart (the reference has only the forward C2C).
 * The last dimension n is a power of two >= 32 (1D: n <= 2^28); the outer
 * dimensions are those a complex plan of the same rank accepts.
 * R2C: `batch` real arrays of shape dims -> outer x (n/2 + 1) complex outputs
 *   each (X[m][k], k = 0..n/2), exec with MCFFT_FORWARD.
 * C2R: outer x (n/2 + 1) complex inputs per transform (the non-redundant half of
 *   a Hermitian spectrum) -> dims reals (unnormalised inverse), exec with
 *   MCFFT_INVERSE.
 * Real data is stored as pairs (x[2j], x[2j+1]) along the last dimension, in the
 * complex format of the same name (FP16/BF16: 2 bytes per real; split: the split
 * layout of the pair).  Computed as the complex FFT over (outer dims, n/2) plus
 * an FP32 pass that pairs bin (m, k) with (-m, n/2 - k) (FP64-generated twiddles
 * W_n^k); the plan owns an outer x n/2-element-per-transform buffer.  The C2R
 * plan reports one more power of two in its output exponent than the complex
 * plan.  mcfft_exec_passes / mcfft_plan_tune do not apply. */
mcfftStatus mcfft_plan_r2c(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                           mcfftFormat out_fmt, mcfftScaling scaling);
mcfftStatus mcfft_plan_c2r(mcfftPlan* plan, int rank, const long long* dims, long long batch, mcfftFormat in_fmt,
                           mcfftFormat out_fmt, mcfftScaling scaling);
/* the rank-1 forms of mcfft_plan_r2c / mcfft_plan_c2r */
mcfftStatus mcfft_plan_r2c_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt,
                              mcfftFormat out_fmt, mcfftScaling scaling);
mcfftStatus mcfft_plan_c2r_1d(mcfftPlan* plan, long long n, long long batch, mcfftFormat in_fmt,
                              mcfftFormat out_fmt, mcfftScaling scaling);

/* Advanced data layouts; no reference counterpart (the reference takes
 * contiguous interleaved data only, source/mcfft_1d_utils.cpp:37-66).
 * Element (b, i_0, .., i_{rank-1}) of transform b of a side is at
 *     b * dist + sum_d i_d * strides[d]
 * counted in elements of that side's form:
 *   planar = 0: interleaved complex elements of the side's format (4 B; split 8 B);
 *   planar = 1: two arrays, real parts and imaginary parts, of scalars
 *               (FP16/BF16: 2 B; split: binary16 hi then lo, 4 B).
 * Strides and dist are >= 1; output positions must not coincide.  A NULL
 * layout is the contiguous interleaved layout of mcfft_plan_many. */
typedef struct mcfftLayout_t {
  long long strides[3]; /* per dimension, dims order (only the first `rank` are read) */
  long long dist;       /* between consecutive transforms */
  int planar;
} mcfftLayout;

/* A plan whose exec reads and writes the given layouts.  When every side is
 * interleaved and dense inside each transform (only `dist` differs from the
 * contiguous layout), the passes read and write the caller's buffers directly,
 * locating transform t at t * dist -- for 2D/3D plans and 1D plans of more
 * than one pass or of one-row tiles (lengths 16384, 32768, 65536 and from 2^17);
 * such a plan is not tunable (MCFFT_UNSUPPORTED).  Otherwise a side that is not
 * the contiguous interleaved layout is staged through a plan-owned contiguous
 * buffer by a gather (input) or scatter (output) kernel: one extra read and
 * write of that side's data; a contiguous side costs nothing extra.  Tuning
 * (mcfft_plan_tune) and mcfft_exec_passes of a staged plan act on the
 * contiguous transform, with contiguous buffers; the tuning cache entry of the
 * same shape applies. */
mcfftStatus mcfft_plan_many_layout(mcfftPlan* plan, int rank, const long long* dims, long long batch,
                                   const mcfftLayout* in_layout, const mcfftLayout* out_layout,
                                   mcfftFormat in_fmt, mcfftFormat out_fmt, mcfftScaling scaling);
/* Execute a layout plan.  *_im are the imaginary-part arrays of planar sides
 * and must be NULL for interleaved sides.  mcfft_exec(plan, in, out, ...) on a
 * layout plan equals mcfft_exec_layout(plan, in, NULL, out, NULL, ...).
 * Input and output may overlap in any way (the staging orders the accesses),
 * except that an interleaved-contiguous input may alias a contiguous output
 * only exactly (in place), and the sides of a plan that stages neither side
 * (above) may alias only exactly, with equal layouts and formats; a partial
 * overlap of those returns MCFFT_INVALID_VALUE. */
mcfftStatus mcfft_exec_layout(mcfftPlan plan, const void* in_re, const void* in_im, void* out_re, void* out_im,
                              int direction, void* stream);

/* Ragged batch: n_items transforms of any shapes mcfft_plan_many accepts, each
 * at its own element offset in the exec input and output buffers; the whole
 * batch is one graph-capturable exec on one stream.
 * - Batched runs: a run of consecutive items of one shape lying back to back in
 *   both buffers with at least 2^23 elements in all (enough to fill the GPU;
 *   environment MCFFT_RAGGED_BULK=<elements> changes the bound, 0 turns it off)
 *   runs as one batched plan.  So do the other items of one shape wherever they
 *   lie (a gathered batch: the plan locates its transforms through per-item
 *   offset tables) when they reach the bound or 1024 items together -- 2D/3D
 *   items and 1D items of 2048 points or more -- and, whatever their size,
 *   items of shapes the queue below cannot run (1D lengths above 2^18, outer
 *   axes narrower than the queue's column tiles).
 * - The rest (1D lengths 16..2^18, sides 16..4096) run by one persistent kernel
 *   launch per pass level (at most three: every item's first passes, then
 *   second, then third) over a device-side work queue, plus a tiny queue-reset
 *   kernel.  Back-to-back 1D items of one length up to 4096 are processed as one
 *   multi-row unit.
 * Items may have different output exponents: see mcfft_get_item_exponents.  No
 * reference counterpart. */
typedef struct mcfftRaggedItem_t {
  int rank;
  long long dims[3];
  long long in_offset;   /* element offset of the transform in the exec input */
  long long out_offset;  /* element offset in the exec output */
} mcfftRaggedItem;

mcfftStatus mcfft_plan_ragged(mcfftPlan* plan, int n_items, const mcfftRaggedItem* items, mcfftFormat in_fmt,
                              mcfftFormat out_fmt, mcfftScaling scaling);
/* Output exponent of every item (n_items entries; 1 entry for a statically
 * scaled regular plan). */
mcfftStatus mcfft_get_item_exponents(mcfftPlan plan, int* exponents);
/* Status bits (MCFFT_STATUS_*): for a regular plan, MCFFT_STATUS_NONFINITE if
 * any exec since the last call stored (static scaling) or read (AUTO) an
 * Inf/NaN -- the bits are cleared by reading them; for a ragged plan,
 * MCFFT_STATUS_WAIT_BOUND if a queue wait of the last exec hit its bound (its
 * result is then wrong), plus its batched runs' bits.  Synchronises with the
 * device. */
mcfftStatus mcfft_get_status_flags(mcfftPlan plan, unsigned* flags);
/* Output exponents of the last exec, one per transform (batch entries; a
 * ragged plan: one per item): true result of transform t = stored * 2^e[t].
 * Synchronises with the device.  No reference counterpart. */
mcfftStatus mcfft_get_transform_exponents(mcfftPlan plan, int* exponents);
/* Device array (batch ints) the plan's execs write their per-transform output
 * exponents to, in stream order: usable by later work on the same stream, or
 * as another exec's in_exp.  Regular plans only. */
mcfftStatus mcfft_get_exponent_buffer(mcfftPlan plan, const int** exponents);

/* Autotuner: for each structure the plan's shape can take -- the radix that
 * splits a 2D plan's outer axis (none, 2, 4, 8), the N1 x N2 split of a two-pass
 * 1D plan (the default and N2 halved or doubled) -- times every compiled kernel
 * variant (tile size, waves per workgroup) of each pass on the given device
 * buffers, `reps` launches each, keeps the fastest per pass, then times whole
 * execs and keeps the fastest structure.  The choice (structure and kernels) is
 * recorded in the tuning cache, keyed by GPU architecture and (rank, dims,
 * batch, formats, scaling); plans created later read it, so a shape is tuned
 * once per machine.  The cache file names the library build that wrote it
 * (mcfft_build_id): one written by another build is ignored.  Blocking; the
 * contents of `out` (and of `in` for an in-place plan) are overwritten, and the
 * plan's device buffers (workspace, exponent buffer) may be replaced, so query
 * mcfft_get_exponent_buffer after tuning.  Cache file: $MCFFT_TUNE_CACHE, default
 * $HOME/.cache/mcfft/tune-<arch>.txt; MCFFT_TUNE_CACHE=off disables it.  Not for
 * ragged or real plans; a layout plan keeps its structure.  No reference
 * counterpart (its size switch, source/mcfft_1d_device.cpp:695-775, is fixed). */
mcfftStatus mcfft_plan_tune(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int reps);

/* Identifier of the kernel sources this library was built from (a hash), as
 * written into the tuning cache. */
mcfftStatus mcfft_build_id(const char** id);

/* Error norms on the GPU in FP64: rel-L2 = ||y - ref||_2 / ||ref||_2 and
 * max-abs = max |y - ref| over re and im, where y is `n` elements stored in
 * `fmt` (device pointer) times 2^exponent and `ref` is n interleaved FP64
 * complex values on the device.  Blocking (synchronises `stream`). */
mcfftStatus mcfft_error_norms(const void* y, mcfftFormat fmt, int exponent, const double* ref, long long n,
                              double* rel_l2, double* max_abs, void* stream);

/* mcfft_error_norms without any allocation: no counterpart in the reference.
 * For callers that cannot call hipMalloc (code under hipGraph capture, tight
 * serving loops, harnesses that forbid allocation).  The per-block partial sums
 * go to a fixed library-owned device array (3 x 1024 doubles = 24 KiB per GPU,
 * held until the process ends) instead of a buffer allocated per call; results
 * are identical to mcfft_error_norms.  The partials do not grow with n, so every
 * n up to 2^40 elements fits; larger n returns MCFFT_INVALID_VALUE, as does
 * mcfft_error_norms.  Calls that overlap in time on one GPU share that array
 * and race: serialise them (one stream, or host order).  Blocking. */
mcfftStatus mcfft_error_norms_noalloc(const void* y, mcfftFormat fmt, int exponent, const double* ref, long long n,
                                      double* rel_l2, double* max_abs, void* stream);

/* Profiling entry points (no reference counterpart): the number of kernel
 * passes of a plan, and an exec that launches only passes [first, first+count)
 * (for every batch chunk), so a caller can time each kernel on its own stream.
 * Running a subset of passes leaves `out` incomplete. */
mcfftStatus mcfft_get_pass_count(mcfftPlan plan, int* count);
mcfftStatus mcfft_exec_passes(mcfftPlan plan, const void* in, void* out, int direction, void* stream, int first,
                              int count);
/* mcfft_exec with every pass's kernel timed on the GPU: pass i is launched
 * with hipExtLaunchKernel, which records start_events[i] and stop_events[i]
 * (hipEvent_t, created by the caller) when that kernel starts and ends, with
 * no extra stream packets between the passes.  Arrays of
 * mcfft_get_pass_count() events; plans whose batch runs in one chunk
 * (MCFFT_UNSUPPORTED otherwise; not for ragged or real plans). */
mcfftStatus mcfft_exec_timed(mcfftPlan plan, const void* in, void* out, int direction, void* stream,
                             void* const* start_events, void* const* stop_events);

/* Statically scaled plans: E such that true result = stored * 2^E (same for
 * both directions).  AUTO plans: MCFFT_UNSUPPORTED (their exponents are per
 * transform: mcfft_get_transform_exponents). */
mcfftStatus mcfft_get_output_exponent(mcfftPlan plan, int* exponent);
/* Workspace the plan uses (bytes; 0 when none). */
mcfftStatus mcfft_get_workspace_size(mcfftPlan plan, size_t* bytes);
/* Human-readable pass list (for logs and tests). */
mcfftStatus mcfft_plan_describe(mcfftPlan plan, char* buf, size_t len);
mcfftStatus mcfft_destroy(mcfftPlan plan);
const char* mcfft_status_string(mcfftStatus status);
int mcfft_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MCFFT_H */
