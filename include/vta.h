/*
 * vta.h -- C ABI of libvta_hip, the MI355X (gfx950) batch-indicator library.
 *
 * Drop-in boundary for VectorTA's GPU batch-indicator path.  Each entry point
 * replaces one method of the reference's Rust CUDA wrappers
 * (the src/cuda/ wrapper files) and keeps its argument meaning:
 *   - *_batch_dev            one series x many params, host input, device-resident
 *                            output, synchronised before return;
 *   - *_batch_device         device input, caller-owned device output, enqueue only
 *                            (no allocation, no synchronisation);
 *   - *_many_series_one_param_time_major_dev
 *                            many time-major series x one param, host input,
 *                            device-resident [rows=T][cols=N] output.
 * Kernels are HIP code objects (one .hsaco per indicator family) loaded with
 * hipModuleLoad from the directory holding libvta_hip.so (or $VTA_KERNEL_DIR).
 *
 * Conventions: every call returns vta_status; on failure vta_last_error()
 * returns a thread-local message.  Integers are size_t / int32_t, device
 * pointers are plain float* / int32_t*.  One-series outputs are row-major
 * [combo][t]; many-series data and outputs are time-major [t][series].
 * A vta_ctx owns one device and one non-blocking stream; like the reference
 * wrappers (alma_wrapper.rs:972-974) it is not thread-safe.
 */
#ifndef VTA_H
#define VTA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error enum, one-to-one with CudaAlmaError (alma_wrapper.rs:72-101) plus
 * CudaEmaError::ArithmeticOverflow (ema_wrapper.rs:21-45). */
typedef enum vta_status {
    VTA_OK = 0,
    VTA_ERR_HIP = 1,                     /* Cuda(CudaError) */
    VTA_ERR_OUT_OF_MEMORY = 2,           /* OutOfMemory{required, free, headroom} */
    VTA_ERR_MISSING_KERNEL_SYMBOL = 3,   /* MissingKernelSymbol{name} */
    VTA_ERR_INVALID_INPUT = 4,           /* InvalidInput(String) */
    VTA_ERR_INVALID_POLICY = 5,          /* InvalidPolicy */
    VTA_ERR_LAUNCH_CONFIG_TOO_LARGE = 6, /* LaunchConfigTooLarge{gx..bz} */
    VTA_ERR_DEVICE_MISMATCH = 7,         /* DeviceMismatch{buf, current} */
    VTA_ERR_NOT_IMPLEMENTED = 8,         /* NotImplemented */
    VTA_ERR_ARITHMETIC_OVERFLOW = 9,     /* ArithmeticOverflow{context} */
    VTA_ERR_NO_DEVICE = 10,              /* cuda_available() == false */
    VTA_ERR_TIMEOUT = 11                 /* vta_server_wait: not finished in time */
} vta_status;

const char *vta_last_error(void);
const char *vta_version(void);

/* Device probe: src/cuda/mod.rs:412-602 (cuda_available, cuda_device_count).
 * Honours VTA_FORCE_SKIP=1 like CUDA_FORCE_SKIP (mod.rs:419-423). */
int vta_device_available(void);
int vta_device_count(void);

/* ------------------------------------------------------------ context */
/* CudaAlma::new (alma_wrapper.rs:133-169): device, modules, NON_BLOCKING stream. */
typedef struct vta_ctx vta_ctx;
vta_status vta_ctx_create(int device_id, vta_ctx **out);
vta_status vta_ctx_destroy(vta_ctx *ctx);
vta_status vta_ctx_synchronize(vta_ctx *ctx);   /* CudaAlma::synchronize (:171) */
void *vta_ctx_stream(vta_ctx *ctx);             /* hipStream_t of the context */
int vta_ctx_device_id(const vta_ctx *ctx);      /* CudaAlma::device_id */
/* Loads every code object eagerly and checks that each expected kernel
 * symbol exists (the reference JITs its PTX in new()). */
vta_status vta_ctx_preload(vta_ctx *ctx);

/* ------------------------------------------------------------ tuning */
/* Launch-shape knobs held per context, the counterpart of the reference's
 * ALMA_TILE environment knob (alma_wrapper.rs:314) and kernel policies
 * (:18-70), extended to every kernel family with a shape choice.  0 selects
 * the built-in heuristic (which an environment variable of the same family
 * may still override); any other value must be in the knob's range. */
typedef enum vta_tune_knob {
    /* split MFMA ALMA sweep kernel: 1 workgroup of 4 N-blocks x 16 tiles,
     * 2 x 8 tiles, 3 x 16 tiles two at a time, 4 x 16 tiles with K > 128 and
     * K <= 128 in separate launches, 5 one wave per N-block x 8 tiles,
     * 6 x 16 tiles, 7 x 32 tiles (1-7: BF16 only); 8 diagonal form, 2-tile
     * sub-runs, LDS-transposed stores, 9 the same with direct stores, 10
     * 4-tile sub-runs, direct stores, 11 2-tile sub-runs whose transposed
     * stores are issued during the next sub-run's MFMAs (runs as 8 when the
     * series length is not a multiple of 4) (8-11: BF16 and F16) */
    VTA_TUNE_ALMA_MFMA_VARIANT = 0,
    VTA_TUNE_SCAN_WAVES = 1,  /* recurrence batch chunked scan: waves per pass, 64..65536 */
    VTA_TUNE_FIR_GROUP = 2,   /* SMA/WMA/SWMA batch: combos per workgroup, 1..256 */
    VTA_TUNE_MS_CHUNK = 3,    /* many-series recurrences: bars per chunk, >= 64 */
    /* many-series percentile_nearest_rank / medium_ad window form: 1 sorted
     * windows in LDS (any length), 2 sorted windows in registers (PNR lengths
     * up to 128, MAD periods up to 64; longer windows run in LDS); the
     * built-in choice is 2 where it applies */
    VTA_TUNE_ROLL_FORM = 4,
    VTA_TUNE_KNOB_COUNT = 5
} vta_tune_knob;
vta_status vta_ctx_set_tuning(vta_ctx *ctx, vta_tune_knob knob, int64_t value);
vta_status vta_ctx_get_tuning(const vta_ctx *ctx, vta_tune_knob knob, int64_t *value);

/* Autotuner (no reference counterpart).  For each candidate value of `knob`
 * it sets the knob and runs enqueue(ctx, user) once untimed; then, in three
 * interleaved rounds over the candidates (so clock drift during the search
 * does not favour the first ones), `reps` times back to back between two HIP
 * events on the context's stream.  The value with the least round mean wins
 * and the knob is left at it.  enqueue
 * must only enqueue work on ctx (an enqueue-only entry point, a sweep plan
 * launch, a graph launch) and return its status.  A candidate whose enqueue
 * fails with VTA_ERR_LAUNCH_CONFIG_TOO_LARGE (e.g. too much LDS for that
 * shape) is skipped; any other failure ends the search with that status and
 * the knob restored.  ms_per_candidate (may be NULL) receives each
 * candidate's mean ms (a negative value for a skipped one). */
typedef vta_status (*vta_enqueue_fn)(vta_ctx *ctx, void *user);
vta_status vta_autotune(vta_ctx *ctx, vta_tune_knob knob, const int64_t *candidates,
                        size_t n_candidates, vta_enqueue_fn enqueue, void *user, int reps,
                        int64_t *best, float *ms_per_candidate);

/* ------------------------------------------------- captured batches */
/* No counterpart in the reference (its wrappers launch eagerly, one launch
 * per call).  A sequence of enqueue-only calls on one context -- the
 * `_device` entry points, vta_alma_sweep_launch, vta_many_series_prepare's
 * launches, the fused many-series plan -- recorded once with stream capture
 * and replayed as one hipGraph launch on the context's stream.
 *
 * vta_ctx_begin_capture loads every code object first (no module load happens
 * while capturing) and starts a relaxed-mode capture of the context's stream.
 * Between begin and end only enqueue-only entry points may be called: an
 * entry point that allocates, copies from the host or synchronises (the `_dev`
 * forms, vta_ctx_synchronize) fails while capturing and invalidates the
 * capture.  vta_ctx_end_capture always ends the capture; it returns the error
 * of an invalidated capture instead of a graph.
 *
 * The graph holds the device pointers and sizes given at capture time, and
 * the context's scratch and chunk-carry buffers: replays read and write those
 * same buffers, in stream order on the context's stream.  Replaying a graph
 * while another stream of the program uses the same output buffers races. */
typedef struct vta_graph vta_graph;
vta_status vta_ctx_begin_capture(vta_ctx *ctx);
vta_status vta_ctx_end_capture(vta_ctx *ctx, vta_graph **out);
vta_status vta_graph_launch(vta_ctx *ctx, const vta_graph *graph);   /* enqueue-only */
size_t vta_graph_num_nodes(const vta_graph *graph);
vta_status vta_graph_destroy(vta_graph *graph);

/* --------------------------------------------------- device arrays */
/* DeviceArrayF32 {buf, rows, cols} (alma_wrapper.rs:103-117): owned device
 * buffer, freed with vta_device_array_free (Rust frees on drop). */
typedef struct vta_device_array_f32 {
    float *ptr;
    size_t rows;
    size_t cols;
} vta_device_array_f32;
vta_status vta_device_array_free(vta_ctx *ctx, vta_device_array_f32 *arr);
/* DeviceBuffer::copy_to -- synchronous device-to-host copy of rows*cols floats. */
vta_status vta_device_array_copy_to_host(vta_ctx *ctx, const vta_device_array_f32 *arr,
                                         float *host, size_t host_len);

/* Raw device memory for the enqueue-only *_device entry points. */
/* DLPack export (export_f32_cuda_dlpack_2d, src/utilities/dlpack_cuda.rs:13-276):
 * moves the buffer of *arr into a new DLManagedTensorVersioned (versioned != 0,
 * DLPack 1.0) or DLManagedTensor (legacy) with device {kDLROCM = 10, the
 * context's device}, float32, shape {rows, cols}, strides {cols, 1}, after the
 * context's stream has drained.  *arr is left empty; the tensor's deleter frees
 * the buffer on its device.  vta_dlpack_delete runs that deleter, for a tensor
 * that no consumer took. */
vta_status vta_device_array_to_dlpack(vta_ctx *ctx, vta_device_array_f32 *arr, int versioned,
                                      void **out_managed);
void vta_dlpack_delete(void *managed, int versioned);
vta_status vta_device_alloc(vta_ctx *ctx, size_t bytes, void **out);
vta_status vta_device_free(vta_ctx *ctx, void *ptr);
vta_status vta_memcpy_htod(vta_ctx *ctx, void *dst, const void *src, size_t bytes);
vta_status vta_memcpy_dtoh(vta_ctx *ctx, void *dst, const void *src, size_t bytes);

/* -------------------------------------------------------- ranges */
typedef struct vta_range_usize { size_t start, end, step; } vta_range_usize;
typedef struct vta_range_f64 { double start, end, step; } vta_range_f64;

/* =============================================================== ALMA */
/* AlmaParams / AlmaBatchRange (alma.rs:60-80, 1067-1082). */
typedef struct vta_alma_params {
    size_t period;
    double offset;
    double sigma;
} vta_alma_params;
typedef struct vta_alma_batch_range {
    vta_range_usize period;
    vta_range_f64 offset;
    vta_range_f64 sigma;
} vta_alma_batch_range;

/* expand_grid (alma_wrapper.rs:1550-1631): period -> offset -> sigma order;
 * f64 axes accumulate x += step while x <= end + 1e-12.  Writes up to cap
 * combos (out may be NULL to query *n_out). */
vta_status vta_alma_expand_grid(const vta_alma_batch_range *range, vta_alma_params *out,
                                size_t cap, size_t *n_out);

/* Batch kernel policy (BatchKernelPolicy / BatchKernelSelected,
 * alma_wrapper.rs:18-70).  FP32 = LDS-staged sliding dot (the reference's
 * Plain/Tiled family); MFMA_BF16X3 = split-BF16 Hankel x weights contraction
 * on the matrix cores with FP32 accumulation (hi/lo, three products);
 * MFMA_F16X3 = the same contraction in split half precision with per-window
 * and per-combo power-of-two scaling, whose split error is bounded per tile
 * and checked: a tile where the bound exceeds 2e-6 of an output is
 * recomputed with the FP32 dot.  AUTO picks MFMA_F16X3 for sweeps of at
 * least 16 combos with periods up to 256, and FP32 otherwise (also for host
 * series that change sign, where the check would send most tiles back to
 * FP32). */
typedef enum vta_alma_batch_kernel {
    VTA_ALMA_BATCH_AUTO = 0,
    VTA_ALMA_BATCH_FP32 = 1,
    VTA_ALMA_BATCH_MFMA_BF16X3 = 2,
    VTA_ALMA_BATCH_MFMA_F16X3 = 3
} vta_alma_batch_kernel;
vta_status vta_alma_set_batch_policy(vta_ctx *ctx, vta_alma_batch_kernel policy);
/* selected_batch_kernel (:239-245): kernel used by the last batch call, 0 if none. */
vta_alma_batch_kernel vta_alma_selected_batch_kernel(const vta_ctx *ctx);

/* alma_batch_dev (alma_wrapper.rs:625-669): host f32 series -> device
 * [combos][len] output. */
vta_status vta_alma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                              const vta_alma_batch_range *range, vta_device_array_f32 *out);
/* alma_batch_from_device_prices (:671-681). */
vta_status vta_alma_batch_from_device_prices(vta_ctx *ctx, const float *d_prices, size_t len,
                                             size_t first_valid,
                                             const vta_alma_batch_range *range,
                                             vta_device_array_f32 *out);
/* alma_batch_into_host_f32 (:783-817): out_len must be combos*len. */
vta_status vta_alma_batch_into_host_f32(vta_ctx *ctx, const float *prices, size_t len,
                                        const vta_alma_batch_range *range, float *out,
                                        size_t out_len);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the sweep of vta_alma_batch_from_device_prices into a
 * caller-owned d_out of out_len = combos*len floats, row-major [combo][t],
 * with the same kernels and results.  The per-combo tables go into a static
 * 16 MiB device array per GPU instead of a per-call allocation; grids whose
 * tables exceed it (about 4,000 MFMA combos) return InvalidInput.  The first
 * call on a context loads the ALMA code object; later calls allocate nothing.
 * Calls that overlap on one GPU race on the array; a call returns after its
 * sweep has completed. */
vta_status vta_alma_batch_from_device_prices_noalloc(vta_ctx *ctx, const float *d_prices,
                                                     size_t len, size_t first_valid,
                                                     const vta_alma_batch_range *range,
                                                     float *d_out, size_t out_len);
/* alma_batch_device (:565-593): enqueue-only FP32 sweep over caller-owned
 * buffers.  d_weights is [n_combos][max_period] f32, multiplied by
 * d_inv_norms[c] (the reference passes pre-normalised weights and 1.0). */
vta_status vta_alma_batch_device(vta_ctx *ctx, const float *d_prices, const float *d_weights,
                                 const int32_t *d_periods, const float *d_inv_norms,
                                 int32_t max_period, int32_t series_len, int32_t n_combos,
                                 int32_t first_valid, float *d_out);
/* alma_batch_device_tm (:595-623): same, output time-major [t][combo]. */
vta_status vta_alma_batch_device_tm(vta_ctx *ctx, const float *d_prices, const float *d_weights,
                                    const int32_t *d_periods, const float *d_inv_norms,
                                    int32_t max_period, int32_t series_len, int32_t n_combos,
                                    int32_t first_valid, float *d_out_tm);

/* Device-resident sweep plan for the MFMA form: the per-combo constants the
 * reference rebuilds on every call (compute_weights_cpu_f32, :721-747) are
 * built and uploaded once -- split-BF16 weight fragments in matrix-core
 * operand order, combos grouped by period.  vta_alma_sweep_launch is the
 * MFMA analogue of alma_batch_device: enqueue only, no allocation, no sync. */
typedef struct vta_alma_sweep_plan vta_alma_sweep_plan;
vta_status vta_alma_sweep_plan_create(vta_ctx *ctx, const vta_alma_batch_range *range,
                                      size_t series_len, vta_alma_batch_kernel kernel,
                                      vta_alma_sweep_plan **out);
vta_status vta_alma_sweep_plan_destroy(vta_ctx *ctx, vta_alma_sweep_plan *plan);
size_t vta_alma_sweep_plan_combos(const vta_alma_sweep_plan *plan);
vta_status vta_alma_sweep_launch(vta_ctx *ctx, const vta_alma_sweep_plan *plan,
                                 const float *d_prices, size_t first_valid, float *d_out);

/* Packed ALMA sweeps (no reference counterpart, whose wrappers run one series
 * per call): one-series batches of uneven lengths, each with its own
 * parameter grid, in one launch through a device table of jobs and work
 * items (a 1,024-bar x 8-combo tile of a job for the FP32 kernel, a 512-bar
 * run of a 128-combo group for the matrix-core forms).  All jobs run one
 * kernel: `kernel` (AUTO: the split-F16 form when every job has 16+ combos
 * and periods <= 256, else FP32).  Each job's outputs are bit-identical to a
 * vta_alma_sweep_launch of its own grid with that kernel (for the split-BF16
 * form, with VTA_TUNE_ALMA_MFMA_VARIANT 8, 9 or 11, the diagonal form the
 * split-F16 form always runs).  vta_alma_pack_create expands and validates
 * every job like vta_alma_batch_from_device_prices and uploads all tables at
 * once; vta_alma_pack_launch is enqueue-only (graph-capturable).  Job
 * pointers are device pointers that must stay valid while the pack is used;
 * a launch overwrites every job's output ([n_combos][series_len], combos in
 * expand_grid order). */
typedef struct vta_alma_job {
    const float *d_prices;
    float *d_out;
    vta_alma_batch_range range;
    size_t series_len;
    size_t first_valid;
} vta_alma_job;
typedef struct vta_alma_pack vta_alma_pack;
vta_status vta_alma_pack_create(vta_ctx *ctx, vta_alma_batch_kernel kernel,
                                const vta_alma_job *jobs, size_t n_jobs, vta_alma_pack **out);
vta_status vta_alma_pack_launch(vta_ctx *ctx, const vta_alma_pack *pack);
/* the kernel the pack runs, and job `job`'s number of combos (0 if out of range) */
vta_alma_batch_kernel vta_alma_pack_kernel(const vta_alma_pack *pack);
size_t vta_alma_pack_job_combos(const vta_alma_pack *pack, size_t job);
vta_status vta_alma_pack_destroy(vta_alma_pack *pack);

/* alma_multi_series_one_param_time_major_dev (:819-899): data_tm is
 * [rows=T][cols=N]; output DeviceArrayF32{rows=T, cols=N}. */
vta_status vta_alma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                         size_t cols, size_t rows,
                                                         const vta_alma_params *params,
                                                         vta_device_array_f32 *out);
/* alma_many_series_one_param_time_major_device_precomputed (:1521-1547). */
vta_status vta_alma_many_series_one_param_time_major_device(
    vta_ctx *ctx, const float *d_prices_tm, const float *d_weights, size_t period,
    float inv_norm, size_t cols, size_t rows, const int32_t *d_first_valids, float *d_out_tm);


/* ====================================================== kernel policies */
/* CudaEmaPolicy (ema_wrapper.rs:49-90), CudaRsiPolicy (rsi_wrapper.rs:92),
 * CudaAdxPolicy (adx_wrapper.rs:49-66, set_policy :116) and CudaPnrPolicy
 * (percentile_nearest_rank_wrapper.rs:51-90, set_policy/policy/
 * selected_*_kernel :143-152): per call family (one-series batch,
 * many-series) Auto or a fixed form, BatchKernelPolicy::Plain { block_x } /
 * ManySeriesKernelPolicy::OneD { block_x }.  Here the fixed form is the plain
 * kernel: one lane per combo (or series) over the whole series from
 * first_valid, no time chunking.  These kernels run 64-lane waves, so
 * block_x must be 0 (default) or 64; other values are InvalidPolicy.  The
 * selected kernel of the last call is PLAIN or CHUNKED (time chunks: the
 * chunked scan or burn-in starts) with block_x 64, NONE before the first
 * call. */
typedef enum vta_kernel_policy_kind {
    VTA_KERNEL_AUTO = 0,
    VTA_KERNEL_PLAIN = 1
} vta_kernel_policy_kind;
typedef struct vta_kernel_policy {
    int32_t batch;               /* vta_kernel_policy_kind */
    uint32_t batch_block_x;
    int32_t many_series;         /* vta_kernel_policy_kind */
    uint32_t many_series_block_x;
} vta_kernel_policy;
typedef enum vta_kernel_selected_kind {
    VTA_SELECTED_NONE = 0,
    VTA_SELECTED_PLAIN = 1,
    VTA_SELECTED_CHUNKED = 2
} vta_kernel_selected_kind;
typedef struct vta_kernel_selected {
    int32_t kind;                /* vta_kernel_selected_kind */
    uint32_t block_x;
} vta_kernel_selected;
vta_status vta_ema_set_policy(vta_ctx *ctx, const vta_kernel_policy *policy);
vta_status vta_ema_policy(const vta_ctx *ctx, vta_kernel_policy *out);
vta_status vta_ema_selected_batch_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_ema_selected_many_series_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_rsi_set_policy(vta_ctx *ctx, const vta_kernel_policy *policy);
vta_status vta_rsi_policy(const vta_ctx *ctx, vta_kernel_policy *out);
vta_status vta_rsi_selected_batch_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_rsi_selected_many_series_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_adx_set_policy(vta_ctx *ctx, const vta_kernel_policy *policy);
vta_status vta_adx_policy(const vta_ctx *ctx, vta_kernel_policy *out);
vta_status vta_adx_selected_batch_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_adx_selected_many_series_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_pnr_set_policy(vta_ctx *ctx, const vta_kernel_policy *policy);
vta_status vta_pnr_policy(const vta_ctx *ctx, vta_kernel_policy *out);
vta_status vta_pnr_selected_batch_kernel(const vta_ctx *ctx, vta_kernel_selected *out);
vta_status vta_pnr_selected_many_series_kernel(const vta_ctx *ctx, vta_kernel_selected *out);

/* ====================================================== recurrences */
/* One-parameter batch ranges are vta_range_usize over the period, expanded
 * like the reference wrappers: EMA/RSI ascending (ema_wrapper.rs:862-873,
 * rsi_wrapper.rs:389-413), ATR (atr_wrapper.rs:286-297), ADX
 * (adx_wrapper.rs:158-178).  Engines run in f64 with the reference's
 * operation order and round to f32 on output. */

/* EMA: ema_wrapper.rs:193 (batch_dev), :256 (batch_device), :309
 * (batch_into_host_f32), :325 (many_series_one_param_time_major_dev), :381
 * (many_series_one_param_device), :422
 * (many_series_one_param_time_major_into_host_f32).
 * The device forms use the caller's smoothing factors (d_alphas per combo,
 * alpha), as the reference kernels do (its host code passes 2/(p+1) in f32);
 * d_alphas NULL derives 2/(p+1) in f64 like the _dev forms.  The many-series
 * device form finds each series' non-finite bars from its first_valid on the
 * device.  *_into_host_f32: out_len must be combos*len (rows*cols). */
vta_status vta_ema_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out);
vta_status vta_ema_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_ema_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                const float *d_alphas, size_t series_len, size_t first_valid,
                                size_t n_combos, float *d_out);
vta_status vta_ema_batch_into_host_f32(vta_ctx *ctx, const float *prices, size_t len,
                                       const vta_range_usize *range, float *out,
                                       size_t out_len);
vta_status vta_ema_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                        size_t cols, size_t rows, size_t period,
                                                        vta_device_array_f32 *out);
vta_status vta_ema_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                const int32_t *d_first_valids, int32_t period,
                                                float alpha, size_t num_series,
                                                size_t series_len, float *d_out_tm);
vta_status vta_ema_many_series_one_param_time_major_into_host_f32(
    vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period, float *out_tm,
    size_t out_len);

/* RSI (Wilder): rsi_wrapper.rs:152 (rsi_batch_dev), :271
 * (rsi_many_series_one_param_time_major_dev). */
vta_status vta_rsi_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out);
vta_status vta_rsi_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_rsi_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out);
vta_status vta_rsi_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *prices_tm,
                                                        size_t cols, size_t rows, size_t period,
                                                        vta_device_array_f32 *out);
vta_status vta_rsi_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                const int32_t *d_first_valids, size_t period,
                                                size_t num_series, size_t series_len,
                                                const int32_t *d_clean, float *d_out_tm);

/* ATR: atr_wrapper.rs:270 (atr_batch_dev), :626 (atr_batch_device_with_tr),
 * :834 (atr_many_series_one_param_time_major_dev). */
vta_status vta_atr_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out);
vta_status vta_atr_batch_dev(vta_ctx *ctx, const float *high, const float *low,
                             const float *close, size_t len, const vta_range_usize *range,
                             vta_device_array_f32 *out);
vta_status vta_atr_batch_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                const float *d_close, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out);
/* tr_from_hlc_device (atr_wrapper.rs:566-624): enqueue only; tr[t] = 0 before
 * first_valid, high - low at first_valid, max(h-l, |h-c[t-1]|, |l-c[t-1]|)
 * after.  atr_batch_device_with_tr (:626-700): ATR of every period over that
 * TR series (seed = mean of the first p TRs at first_valid + p - 1); the
 * reference's alphas and warm indices are derived from the periods. */
vta_status vta_tr_from_hlc_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                  const float *d_close, int32_t series_len, int32_t first_valid,
                                  float *d_tr_out);
vta_status vta_atr_batch_device_with_tr(vta_ctx *ctx, const float *d_tr, const int32_t *d_periods,
                                        int32_t max_period, int32_t series_len,
                                        int32_t first_valid, int32_t n_combos, float *d_out);

/* ------------------------------------------------- packed one-series batches */
/* No counterpart in the reference, whose wrappers run one series per call.
 * Many one-series batches of one recurrence indicator -- series of uneven
 * lengths, each with its own period table -- in one launch per scan pass
 * through a device job table (offsets of every job's blocks and combos).
 * Each job's outputs are bit-identical to its own vta_<ind>_batch_device call
 * on a context with the same tuning (same chunking, same kernels), except
 * where that call would skip the scan for a grid of more than ~260k combos
 * (the pack then runs the job exactly in one chunk: within the parity bound).
 * vta_rec_pack_create validates the jobs like the _device forms, uploads the
 * table and allocates the pack's own chunk-carry buffer and flags (so
 * launches allocate nothing); vta_rec_pack_launch is enqueue-only (graph-
 * capturable).  Job pointers are device pointers that must stay valid while
 * the pack is used; a launch overwrites every job's output. */
typedef enum vta_rec_kind {
    VTA_REC_EMA = 0,
    VTA_REC_RSI = 1,
    VTA_REC_ATR = 2,    /* high, low, close */
    VTA_REC_ADX = 3,    /* high, low, close */
    VTA_REC_SMMA = 4,
    VTA_REC_WILDERS = 5
} vta_rec_kind;
typedef struct vta_rec_job {
    const float *d_close;     /* prices (EMA, RSI, SMMA, Wilders) or close (ATR, ADX) */
    const float *d_high;      /* ATR, ADX; NULL otherwise */
    const float *d_low;       /* ATR, ADX; NULL otherwise */
    const int32_t *d_periods; /* n_combos periods */
    float *d_out;             /* [n_combos][series_len] */
    int32_t series_len;
    int32_t first_valid;      /* of close (ATR, ADX: of high, low and close) */
    int32_t n_combos;
    int32_t max_period;
} vta_rec_job;
typedef struct vta_rec_pack vta_rec_pack;
vta_status vta_rec_pack_create(vta_ctx *ctx, vta_rec_kind kind, const vta_rec_job *jobs,
                               size_t n_jobs, vta_rec_pack **out);
vta_status vta_rec_pack_launch(vta_ctx *ctx, const vta_rec_pack *pack);
vta_status vta_rec_pack_destroy(vta_rec_pack *pack);
/* Packed one-series Bollinger Bands batches (no reference counterpart), the
 * same idea for vta_bollinger_bands_batch_device: each job's three outputs
 * are bit-identical to its own _device call; one launch for every job. */
typedef struct vta_bb_job {
    const float *d_prices;
    const int32_t *d_periods;
    const double *d_devup, *d_devdn;
    float *d_upper, *d_middle, *d_lower;   /* [n_combos][series_len] each */
    int32_t series_len;
    int32_t first_valid;
    int32_t first_nonfinite;               /* first NaN/Inf at or after first_valid; series_len if none */
    int32_t n_combos;
    int32_t max_period;
} vta_bb_job;
typedef struct vta_bb_pack vta_bb_pack;
vta_status vta_bb_pack_create(vta_ctx *ctx, const vta_bb_job *jobs, size_t n_jobs,
                              vta_bb_pack **out);
vta_status vta_bb_pack_launch(vta_ctx *ctx, const vta_bb_pack *pack);
vta_status vta_bb_pack_destroy(vta_bb_pack *pack);
vta_status vta_atr_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *high_tm,
                                                        const float *low_tm,
                                                        const float *close_tm, size_t cols,
                                                        size_t rows, size_t period,
                                                        vta_device_array_f32 *out);
vta_status vta_atr_many_series_one_param_device(vta_ctx *ctx, const float *d_high_tm,
                                                const float *d_low_tm, const float *d_close_tm,
                                                const int32_t *d_first_valids, size_t period,
                                                size_t num_series, size_t series_len,
                                                const int32_t *d_clean, float *d_out_tm);

/* ADX: adx_wrapper.rs:199 (adx_batch_dev), :340
 * (adx_many_series_one_param_time_major_dev). */
vta_status vta_adx_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out);
vta_status vta_adx_batch_dev(vta_ctx *ctx, const float *high, const float *low,
                             const float *close, size_t len, const vta_range_usize *range,
                             vta_device_array_f32 *out);
/* Enqueue-only device form (companion of vta_atr_batch_device; the reference
 * has no ADX device form).  first_valid: the first bar where high, low and
 * close are all non-NaN (adx.rs:190-198). */
vta_status vta_adx_batch_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                const float *d_close, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out);
vta_status vta_adx_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *high_tm,
                                                        const float *low_tm,
                                                        const float *close_tm, size_t cols,
                                                        size_t rows, size_t period,
                                                        vta_device_array_f32 *out);
/* adx_batch_into_host_f32 (:253-280) and
 * adx_many_series_one_param_time_major_into_host_f32 (:417-440): out_len must
 * be combos*len (rows*cols). */
vta_status vta_adx_batch_into_host_f32(vta_ctx *ctx, const float *high, const float *low,
                                       const float *close, size_t len,
                                       const vta_range_usize *range, float *out,
                                       size_t out_len);
vta_status vta_adx_many_series_one_param_time_major_into_host_f32(
    vta_ctx *ctx, const float *high_tm, const float *low_tm, const float *close_tm, size_t cols,
    size_t rows, size_t period, float *out_tm, size_t out_len);

/* SMMA: smma_wrapper.rs:398 (smma_batch_dev), :377 (smma_batch_device), :623
 * (smma_multi_series_one_param_time_major_dev), :599
 * (smma_multi_series_one_param_device).  Grid: smma.rs:430-471.  Seed = mean
 * of the first p bars at fv + p - 1, then prev = (prev (p-1) + x) / p.
 *
 * Warm indices (SMMA, Wilders, SWMA device forms): the reference's host code
 * fills d_warms[c] = first_valid + period[c] - 1, and the kernels here derive
 * the warm-up from the periods, so the buffers are checked on the device: a
 * mismatch makes the next vta_ctx_synchronize return InvalidInput (the
 * enqueue-only call itself cannot wait for the check), and that call's
 * outputs are undefined; a NULL warm buffer is not checked.  The many-series
 * device forms find each series' non-finite bars from its first_valid on the
 * device. */
vta_status vta_smma_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                                size_t *n_out);
vta_status vta_smma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                              const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_smma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                 const int32_t *d_warms, size_t first_valid, size_t series_len,
                                 size_t n_combos, float *d_out);
vta_status vta_smma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                         size_t cols, size_t rows, size_t period,
                                                         vta_device_array_f32 *out);
vta_status vta_smma_multi_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                  int32_t period, int32_t num_series,
                                                  int32_t series_len,
                                                  const int32_t *d_first_valids,
                                                  float *d_out_tm);

/* Wilders: wilders_wrapper.rs:134 (wilders_batch_dev; the first max_period
 * bars after first_valid must be finite, :391-400), :252
 * (wilders_batch_device: d_alphas used, NULL derives 1/p in f64; d_warm
 * checked as above),
 * :600 (wilders_many_series_one_param_time_major_dev; first_valid = first
 * finite bar), :662 (wilders_many_series_one_param_device: alpha used).
 * Grid: wilders_wrapper.rs:909-940. */
vta_status vta_wilders_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                                   size_t *n_out);
vta_status vta_wilders_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                                 const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_wilders_batch_device(vta_ctx *ctx, const float *d_prices,
                                    const int32_t *d_periods, const float *d_alphas,
                                    const int32_t *d_warm, int32_t series_len,
                                    int32_t first_valid, int32_t n_combos, float *d_out);
vta_status vta_wilders_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                            size_t cols, size_t rows,
                                                            size_t period,
                                                            vta_device_array_f32 *out);
vta_status vta_wilders_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                    const int32_t *d_first_valids,
                                                    int32_t period, float alpha,
                                                    size_t num_series, size_t series_len,
                                                    float *d_out_tm);

/* No reference counterpart.  The one-input many-series device forms (EMA,
 * SMMA, Wilders, SMA, WMA, SWMA) over per-series metadata prepared once with
 * vta_many_series_prepare on the same device data -- d_meta is the clean flag
 * (mode 0) for EMA / SMMA / Wilders and the first non-finite bar (mode 4) for
 * SMA / WMA / SWMA -- for callers that launch repeatedly over that data; the
 * reference-signature device forms above find the metadata on every call
 * (one extra read of the matrix).  alpha <= 0 derives the smoothing factor
 * from the period (EMA 2/(p+1), Wilders 1/p, in f64); other kinds ignore it. */
typedef enum vta_ma_kind {
    VTA_MA_EMA = 0,
    VTA_MA_SMMA = 1,
    VTA_MA_WILDERS = 2,
    VTA_MA_SMA = 3,
    VTA_MA_WMA = 4,
    VTA_MA_SWMA = 5
} vta_ma_kind;
vta_status vta_ma_many_series_prepared_device(vta_ctx *ctx, vta_ma_kind kind,
                                              const float *d_prices_tm,
                                              const int32_t *d_first_valids,
                                              const int32_t *d_meta, size_t period, float alpha,
                                              size_t num_series, size_t series_len,
                                              float *d_out_tm);

/* Per-series metadata of device-resident time-major inputs, enqueue only:
 * first_valid (mode 0: x; 1: first bar with high, low, x all non-NaN; 2: ADX
 * triple rule; 3: x, with the mode-1 bar in d_first_valids_hlc; 4: x, with
 * d_clean receiving the first non-finite bar at or after first_valid, T if
 * none, for the finite-window averages) and the clean flag.  The reference computes first_valids on the host
 * (ema_wrapper.rs:637-700); this is its device-side counterpart. */
vta_status vta_many_series_prepare(vta_ctx *ctx, const float *d_x_tm, const float *d_high_tm,
                                   const float *d_low_tm, size_t cols, size_t rows, int mode,
                                   int32_t *d_first_valids, int32_t *d_first_valids_hlc,
                                   int32_t *d_clean);

/* Fused EMA(p1) + RSI(p2) of close and ATR(p3) of (high, low, close) over
 * time-major [rows][cols] series in one launch (no reference counterpart:
 * the reference runs three wrappers, three uploads and three launches,
 * ema_wrapper.rs:325, rsi_wrapper.rs:271, atr_wrapper.rs:834).  _device is
 * enqueue only with metadata from vta_many_series_prepare(mode 3); _dev takes
 * host data and returns three device arrays {ema, rsi, atr}. */
vta_status vta_ema_rsi_atr_many_series_device(vta_ctx *ctx, const float *d_close_tm,
                                              const float *d_high_tm, const float *d_low_tm,
                                              size_t cols, size_t rows, size_t ema_period,
                                              size_t rsi_period, size_t atr_period,
                                              const int32_t *d_fv_close,
                                              const int32_t *d_fv_hlc, const int32_t *d_clean,
                                              float *d_ema_tm, float *d_rsi_tm,
                                              float *d_atr_tm);
vta_status vta_ema_rsi_atr_many_series_dev(vta_ctx *ctx, const float *close_tm,
                                           const float *high_tm, const float *low_tm,
                                           size_t cols, size_t rows, size_t ema_period,
                                           size_t rsi_period, size_t atr_period,
                                           vta_device_array_f32 *out3);

/* ------------------------------------------------------------------ finite-window averages */
/* SMA (sma_wrapper.rs:548 sma_batch_dev, :830 sma_batch_dev_from_device, :750
 * sma_multi_series_one_param_time_major_dev, :982 ..._dev_from_device), WMA
 * (wma_wrapper.rs:713 wma_batch_dev, :747 wma_batch_device, :1014
 * wma_multi_series_one_param_time_major_dev, :1049 ..._device) and SWMA
 * (swma_wrapper.rs:678 swma_batch_dev, :663 swma_batch_device, :825
 * swma_multi_series_one_param_time_major_dev, :787 swma_multi_series_one_param_device).
 * Grids: sma.rs:705-755, wma_wrapper.rs:324-350, swma_wrapper.rs:275-301.
 * Periods: SMA/SWMA >= 1, WMA >= 2.  The device forms find the first
 * non-finite bar (of each series) on the device.  swma_batch_device derives
 * first_valid from the warm indices on the device and checks them (see the
 * warm-index note above); swma_multi_series_one_param_device derives the
 * triangular weights from the period (d_weights is not read: the reference's
 * host code uploads those same normalised weights).  vta_sma_batch_device and
 * vta_sma_many_series_one_param_device are enqueue-only SMA forms with no
 * reference counterpart (the reference's SMA device-input forms allocate
 * their output: vta_sma_batch_dev_from_device, where periods[] is the host
 * array of SmaParams periods, and ..._dev_from_device). */
vta_status vta_sma_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                              size_t *n_out);
vta_status vta_sma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                            const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_sma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                               int32_t series_len, int32_t n_combos, int32_t first_valid,
                               int32_t max_period, float *d_out);
vta_status vta_sma_batch_dev_from_device(vta_ctx *ctx, const float *d_prices, size_t len,
                                         const size_t *periods, size_t n_combos,
                                         size_t first_valid, vta_device_array_f32 *out);
vta_status vta_sma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                       size_t cols, size_t rows, size_t period,
                                                       vta_device_array_f32 *out);
vta_status vta_sma_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                               const int32_t *d_first_valids, size_t cols,
                                               size_t rows, size_t period, float *d_out_tm);
vta_status vta_sma_multi_series_one_param_time_major_dev_from_device(
    vta_ctx *ctx, const float *d_prices_tm, const int32_t *d_first_valids, size_t cols,
    size_t rows, size_t period, vta_device_array_f32 *out);

vta_status vta_wma_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                              size_t *n_out);
vta_status vta_wma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                            const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_wma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                               int32_t series_len, int32_t n_combos, int32_t first_valid,
                               int32_t max_period, float *d_out);
vta_status vta_wma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                       size_t cols, size_t rows, size_t period,
                                                       vta_device_array_f32 *out);
vta_status vta_wma_multi_series_one_param_time_major_device(vta_ctx *ctx,
                                                            const float *d_prices_tm,
                                                            int32_t period, int32_t num_series,
                                                            int32_t series_len,
                                                            const int32_t *d_first_valids,
                                                            float *d_out_tm);

vta_status vta_swma_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                              size_t *n_out);
vta_status vta_swma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                            const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_swma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                 const int32_t *d_warms, size_t series_len, size_t n_combos,
                                 size_t max_period, float *d_out);
vta_status vta_swma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                       size_t cols, size_t rows, size_t period,
                                                       vta_device_array_f32 *out);
vta_status vta_swma_multi_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                  const float *d_weights, int32_t period,
                                                  int32_t num_series, int32_t series_len,
                                                  const int32_t *d_first_valids,
                                                  float *d_out_tm);

/* ------------------------------------------------------------------ rolling windows */

/* Bollinger Bands, SMA middle band and population standard deviation: the
 * combinations the reference's GPU path accepts (ma_type "sma", devtype 0,
 * bollinger_bands_wrapper.rs:284-350); the grid is period x devup x devdn in
 * that nesting order (bollinger_bands_wrapper.rs:179-281).  devup/devdn stay
 * f64 here (the reference narrows them to f32, :330-334). */
typedef struct vta_bb_params {
    size_t period;
    double devup;
    double devdn;
} vta_bb_params;
typedef struct vta_bb_batch_range {
    vta_range_usize period;
    vta_range_f64 devup;
    vta_range_f64 devdn;
} vta_bb_batch_range;

vta_status vta_bb_expand_grid(const vta_bb_batch_range *range, vta_bb_params *out, size_t cap,
                              size_t *n_out);
/* bollinger_bands_wrapper.rs:567 (bollinger_bands_batch_dev): out3 receives
 * {upper, middle, lower}, each [combo][len]. */
vta_status vta_bollinger_bands_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                                         const vta_bb_batch_range *range,
                                         vta_device_array_f32 *out3);
/* Enqueue-only one-series batch on device-resident prices (no reference
 * counterpart): per-combo d_periods / d_devup / d_devdn, outputs
 * [combo][series_len] each; first_nonfinite = the first bar at or after
 * first_valid that is NaN or +-Inf (series_len if none). */
vta_status vta_bollinger_bands_batch_device(vta_ctx *ctx, const float *d_prices,
                                            int32_t series_len, int32_t first_valid,
                                            int32_t first_nonfinite, const int32_t *d_periods,
                                            const double *d_devup, const double *d_devdn,
                                            int32_t n_combos, int32_t max_period,
                                            float *d_upper, float *d_middle, float *d_lower);
/* bollinger_bands_wrapper.rs:605
 * (bollinger_bands_many_series_one_param_time_major_dev): out3 receives
 * {upper, middle, lower}, each time-major [rows][cols]. */
vta_status vta_bollinger_bands_many_series_one_param_time_major_dev(
    vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period, double devup,
    double devdn, vta_device_array_f32 *out3);
/* Enqueue-only device form (no reference counterpart; the companion of the
 * *_device entry points of the other wrappers).  d_first_nonfinite[s] is the
 * first bar at or after first_valid whose value is NaN or +-Inf (rows if
 * none); NULL runs each series as one exact sequential pass. */
vta_status vta_bollinger_bands_many_series_one_param_device(
    vta_ctx *ctx, const float *d_data_tm, const int32_t *d_first_valids,
    const int32_t *d_first_nonfinite, size_t cols, size_t rows, size_t period, double devup,
    double devdn, float *d_upper_tm, float *d_middle_tm, float *d_lower_tm);

/* percentile_nearest_rank: grid length x percentage
 * (percentile_nearest_rank_wrapper.rs:200-290). */
typedef struct vta_pnr_params {
    size_t length;
    double percentage;
} vta_pnr_params;
typedef struct vta_pnr_batch_range {
    vta_range_usize length;
    vta_range_f64 percentage;
} vta_pnr_batch_range;

vta_status vta_pnr_expand_grid(const vta_pnr_batch_range *range, vta_pnr_params *out,
                               size_t cap, size_t *n_out);
/* percentile_nearest_rank_wrapper.rs:365 (pnr_batch_dev), :623
 * (pnr_many_series_one_param_time_major_dev). */
vta_status vta_pnr_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_pnr_batch_range *range, vta_device_array_f32 *out);
vta_status vta_pnr_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *prices_tm,
                                                        size_t cols, size_t rows, size_t length,
                                                        double percentage,
                                                        vta_device_array_f32 *out);
/* Enqueue-only device form (no reference counterpart): per-series
 * d_first_valids must leave each series at least `length` bars. */
vta_status vta_pnr_many_series_one_param_device(vta_ctx *ctx, const float *d_data_tm,
                                                const int32_t *d_first_valids, size_t cols,
                                                size_t rows, size_t length, double percentage,
                                                float *d_out_tm);

/* Rolling median absolute deviation: medium_ad_wrapper.rs:120 (expand_grid),
 * :366 (medium_ad_batch_dev), :470
 * (medium_ad_many_series_one_param_time_major_dev). */
vta_status vta_medium_ad_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                                     size_t *n_out);
vta_status vta_medium_ad_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                                   const vta_range_usize *range, vta_device_array_f32 *out);
vta_status vta_medium_ad_many_series_one_param_time_major_dev(vta_ctx *ctx,
                                                              const float *data_tm, size_t cols,
                                                              size_t rows, size_t period,
                                                              vta_device_array_f32 *out);
/* Enqueue-only device form (no reference counterpart). */
vta_status vta_medium_ad_many_series_one_param_device(vta_ctx *ctx, const float *d_data_tm,
                                                      const int32_t *d_first_valids, size_t cols,
                                                      size_t rows, size_t period,
                                                      float *d_out_tm);

/* ------------------------------------------------------- double-ALMA crossover */

/* Backtest reduction of a double-ALMA crossover sweep (no reference
 * counterpart: README.md:9 names the workload, 58,300 fast/slow pairs over
 * 200,000 bars).  Pairs are fast x slow in that nesting order (pair
 * i * n_slow + j); for each, pos_t = sign(ALMA_fast[t] - ALMA_slow[t]) with
 * the batch ALMA weights (alma.rs:1426-1436), from the first bar where both
 * are valid, reduced to
 *   stats4[4p + 0] trades    (bars where the position changes),
 *   stats4[4p + 1] n_long    (bars with pos = +1),
 *   stats4[4p + 2] n_short   (bars with pos = -1),
 *   stats4[4p + 3] checksum  = sum_t pos_t * (t + 1),
 *   pnl[p]         = sum_{t <= len-2} pos_t * (price[t+1] - price[t]).
 * Positions equal the f64 ALMA's signs exactly: the rows of the distinct
 * parameter sets are computed on the device in f64 in the scalar kernel's
 * operation order (alma.rs:1505-1527).  Prices must be finite after
 * first_valid; periods <= 294.  stats4 and pnl NULL: *n_pairs only. */
vta_status vta_double_alma_crossover(vta_ctx *ctx, const float *prices, size_t len,
                                     const vta_alma_batch_range *fast,
                                     const vta_alma_batch_range *slow, int64_t *stats4,
                                     double *pnl, size_t cap, size_t *n_pairs);
/* Prepared form: prices resident on the device, tables built once; launch is
 * enqueue only (f64 rows, pair walk, segment reduction); results
 * synchronises and copies. */
typedef struct vta_double_alma_plan vta_double_alma_plan;
vta_status vta_double_alma_plan_create(vta_ctx *ctx, const float *prices, size_t len,
                                       const vta_alma_batch_range *fast,
                                       const vta_alma_batch_range *slow,
                                       vta_double_alma_plan **out);
size_t vta_double_alma_plan_pairs(const vta_double_alma_plan *plan);
vta_status vta_double_alma_plan_launch(vta_double_alma_plan *plan);
vta_status vta_double_alma_plan_results(vta_double_alma_plan *plan, int64_t *stats4,
                                        double *pnl, size_t cap);
vta_status vta_double_alma_plan_destroy(vta_double_alma_plan *plan);

/* ------------------------------------------------------------- request server */

/* Continuous-batching request server (no reference counterpart).  n_slots
 * worker threads, each with its own context and stream on device_id, take
 * one-series batch requests from a FIFO queue; a request is admitted when a
 * slot is free and hbm_budget bytes hold its outputs plus, on a series-cache
 * miss, its series.  Results stay on the device (counted against the budget)
 * until vta_server_release.  Series are cached under the client's key
 * (series_key != 0 and cache = 1), evicted least-recently-used when admission
 * needs room; a request always carries its host data, used on a miss.  Results
 * are the library's batch kernels on the same inputs: identical whether a
 * request runs alone or among others. */
typedef enum vta_req_kind {
    VTA_REQ_ALMA = 0,   /* alma: grid alma */
    VTA_REQ_EMA = 1,    /* ema: grid period */
    VTA_REQ_RSI = 2,    /* rsi: grid period */
    VTA_REQ_BB = 3,     /* bollinger_bands: grid bb; three outputs */
    VTA_REQ_ADX = 4     /* adx: grid period; needs high and low */
} vta_req_kind;

typedef struct vta_request {
    int32_t kind;                /* vta_req_kind */
    int32_t cache;               /* keep the series in the device cache */
    uint64_t series_key;         /* 0: not cacheable */
    const float *close;          /* host data, copied at submit */
    const float *high;           /* NULL unless the indicator needs it */
    const float *low;
    size_t len;
    vta_alma_batch_range alma;
    vta_range_usize period;
    vta_bb_batch_range bb;
} vta_request;

typedef struct vta_result {
    int32_t status;              /* vta_status of the request */
    int32_t n_outputs;           /* 1, or 3 for Bollinger Bands (upper, middle, lower) */
    size_t n_combos;
    vta_device_array_f32 out[3]; /* [combo][len], owned by the server until release */
    double t_submit, t_start, t_done;   /* seconds on the server's clock */
    double t_series, t_launch;          /* series resident; kernels enqueued */
} vta_result;

typedef struct vta_server_stats {
    size_t budget, in_use, peak_in_use;
    size_t queued, in_flight, peak_in_flight;
    size_t cached_series;
    /* small ALMA / EMA / RSI / ADX / Bollinger Bands requests served together in packed
     * launches: launches (packs) and the requests they served */
    size_t packs, packed_requests;
} vta_server_stats;

typedef struct vta_server vta_server;
vta_status vta_server_create(int device_id, int n_slots, size_t hbm_budget, vta_server **out);
vta_status vta_server_destroy(vta_server *server);
vta_status vta_server_submit(vta_server *server, const vta_request *req, uint64_t *req_id);
/* timeout_ms < 0 waits for ever; VTA_ERR_TIMEOUT if not finished in time. */
vta_status vta_server_wait(vta_server *server, uint64_t req_id, int timeout_ms, vta_result *res);
vta_status vta_server_release(vta_server *server, uint64_t req_id);
vta_status vta_server_get_stats(vta_server *server, vta_server_stats *out);
double vta_server_now(const vta_server *server);

#ifdef __cplusplus
}
#endif

#endif /* VTA_H */
