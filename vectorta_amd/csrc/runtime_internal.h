// Internal helpers shared by the libvta_hip host runtime translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "vta.h"

struct vta_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    std::string kernel_dir;
    std::unordered_map<std::string, hipModule_t> modules;
    std::unordered_map<std::string, hipFunction_t> funcs;
    // per-indicator policies and last selections (CudaAlmaPolicy etc.)
    vta_alma_batch_kernel alma_policy = VTA_ALMA_BATCH_AUTO;
    vta_alma_batch_kernel alma_selected = VTA_ALMA_BATCH_AUTO;
    // small device buffer owned by the context (flags written by prep kernels
    // of the enqueue-only entry points; reused in stream order)
    void *scratch = nullptr;
    static constexpr size_t kScratchBytes = 64 * 1024;
    // chunk-end states of the recurrences' chunked scan ([combo][chunk][state]
    // f64), reused in stream order by every batch launch of the context;
    // 64 MiB holds 1,250 ADX combos (four states) at 1,024 chunks
    void *carry = nullptr;
    static constexpr size_t kCarryBytes = 64u << 20;
    // launch-shape knobs (vta_tune_knob); 0 = built-in heuristic
    int64_t tune[VTA_TUNE_KNOB_COUNT] = {};
    // EMA / RSI / ADX / PNR kernel policies and the forms their last batch and
    // many-series calls ran (indexed by vta::PolicySlot)
    vta_kernel_policy policy[4] = {};
    vta_kernel_selected sel_batch[4] = {};
    vta_kernel_selected sel_ms[4] = {};
    // an enqueue-only call with device-side argument checks (warm indices) has
    // run since the last vta_ctx_synchronize, which then reads the error word
    bool arg_check_pending = false;
    // epoch of the last device-resident one-series batch (batch_prepare)
    uint32_t prep_epoch = 0;
};

namespace vta {

enum PolicySlot {
    VTA_POLICY_SLOT_EMA = 0,
    VTA_POLICY_SLOT_RSI = 1,
    VTA_POLICY_SLOT_ADX = 2,
    VTA_POLICY_SLOT_PNR = 3
};

// int32 word of the context scratch that device-side argument checks set to 1
// on a mismatch; vta_ctx_synchronize reports and clears it.
constexpr size_t kArgErrorWord = 1024;

vta_status fail(vta_status st, const char *fmt, ...);
void clear_error();

#define VTA_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t vta_e_ = (expr);                                                      \
        if (vta_e_ != hipSuccess)                                                        \
            return ::vta::fail(VTA_ERR_HIP, "%s failed: %s", #expr,                      \
                               hipGetErrorString(vta_e_));                               \
    } while (0)

#define VTA_TRY(expr)                                                                    \
    do {                                                                                 \
        vta_status vta_s_ = (expr);                                                      \
        if (vta_s_ != VTA_OK) return vta_s_;                                             \
    } while (0)

// Module / symbol lookup: <kernel_dir>/vta_<module>.hsaco, hipModuleGetFunction.
vta_status get_function(vta_ctx *ctx, const char *module, const char *name, hipFunction_t *out);
bool has_function(vta_ctx *ctx, const char *module, const char *name);

// Launch-limit check against the device (validate_launch, alma_wrapper.rs:365-417).
vta_status validate_launch(vta_ctx *ctx, dim3 grid, dim3 block);

// Raw launch: args is the kernelParams array.
vta_status launch_raw(vta_ctx *ctx, hipFunction_t f, dim3 grid, dim3 block, unsigned shmem,
                      void **args);

template <typename... A>
vta_status launch(vta_ctx *ctx, hipFunction_t f, dim3 grid, dim3 block, unsigned shmem,
                  A... a) {
    void *args[] = {static_cast<void *>(&a)...};
    return launch_raw(ctx, f, grid, block, shmem, args);
}

// will_fit with 64 MiB headroom (alma_wrapper.rs:287-310, 645-664);
// VTA_MEM_CHECK=0 disables it like CUDA_MEM_CHECK=0.
vta_status check_fits(size_t required_bytes);

// checked size arithmetic
bool mul_ok(size_t a, size_t b, size_t *out);
bool add_ok(size_t a, size_t b, size_t *out);

// RAII device buffer used for temporaries.
struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    vta_status alloc(size_t bytes);
    void *release() {
        void *q = p;
        p = nullptr;
        return q;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

vta_status upload(vta_ctx *ctx, DevBuf &buf, const void *src, size_t bytes);

size_t first_valid_f32(const float *x, size_t n);  // n if all NaN

// Parameter axis expansions used by the reference wrappers:
//   AXIS_SORTED  start..=end ascending, either order of start/end (ema_wrapper.rs:862-873,
//                rsi_wrapper.rs:389-413)
//   AXIS_DOWN    ascending if start < end, else start, start-step, ... >= end
//                (alma_wrapper.rs:1551-1575, adx_wrapper.rs:158-178)
//   AXIS_ATR     ascending if start < end, else (end..=start step) reversed (atr_wrapper.rs:286-297)
//   AXIS_DOWN_END like AXIS_DOWN, then end appended when the countdown misses it
//                (wma_wrapper.rs:324-350)
enum AxisKind { AXIS_SORTED, AXIS_DOWN, AXIS_ATR, AXIS_DOWN_END };
// Largest axis, and largest grid (product of axes), an expansion builds:
// 16,777,216 values.  Wider ranges, and f64 axes with an infinite bound that
// survives the single-value rule, are refused (the expansion returns false,
// the caller InvalidInput) instead of exhausting host memory or looping
// forever, which is what the reference's collecting loops would do.
constexpr size_t kMaxGridValues = size_t(1) << 24;
// True when n values stay within kMaxGridValues.
inline bool grid_fits(size_t n) { return n <= kMaxGridValues; }
bool expand_usize(vta_range_usize r, AxisKind kind, std::vector<size_t> &v);
// f64 axis: values accumulate x += step while x <= end + 1e-12 (or descend
// while x + 1e-12 >= end), alma_wrapper.rs:1576-1612.
bool expand_f64(vta_range_f64 r, std::vector<double> &v);

// Per-series first_valid of a time-major [rows][cols] host matrix; fails with
// InvalidInput if a series is all NaN or has fewer than `need` valid bars.
vta_status host_first_valids(const float *tm, size_t cols, size_t rows, size_t need,
                             std::vector<int32_t> &fvs);

// Device tables of one ALMA sweep (alma.cpp), built on the host and laid out
// in one buffer: each table starts at a 256-byte boundary at off[section] of
// blob.  A sweep plan uploads the blob to a buffer of its own; the
// allocation-free entry point copies it into a static device array.
struct AlmaTables {
    enum Section {
        W, PERIODS, INV,                                                  // FP32 form
        FRAGS, NBLOCKS, GROUPS, SORTED_PERIOD, SORTED_ROW, W_F32,         // MFMA forms
        CSCALE,                                                           // MFMA F16 form
        COUNT
    };
    vta_alma_batch_kernel kernel = VTA_ALMA_BATCH_FP32;
    size_t n_combos = 0;
    size_t max_period = 0;
    int n_nblocks = 0;
    int n_groups = 0;
    int n_groups_klo = 0;   // leading groups with K <= 128
    int kmax = 0;
    size_t off[COUNT] = {};
    std::vector<unsigned char> blob;
};
// Expands and validates `range` against a series of len bars whose first
// valid bar is fv (as alma_batch_from_device_prices does) and builds its
// tables for the context's batch policy.
vta_status alma_tables_for_range(vta_ctx *ctx, const vta_alma_batch_range *range, size_t len,
                                 size_t fv, AlmaTables &t);
// Enqueues the sweep on the context's stream with the tables at device
// address d_tables (t.blob's layout); d_out holds n_combos x len floats.
vta_status alma_tables_launch(vta_ctx *ctx, const AlmaTables &t, const void *d_tables,
                              const float *d_prices, size_t len, size_t fv, float *d_out);

// Packed sweeps over existing tables: one job per one-series sweep, its tables
// (t, at device address d_tables, all of one kernel), prices, output, length
// and first_valid.  alma_pack_build fills the kernel's job records and work
// items (4 int32 each, heaviest first) and the launch's dynamic LDS;
// alma_pack_enqueue launches them from device copies of records and items.
struct AlmaPackSrc {
    const AlmaTables *t;
    const unsigned char *d_tables;
    const float *d_prices;
    float *d_out;
    size_t len;
    size_t fv;
};
struct AlmaPackImage {
    vta_alma_batch_kernel kernel = VTA_ALMA_BATCH_FP32;
    std::vector<unsigned char> records;
    std::vector<int32_t> items;
    unsigned n_items = 0;
    size_t smem = 0;
};
vta_status alma_pack_build(vta_ctx *ctx, vta_alma_batch_kernel kernel,
                           const std::vector<AlmaPackSrc> &jobs, AlmaPackImage &img);
vta_status alma_pack_enqueue(vta_ctx *ctx, const AlmaPackImage &img, const void *d_records,
                             const void *d_items);
// A sweep plan's tables and their device copy.
const AlmaTables &alma_plan_tables(const vta_alma_sweep_plan *plan);
const unsigned char *alma_plan_device(const vta_alma_sweep_plan *plan);

// Enqueues the check of the reference's per-combo warm indices against
// first_valid + period - 1 (kernel vta_check_warms); first_valid < 0 derives
// first_valid from combo 0 into *d_fv_out (SWMA's device form has no
// first_valid argument).  A mismatch is reported by the next synchronize.
vta_status check_warms_device(vta_ctx *ctx, const int32_t *d_periods, const int32_t *d_warms,
                              int n, int T, int first_valid, int max_period,
                              int32_t *d_fv_out);

// Many-series launch of SMA / WMA / SWMA (fir.cpp; kind 0 / 1 / 2) over
// caller-prepared first non-finite bars (d_zs, NULL: every bar finite).
vta_status fir_ms_prepared(vta_ctx *ctx, int kind, const float *d_x_tm, const int32_t *d_fvs,
                           const int32_t *d_zs, size_t cols, size_t rows, size_t period,
                           float *d_out_tm);

// Looks up, once per device, the device address of the static scratch array
// of the allocation-free entry point (noalloc.hip); `device` must be current.
// A failure is left for that entry point to report.
void noalloc_setup(int device);

// ALMA grid expansion (alma.cpp), shared with the crossover module.
vta_status alma_expand_range(const vta_alma_batch_range *range,
                             std::vector<vta_alma_params> &out);
// A sweep plan for a grid, independent of the series length (the server's
// per-grid table cache): create once, launch over any long-enough series
// (enqueue only), destroy without a context.
vta_status alma_grid_plan_create(vta_ctx *ctx, const vta_alma_batch_range *range,
                                 vta_alma_sweep_plan **out, size_t *max_period,
                                 size_t *n_combos);
vta_status alma_grid_plan_launch(vta_ctx *ctx, const vta_alma_sweep_plan *plan,
                                 const float *d_prices, size_t series_len, size_t fv,
                                 float *d_out);
void alma_grid_plan_destroy(vta_alma_sweep_plan *plan);

// Packed one-series recurrence batches (rec.cpp), for callers that bring their
// own device memory (the request server's arena): the job table (device
// layout of rec_job in rec.hip) and launch dimensions of a pack, and the
// enqueue-only launch over a table already on the device.
struct PackJob {
    const float *x, *h, *l;
    const int32_t *periods;
    float *out;
    double *carry;
    int T, nlanes, fv, chunk, nch, groups, block0, combo0;
};
struct PackDims {
    int kind = 0, n_jobs = 0;
    unsigned total_blocks = 0, total_combos = 0, flag_blocks = 1;
};
// Validates the jobs like the _device forms and lays out the table; each
// job's carry pointer is left null, its offset (in doubles) in carry_off.
vta_status pack_layout(const vta_ctx *ctx, int kind, const vta_rec_job *jobs, size_t n_jobs,
                       std::vector<PackJob> &tab, std::vector<size_t> &carry_off,
                       size_t *carry_elems, PackDims *dims);
// d_flags: n_jobs int32 words (cleared by the launch itself).
vta_status pack_launch(vta_ctx *ctx, const PackDims &d, const void *d_jobs, int32_t *d_flags);

// Packed one-series BB batches (rolling.cpp): bb_job of rolling.hip.
struct BbPackJob {
    const float *x;
    const int32_t *p_len;
    const double *p_a, *p_b;
    float *out0, *out1, *out2;
    int T, nlanes, fv, z, chunk, groups, block0, pad_;
};
vta_status bb_pack_layout(const vta_ctx *ctx, const vta_bb_job *jobs, size_t n_jobs,
                          std::vector<BbPackJob> &tab, unsigned *total_blocks);
vta_status bb_pack_launch(vta_ctx *ctx, const void *d_jobs, int n_jobs, unsigned total_blocks);

// bf16 round-to-nearest-even of an f32 (bit pattern).
uint16_t bf16_rne(float f);
float bf16_to_f32(uint16_t h);

}  // namespace vta
