// Host entry points for the finite-window moving averages: CudaSma
// (src/cuda/moving_averages/sma_wrapper.rs), CudaWma (wma_wrapper.rs) and
// CudaSwma (swma_wrapper.rs).  Kernels: kernels/fir.hip.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

// Host mirror of struct fir_args in kernels/fir.hip (same field order/types).
struct fir_args_host {
    const float *x;
    const int32_t *periods;
    const int32_t *meta;
    const int32_t *fvs;
    const int32_t *zs;
    float *out;
    int T;
    int nlanes;
    int kind;
    int tile;
    int span_cap;
    int chunk;
    int period;
    int group;
    int ring;
    int pad_;
};
static_assert(sizeof(fir_args_host) == 88, "fir_args layout");

enum Fir { FIR_SMA = 0, FIR_WMA = 1, FIR_SWMA = 2 };

// Combos per workgroup of the batch kernel (the context's VTA_TUNE_FIR_GROUP
// knob, else VTA_FIR_GROUP, 1..256).
// Measured on 1,250 periods x 200k bars: 64 per workgroup amortise the tile's
// prefix scan best (SMA 0.24 ms vs 0.30 ms at 16).  Lane groups of four bars
// with 16-byte stores were 5-15% slower than one 4-byte store per lane (their
// strided LDS reads conflict) and were dropped.
unsigned fir_group(const vta_ctx *ctx) {
    if (ctx->tune[VTA_TUNE_FIR_GROUP] > 0) return (unsigned)ctx->tune[VTA_TUNE_FIR_GROUP];
    if (const char *e = getenv("VTA_FIR_GROUP")) {
        long v = atol(e);
        if (v >= 1 && v <= 256) return (unsigned)v;
    }
    return 64;
}

// Grids: SMA sma.rs:705-755 (expand_grid_sma), WMA wma_wrapper.rs:324-350
// (countdown plus end), SWMA swma_wrapper.rs:275-301.
vta_status expand(Fir k, const vta_range_usize *r, std::vector<size_t> &ps) {
    if (!r) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    if (!expand_usize(*r, k == FIR_WMA ? AXIS_DOWN_END : AXIS_DOWN, ps) || ps.empty())
        return fail(VTA_ERR_INVALID_INPUT, "invalid range: start=%zu end=%zu step=%zu", r->start,
                    r->end, r->step);
    return VTA_OK;
}

// Period validation of the wrappers' prepare_batch_inputs: SMA and SWMA need
// period >= 1, WMA period > 1 (wma_wrapper.rs:379-384); all need period <= len
// and len - first_valid >= period.
size_t min_period(Fir k) { return k == FIR_WMA ? 2 : 1; }

vta_status check_period(Fir k, size_t p, size_t len, size_t fv) {
    if (p < min_period(k))
        return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu (must be >= %zu)", p,
                    min_period(k));
    if (p > len) return fail(VTA_ERR_INVALID_INPUT, "period %zu exceeds data length %zu", p, len);
    if (len - fv < p)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data: needed >= %zu, valid = %zu", p,
                    len - fv);
    return VTA_OK;
}

// Batch launch: tile of 1024 bars unless the prefix arrays of the tile plus
// the longest period would leave fewer than two workgroups per CU (then 512,
// 256); the LDS holds (tile + max_p + 2) doubles per prefix array.
vta_status batch_launch(vta_ctx *ctx, Fir k, const float *d_x, const int32_t *d_periods,
                        const int32_t *d_meta, size_t T, size_t C, size_t max_p, float *d_out) {
    const size_t narr = k == FIR_SMA ? 1 : 2;
    const size_t limit = ctx->props.sharedMemPerBlock;
    size_t tile = 1024, smem = 0;
    for (;; tile >>= 1) {
        smem = narr * (tile + max_p + 3) * sizeof(double);
        if (smem <= 64 * 1024 || tile == 256) break;
    }
    if (smem > limit)
        return fail(VTA_ERR_INVALID_INPUT,
                    "period %zu exceeds the LDS window of one workgroup (%zu bytes)", max_p,
                    limit);
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "fir", "fir_batch_f32", &f));
    fir_args_host a{};
    a.x = d_x;
    a.periods = d_periods;
    a.meta = d_meta;
    a.out = d_out;
    a.T = (int)T;
    a.nlanes = (int)C;
    a.kind = (int)k;
    a.tile = (int)tile;
    a.span_cap = (int)(tile + max_p + 1);
    const unsigned group = fir_group(ctx);
    a.group = (int)group;
    unsigned gx = (unsigned)((T + tile - 1) / tile);
    unsigned gy = (unsigned)((C + group - 1) / group);
    if (gy > (unsigned)ctx->props.maxGridSize[1])
        return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "too many combos for one launch (%zu)", C);
    return launch(ctx, f, dim3(gx, gy, 1), dim3(256), (unsigned)smem, a);
}

// Many-series chunk: at least 12x the replayed look-back (p bars) and 256
// bars, small enough for ~8192 waves (32 per CU: each wave's bars are a
// dependent chain, so latency is hidden by many waves in flight).  4,096 x
// 50,000 at period 64: 0.52 ms with 448-bar chunks (8,192 waves), 0.44 ms
// with ~830 (4,096); periods 14 / 30 keep their 448-bar chunks.
// Rows of the many-series LDS ring at most (24 KiB per wave); longer periods
// read the lagged bars from global memory.  VTA_FIR_NO_RING=1 turns the ring
// off.
constexpr size_t kRingMaxRows = 96;
constexpr size_t kSwmaMsBlock = 16;   // bars per block of fir_ms1p_swma_f32

bool ring_disabled() {
    static const bool off = [] {
        const char *e = getenv("VTA_FIR_NO_RING");
        return e && e[0] == '1';
    }();
    return off;
}

int pick_chunk(size_t p, size_t T, size_t groups) {
    size_t want = std::max<size_t>(1, 8192 / std::max<size_t>(groups, 1));
    size_t c = std::max<size_t>({12 * p, 256, (T + want - 1) / want});
    c = (c + 63) & ~(size_t)63;
    return (int)std::min(c, T);
}

vta_status ms_launch(vta_ctx *ctx, Fir k, const float *d_x, const int32_t *d_fvs,
                     const int32_t *d_zs, size_t cols, size_t rows, size_t period, float *d_out) {
    hipFunction_t f;
    static const char *const kNames[] = {"fir_ms1p_sma_f32", "fir_ms1p_wma_f32",
                                         "fir_ms1p_swma_f32"};
    VTA_TRY(get_function(ctx, "fir", kNames[k], &f));
    fir_args_host a{};
    a.x = d_x;
    a.fvs = d_fvs;
    a.zs = d_zs;
    a.out = d_out;
    a.T = (int)rows;
    a.nlanes = (int)cols;
    a.kind = (int)k;
    a.period = (int)period;
    unsigned groups = (unsigned)((cols + 63) / 64);
    a.chunk = pick_chunk(period, rows, groups);
    unsigned chunks = (unsigned)((rows + a.chunk - 1) / a.chunk);
    // SWMA's three lagged streams come from a per-wave LDS ring of the last
    // bars (the longest lag is at most the period, plus the block being
    // read).  SMA and WMA have one lagged stream each: re-reading it through
    // the cache measured as fast as the ring at the occupancy the ring's LDS
    // allows, so they keep the cache.
    a.ring = (k == FIR_SWMA && period > 2 && period + kSwmaMsBlock <= kRingMaxRows)
                 ? (int)(period + kSwmaMsBlock) : 0;
    if (ring_disabled()) a.ring = 0;
    const unsigned smem = (unsigned)a.ring * 64u * (unsigned)sizeof(float);
    return launch(ctx, f, dim3(groups, chunks, 1), dim3(64), smem, a);
}

int32_t first_nonfinite(const float *x, size_t n, size_t fv, size_t ld) {
    for (size_t i = fv; i < n; ++i)
        if (!std::isfinite(x[i * ld])) return (int32_t)i;
    return (int32_t)n;
}

vta_status batch_dev(vta_ctx *ctx, Fir k, const float *x, size_t len,
                     const vta_range_usize *range, vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !x || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "empty data");
    if (len > (size_t)INT32_MAX) return fail(VTA_ERR_INVALID_INPUT, "series too long");
    size_t fv = first_valid_f32(x, len);
    if (fv >= len) return fail(VTA_ERR_INVALID_INPUT, "all values are NaN");
    std::vector<size_t> ps;
    VTA_TRY(expand(k, range, ps));
    for (size_t p : ps) VTA_TRY(check_period(k, p, len, fv));
    const size_t C = ps.size();
    const size_t max_p = *std::max_element(ps.begin(), ps.end());
    size_t n_out;
    if (!mul_ok(C, len, &n_out) || !mul_ok(n_out, sizeof(float), &n_out))
        return fail(VTA_ERR_INVALID_INPUT, "output bytes overflow");
    VTA_TRY(check_fits(n_out + len * sizeof(float)));
    std::vector<int32_t> per(ps.begin(), ps.end());
    int32_t meta[2] = {(int32_t)fv, first_nonfinite(x, len, fv, 1)};
    DevBuf dx, dp, dm, dout;
    VTA_TRY(upload(ctx, dx, x, len * sizeof(float)));
    VTA_TRY(upload(ctx, dp, per.data(), C * sizeof(int32_t)));
    VTA_TRY(upload(ctx, dm, meta, sizeof(meta)));
    VTA_TRY(dout.alloc(n_out));
    VTA_TRY(batch_launch(ctx, k, dx.as<float>(), dp.as<int32_t>(), dm.as<int32_t>(), len, C,
                         max_p, dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = C;
    out->cols = len;
    return VTA_OK;
}

// Device-resident one-series batch, enqueue only: first_valid from the
// caller, the first non-finite bar found on the device into the context
// scratch.  With d_warms (SWMA's device form, which passes warm indices
// first_valid + period - 1 and no first_valid) first_valid is derived from
// combo 0 on the device and every warm index checked against it.
vta_status batch_device(vta_ctx *ctx, Fir k, const float *d_x, const int32_t *d_periods,
                        int64_t series_len, int64_t n_combos, int64_t first_valid,
                        int64_t max_period, float *d_out, const int32_t *d_warms = nullptr) {
    clear_error();
    if (!ctx || !d_x || !d_periods || !d_out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (d_warms) first_valid = 0;   // derived on the device
    if (series_len <= 0 || n_combos <= 0 || first_valid < 0 || first_valid >= series_len ||
        max_period < (int64_t)min_period(k) || max_period > series_len - first_valid ||
        series_len > INT32_MAX || n_combos > INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT,
                    "series_len, n_combos must be positive, first_valid < series_len and "
                    "%zu <= max_period <= series_len - first_valid", min_period(k));
    int32_t *meta = static_cast<int32_t *>(ctx->scratch) + 192;
    hipFunction_t fill;
    VTA_TRY(get_function(ctx, "rec", "vta_fill_i32", &fill));
    VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, meta + 1, 1, (int)series_len));
    if (d_warms) {
        VTA_TRY(check_warms_device(ctx, d_periods, d_warms, (int)n_combos, (int)series_len, -1,
                                   (int)max_period, meta));
        hipFunction_t fnf;
        VTA_TRY(get_function(ctx, "fir", "series_first_nonfinite_dfv", &fnf));
        VTA_TRY(launch(ctx, fnf, dim3((unsigned)((series_len + 255) / 256)), dim3(256), 0, d_x,
                       (int)series_len, (const int32_t *)meta, meta + 1));
    } else {
        hipFunction_t fnf;
        VTA_TRY(get_function(ctx, "fir", "series_first_nonfinite", &fnf));
        VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, meta, 1, (int)first_valid));
        unsigned nb = (unsigned)((series_len - first_valid + 255) / 256);
        VTA_TRY(launch(ctx, fnf, dim3(nb), dim3(256), 0, d_x, (int)series_len, (int)first_valid,
                       meta + 1));
    }
    return batch_launch(ctx, k, d_x, d_periods, meta, (size_t)series_len, (size_t)n_combos,
                        (size_t)max_period, d_out);
}

// sma_batch_dev_from_device (sma_wrapper.rs:830-860): device prices, host
// combos, device-resident output.
vta_status batch_dev_from_device(vta_ctx *ctx, Fir k, const float *d_x, size_t len,
                                 const size_t *periods, size_t n_combos, size_t first_valid,
                                 vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !d_x || !periods || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0 || n_combos == 0 || first_valid >= len || len > (size_t)INT32_MAX ||
        n_combos > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "len and n_combos must be positive, first_valid < len");
    size_t max_p = 0;
    for (size_t i = 0; i < n_combos; ++i) {
        VTA_TRY(check_period(k, periods[i], len, first_valid));
        max_p = std::max(max_p, periods[i]);
    }
    size_t n_out;
    if (!mul_ok(n_combos, len, &n_out) || !mul_ok(n_out, sizeof(float), &n_out))
        return fail(VTA_ERR_INVALID_INPUT, "output bytes overflow");
    VTA_TRY(check_fits(n_out));
    std::vector<int32_t> per(periods, periods + n_combos);
    DevBuf dp, dout;
    VTA_TRY(upload(ctx, dp, per.data(), n_combos * sizeof(int32_t)));
    VTA_TRY(dout.alloc(n_out));
    VTA_TRY(batch_device(ctx, k, d_x, dp.as<int32_t>(), (int64_t)len, (int64_t)n_combos,
                         (int64_t)first_valid, (int64_t)max_p, dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = n_combos;
    out->cols = len;
    return VTA_OK;
}

vta_status ms_dev(vta_ctx *ctx, Fir k, const float *x_tm, size_t cols, size_t rows,
                  size_t period, vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !x_tm || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    if (period < min_period(k) || period > rows)
        return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", period);
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    size_t elems;
    if (!mul_ok(cols, rows, &elems)) return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    std::vector<int32_t> fvs, zs(cols);
    VTA_TRY(host_first_valids(x_tm, cols, rows, period, fvs));
    for (size_t s = 0; s < cols; ++s) zs[s] = first_nonfinite(x_tm + s, rows, fvs[s], cols);
    VTA_TRY(check_fits(2 * elems * sizeof(float)));
    DevBuf dx, dfv, dz, dout;
    VTA_TRY(upload(ctx, dx, x_tm, elems * sizeof(float)));
    VTA_TRY(upload(ctx, dfv, fvs.data(), cols * sizeof(int32_t)));
    VTA_TRY(upload(ctx, dz, zs.data(), cols * sizeof(int32_t)));
    VTA_TRY(dout.alloc(elems * sizeof(float)));
    VTA_TRY(ms_launch(ctx, k, dx.as<float>(), dfv.as<int32_t>(), dz.as<int32_t>(), cols, rows,
                      period, dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = rows;
    out->cols = cols;
    return VTA_OK;
}

// Many-series device form: the first non-finite bar of each series (from the
// caller's first_valid) is found on the device, into the context's carry
// buffer (stream-ordered with its other uses).
vta_status ms_device(vta_ctx *ctx, Fir k, const float *d_x_tm, const int32_t *d_fvs,
                     int64_t cols, int64_t rows, int64_t period, float *d_out_tm) {
    clear_error();
    if (!ctx || !d_x_tm || !d_fvs || !d_out_tm || cols <= 0 || rows <= 0 ||
        period < (int64_t)min_period(k) || period > rows)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    if (rows > INT32_MAX || cols > INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    if ((size_t)cols * sizeof(int32_t) > vta_ctx::kCarryBytes)
        return fail(VTA_ERR_INVALID_INPUT, "more than %zu series in one launch",
                    vta_ctx::kCarryBytes / sizeof(int32_t));
    int32_t *zs = static_cast<int32_t *>(ctx->carry);
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rec", "ms_first_nonfinite_from_first_valids", &f));
    VTA_TRY(launch(ctx, f, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, d_x_tm, (int)rows,
                   (int)cols, d_fvs, zs));
    return ms_launch(ctx, k, d_x_tm, d_fvs, zs, (size_t)cols, (size_t)rows, (size_t)period,
                     d_out_tm);
}

// sma_multi_series_one_param_time_major_dev_from_device (sma_wrapper.rs:982-1000).
vta_status ms_dev_from_device(vta_ctx *ctx, Fir k, const float *d_x_tm, const int32_t *d_fvs,
                              size_t cols, size_t rows, size_t period,
                              vta_device_array_f32 *out) {
    clear_error();
    if (!out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    size_t elems;
    if (!mul_ok(cols, rows, &elems) || !mul_ok(elems, sizeof(float), &elems))
        return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    if (cols > (size_t)INT32_MAX || rows > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    VTA_TRY(check_fits(elems));
    DevBuf dout;
    VTA_TRY(dout.alloc(std::max<size_t>(elems, 4)));
    VTA_TRY(ms_device(ctx, k, d_x_tm, d_fvs, (int64_t)cols, (int64_t)rows, (int64_t)period,
                      dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = rows;
    out->cols = cols;
    return VTA_OK;
}

vta_status expand_out(Fir k, const vta_range_usize *range, size_t *out, size_t cap,
                      size_t *n_out) {
    clear_error();
    std::vector<size_t> ps;
    VTA_TRY(expand(k, range, ps));
    if (n_out) *n_out = ps.size();
    if (out) memcpy(out, ps.data(), std::min(cap, ps.size()) * sizeof(size_t));
    return VTA_OK;
}

}  // namespace

namespace vta {

vta_status fir_ms_prepared(vta_ctx *ctx, int kind, const float *d_x_tm, const int32_t *d_fvs,
                           const int32_t *d_zs, size_t cols, size_t rows, size_t period,
                           float *d_out_tm) {
    const Fir k = (Fir)kind;
    if (!ctx || !d_x_tm || !d_fvs || !d_out_tm || cols == 0 || rows == 0 ||
        period < min_period(k) || period > rows)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    return ms_launch(ctx, k, d_x_tm, d_fvs, d_zs, cols, rows, period, d_out_tm);
}

}  // namespace vta

extern "C" {

#define VTA_FIR_ENTRY_POINTS(NAME, KIND)                                                        \
    vta_status vta_##NAME##_expand_grid(const vta_range_usize *range, size_t *periods,        \
                                        size_t cap, size_t *n_out) {                          \
        return expand_out(KIND, range, periods, cap, n_out);                                  \
    }                                                                                         \
    vta_status vta_##NAME##_batch_dev(vta_ctx *ctx, const float *prices, size_t len,          \
                                      const vta_range_usize *range,                           \
                                      vta_device_array_f32 *out) {                            \
        return batch_dev(ctx, KIND, prices, len, range, out);                                 \
    }                                                                                         \
    vta_status vta_##NAME##_many_series_one_param_time_major_dev(                             \
        vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period,          \
        vta_device_array_f32 *out) {                                                          \
        return ms_dev(ctx, KIND, data_tm, cols, rows, period, out);                           \
    }

VTA_FIR_ENTRY_POINTS(sma, FIR_SMA)
VTA_FIR_ENTRY_POINTS(wma, FIR_WMA)
VTA_FIR_ENTRY_POINTS(swma, FIR_SWMA)
#undef VTA_FIR_ENTRY_POINTS

// ---- SMA: enqueue-only forms (no reference counterpart) and the reference's
// device-input forms
vta_status vta_sma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                int32_t series_len, int32_t n_combos, int32_t first_valid,
                                int32_t max_period, float *d_out) {
    return batch_device(ctx, FIR_SMA, d_prices, d_periods, series_len, n_combos, first_valid,
                        max_period, d_out);
}
vta_status vta_sma_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                const int32_t *d_first_valids, size_t cols,
                                                size_t rows, size_t period, float *d_out_tm) {
    if (cols > (size_t)INT32_MAX || rows > (size_t)INT32_MAX || period > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    return ms_device(ctx, FIR_SMA, d_prices_tm, d_first_valids, (int64_t)cols, (int64_t)rows,
                     (int64_t)period, d_out_tm);
}
vta_status vta_sma_batch_dev_from_device(vta_ctx *ctx, const float *d_prices, size_t len,
                                         const size_t *periods, size_t n_combos,
                                         size_t first_valid, vta_device_array_f32 *out) {
    return batch_dev_from_device(ctx, FIR_SMA, d_prices, len, periods, n_combos, first_valid,
                                 out);
}
vta_status vta_sma_multi_series_one_param_time_major_dev_from_device(
    vta_ctx *ctx, const float *d_prices_tm, const int32_t *d_first_valids, size_t cols,
    size_t rows, size_t period, vta_device_array_f32 *out) {
    return ms_dev_from_device(ctx, FIR_SMA, d_prices_tm, d_first_valids, cols, rows, period,
                              out);
}

// ---- WMA: wma_batch_device (wma_wrapper.rs:747-790),
// wma_multi_series_one_param_time_major_device (:1049-1080)
vta_status vta_wma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                int32_t series_len, int32_t n_combos, int32_t first_valid,
                                int32_t max_period, float *d_out) {
    return batch_device(ctx, FIR_WMA, d_prices, d_periods, series_len, n_combos, first_valid,
                        max_period, d_out);
}
vta_status vta_wma_multi_series_one_param_time_major_device(vta_ctx *ctx,
                                                            const float *d_prices_tm,
                                                            int32_t period, int32_t num_series,
                                                            int32_t series_len,
                                                            const int32_t *d_first_valids,
                                                            float *d_out_tm) {
    return ms_device(ctx, FIR_WMA, d_prices_tm, d_first_valids, num_series, series_len, period,
                     d_out_tm);
}

// ---- SWMA: swma_batch_device (swma_wrapper.rs:663-676): first_valid derived
// from the warm indices on the device; swma_multi_series_one_param_device
// (:787-830): the triangular weights are derived from the period (the
// reference's host code uploads the same normalised weights), d_weights is
// accepted for the binding and not read.
vta_status vta_swma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                 const int32_t *d_warms, size_t series_len, size_t n_combos,
                                 size_t max_period, float *d_out) {
    if (!d_warms) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (series_len > (size_t)INT32_MAX || n_combos > (size_t)INT32_MAX ||
        max_period > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "series_len, n_combos or max_period out of range");
    return batch_device(ctx, FIR_SWMA, d_prices, d_periods, (int64_t)series_len,
                        (int64_t)n_combos, 0, (int64_t)max_period, d_out, d_warms);
}
vta_status vta_swma_multi_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                  const float *d_weights, int32_t period,
                                                  int32_t num_series, int32_t series_len,
                                                  const int32_t *d_first_valids,
                                                  float *d_out_tm) {
    (void)d_weights;
    return ms_device(ctx, FIR_SWMA, d_prices_tm, d_first_valids, num_series, series_len, period,
                     d_out_tm);
}

}  // extern "C"
