// Host entry points for the rolling-window indicators: CudaBollingerBands
// (src/cuda/bollinger_bands_wrapper.rs), CudaPercentileNearestRank
// (src/cuda/percentile_nearest_rank_wrapper.rs) and CudaMediumAd
// (src/cuda/medium_ad_wrapper.rs).  Kernels: kernels/rolling.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

struct roll_args_host {
    const float *x;
    const int32_t *p_len;
    const double *p_a;
    const double *p_b;
    const int32_t *fvs;
    const int32_t *poison;
    float *out0;
    float *out1;
    float *out2;
    double a1;
    double b1;
    int T;
    int nlanes;
    int chunk;
    int wcap;
    int p1;
    int pad_;
};
static_assert(sizeof(roll_args_host) == 112, "roll_args layout");

enum RInd { R_BB = 0, R_PNR = 1, R_MAD = 2 };

// Chunk length: enough chunks for ~8192 waves (BB), or for the PNR / MAD kernels
// (whose sorted windows take smem bytes of LDS per wave) as many waves as
// the CUs hold at once, so the launch runs in one round (PNR, window 100,
// 1,024 x 50,000: 25.9 KB per wave, 6 waves per CU; 400-bar chunks = 2,000
// waves in 1.3 rounds took 4.27 ms, 521-bar chunks = 1,536 waves 2.81 ms;
// these kernels are bound by the latency of each lane's LDS shift chain, so a
// second round costs a whole chunk's time); at least 4x the
// window so the per-chunk rebuild stays a small fraction, at most 4096 bars
// (BB re-anchors its f64 running sums at every chunk start).  `chunked`
// false: one chunk.
int roll_chunk(const vta_ctx *ctx, size_t T, size_t lanes, size_t max_w, bool chunked,
               size_t smem, size_t bb_waves = 8192) {
    if (!chunked) return (int)T;
    if (const char *e = getenv("VTA_ROLL_CHUNK")) {
        long v = atol(e);
        if (v >= 64) return (int)std::min<size_t>((size_t)v, T);
    }
    // Bollinger Bands: bb_waves waves, chunks down to 128 bars.  Read from
    // HBM, the outgoing bars favoured short chunks (1,024 x 50,000, period
    // 20: 391-bar chunks = 2,048 waves 0.291 ms, 256 0.254, 192 0.222, 128
    // 0.216 ms at ~8,192 waves); the many-series kernel's LDS ring of the
    // last p bars, filled once per chunk, favours long ones (128-bar chunks
    // 0.194 ms, 196 0.191, 391 = 2,048 waves 0.182, 520 = 1,552 waves 0.215
    // ms), so its launches ask for ~2,048 waves.
    size_t waves = bb_waves, min_chunk = 128;
    if (smem > 0) {
        const size_t per_cu = std::max<size_t>(
            1, (size_t)ctx->props.maxSharedMemoryPerMultiProcessor / smem);
        waves = std::max<size_t>(256, per_cu * (size_t)ctx->props.multiProcessorCount);
        min_chunk = 256;
    }
    size_t groups = (lanes + 63) / 64;
    size_t want = std::max<size_t>(1, waves / std::max<size_t>(groups, 1));
    size_t c = std::max<size_t>({(T + want - 1) / want, 4 * max_w, min_chunk});
    c = std::min<size_t>(c, 4096);
    return (int)std::min(c, T);
}

const char *kname(RInd k, bool batch) {
    static const char *b[] = {"bb_batch_f32", "pnr_batch_f32", "mad_batch_f32"};
    static const char *m[] = {"bb_ms1p_f32", "pnr_ms1p_f32", "mad_ms1p_f32"};
    return batch ? b[k] : m[k];
}

// Register-window capacity (the window length rounded up to 16 slots) of the
// many-series PNR / MAD kernels, 0 when the LDS form runs: the knob's choice,
// else the register form wherever a kernel exists (PNR up to 128 slots, MAD
// up to 64: a longer MAD window's f64 deviation paths no longer fit the
// VGPRs beside the window).
int reg_window_cap(const vta_ctx *ctx, RInd k, size_t len) {
    if (k == R_BB || ctx->tune[VTA_TUNE_ROLL_FORM] == 1) return 0;
    const size_t cap = (len + 15) / 16 * 16;
    if (cap == 0 || cap > (k == R_PNR ? 128u : 64u)) return 0;
    return (int)cap;
}

// Chunk length of the register-window form.  Its steps are VALU runs, so
// the launch time is (rounds of resident waves) x (chunk length): the chunk
// count per series is a whole number of rounds of the CUs' wave slots (waves
// per SIMD from the kernel's VGPR count), the round count chosen to minimise
// rounds x (chunk + half a window for the chunk's fill by insertions)
// (1,024 x 50,000, PNR window 100 at 2 waves per SIMD: one round, 391-bar
// chunks 0.846 ms; 2.0 rounds of 200 bars 0.90; 1.5 rounds of 260 0.89; 1.2
// rounds of 330 1.09 ms).
int reg_window_chunk(const vta_ctx *ctx, hipFunction_t f, size_t T, size_t lanes, size_t len) {
    if (const char *e = getenv("VTA_ROLL_CHUNK")) {
        long v = atol(e);
        if (v >= 64) return (int)std::min<size_t>((size_t)v, T);
    }
    int regs = 0;
    if (hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f) != hipSuccess || regs <= 0)
        regs = 256;
    const size_t per_simd = std::max<size_t>(1, std::min<size_t>(8, 512 / (size_t)((regs + 7) / 8 * 8)));
    const size_t slots = per_simd * 4 * (size_t)std::max(1, ctx->props.multiProcessorCount);
    const size_t groups = std::max<size_t>(1, (lanes + 63) / 64);
    size_t best = T;
    double best_cost = 0.0;
    for (size_t r = 1; r <= 4; ++r) {
        const size_t per_series = std::max<size_t>(1, r * slots / groups);
        const size_t c = std::max<size_t>((T + per_series - 1) / per_series, 64);
        const size_t chunks = (T + c - 1) / c;
        const size_t rounds = (chunks * groups + slots - 1) / slots;
        const double cost = (double)rounds * ((double)std::min(c, T) + 0.5 * (double)len);
        if (r == 1 || cost < best_cost) {
            best_cost = cost;
            best = c;
        }
    }
    return (int)std::min(best, T);
}

vta_status roll_launch(vta_ctx *ctx, RInd k, bool batch, roll_args_host &a, size_t max_w,
                       bool chunked) {
    hipFunction_t f;
    if (!batch) {
        if (const int cap = reg_window_cap(ctx, k, max_w)) {
            // MAD: one kernel per period from 4 (its ranks and selection
            // paths are compile-time); PNR and MAD periods 1-3: per capacity
            char name[32];
            if (k == R_MAD && max_w >= 4)
                snprintf(name, sizeof name, "mad_ms1p_len%d", (int)max_w);
            else
                snprintf(name, sizeof name, "%s_ms1p_reg%d", k == R_PNR ? "pnr" : "mad", cap);
            VTA_TRY(get_function(ctx, "rolling", name, &f));
            if (k == R_PNR) {
                const vta_kernel_policy &pol = ctx->policy[VTA_POLICY_SLOT_PNR];
                if (pol.many_series == VTA_KERNEL_PLAIN) chunked = false;
            }
            a.chunk = chunked ? reg_window_chunk(ctx, f, (size_t)a.T, (size_t)a.nlanes, max_w) : a.T;
            if (k == R_PNR)
                ctx->sel_ms[VTA_POLICY_SLOT_PNR] = {
                    a.chunk < a.T ? VTA_SELECTED_CHUNKED : VTA_SELECTED_PLAIN, 64};
            a.wcap = (int)max_w;
            const unsigned groups = (unsigned)((a.nlanes + 63) / 64);
            const unsigned chunks = (unsigned)((a.T + a.chunk - 1) / a.chunk);
            return launch(ctx, f, dim3(groups, chunks, 1), dim3(64), 0, a);
        }
    }
    VTA_TRY(get_function(ctx, "rolling", kname(k, batch), &f));
    // many-series BB: the lanes' rings of their last `period` bars (windows up
    // to 256, 64 KiB); wcap 0 reads the outgoing bars from HBM instead
    size_t ring = 0;
    if (k == R_BB && !batch && max_w <= 256) ring = (size_t)64 * max_w * sizeof(float);
    // PNR / MAD: the lanes' sorted windows, plus (batch) the 64 x 33 output staging
    size_t smem = (k == R_BB) ? 0
                              : (size_t)64 * (max_w + 1) * sizeof(float) +
                                    (batch ? (size_t)64 * 33 * sizeof(float) : 0);
    if (k == R_PNR) {
        // CudaPnrPolicy: the plain form runs each lane over the whole series
        const vta_kernel_policy &pol = ctx->policy[VTA_POLICY_SLOT_PNR];
        if ((batch ? pol.batch : pol.many_series) == VTA_KERNEL_PLAIN) chunked = false;
    }
    a.chunk = roll_chunk(ctx, (size_t)a.T, (size_t)a.nlanes, max_w, chunked, smem,
                         ring ? 2048 : 8192);
    if (k == R_PNR) {
        const vta_kernel_selected sel = {a.chunk < a.T ? VTA_SELECTED_CHUNKED : VTA_SELECTED_PLAIN,
                                         64};
        (batch ? ctx->sel_batch : ctx->sel_ms)[VTA_POLICY_SLOT_PNR] = sel;
    }
    a.wcap = k == R_BB ? (ring ? (int)max_w : 0) : (int)max_w;
    unsigned groups = (unsigned)((a.nlanes + 63) / 64);
    unsigned chunks = (unsigned)((a.T + a.chunk - 1) / a.chunk);
    if (smem > ctx->props.sharedMemPerBlock)
        return fail(VTA_ERR_INVALID_INPUT, "window length %zu exceeds the LDS capacity", max_w);
    return launch(ctx, f, dim3(groups, chunks, 1), dim3(64), (unsigned)(smem + ring), a);
}

// First non-finite element at or after fv (n if none), stride `ld` apart.
int32_t first_nonfinite(const float *x, size_t n, size_t fv, size_t ld) {
    for (size_t i = fv; i < n; ++i)
        if (!std::isfinite(x[i * ld])) return (int32_t)i;
    return (int32_t)n;
}

struct Combo {
    size_t len;
    double a, b;
};

// Grids: BB period x devup x devdn (bollinger_bands_wrapper.rs:179-265), PNR
// length x percentage (percentile_nearest_rank_wrapper.rs:200-290), MAD period
// (medium_ad_wrapper.rs:120-150).
vta_status expand_bb(const vta_bb_batch_range *r, std::vector<Combo> &out) {
    if (!r) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    std::vector<size_t> ps;
    std::vector<double> us, ds;
    if (!expand_usize(r->period, AXIS_DOWN, ps) || !expand_f64(r->devup, us) ||
        !expand_f64(r->devdn, ds))
        return fail(VTA_ERR_INVALID_INPUT, "invalid range");
    size_t n;
    if (!mul_ok(ps.size(), us.size(), &n) || !mul_ok(n, ds.size(), &n) || !grid_fits(n))
        return fail(VTA_ERR_INVALID_INPUT, "parameter grid too large");
    out.clear();
    for (size_t p : ps)
        for (double u : us)
            for (double d : ds) out.push_back({p, u, d});
    return VTA_OK;
}

vta_status expand_pnr(const vta_pnr_batch_range *r, std::vector<Combo> &out) {
    if (!r) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    std::vector<size_t> ls;
    std::vector<double> ps;
    if (!expand_usize(r->length, AXIS_DOWN, ls) || !expand_f64(r->percentage, ps))
        return fail(VTA_ERR_INVALID_INPUT, "invalid range");
    size_t n;
    if (!mul_ok(ls.size(), ps.size(), &n) || !grid_fits(n))
        return fail(VTA_ERR_INVALID_INPUT, "parameter grid too large");
    out.clear();
    for (size_t l : ls)
        for (double p : ps) out.push_back({l, p, 0.0});
    return VTA_OK;
}

vta_status expand_mad(const vta_range_usize *r, std::vector<Combo> &out) {
    if (!r) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    std::vector<size_t> ps;
    if (!expand_usize(*r, AXIS_DOWN, ps)) return fail(VTA_ERR_INVALID_INPUT, "invalid range");
    out.clear();
    for (size_t p : ps) out.push_back({p, 0.0, 0.0});
    return VTA_OK;
}

vta_status validate(RInd k, const std::vector<Combo> &cs, size_t len, size_t fv,
                    size_t *max_len) {
    if (cs.empty()) return fail(VTA_ERR_INVALID_INPUT, "no parameter combinations");
    size_t mx = 0;
    for (const auto &c : cs) {
        if (c.len == 0 || c.len > len)
            return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu (data length %zu)", c.len, len);
        if (k == R_PNR && !(c.a >= 0.0 && c.a <= 100.0))
            return fail(VTA_ERR_INVALID_INPUT, "invalid percentage %g", c.a);
        if (c.len > (size_t)INT32_MAX) return fail(VTA_ERR_INVALID_INPUT, "period too large");
        mx = std::max(mx, c.len);
    }
    if (fv >= len) return fail(VTA_ERR_INVALID_INPUT, "all values are NaN");
    if (len - fv < mx)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu, valid = %zu)",
                    mx, len - fv);
    *max_len = mx;
    return VTA_OK;
}

vta_status batch_dev(vta_ctx *ctx, RInd k, const float *x, size_t len,
                     const std::vector<Combo> &cs, int nout, vta_device_array_f32 *out) {
    if (!ctx || !out || !x) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "empty data");
    if (len > (size_t)INT32_MAX) return fail(VTA_ERR_INVALID_INPUT, "series too long");
    size_t fv = first_valid_f32(x, len), mx;
    VTA_TRY(validate(k, cs, len, fv, &mx));
    const size_t C = cs.size();
    size_t n_out;
    if (!mul_ok(C, len, &n_out) || !mul_ok(n_out, sizeof(float), &n_out))
        return fail(VTA_ERR_INVALID_INPUT, "output bytes overflow");
    VTA_TRY(check_fits((size_t)nout * n_out + len * sizeof(float)));
    std::vector<int32_t> pl(C);
    std::vector<double> pa(C), pb(C);
    for (size_t i = 0; i < C; ++i) {
        pl[i] = (int32_t)cs[i].len;
        pa[i] = cs[i].a;
        pb[i] = cs[i].b;
    }
    int32_t meta[2] = {(int32_t)fv, first_nonfinite(x, len, fv, 1)};
    DevBuf dx, dl, da, db, dm, o[3];
    VTA_TRY(upload(ctx, dx, x, len * sizeof(float)));
    VTA_TRY(upload(ctx, dl, pl.data(), C * sizeof(int32_t)));
    VTA_TRY(upload(ctx, da, pa.data(), C * sizeof(double)));
    VTA_TRY(upload(ctx, db, pb.data(), C * sizeof(double)));
    VTA_TRY(upload(ctx, dm, meta, sizeof(meta)));
    for (int i = 0; i < nout; ++i) VTA_TRY(o[i].alloc(n_out));
    roll_args_host a{};
    a.x = dx.as<float>();
    a.p_len = dl.as<int32_t>();
    a.p_a = da.as<double>();
    a.p_b = db.as<double>();
    a.fvs = dm.as<int32_t>();
    a.poison = dm.as<int32_t>() + 1;
    a.out0 = o[0].as<float>();
    a.out1 = o[1].as<float>();
    a.out2 = o[2].as<float>();
    a.T = (int)len;
    a.nlanes = (int)C;
    VTA_TRY(roll_launch(ctx, k, true, a, mx, true));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nout; ++i) {
        out[i].ptr = static_cast<float *>(o[i].release());
        out[i].rows = C;
        out[i].cols = len;
    }
    return VTA_OK;
}

vta_status ms_dev(vta_ctx *ctx, RInd k, const float *x_tm, size_t cols, size_t rows, size_t len,
                  double pa, double pb, int nout, vta_device_array_f32 *out) {
    if (!ctx || !out || !x_tm) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    if (len == 0 || len > rows) return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", len);
    if (k == R_PNR && !(pa >= 0.0 && pa <= 100.0))
        return fail(VTA_ERR_INVALID_INPUT, "invalid percentage %g", pa);
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    size_t elems;
    if (!mul_ok(cols, rows, &elems)) return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    std::vector<int32_t> fvs, zs(cols, (int32_t)rows);
    VTA_TRY(host_first_valids(x_tm, cols, rows, len, fvs));
    if (k == R_BB)
        for (size_t s = 0; s < cols; ++s) zs[s] = first_nonfinite(x_tm + s, rows, fvs[s], cols);
    VTA_TRY(check_fits((size_t)(nout + 1) * elems * sizeof(float)));
    DevBuf dx, dfv, dz, o[3];
    VTA_TRY(upload(ctx, dx, x_tm, elems * sizeof(float)));
    VTA_TRY(upload(ctx, dfv, fvs.data(), cols * sizeof(int32_t)));
    VTA_TRY(upload(ctx, dz, zs.data(), cols * sizeof(int32_t)));
    for (int i = 0; i < nout; ++i) VTA_TRY(o[i].alloc(elems * sizeof(float)));
    roll_args_host a{};
    a.x = dx.as<float>();
    a.p1 = (int)len;
    a.a1 = pa;
    a.b1 = pb;
    a.fvs = dfv.as<int32_t>();
    a.poison = dz.as<int32_t>();
    a.out0 = o[0].as<float>();
    a.out1 = o[1].as<float>();
    a.out2 = o[2].as<float>();
    a.T = (int)rows;
    a.nlanes = (int)cols;
    VTA_TRY(roll_launch(ctx, k, false, a, len, true));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nout; ++i) {
        out[i].ptr = static_cast<float *>(o[i].release());
        out[i].rows = rows;
        out[i].cols = cols;
    }
    return VTA_OK;
}

// Enqueue-only many-series launch of PNR / MAD on device-resident data.
vta_status ms_device(vta_ctx *ctx, RInd k, const float *d_x, const int32_t *d_fv, size_t cols,
                     size_t rows, size_t len, double pa, float *d_out) {
    if (!ctx || !d_x || !d_fv || !d_out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    if (len == 0 || len > rows) return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", len);
    if (k == R_PNR && !(pa >= 0.0 && pa <= 100.0))
        return fail(VTA_ERR_INVALID_INPUT, "invalid percentage %g", pa);
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    roll_args_host a{};
    a.x = d_x;
    a.p1 = (int)len;
    a.a1 = pa;
    a.fvs = d_fv;
    a.out0 = d_out;
    a.T = (int)rows;
    a.nlanes = (int)cols;
    return roll_launch(ctx, k, false, a, len, true);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ Bollinger Bands
vta_status vta_bb_expand_grid(const vta_bb_batch_range *range, vta_bb_params *out, size_t cap,
                              size_t *n_out) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_bb(range, cs));
    if (n_out) *n_out = cs.size();
    if (out)
        for (size_t i = 0; i < std::min(cap, cs.size()); ++i)
            out[i] = {cs[i].len, cs[i].a, cs[i].b};
    return VTA_OK;
}

vta_status vta_bollinger_bands_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                                         const vta_bb_batch_range *range,
                                         vta_device_array_f32 *out3) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_bb(range, cs));
    return batch_dev(ctx, R_BB, prices, len, cs, 3, out3);
}

vta_status vta_bollinger_bands_batch_device(vta_ctx *ctx, const float *d_prices,
                                            int32_t series_len, int32_t first_valid,
                                            int32_t first_nonfinite, const int32_t *d_periods,
                                            const double *d_devup, const double *d_devdn,
                                            int32_t n_combos, int32_t max_period,
                                            float *d_upper, float *d_middle, float *d_lower) {
    clear_error();
    if (!ctx || !d_prices || !d_periods || !d_devup || !d_devdn || !d_upper || !d_middle ||
        !d_lower)
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (series_len <= 0 || n_combos <= 0 || max_period <= 0 || first_valid < 0 ||
        first_valid >= series_len || first_nonfinite < first_valid ||
        first_nonfinite > series_len || series_len - first_valid < max_period)
        return fail(VTA_ERR_INVALID_INPUT, "invalid sizes");
    // per-call metadata {first_valid, first_nonfinite} in the context scratch,
    // written by a fill kernel in stream order
    int32_t *meta = static_cast<int32_t *>(ctx->scratch) + 128;
    hipFunction_t fill;
    VTA_TRY(get_function(ctx, "rec", "vta_fill_i32", &fill));
    VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, meta, 1, (int)first_valid));
    VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, meta + 1, 1, (int)first_nonfinite));
    roll_args_host a{};
    a.x = d_prices;
    a.p_len = d_periods;
    a.p_a = d_devup;
    a.p_b = d_devdn;
    a.fvs = meta;
    a.poison = meta + 1;
    a.out0 = d_upper;
    a.out1 = d_middle;
    a.out2 = d_lower;
    a.T = series_len;
    a.nlanes = n_combos;
    return roll_launch(ctx, R_BB, true, a, (size_t)max_period, true);
}

vta_status vta_bollinger_bands_many_series_one_param_time_major_dev(
    vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period, double devup,
    double devdn, vta_device_array_f32 *out3) {
    clear_error();
    return ms_dev(ctx, R_BB, data_tm, cols, rows, period, devup, devdn, 3, out3);
}

vta_status vta_bollinger_bands_many_series_one_param_device(
    vta_ctx *ctx, const float *d_data_tm, const int32_t *d_first_valids,
    const int32_t *d_first_nonfinite,
    size_t cols, size_t rows, size_t period, double devup, double devdn, float *d_upper_tm,
    float *d_middle_tm, float *d_lower_tm) {
    clear_error();
    if (!ctx || !d_data_tm || !d_first_valids || !d_upper_tm || !d_middle_tm || !d_lower_tm ||
        cols == 0 || rows == 0 || period == 0 || period > rows)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    roll_args_host a{};
    a.x = d_data_tm;
    a.p1 = (int)period;
    a.a1 = devup;
    a.b1 = devdn;
    a.fvs = d_first_valids;
    a.poison = d_first_nonfinite;
    a.out0 = d_upper_tm;
    a.out1 = d_middle_tm;
    a.out2 = d_lower_tm;
    a.T = (int)rows;
    a.nlanes = (int)cols;
    return roll_launch(ctx, R_BB, false, a, period, d_first_nonfinite != nullptr);
}

// ------------------------------------------------------------------ percentile nearest rank
vta_status vta_pnr_expand_grid(const vta_pnr_batch_range *range, vta_pnr_params *out,
                               size_t cap, size_t *n_out) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_pnr(range, cs));
    if (n_out) *n_out = cs.size();
    if (out)
        for (size_t i = 0; i < std::min(cap, cs.size()); ++i) out[i] = {cs[i].len, cs[i].a};
    return VTA_OK;
}

vta_status vta_pnr_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_pnr_batch_range *range, vta_device_array_f32 *out) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_pnr(range, cs));
    return batch_dev(ctx, R_PNR, prices, len, cs, 1, out);
}

vta_status vta_pnr_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *prices_tm,
                                                        size_t cols, size_t rows, size_t length,
                                                        double percentage,
                                                        vta_device_array_f32 *out) {
    clear_error();
    return ms_dev(ctx, R_PNR, prices_tm, cols, rows, length, percentage, 0.0, 1, out);
}

vta_status vta_pnr_many_series_one_param_device(vta_ctx *ctx, const float *d_data_tm,
                                                const int32_t *d_first_valids, size_t cols,
                                                size_t rows, size_t length, double percentage,
                                                float *d_out_tm) {
    clear_error();
    return ms_device(ctx, R_PNR, d_data_tm, d_first_valids, cols, rows, length, percentage,
                     d_out_tm);
}

// ------------------------------------------------------------------ rolling MAD
vta_status vta_medium_ad_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                                     size_t *n_out) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_mad(range, cs));
    if (n_out) *n_out = cs.size();
    if (periods)
        for (size_t i = 0; i < std::min(cap, cs.size()); ++i) periods[i] = cs[i].len;
    return VTA_OK;
}

vta_status vta_medium_ad_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                                   const vta_range_usize *range, vta_device_array_f32 *out) {
    clear_error();
    std::vector<Combo> cs;
    VTA_TRY(expand_mad(range, cs));
    return batch_dev(ctx, R_MAD, prices, len, cs, 1, out);
}

vta_status vta_medium_ad_many_series_one_param_time_major_dev(vta_ctx *ctx,
                                                              const float *data_tm, size_t cols,
                                                              size_t rows, size_t period,
                                                              vta_device_array_f32 *out) {
    clear_error();
    return ms_dev(ctx, R_MAD, data_tm, cols, rows, period, 0.0, 0.0, 1, out);
}

vta_status vta_medium_ad_many_series_one_param_device(vta_ctx *ctx, const float *d_data_tm,
                                                      const int32_t *d_first_valids, size_t cols,
                                                      size_t rows, size_t period,
                                                      float *d_out_tm) {
    clear_error();
    return ms_device(ctx, R_MAD, d_data_tm, d_first_valids, cols, rows, period, 0.0, d_out_tm);
}

}  // extern "C"

// ---------------------------------------------------------------- packed BB batches

static_assert(sizeof(BbPackJob) == 88, "bb_job layout");

namespace vta {

vta_status bb_pack_layout(const vta_ctx *ctx, const vta_bb_job *jobs, size_t n_jobs,
                          std::vector<BbPackJob> &tab, unsigned *total_blocks) {
    if (!ctx || !jobs || !total_blocks) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (n_jobs == 0 || n_jobs > 65535)
        return fail(VTA_ERR_INVALID_INPUT, "n_jobs must be in 1..65535 (got %zu)", n_jobs);
    tab.assign(n_jobs, BbPackJob{});
    size_t blocks = 0;
    for (size_t j = 0; j < n_jobs; ++j) {
        const vta_bb_job &q = jobs[j];
        // the checks of vta_bollinger_bands_batch_device
        if (!q.d_prices || !q.d_periods || !q.d_devup || !q.d_devdn || !q.d_upper ||
            !q.d_middle || !q.d_lower)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: null pointer", j);
        if (q.series_len <= 0 || q.n_combos <= 0 || q.max_period <= 0 || q.first_valid < 0 ||
            q.first_valid >= q.series_len || q.first_nonfinite < q.first_valid ||
            q.first_nonfinite > q.series_len || q.series_len - q.first_valid < q.max_period)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: invalid sizes", j);
        const size_t T = (size_t)q.series_len, C = (size_t)q.n_combos;
        const size_t chunk = (size_t)roll_chunk(ctx, T, C, (size_t)q.max_period, true, 0);
        const size_t nch = (T + chunk - 1) / chunk;
        const size_t groups = (C + 63) / 64;
        BbPackJob &e = tab[j];
        e.x = q.d_prices;
        e.p_len = q.d_periods;
        e.p_a = q.d_devup;
        e.p_b = q.d_devdn;
        e.out0 = q.d_upper;
        e.out1 = q.d_middle;
        e.out2 = q.d_lower;
        e.T = (int)T;
        e.nlanes = (int)C;
        e.fv = q.first_valid;
        e.z = q.first_nonfinite;
        e.chunk = (int)chunk;
        e.groups = (int)groups;
        e.block0 = (int)blocks;
        blocks += groups * nch;
        if (blocks > (size_t)INT32_MAX)
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "packed grid exceeds 2^31 blocks");
    }
    *total_blocks = (unsigned)blocks;
    return VTA_OK;
}

vta_status bb_pack_launch(vta_ctx *ctx, const void *d_jobs, int n_jobs, unsigned total_blocks) {
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rolling", "bb_packed_f32", &f));
    return launch(ctx, f, dim3(total_blocks), dim3(64), 0, d_jobs, n_jobs);
}

}  // namespace vta

struct vta_bb_pack {
    int device = 0;
    int n_jobs = 0;
    unsigned total_blocks = 0;
    DevBuf jobs;
};

extern "C" {

vta_status vta_bb_pack_create(vta_ctx *ctx, const vta_bb_job *jobs, size_t n_jobs,
                              vta_bb_pack **out) {
    clear_error();
    if (!ctx || !jobs || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = nullptr;
    std::vector<BbPackJob> tab;
    auto pk = std::make_unique<vta_bb_pack>();
    VTA_TRY(bb_pack_layout(ctx, jobs, n_jobs, tab, &pk->total_blocks));
    pk->device = ctx->device;
    pk->n_jobs = (int)n_jobs;
    VTA_HIP(hipSetDevice(ctx->device));
    VTA_TRY(upload(ctx, pk->jobs, tab.data(), n_jobs * sizeof(BbPackJob)));
    *out = pk.release();
    return VTA_OK;
}

vta_status vta_bb_pack_launch(vta_ctx *ctx, const vta_bb_pack *pk) {
    clear_error();
    if (!ctx || !pk) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (pk->device != ctx->device)
        return fail(VTA_ERR_DEVICE_MISMATCH, "pack built on device %d, context on %d", pk->device,
                    ctx->device);
    return bb_pack_launch(ctx, pk->jobs.p, pk->n_jobs, pk->total_blocks);
}

vta_status vta_bb_pack_destroy(vta_bb_pack *pk) {
    if (!pk) return VTA_OK;
    (void)hipSetDevice(pk->device);
    delete pk;
    return VTA_OK;
}

}  // extern "C"
