// Host entry points for the recurrence indicators: CudaEma
// (src/cuda/moving_averages/ema_wrapper.rs), CudaRsi (src/cuda/rsi_wrapper.rs),
// CudaAtr (src/cuda/atr_wrapper.rs), CudaAdx (src/cuda/adx_wrapper.rs), plus the
// fused EMA + RSI + ATR many-series launch.  Kernels: kernels/rec.hip.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

// Host mirror of struct rec_args in kernels/rec.hip (same field order/types).
struct rec_args_host {
    const float *x;
    const float *h;
    const float *l;
    const int32_t *periods;
    const int32_t *fvs;
    const int32_t *clean;
    const int32_t *fvs2;
    float *out0;
    float *out1;
    float *out2;
    double *carry;
    const float *alphas;
    double *seeds;
    int T;
    int nlanes;
    int chunk;
    int p1, p2, p3;
    int phase;
    int clean_epoch;
};
static_assert(sizeof(rec_args_host) == 136, "rec_args layout");

enum Ind {
    IND_EMA = 0,
    IND_RSI = 1,
    IND_ATR = 2,
    IND_ADX = 3,
    IND_SMMA = 4,
    IND_WILDERS = 5,
    IND_ATR_TR = 6  // ATR over a precomputed true-range series
};

// Indicators computed from (high, low, close).
bool hlc(Ind k) { return k == IND_ATR || k == IND_ADX; }

const char *batch_kernel(Ind k) {
    static const char *n[] = {"ema_batch_f32",  "rsi_batch_f32",     "atr_batch_f32",
                              "adx_batch_f32",  "smma_batch_f32",    "wilders_batch_f32",
                              "atr_tr_batch_f32"};
    return n[k];
}
const char *ms_kernel(Ind k) {
    static const char *n[] = {"ema_ms1p_f32",  "rsi_ms1p_f32",  "atr_ms1p_f32",
                              "adx_ms1p_f32",  "smma_ms1p_f32", "wilders_ms1p_f32",
                              nullptr};
    return n[k];
}

// Bars of burn-in each engine needs (same formula as burn_bars in rec.hip).
double burn_bars(Ind k, size_t p) {
    double beta = k == IND_EMA ? 1.0 - 2.0 / ((double)p + 1.0) : 1.0 - 1.0 / (double)p;
    double nats = k == IND_ADX ? 32.0 : 20.0;
    if (!(beta > 0.0)) return 1.0;
    return std::ceil(nats / -std::log(beta));
}

// Chunk length: about four times the longest burn-in (so burn-in costs at
// most ~25% extra bars), at least 1024 bars, small enough that the launch
// has roughly 2048+ waves; the context's VTA_TUNE_MS_CHUNK knob, else
// VTA_REC_CHUNK, overrides.
int ms_chunk_override(const vta_ctx *ctx, size_t T) {
    if (ctx->tune[VTA_TUNE_MS_CHUNK] > 0)
        return (int)std::min<size_t>((size_t)ctx->tune[VTA_TUNE_MS_CHUNK], T);
    if (const char *e = getenv("VTA_REC_CHUNK")) {
        long v = atol(e);
        if (v >= 64) return (int)std::min<size_t>((size_t)v, T);
    }
    return 0;
}

int pick_chunk(const vta_ctx *ctx, Ind k, size_t max_p, size_t T, size_t lane_groups,
               bool clean) {
    if (!clean) return (int)T;  // exact from first_valid in one pass
    if (int c = ms_chunk_override(ctx, T)) return c;
    double kb = burn_bars(k, max_p);
    size_t want_chunks = std::max<size_t>(1, 2048 / std::max<size_t>(lane_groups, 1));
    size_t by_par = (T + want_chunks - 1) / want_chunks;
    size_t c = std::max<size_t>({(size_t)(4.0 * kb), 1024, by_par});
    c = (c + 63) & ~(size_t)63;
    return (int)std::min(c, T);
}

// One-series batches: a chunk of at least the burn-in (so the burn-in at most
// doubles the work) and small enough for ~2048 waves.  With long periods the
// burn-in is tens of thousands of bars (25-32 nats at p ~ 2000); the 4x-burn-in
// floor of the many-series path would leave only a few waves per lane group,
// while smaller chunks would multiply the work, which a device shared by
// concurrent requests pays for in full.
int pick_chunk_batch(Ind k, size_t max_p, size_t T, size_t lane_groups, bool clean) {
    if (!clean) return (int)T;
    if (const char *e = getenv("VTA_REC_CHUNK")) {
        long v = atol(e);
        if (v >= 64) return (int)std::min<size_t>((size_t)v, T);
    }
    const size_t kb = (size_t)burn_bars(k, max_p);
    const size_t want_chunks = std::max<size_t>(1, 2048 / std::max<size_t>(lane_groups, 1));
    const size_t by_par = (T + want_chunks - 1) / want_chunks;
    size_t c = std::max<size_t>({kb, by_par, 256});
    c = (c + 63) & ~(size_t)63;
    return (int)std::min(c, T);
}

// Grid expansion for the one-parameter indicators.
vta_status expand_periods(Ind k, const vta_range_usize *r, std::vector<size_t> &ps) {
    if (!r) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    // SMMA: smma.rs:430-471; Wilders: wilders_wrapper.rs:909-940 (both count
    // down to end when start > end, like AXIS_DOWN)
    AxisKind ak = (k == IND_ATR)                                        ? AXIS_ATR
                  : (k == IND_ADX || k == IND_SMMA || k == IND_WILDERS) ? AXIS_DOWN
                                                                        : AXIS_SORTED;
    if (!expand_usize(*r, ak, ps) || ps.empty())
        return fail(VTA_ERR_INVALID_INPUT, "invalid period range: start=%zu end=%zu step=%zu",
                    r->start, r->end, r->step);
    for (size_t p : ps)
        if (p == 0 || p > (size_t)INT32_MAX)
            return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", p);
    return VTA_OK;
}

// Minimum valid bars after first_valid each indicator needs (the CPU
// validation: ema.rs:307-312, rsi.rs:211-216, atr.rs:218-225, adx.rs:252-257).
size_t need_bars(Ind k, size_t p) { return k == IND_ADX ? p + 1 : p; }

size_t first_valid_of(Ind k, const float *x, const float *h, const float *l, size_t n) {
    if (!hlc(k)) return first_valid_f32(x, n);
    if (k == IND_ATR) {
        size_t i = 0;
        while (i < n && (std::isnan(h[i]) || std::isnan(l[i]) || std::isnan(x[i]))) ++i;
        return i;
    }
    size_t a = first_valid_f32(x, n), b = first_valid_f32(h, n), c = first_valid_f32(l, n);
    return std::max(a, std::max(b, c));
}

bool host_clean(const float *x, const float *h, const float *l, size_t n, size_t fv) {
    for (size_t i = fv; i < n; ++i) {
        if (!std::isfinite(x[i])) return false;
        if (h && (!std::isfinite(h[i]) || !std::isfinite(l[i]))) return false;
    }
    return true;
}

// Output staging of the one-series batch kernels, needed by the passes that
// write rows (phase 0 and 3): 64 rows x (W + 1) floats of dynamic LDS, W the
// bars staged between flushes (32 for RSI and ADX, whose longer steps gain
// from the extra resident waves; 64 for the others; kStageW in kernels/rec.hip).
unsigned batch_stage_bytes(Ind k) {
    const unsigned w = (k == IND_RSI || k == IND_ADX) ? 32u : 64u;
    return 64u * (w + 1u) * (unsigned)sizeof(float);
}

// Wilders seeds: doubles from byte 32,768 of the context scratch, one per
// combo, for batches of up to 4,096 combos (larger ones replay the warm-ups).
constexpr size_t kSeedOffset = 32 * 1024;
constexpr size_t kSeedCombos = (vta_ctx::kScratchBytes - kSeedOffset) / sizeof(double);
static_assert(kSeedOffset >= (kArgErrorWord + 1) * sizeof(int32_t), "seeds after the error word");

vta_status launch_rec(vta_ctx *ctx, const char *kname, rec_args_host &a, unsigned groups,
                      unsigned chunks, unsigned block = 64, unsigned smem = 0) {
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rec", kname, &f));
    return launch(ctx, f, dim3(groups, chunks, 1), dim3(block), smem, a);
}

// Chunked-scan launch: chunks sized for ~kScanWaves waves, at least 64 bars
// and at most 1,024 chunks (the fold walks them 64 at a time; a Wilders lane
// whose warm-up spans several chunks starts from its seed at the last
// warm-up bar, or, past kSeedCombos, replays the warm-up exactly from
// first_valid in each), and as many as the
// context's carry buffer holds for C
// combos; pass 1, the per-combo fold (one wave per combo; two rounds for
// ADX), pass 3.  Small requests are latency-bound, so their chunk is short:
// 16 one-series requests of 10,000 bars x 32 periods (EMA, RSI, ADX, SMA,
// BB) took 2.69 / 1.96 / 1.34 ms with 256- / 128- / 64-bar chunks.  Sets
// *done = false (nothing launched) when one chunk would cover the series.
// Measured on 1,250 periods x 200k bars: EMA 0.395 / 0.365 / 0.336 ms and ADX
// 1.14 / 1.24 / 1.09 ms at 2,048 / 4,096 / 8,192 waves (an earlier build);
// later EMA 0.298 / 0.285 and RSI 0.442 / 0.398 ms at 8,192 / 16,384 waves
// (256-bar chunks, the 1,024-chunk cap); flat on small shapes, whose chunks
// sit at the 64-bar floor either way.
constexpr size_t kScanWaves = 16384;

size_t scan_states(Ind k) { return k == IND_RSI ? 2 : k == IND_ADX ? 4 : 1; }

const char *combine_kernel(Ind k) {
    static const char *n[] = {"ema_scan_combine",  "rsi_scan_combine",     "atr_scan_combine",
                              "adx_scan_combine",  "smma_scan_combine",    "wilders_scan_combine",
                              "atr_tr_scan_combine"};
    return n[k];
}

// The chunked scan's chunk length for T bars x C combos of indicator k (the
// rule above), or 0 when the series is not scanned (one chunk would cover
// it, or the carry buffer holds fewer than two chunk-end states per combo).
size_t scan_chunk(const vta_ctx *ctx, Ind k, size_t T, size_t C, size_t max_p) {
    if (getenv("VTA_REC_NO_SCAN")) return 0;
    // states per (chunk, combo): RSI 2; ADX 4 (three first-stage states and
    // the ADX itself, folded in two rounds); the others 1
    const size_t S = scan_states(k);
    const size_t groups = (C + 63) / 64;
    size_t waves = kScanWaves;
    if (ctx->tune[VTA_TUNE_SCAN_WAVES] > 0) {
        waves = (size_t)ctx->tune[VTA_TUNE_SCAN_WAVES];
    } else if (const char *e = getenv("VTA_SCAN_WAVES")) {
        long v = atol(e);
        if (v >= 64) waves = (size_t)v;
    }
    const size_t want = std::max<size_t>(1, waves / groups);
    size_t chunk = std::max<size_t>({64, (T + want - 1) / want, (T + 1023) / 1024});
    if (const char *e = getenv("VTA_SCAN_CHUNK")) {
        long v = atol(e);
        if (v >= 64) chunk = (size_t)v;
    }
    const size_t max_chunks = vta_ctx::kCarryBytes / (C * S * sizeof(double));
    if (max_chunks < 2) return 0;
    if ((T + chunk - 1) / chunk > max_chunks) chunk = (T + max_chunks - 1) / max_chunks;
    chunk = (chunk + 63) & ~(size_t)63;
    return (T + chunk - 1) / chunk < 2 ? 0 : chunk;
}

vta_status scan_launch(vta_ctx *ctx, Ind k, rec_args_host a, size_t T, size_t C, size_t max_p,
                       bool *done) {
    *done = false;
    const size_t chunk = scan_chunk(ctx, k, T, C, max_p);
    if (chunk == 0) return VTA_OK;
    const size_t groups = (C + 63) / 64;
    const size_t nch = (T + chunk - 1) / chunk;
    a.chunk = (int)chunk;
    a.carry = static_cast<double *>(ctx->carry);
    hipFunction_t fb, fc;
    VTA_TRY(get_function(ctx, "rec", batch_kernel(k), &fb));
    VTA_TRY(get_function(ctx, "rec", combine_kernel(k), &fc));
    a.phase = 1;
    VTA_TRY(launch(ctx, fb, dim3((unsigned)groups, (unsigned)nch, 1), dim3(64), 0, a));
    VTA_TRY(launch(ctx, fc, dim3((unsigned)C), dim3(64), 0, a));
    if (k == IND_ADX) {
        // second stage: DX from the exact first-stage states, ADX from zero
        a.phase = 2;
        VTA_TRY(launch(ctx, fb, dim3((unsigned)groups, (unsigned)nch, 1), dim3(64), 0, a));
        VTA_TRY(launch(ctx, fc, dim3((unsigned)C), dim3(64), 0, a));
    }
    a.phase = 3;
    VTA_TRY(launch(ctx, fb, dim3((unsigned)groups, (unsigned)nch, 1), dim3(64),
                   batch_stage_bytes(k), a));
    *done = true;
    return VTA_OK;
}

// Index of the indicators with a kernel policy (CudaEmaPolicy, CudaRsiPolicy,
// CudaAdxPolicy), -1 for the others.
int policy_slot(Ind k) {
    return k == IND_EMA ? VTA_POLICY_SLOT_EMA
           : k == IND_RSI ? VTA_POLICY_SLOT_RSI
           : k == IND_ADX ? VTA_POLICY_SLOT_ADX
                          : -1;
}

bool plain_batch(const vta_ctx *ctx, Ind k) {
    const int s = policy_slot(k);
    return s >= 0 && ctx->policy[s].batch == VTA_KERNEL_PLAIN;
}

void record_batch(vta_ctx *ctx, Ind k, bool chunked) {
    const int s = policy_slot(k);
    if (s >= 0) ctx->sel_batch[s] = {chunked ? VTA_SELECTED_CHUNKED : VTA_SELECTED_PLAIN, 64};
}

// One series x many periods, device input and output, fv and clean known.
// max_p = 0: the longest period is not known on the host (the reference's
// device forms pass none), so no burn-in is planned from it: the series runs
// as a chunked scan (Wilders chunks then are not stretched to the longest
// warm-up), or else exactly from first_valid in one chunk.  d_alphas
// (per combo, may be NULL) overrides the EMA / Wilders smoothing factor.
vta_status batch_launch(vta_ctx *ctx, Ind k, const float *d_x, const float *d_h,
                        const float *d_l, const int32_t *d_periods, size_t max_p, size_t T,
                        size_t C, const int32_t *d_fv1, const int32_t *d_clean1, bool clean,
                        float *d_out, const float *d_alphas = nullptr, int clean_epoch = 0) {
    rec_args_host a{};
    a.x = d_x;
    a.h = d_h;
    a.l = d_l;
    a.periods = d_periods;
    a.fvs = d_fv1;
    a.clean = d_clean1;
    a.clean_epoch = clean_epoch;
    a.out0 = d_out;
    a.alphas = d_alphas;
    a.T = (int)T;
    a.nlanes = (int)C;
    const bool plain = plain_batch(ctx, k);
    if (k == IND_WILDERS && C <= kSeedCombos) {
        // per-combo seeds (the second half of the context scratch): an exact
        // lane then starts at its last warm-up bar instead of replaying the
        // warm-up from first_valid in every chunk the warm-up spans
        a.seeds = reinterpret_cast<double *>(static_cast<char *>(ctx->scratch) + kSeedOffset);
        VTA_TRY(launch_rec(ctx, "wilders_seed", a, (unsigned)((C + 3) / 4), 1, 256));
    }
    // (a Wilders warm-up longer than a chunk without seeds is still exact:
    // its lanes replay it from first_valid, see test_scan_short_chunks_long_lead)
    if (clean && !plain) {
        bool done;
        VTA_TRY(scan_launch(ctx, k, a, T, C, max_p, &done));
        if (done) {
            record_batch(ctx, k, true);
            return VTA_OK;
        }
    }
    unsigned groups = (unsigned)((C + 63) / 64);
    // a caller-supplied smoothing factor is not the one the burn-in length
    // would be planned from, so those batches also run in one chunk
    a.chunk = (plain || max_p == 0 || d_alphas) ? (int)T
                                                : pick_chunk_batch(k, max_p, T, groups, clean);
    unsigned chunks = (unsigned)((T + a.chunk - 1) / a.chunk);
    record_batch(ctx, k, chunks > 1);
    return launch_rec(ctx, batch_kernel(k), a, groups, chunks, 64, batch_stage_bytes(k));
}

vta_status batch_dev(vta_ctx *ctx, Ind k, const float *x, const float *h, const float *l,
                     size_t len, const vta_range_usize *range, vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !out || !x || (hlc(k) && (!h || !l)))
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "empty data");
    std::vector<size_t> ps;
    VTA_TRY(expand_periods(k, range, ps));
    size_t fv = first_valid_of(k, x, h, l, len);
    if (fv >= len) return fail(VTA_ERR_INVALID_INPUT, "all values are NaN");
    size_t max_p = *std::max_element(ps.begin(), ps.end());
    if (len - fv < need_bars(k, max_p))
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu, valid = %zu)",
                    need_bars(k, max_p), len - fv);
    if (len > (size_t)INT32_MAX) return fail(VTA_ERR_INVALID_INPUT, "series too long");
    if (k == IND_WILDERS) {
        // wilders_wrapper.rs:391-400: every bar of the longest warm window finite
        for (size_t i = 0; i < max_p; ++i)
            if (!std::isfinite(x[fv + i]))
                return fail(VTA_ERR_INVALID_INPUT, "non-finite value in warm window at offset %zu",
                            i);
    }
    bool clean = host_clean(x, hlc(k) ? h : nullptr, l, len, fv);
    size_t C = ps.size(), n_out;
    if (!mul_ok(C, len, &n_out) || !mul_ok(n_out, sizeof(float), &n_out))
        return fail(VTA_ERR_INVALID_INPUT, "output bytes overflow");
    size_t nin = hlc(k) ? 3 : 1;
    VTA_TRY(check_fits(n_out + nin * len * sizeof(float)));
    std::vector<int32_t> per(ps.begin(), ps.end());
    int32_t meta[2] = {(int32_t)fv, clean ? 1 : 0};
    DevBuf dx, dh, dl, dp, dm, dout;
    VTA_TRY(upload(ctx, dx, x, len * sizeof(float)));
    if (nin == 3) {
        VTA_TRY(upload(ctx, dh, h, len * sizeof(float)));
        VTA_TRY(upload(ctx, dl, l, len * sizeof(float)));
    }
    VTA_TRY(upload(ctx, dp, per.data(), per.size() * sizeof(int32_t)));
    VTA_TRY(upload(ctx, dm, meta, sizeof(meta)));
    VTA_TRY(dout.alloc(n_out));
    VTA_TRY(batch_launch(ctx, k, dx.as<float>(), dh.as<float>(), dl.as<float>(),
                         dp.as<int32_t>(), max_p, len, C, dm.as<int32_t>(),
                         dm.as<int32_t>() + 1, clean, dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = C;
    out->cols = len;
    return VTA_OK;
}

// Device-resident one-series batch: the clean flag is computed on the device
// into the context scratch.  max_period = 0 when the caller's signature has
// none (see batch_launch); d_alphas and d_warms (per combo, may be NULL) are
// the reference's EMA / SMMA / Wilders device arguments: the alphas are used,
// the warm indices checked on the device against first_valid + period - 1.
vta_status batch_device(vta_ctx *ctx, Ind k, const float *d_x, const float *d_h,
                        const float *d_l, const int32_t *d_periods, int64_t max_period,
                        int64_t series_len, int64_t first_valid, int64_t n_combos,
                        float *d_out, const float *d_alphas = nullptr,
                        const int32_t *d_warms = nullptr) {
    clear_error();
    if (!ctx || !d_x || !d_periods || !d_out || (hlc(k) && (!d_h || !d_l)))
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (series_len <= 0 || n_combos <= 0 || max_period < 0 || first_valid < 0 ||
        first_valid >= series_len)
        return fail(VTA_ERR_INVALID_INPUT, "series_len and n_combos must be positive "
                                           "and first_valid < series_len");
    if (series_len > INT32_MAX || n_combos > INT32_MAX || max_period > series_len)
        return fail(VTA_ERR_INVALID_INPUT, "series_len, n_combos or max_period out of range");
    // context scratch: [0] first_valid, [2] the non-finite word; one launch
    // (batch_prepare) stores both and checks the warm indices.  The word
    // holds this call's epoch if a bar is not finite, so it is never
    // cleared; an epoch repeated after 2^31 calls can only read a clean
    // series as unclean, which runs exactly (and correctly) from first_valid.
    int32_t *meta = static_cast<int32_t *>(ctx->scratch);
    if (++ctx->prep_epoch > INT32_MAX || ctx->prep_epoch == 0) ctx->prep_epoch = 1;
    const int epoch = (int)ctx->prep_epoch;
    hipFunction_t prep;
    VTA_TRY(get_function(ctx, "rec", "batch_prepare", &prep));
    int32_t *err = static_cast<int32_t *>(ctx->scratch) + kArgErrorWord;
    if (d_warms) ctx->arg_check_pending = true;
    const int64_t span = std::max<int64_t>(series_len - first_valid, d_warms ? n_combos : 0);
    VTA_TRY(launch(ctx, prep, dim3((unsigned)((span + 255) / 256)), dim3(256), 0, d_x,
                   hlc(k) ? d_h : nullptr, hlc(k) ? d_l : nullptr, (int)series_len,
                   (int)first_valid, epoch, meta, d_periods, d_warms, (int)n_combos, err));
    // Chunked with burn-in: a series found unclean still runs correctly, since
    // every lane of an unclean stream starts exactly at first_valid.
    return batch_launch(ctx, k, d_x, d_h, d_l, d_periods, (size_t)max_period,
                        (size_t)series_len, (size_t)n_combos, meta, meta + 2, true, d_out,
                        d_alphas, epoch);
}

vta_status ms_launch(vta_ctx *ctx, Ind k, const float *d_x, const float *d_h, const float *d_l,
                     size_t cols, size_t rows, size_t period, const int32_t *d_fvs,
                     const int32_t *d_clean, float *d_out, const float *alpha = nullptr) {
    int32_t *pp = static_cast<int32_t *>(ctx->scratch) + 64;
    hipFunction_t fill;
    VTA_TRY(get_function(ctx, "rec", "vta_fill_i32", &fill));
    VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, pp, 1, (int)period));
    const float *d_alpha = nullptr;
    if (alpha) {
        // the caller's smoothing factor, as one f32 word after the period
        int32_t bits;
        memcpy(&bits, alpha, sizeof(bits));
        VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, pp + 1, 1, (int)bits));
        d_alpha = reinterpret_cast<const float *>(pp + 1);
    }
    const int slot = policy_slot(k);
    const bool plain = slot >= 0 && ctx->policy[slot].many_series == VTA_KERNEL_PLAIN;
    rec_args_host a{};
    a.alphas = d_alpha;
    a.x = d_x;
    a.h = d_h;
    a.l = d_l;
    a.periods = pp;
    a.fvs = d_fvs;
    a.clean = d_clean;
    a.out0 = d_out;
    a.T = (int)rows;
    a.nlanes = (int)cols;
    unsigned groups = (unsigned)((cols + 63) / 64);
    a.chunk = pick_chunk(ctx, k, period, rows, groups, d_clean != nullptr);
    if (alpha && d_clean && !plain) {
        // burn-in planned from the caller's smoothing factor (the engines'
        // own burn-in follows the factor they run with)
        const double beta = 1.0 - (double)*alpha;
        if (!(beta > 0.0 && beta < 1.0)) {
            a.chunk = (int)rows;
        } else if (!ms_chunk_override(ctx, rows)) {
            const double kb = std::ceil(20.0 / -std::log(beta));
            const size_t want = std::max<size_t>(1, 2048 / std::max<size_t>(groups, 1));
            size_t c = std::max<size_t>({(size_t)std::min(4.0 * kb, 1e9), 1024,
                                         (rows + want - 1) / want});
            a.chunk = (int)std::min(((c + 63) & ~(size_t)63), rows);
        }
    }
    // without clean flags, or under the plain policy, every series runs
    // exactly from first_valid in one chunk
    if (!d_clean || plain) a.chunk = (int)rows;
    unsigned chunks = (unsigned)((rows + a.chunk - 1) / a.chunk);
    if (slot >= 0)
        ctx->sel_ms[slot] = {chunks > 1 ? VTA_SELECTED_CHUNKED : VTA_SELECTED_PLAIN, 64};
    return launch_rec(ctx, ms_kernel(k), a, groups, chunks);
}

vta_status ms_dev(vta_ctx *ctx, Ind k, const float *x_tm, const float *h_tm, const float *l_tm,
                  size_t cols, size_t rows, size_t period, vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !out || !x_tm || (hlc(k) && (!h_tm || !l_tm)))
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    if (period == 0) return fail(VTA_ERR_INVALID_INPUT, "period must be > 0");
    size_t elems;
    if (!mul_ok(cols, rows, &elems) || elems > (size_t)INT32_MAX * 16)
        return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    // per-series first_valid and clean flag on the host (the reference scans
    // each column on the host too, e.g. ema_wrapper.rs:637-700)
    std::vector<int32_t> fvs(cols), cl(cols);
    std::vector<float> col_x(rows), col_h, col_l;
    if (hlc(k)) {
        col_h.resize(rows);
        col_l.resize(rows);
    }
    for (size_t s = 0; s < cols; ++s) {
        for (size_t t = 0; t < rows; ++t) col_x[t] = x_tm[t * cols + s];
        if (hlc(k))
            for (size_t t = 0; t < rows; ++t) {
                col_h[t] = h_tm[t * cols + s];
                col_l[t] = l_tm[t * cols + s];
            }
        size_t fv = first_valid_of(k, col_x.data(), col_h.data(), col_l.data(), rows);
        if (k == IND_WILDERS) {
            // the Wilders many-series wrapper starts at the first finite value
            // (wilders_wrapper.rs:329-349)
            fv = 0;
            while (fv < rows && !std::isfinite(col_x[fv])) ++fv;
        }
        if (fv >= rows) return fail(VTA_ERR_INVALID_INPUT, "series %zu all NaN", s);
        if (rows - fv < need_bars(k, period))
            return fail(VTA_ERR_INVALID_INPUT,
                        "series %zu not enough valid data (needed >= %zu, valid = %zu)", s,
                        need_bars(k, period), rows - fv);
        fvs[s] = (int32_t)fv;
        cl[s] = host_clean(col_x.data(), hlc(k) ? col_h.data() : nullptr, col_l.data(),
                           rows, fv) ? 1 : 0;
    }
    size_t nin = hlc(k) ? 3 : 1;
    VTA_TRY(check_fits((nin + 1) * elems * sizeof(float)));
    DevBuf dx, dh, dl, dfv, dcl, dout;
    VTA_TRY(upload(ctx, dx, x_tm, elems * sizeof(float)));
    if (nin == 3) {
        VTA_TRY(upload(ctx, dh, h_tm, elems * sizeof(float)));
        VTA_TRY(upload(ctx, dl, l_tm, elems * sizeof(float)));
    }
    VTA_TRY(upload(ctx, dfv, fvs.data(), cols * sizeof(int32_t)));
    VTA_TRY(upload(ctx, dcl, cl.data(), cols * sizeof(int32_t)));
    VTA_TRY(dout.alloc(elems * sizeof(float)));
    VTA_TRY(ms_launch(ctx, k, dx.as<float>(), dh.as<float>(), dl.as<float>(), cols, rows, period,
                      dfv.as<int32_t>(), dcl.as<int32_t>(), dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = rows;
    out->cols = cols;
    return VTA_OK;
}

// *_batch_into_host_f32 (ema_wrapper.rs:309-323, adx_wrapper.rs:253-280): the
// batch into a caller-owned host slice of combos x len floats, [combo][t].
vta_status batch_into_host(vta_ctx *ctx, Ind k, const float *x, const float *h, const float *l,
                           size_t len, const vta_range_usize *range, float *out,
                           size_t out_len) {
    clear_error();
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::vector<size_t> ps;
    VTA_TRY(expand_periods(k, range, ps));
    size_t expected;
    if (!mul_ok(ps.size(), len, &expected))
        return fail(VTA_ERR_INVALID_INPUT, "combos*series_len overflow");
    if (out_len != expected)
        return fail(VTA_ERR_INVALID_INPUT, "output slice wrong length: got %zu, need %zu",
                    out_len, expected);
    vta_device_array_f32 arr{};
    VTA_TRY(batch_dev(ctx, k, x, h, l, len, range, &arr));
    vta_status st = vta_device_array_copy_to_host(ctx, &arr, out, out_len);
    vta_device_array_free(ctx, &arr);
    return st;
}

// *_many_series_one_param_time_major_into_host_f32 (ema_wrapper.rs:422-441,
// adx_wrapper.rs:417-440): [rows][cols] into a caller-owned host slice.
vta_status ms_into_host(vta_ctx *ctx, Ind k, const float *x_tm, const float *h_tm,
                        const float *l_tm, size_t cols, size_t rows, size_t period,
                        float *out_tm, size_t out_len) {
    clear_error();
    if (!ctx || !out_tm) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    size_t expected;
    if (!mul_ok(cols, rows, &expected)) return fail(VTA_ERR_INVALID_INPUT, "rows*cols overflow");
    if (out_len != expected) return fail(VTA_ERR_INVALID_INPUT, "out slice wrong length");
    vta_device_array_f32 arr{};
    VTA_TRY(ms_dev(ctx, k, x_tm, h_tm, l_tm, cols, rows, period, &arr));
    vta_status st = vta_device_array_copy_to_host(ctx, &arr, out_tm, out_len);
    vta_device_array_free(ctx, &arr);
    return st;
}

vta_status expand_out(Ind k, const vta_range_usize *range, size_t *out, size_t cap,
                      size_t *n_out) {
    clear_error();
    std::vector<size_t> ps;
    VTA_TRY(expand_periods(k, range, ps));
    if (n_out) *n_out = ps.size();
    if (out) memcpy(out, ps.data(), std::min(cap, ps.size()) * sizeof(size_t));
    return VTA_OK;
}

// Many-series device forms of EMA, SMMA and Wilders: per-series clean flags
// against the caller's first_valids computed on the device (into the
// context's carry buffer) so clean series run in time chunks.
vta_status ms_device_one_input(vta_ctx *ctx, Ind k, const float *d_prices_tm,
                               const int32_t *d_first_valids, int64_t period,
                               int64_t num_series, int64_t series_len, float *d_out_tm,
                               const float *alpha) {
    clear_error();
    if (!ctx || !d_prices_tm || !d_first_valids || !d_out_tm || period <= 0 ||
        num_series <= 0 || series_len <= 0 || period > series_len)
        return fail(VTA_ERR_INVALID_INPUT,
                    "period, num_series and series_len must be positive, period <= series_len");
    if (num_series > INT32_MAX || series_len > INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    const int32_t *d_clean = nullptr;
    if ((size_t)num_series * sizeof(int32_t) <= vta_ctx::kCarryBytes) {
        int32_t *cl = static_cast<int32_t *>(ctx->carry);
        hipFunction_t f;
        VTA_TRY(get_function(ctx, "rec", "ms_clean_from_first_valids", &f));
        VTA_TRY(launch(ctx, f, dim3((unsigned)((num_series + 63) / 64)), dim3(64), 0,
                       d_prices_tm, (int)series_len, (int)num_series, d_first_valids, cl));
        d_clean = cl;
    }
    return ms_launch(ctx, k, d_prices_tm, nullptr, nullptr, (size_t)num_series,
                     (size_t)series_len, (size_t)period, d_first_valids, d_clean, d_out_tm,
                     alpha);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ ATR from a TR series
vta_status vta_tr_from_hlc_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                  const float *d_close, int32_t series_len, int32_t first_valid,
                                  float *d_tr_out) {
    clear_error();
    if (!ctx || !d_high || !d_low || !d_close || !d_tr_out)
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (series_len <= 0) return fail(VTA_ERR_INVALID_INPUT, "empty input");
    if (first_valid < 0 || first_valid >= series_len)
        return fail(VTA_ERR_INVALID_INPUT, "first_valid out of range");
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rec", "tr_from_hlc_f32", &f));
    return launch(ctx, f, dim3((unsigned)((series_len + 255) / 256)), dim3(256), 0, d_high,
                  d_low, d_close, (int)series_len, (int)first_valid, d_tr_out);
}

vta_status vta_atr_batch_device_with_tr(vta_ctx *ctx, const float *d_tr, const int32_t *d_periods,
                                        int32_t max_period, int32_t series_len,
                                        int32_t first_valid, int32_t n_combos, float *d_out) {
    return batch_device(ctx, IND_ATR_TR, d_tr, nullptr, nullptr, d_periods, max_period,
                        series_len, first_valid, n_combos, d_out);
}

// ------------------------------------------------------------------ SMMA, Wilders
#define VTA_ONE_INPUT_REC(NAME, IND)                                                              \
    vta_status vta_##NAME##_expand_grid(const vta_range_usize *range, size_t *periods,          \
                                        size_t cap, size_t *n_out) {                            \
        return expand_out(IND, range, periods, cap, n_out);                                     \
    }                                                                                           \
    vta_status vta_##NAME##_batch_dev(vta_ctx *ctx, const float *prices, size_t len,            \
                                      const vta_range_usize *range, vta_device_array_f32 *out) { \
        return batch_dev(ctx, IND, prices, nullptr, nullptr, len, range, out);                  \
    }                                                                                           \
    vta_status vta_##NAME##_many_series_one_param_time_major_dev(                               \
        vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period,            \
        vta_device_array_f32 *out) {                                                            \
        return ms_dev(ctx, IND, data_tm, nullptr, nullptr, cols, rows, period, out);            \
    }

VTA_ONE_INPUT_REC(smma, IND_SMMA)
VTA_ONE_INPUT_REC(wilders, IND_WILDERS)
#undef VTA_ONE_INPUT_REC

// smma_batch_device (smma_wrapper.rs:377-396): the warm indices are checked
// against first_valid + period - 1 on the device (NULL: no check).
vta_status vta_smma_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                 const int32_t *d_warms, size_t first_valid, size_t series_len,
                                 size_t n_combos, float *d_out) {
    if (series_len > (size_t)INT32_MAX || n_combos > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "series_len or n_combos out of range");
    return batch_device(ctx, IND_SMMA, d_prices, nullptr, nullptr, d_periods, 0,
                        (int64_t)series_len, (int64_t)first_valid, (int64_t)n_combos, d_out,
                        nullptr, d_warms);
}

// wilders_batch_device (wilders_wrapper.rs:252-330): the alphas are used (NULL:
// 1/p in f64), the warm indices checked against first_valid + period - 1 on
// the device (NULL: no check).
vta_status vta_wilders_batch_device(vta_ctx *ctx, const float *d_prices,
                                    const int32_t *d_periods, const float *d_alphas,
                                    const int32_t *d_warm, int32_t series_len,
                                    int32_t first_valid, int32_t n_combos, float *d_out) {
    return batch_device(ctx, IND_WILDERS, d_prices, nullptr, nullptr, d_periods, 0, series_len,
                        first_valid, n_combos, d_out, d_alphas, d_warm);
}

// smma_multi_series_one_param_device (smma_wrapper.rs:599-621).
vta_status vta_smma_multi_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                  int32_t period, int32_t num_series,
                                                  int32_t series_len,
                                                  const int32_t *d_first_valids,
                                                  float *d_out_tm) {
    return ms_device_one_input(ctx, IND_SMMA, d_prices_tm, d_first_valids, period, num_series,
                               series_len, d_out_tm, nullptr);
}

// wilders_many_series_one_param_device (wilders_wrapper.rs:662-700): alpha is used.
vta_status vta_wilders_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                    const int32_t *d_first_valids,
                                                    int32_t period, float alpha,
                                                    size_t num_series, size_t series_len,
                                                    float *d_out_tm) {
    if (num_series > (size_t)INT32_MAX || series_len > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    return ms_device_one_input(ctx, IND_WILDERS, d_prices_tm, d_first_valids, period,
                               (int64_t)num_series, (int64_t)series_len, d_out_tm, &alpha);
}

// ------------------------------------------------------------------ EMA
vta_status vta_ema_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out) {
    return expand_out(IND_EMA, range, periods, cap, n_out);
}
vta_status vta_ema_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_range_usize *range, vta_device_array_f32 *out) {
    return batch_dev(ctx, IND_EMA, prices, nullptr, nullptr, len, range, out);
}
vta_status vta_ema_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                const float *d_alphas, size_t series_len, size_t first_valid,
                                size_t n_combos, float *d_out) {
    if (series_len > (size_t)INT32_MAX || n_combos > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "series_len or n_combos out of range");
    return batch_device(ctx, IND_EMA, d_prices, nullptr, nullptr, d_periods, 0,
                        (int64_t)series_len, (int64_t)first_valid, (int64_t)n_combos, d_out,
                        d_alphas);
}
vta_status vta_ema_batch_into_host_f32(vta_ctx *ctx, const float *prices, size_t len,
                                       const vta_range_usize *range, float *out,
                                       size_t out_len) {
    return batch_into_host(ctx, IND_EMA, prices, nullptr, nullptr, len, range, out, out_len);
}
vta_status vta_ema_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                        size_t cols, size_t rows, size_t period,
                                                        vta_device_array_f32 *out) {
    return ms_dev(ctx, IND_EMA, data_tm, nullptr, nullptr, cols, rows, period, out);
}
vta_status vta_ema_many_series_one_param_time_major_into_host_f32(
    vta_ctx *ctx, const float *data_tm, size_t cols, size_t rows, size_t period, float *out_tm,
    size_t out_len) {
    return ms_into_host(ctx, IND_EMA, data_tm, nullptr, nullptr, cols, rows, period, out_tm,
                        out_len);
}
vta_status vta_ema_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                const int32_t *d_first_valids, int32_t period,
                                                float alpha, size_t num_series,
                                                size_t series_len, float *d_out_tm) {
    if (num_series > (size_t)INT32_MAX || series_len > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    return ms_device_one_input(ctx, IND_EMA, d_prices_tm, d_first_valids, period,
                               (int64_t)num_series, (int64_t)series_len, d_out_tm, &alpha);
}

// ------------------------------------------------------------------ RSI
vta_status vta_rsi_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out) {
    return expand_out(IND_RSI, range, periods, cap, n_out);
}
vta_status vta_rsi_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                             const vta_range_usize *range, vta_device_array_f32 *out) {
    return batch_dev(ctx, IND_RSI, prices, nullptr, nullptr, len, range, out);
}
vta_status vta_rsi_batch_device(vta_ctx *ctx, const float *d_prices, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out) {
    return batch_device(ctx, IND_RSI, d_prices, nullptr, nullptr, d_periods, max_period,
                        series_len, first_valid, n_combos, d_out);
}
vta_status vta_rsi_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *prices_tm,
                                                        size_t cols, size_t rows, size_t period,
                                                        vta_device_array_f32 *out) {
    return ms_dev(ctx, IND_RSI, prices_tm, nullptr, nullptr, cols, rows, period, out);
}
vta_status vta_rsi_many_series_one_param_device(vta_ctx *ctx, const float *d_prices_tm,
                                                const int32_t *d_first_valids, size_t period,
                                                size_t num_series, size_t series_len,
                                                const int32_t *d_clean, float *d_out_tm) {
    clear_error();
    if (!ctx || !d_prices_tm || !d_first_valids || !d_out_tm || period == 0 || num_series == 0 ||
        series_len == 0)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    return ms_launch(ctx, IND_RSI, d_prices_tm, nullptr, nullptr, num_series, series_len, period,
                     d_first_valids, d_clean, d_out_tm);
}

// ------------------------------------------------------------------ ATR
vta_status vta_atr_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out) {
    return expand_out(IND_ATR, range, periods, cap, n_out);
}
vta_status vta_atr_batch_dev(vta_ctx *ctx, const float *high, const float *low,
                             const float *close, size_t len, const vta_range_usize *range,
                             vta_device_array_f32 *out) {
    return batch_dev(ctx, IND_ATR, close, high, low, len, range, out);
}
vta_status vta_atr_batch_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                const float *d_close, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out) {
    return batch_device(ctx, IND_ATR, d_close, d_high, d_low, d_periods, max_period, series_len,
                        first_valid, n_combos, d_out);
}
vta_status vta_atr_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *high_tm,
                                                        const float *low_tm,
                                                        const float *close_tm, size_t cols,
                                                        size_t rows, size_t period,
                                                        vta_device_array_f32 *out) {
    return ms_dev(ctx, IND_ATR, close_tm, high_tm, low_tm, cols, rows, period, out);
}
vta_status vta_atr_many_series_one_param_device(vta_ctx *ctx, const float *d_high_tm,
                                                const float *d_low_tm, const float *d_close_tm,
                                                const int32_t *d_first_valids, size_t period,
                                                size_t num_series, size_t series_len,
                                                const int32_t *d_clean, float *d_out_tm) {
    clear_error();
    if (!ctx || !d_high_tm || !d_low_tm || !d_close_tm || !d_first_valids || !d_out_tm ||
        period == 0 || num_series == 0 || series_len == 0)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    return ms_launch(ctx, IND_ATR, d_close_tm, d_high_tm, d_low_tm, num_series, series_len,
                     period, d_first_valids, d_clean, d_out_tm);
}

// ------------------------------------------------------------------ ADX
vta_status vta_adx_expand_grid(const vta_range_usize *range, size_t *periods, size_t cap,
                               size_t *n_out) {
    return expand_out(IND_ADX, range, periods, cap, n_out);
}
vta_status vta_adx_batch_dev(vta_ctx *ctx, const float *high, const float *low,
                             const float *close, size_t len, const vta_range_usize *range,
                             vta_device_array_f32 *out) {
    return batch_dev(ctx, IND_ADX, close, high, low, len, range, out);
}
vta_status vta_adx_batch_device(vta_ctx *ctx, const float *d_high, const float *d_low,
                                const float *d_close, const int32_t *d_periods,
                                int32_t max_period, int32_t series_len, int32_t first_valid,
                                int32_t n_combos, float *d_out) {
    return batch_device(ctx, IND_ADX, d_close, d_high, d_low, d_periods, max_period, series_len,
                        first_valid, n_combos, d_out);
}
vta_status vta_adx_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *high_tm,
                                                        const float *low_tm,
                                                        const float *close_tm, size_t cols,
                                                        size_t rows, size_t period,
                                                        vta_device_array_f32 *out) {
    return ms_dev(ctx, IND_ADX, close_tm, high_tm, low_tm, cols, rows, period, out);
}
vta_status vta_adx_batch_into_host_f32(vta_ctx *ctx, const float *high, const float *low,
                                       const float *close, size_t len,
                                       const vta_range_usize *range, float *out,
                                       size_t out_len) {
    return batch_into_host(ctx, IND_ADX, close, high, low, len, range, out, out_len);
}
vta_status vta_adx_many_series_one_param_time_major_into_host_f32(
    vta_ctx *ctx, const float *high_tm, const float *low_tm, const float *close_tm, size_t cols,
    size_t rows, size_t period, float *out_tm, size_t out_len) {
    return ms_into_host(ctx, IND_ADX, close_tm, high_tm, low_tm, cols, rows, period, out_tm,
                        out_len);
}

// ------------------------------------------------------------------ prepared many-series
vta_status vta_ma_many_series_prepared_device(vta_ctx *ctx, vta_ma_kind kind,
                                              const float *d_prices_tm,
                                              const int32_t *d_first_valids,
                                              const int32_t *d_meta, size_t period, float alpha,
                                              size_t num_series, size_t series_len,
                                              float *d_out_tm) {
    clear_error();
    if (!ctx || !d_prices_tm || !d_first_valids || !d_out_tm || period == 0 || num_series == 0 ||
        series_len == 0 || period > series_len)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    if (num_series > (size_t)INT32_MAX || series_len > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    switch (kind) {
        case VTA_MA_EMA:
        case VTA_MA_SMMA:
        case VTA_MA_WILDERS: {
            const Ind k = kind == VTA_MA_EMA ? IND_EMA : kind == VTA_MA_SMMA ? IND_SMMA
                                                                             : IND_WILDERS;
            const bool own = alpha > 0.0f && k != IND_SMMA;
            return ms_launch(ctx, k, d_prices_tm, nullptr, nullptr, num_series, series_len,
                             period, d_first_valids, d_meta, d_out_tm, own ? &alpha : nullptr);
        }
        case VTA_MA_SMA:
        case VTA_MA_WMA:
        case VTA_MA_SWMA:
            return fir_ms_prepared(ctx, (int)kind - (int)VTA_MA_SMA, d_prices_tm, d_first_valids,
                                   d_meta, num_series, series_len, period, d_out_tm);
    }
    return fail(VTA_ERR_INVALID_INPUT, "unknown moving-average kind %d", (int)kind);
}

// ------------------------------------------------------------------ shared prep
vta_status vta_many_series_prepare(vta_ctx *ctx, const float *d_x_tm, const float *d_high_tm,
                                   const float *d_low_tm, size_t cols, size_t rows, int mode,
                                   int32_t *d_first_valids, int32_t *d_first_valids_hlc,
                                   int32_t *d_clean) {
    clear_error();
    if (!ctx || !d_x_tm || !d_first_valids || !d_clean || cols == 0 || rows == 0 || mode < 0 ||
        mode > 4 || (mode != 0 && mode != 4 && (!d_high_tm || !d_low_tm)))
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rec", "ms_prepare_meta", &f));
    return launch(ctx, f, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, d_x_tm, d_high_tm,
                  d_low_tm, (int)rows, (int)cols, mode, d_first_valids, d_first_valids_hlc,
                  d_clean);
}

// ------------------------------------------------------------------ fused EMA + RSI + ATR
vta_status vta_ema_rsi_atr_many_series_device(vta_ctx *ctx, const float *d_close_tm,
                                              const float *d_high_tm, const float *d_low_tm,
                                              size_t cols, size_t rows, size_t ema_period,
                                              size_t rsi_period, size_t atr_period,
                                              const int32_t *d_fv_close,
                                              const int32_t *d_fv_hlc, const int32_t *d_clean,
                                              float *d_ema_tm, float *d_rsi_tm,
                                              float *d_atr_tm) {
    clear_error();
    if (!ctx || !d_close_tm || !d_high_tm || !d_low_tm || !d_fv_close || !d_fv_hlc ||
        !d_clean || !d_ema_tm || !d_rsi_tm || !d_atr_tm || cols == 0 || rows == 0 ||
        ema_period == 0 || rsi_period == 0 || atr_period == 0)
        return fail(VTA_ERR_INVALID_INPUT, "invalid arguments");
    if (rows > (size_t)INT32_MAX || cols > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "dimensions exceed int32");
    rec_args_host a{};
    a.x = d_close_tm;
    a.h = d_high_tm;
    a.l = d_low_tm;
    a.fvs = d_fv_close;
    a.fvs2 = d_fv_hlc;
    a.clean = d_clean;
    a.out0 = d_ema_tm;
    a.out1 = d_rsi_tm;
    a.out2 = d_atr_tm;
    a.T = (int)rows;
    a.nlanes = (int)cols;
    a.p1 = (int)ema_period;
    a.p2 = (int)rsi_period;
    a.p3 = (int)atr_period;
    unsigned groups = (unsigned)((cols + 63) / 64);
    size_t kb = std::max({(size_t)burn_bars(IND_EMA, ema_period),
                          (size_t)burn_bars(IND_RSI, rsi_period),
                          (size_t)burn_bars(IND_ATR, atr_period)});
    a.chunk = pick_chunk(ctx, IND_ATR, 1, rows, groups, true);
    if (!ms_chunk_override(ctx, rows)) {
        size_t c = std::max<size_t>({4 * kb, 1024, (size_t)a.chunk});
        a.chunk = (int)std::min(((c + 63) & ~(size_t)63), rows);
    }
    unsigned chunks = (unsigned)((rows + a.chunk - 1) / a.chunk);
    return launch_rec(ctx, "ms_ema_rsi_atr_f32", a, groups, chunks);
}

vta_status vta_ema_rsi_atr_many_series_dev(vta_ctx *ctx, const float *close_tm,
                                           const float *high_tm, const float *low_tm,
                                           size_t cols, size_t rows, size_t ema_period,
                                           size_t rsi_period, size_t atr_period,
                                           vta_device_array_f32 *out3) {
    clear_error();
    if (!ctx || !out3 || !close_tm || !high_tm || !low_tm)
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    size_t elems;
    if (!mul_ok(cols, rows, &elems)) return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    VTA_TRY(check_fits(6 * elems * sizeof(float)));
    DevBuf dc, dh, dl, dfv, dfv3, dcl, o0, o1, o2;
    VTA_TRY(upload(ctx, dc, close_tm, elems * sizeof(float)));
    VTA_TRY(upload(ctx, dh, high_tm, elems * sizeof(float)));
    VTA_TRY(upload(ctx, dl, low_tm, elems * sizeof(float)));
    VTA_TRY(dfv.alloc(cols * sizeof(int32_t)));
    VTA_TRY(dfv3.alloc(cols * sizeof(int32_t)));
    VTA_TRY(dcl.alloc(cols * sizeof(int32_t)));
    VTA_TRY(vta_many_series_prepare(ctx, dc.as<float>(), dh.as<float>(), dl.as<float>(), cols,
                                    rows, 3, dfv.as<int32_t>(), dfv3.as<int32_t>(),
                                    dcl.as<int32_t>()));
    // validation against the CPU rules needs the first_valids on the host
    std::vector<int32_t> fvc(cols), fv3(cols);
    VTA_HIP(hipMemcpyAsync(fvc.data(), dfv.p, cols * 4, hipMemcpyDeviceToHost, ctx->stream));
    VTA_HIP(hipMemcpyAsync(fv3.data(), dfv3.p, cols * 4, hipMemcpyDeviceToHost, ctx->stream));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    size_t need = std::max({ema_period, rsi_period, atr_period});
    for (size_t s = 0; s < cols; ++s) {
        if ((size_t)fv3[s] >= rows || (size_t)fvc[s] >= rows)
            return fail(VTA_ERR_INVALID_INPUT, "series %zu all NaN", s);
        if (rows - (size_t)std::max(fvc[s], fv3[s]) < need)
            return fail(VTA_ERR_INVALID_INPUT, "series %zu not enough valid data", s);
    }
    VTA_TRY(o0.alloc(elems * sizeof(float)));
    VTA_TRY(o1.alloc(elems * sizeof(float)));
    VTA_TRY(o2.alloc(elems * sizeof(float)));
    VTA_TRY(vta_ema_rsi_atr_many_series_device(
        ctx, dc.as<float>(), dh.as<float>(), dl.as<float>(), cols, rows, ema_period, rsi_period,
        atr_period, dfv.as<int32_t>(), dfv3.as<int32_t>(), dcl.as<int32_t>(), o0.as<float>(),
        o1.as<float>(), o2.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    DevBuf *outs[3] = {&o0, &o1, &o2};
    for (int i = 0; i < 3; ++i) {
        out3[i].ptr = static_cast<float *>(outs[i]->release());
        out3[i].rows = rows;
        out3[i].cols = cols;
    }
    return VTA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- packed batches

static_assert(sizeof(PackJob) == 80, "rec_job layout");

namespace {

struct rec_pack_args_host {
    const void *jobs;
    const int32_t *flags;
    int n_jobs;
    int phase;
};
static_assert(sizeof(rec_pack_args_host) == 24, "rec_pack_args layout");

const char *packed_kernel(Ind k, bool fold) {
    static const char *p[] = {"ema_packed_f32", "rsi_packed_f32", "atr_packed_f32",
                              "adx_packed_f32", "smma_packed_f32", "wilders_packed_f32"};
    static const char *f[] = {"ema_packed_fold", "rsi_packed_fold", "atr_packed_fold",
                              "adx_packed_fold", "smma_packed_fold", "wilders_packed_fold"};
    return fold ? f[k] : p[k];
}

}  // namespace

namespace vta {

vta_status pack_layout(const vta_ctx *ctx, int kind, const vta_rec_job *jobs, size_t n_jobs,
                       std::vector<PackJob> &tab, std::vector<size_t> &carry_off,
                       size_t *carry_elems, PackDims *dims) {
    if (!ctx || !jobs || !dims || !carry_elems)
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (kind < (int)VTA_REC_EMA || kind > (int)VTA_REC_WILDERS)
        return fail(VTA_ERR_INVALID_INPUT, "unknown indicator kind %d", kind);
    if (n_jobs == 0 || n_jobs > 65535)
        return fail(VTA_ERR_INVALID_INPUT, "n_jobs must be in 1..65535 (got %zu)", n_jobs);
    const Ind k = (Ind)kind;
    const size_t S = scan_states(k);
    tab.assign(n_jobs, PackJob{});
    carry_off.assign(n_jobs, 0);
    size_t blocks = 0, combos = 0, elems = 0, span = 1;
    for (size_t j = 0; j < n_jobs; ++j) {
        const vta_rec_job &q = jobs[j];
        if (!q.d_close || !q.d_periods || !q.d_out || (hlc(k) && (!q.d_high || !q.d_low)))
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: null pointer", j);
        if (q.series_len <= 0 || q.n_combos <= 0 || q.max_period <= 0 || q.first_valid < 0 ||
            q.first_valid >= q.series_len)
            return fail(VTA_ERR_INVALID_INPUT,
                        "job %zu: series_len, n_combos and max_period must be positive and "
                        "first_valid < series_len", j);
        const size_t T = (size_t)q.series_len, C = (size_t)q.n_combos;
        // the chunking of the job's own _device call (one exact chunk where
        // that call would not scan)
        size_t chunk = scan_chunk(ctx, k, T, C, (size_t)q.max_period);
        if (chunk == 0) chunk = T;
        const size_t nch = (T + chunk - 1) / chunk;
        const size_t groups = (C + 63) / 64;
        PackJob &e = tab[j];
        e.x = q.d_close;
        e.h = hlc(k) ? q.d_high : nullptr;
        e.l = hlc(k) ? q.d_low : nullptr;
        e.periods = q.d_periods;
        e.out = q.d_out;
        e.carry = nullptr;                      // set once the buffer exists
        e.T = (int)T;
        e.nlanes = (int)C;
        e.fv = q.first_valid;
        e.chunk = (int)chunk;
        e.nch = (int)nch;
        e.groups = (int)groups;
        e.block0 = (int)blocks;
        e.combo0 = (int)combos;
        carry_off[j] = elems;
        blocks += groups * nch;
        combos += C;
        elems += C * nch * S;
        span = std::max(span, T - (size_t)q.first_valid);
        if (blocks > (size_t)INT32_MAX || combos > (size_t)INT32_MAX)
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "packed grid exceeds 2^31 blocks");
    }
    dims->kind = kind;
    dims->n_jobs = (int)n_jobs;
    dims->total_blocks = (unsigned)blocks;
    dims->total_combos = (unsigned)combos;
    dims->flag_blocks = (unsigned)std::min<size_t>((span + 255) / 256, 1024);
    *carry_elems = elems;
    return VTA_OK;
}

vta_status pack_launch(vta_ctx *ctx, const PackDims &d, const void *d_jobs, int32_t *d_flags) {
    const Ind k = (Ind)d.kind;
    hipFunction_t fill, flag, fp, ff;
    VTA_TRY(get_function(ctx, "rec", "vta_fill_i32", &fill));
    VTA_TRY(get_function(ctx, "rec", "packed_nonfinite_flag", &flag));
    VTA_TRY(get_function(ctx, "rec", packed_kernel(k, false), &fp));
    VTA_TRY(get_function(ctx, "rec", packed_kernel(k, true), &ff));
    VTA_TRY(launch(ctx, fill, dim3((unsigned)((d.n_jobs + 255) / 256)), dim3(256), 0, d_flags,
                   d.n_jobs, 0));
    VTA_TRY(launch(ctx, flag, dim3(d.flag_blocks, (unsigned)d.n_jobs), dim3(256), 0, d_jobs,
                   d_flags, hlc(k) ? 1 : 0));
    rec_pack_args_host a{d_jobs, d_flags, d.n_jobs, 1};
    VTA_TRY(launch(ctx, fp, dim3(d.total_blocks), dim3(64), 0, a));
    VTA_TRY(launch(ctx, ff, dim3(d.total_combos), dim3(64), 0, a));
    if (k == IND_ADX) {
        a.phase = 2;
        VTA_TRY(launch(ctx, fp, dim3(d.total_blocks), dim3(64), 0, a));
        VTA_TRY(launch(ctx, ff, dim3(d.total_combos), dim3(64), 0, a));
    }
    a.phase = 3;
    VTA_TRY(launch(ctx, fp, dim3(d.total_blocks), dim3(64), batch_stage_bytes(k), a));
    return VTA_OK;
}

}  // namespace vta

struct vta_rec_pack {
    int device = 0;
    PackDims dims{};
    DevBuf jobs, flags, carry;
};

extern "C" {

vta_status vta_rec_pack_create(vta_ctx *ctx, vta_rec_kind kind, const vta_rec_job *jobs,
                               size_t n_jobs, vta_rec_pack **out) {
    clear_error();
    if (!ctx || !jobs || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = nullptr;
    std::vector<PackJob> tab;
    std::vector<size_t> off;
    size_t elems = 0;
    auto pk = std::make_unique<vta_rec_pack>();
    VTA_TRY(pack_layout(ctx, (int)kind, jobs, n_jobs, tab, off, &elems, &pk->dims));
    pk->device = ctx->device;
    VTA_HIP(hipSetDevice(ctx->device));
    VTA_TRY(pk->carry.alloc(std::max<size_t>(elems, 1) * sizeof(double)));
    VTA_TRY(pk->flags.alloc(n_jobs * sizeof(int32_t)));
    for (size_t j = 0; j < n_jobs; ++j) tab[j].carry = pk->carry.as<double>() + off[j];
    VTA_TRY(upload(ctx, pk->jobs, tab.data(), n_jobs * sizeof(PackJob)));
    *out = pk.release();
    return VTA_OK;
}

vta_status vta_rec_pack_launch(vta_ctx *ctx, const vta_rec_pack *pk) {
    clear_error();
    if (!ctx || !pk) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (pk->device != ctx->device)
        return fail(VTA_ERR_DEVICE_MISMATCH, "pack built on device %d, context on %d", pk->device,
                    ctx->device);
    return pack_launch(ctx, pk->dims, pk->jobs.p, pk->flags.as<int32_t>());
}

vta_status vta_rec_pack_destroy(vta_rec_pack *pk) {
    if (!pk) return VTA_OK;
    (void)hipSetDevice(pk->device);
    delete pk;
    return VTA_OK;
}

}  // extern "C"
