// Launch-shape knobs per context and the autotuner that picks them by timing
// the caller's own workload.
//
// The reference exposes one such knob, ALMA_TILE read from the environment
// (src/cuda/moving_averages/alma_wrapper.rs:314), and fixed kernel policies
// (:18-70).  Here every kernel family with a shape choice has a knob on the
// context (0 = built-in heuristic), and vta_autotune measures candidates with
// HIP events on the context's stream.
#include <cmath>
#include <limits>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

bool knob_value_ok(vta_tune_knob knob, int64_t v) {
    if (v == 0) return true;
    switch (knob) {
        case VTA_TUNE_ALMA_MFMA_VARIANT: return v >= 1 && v <= 11;
        case VTA_TUNE_SCAN_WAVES: return v >= 64 && v <= 65536;
        case VTA_TUNE_FIR_GROUP: return v >= 1 && v <= 256;
        case VTA_TUNE_MS_CHUNK: return v >= 64 && v <= (int64_t)1 << 30;
        case VTA_TUNE_ROLL_FORM: return v == 1 || v == 2;
        default: return false;
    }
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" {

vta_status vta_ctx_set_tuning(vta_ctx *ctx, vta_tune_knob knob, int64_t value) {
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if ((int)knob < 0 || (int)knob >= VTA_TUNE_KNOB_COUNT)
        return fail(VTA_ERR_INVALID_INPUT, "unknown tuning knob %d", (int)knob);
    if (!knob_value_ok(knob, value))
        return fail(VTA_ERR_INVALID_INPUT, "value %lld out of range for knob %d",
                    (long long)value, (int)knob);
    ctx->tune[knob] = value;
    return VTA_OK;
}

vta_status vta_ctx_get_tuning(const vta_ctx *ctx, vta_tune_knob knob, int64_t *value) {
    clear_error();
    if (!ctx || !value) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if ((int)knob < 0 || (int)knob >= VTA_TUNE_KNOB_COUNT)
        return fail(VTA_ERR_INVALID_INPUT, "unknown tuning knob %d", (int)knob);
    *value = ctx->tune[knob];
    return VTA_OK;
}

vta_status vta_autotune(vta_ctx *ctx, vta_tune_knob knob, const int64_t *candidates,
                        size_t n_candidates, vta_enqueue_fn enqueue, void *user, int reps,
                        int64_t *best, float *ms_per_candidate) {
    clear_error();
    if (!ctx || !candidates || !enqueue || !best)
        return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if ((int)knob < 0 || (int)knob >= VTA_TUNE_KNOB_COUNT)
        return fail(VTA_ERR_INVALID_INPUT, "unknown tuning knob %d", (int)knob);
    if (n_candidates == 0 || reps <= 0)
        return fail(VTA_ERR_INVALID_INPUT, "need at least one candidate and one rep");
    for (size_t i = 0; i < n_candidates; ++i)
        if (!knob_value_ok(knob, candidates[i]))
            return fail(VTA_ERR_INVALID_INPUT, "candidate %lld out of range for knob %d",
                        (long long)candidates[i], (int)knob);
    VTA_HIP(hipSetDevice(ctx->device));
    const int64_t saved = ctx->tune[knob];
    Events ev;
    VTA_HIP(hipEventCreate(&ev.a));
    VTA_HIP(hipEventCreate(&ev.b));
    const float inf = std::numeric_limits<float>::infinity();
    std::vector<float> ms(n_candidates, inf);
    std::vector<char> skip(n_candidates, 0);
    auto run = [&](size_t i, int n, float *out_ms) -> vta_status {
        ctx->tune[knob] = candidates[i];
        VTA_HIP(hipEventRecord(ev.a, ctx->stream));
        for (int r = 0; r < n; ++r) VTA_TRY(enqueue(ctx, user));
        VTA_HIP(hipEventRecord(ev.b, ctx->stream));
        VTA_HIP(hipEventSynchronize(ev.b));
        VTA_HIP(hipEventElapsedTime(out_ms, ev.a, ev.b));
        return VTA_OK;
    };
    // one untimed run per candidate; then kRounds interleaved timed rounds, so
    // a clock drift over the search (sustained MFMA + HBM load lowers the
    // clocks within milliseconds) does not favour the candidates timed first;
    // each candidate keeps its least round mean
    constexpr int kRounds = 3;
    vta_status st = VTA_OK;
    for (int round = -1; round < kRounds && st == VTA_OK; ++round) {
        for (size_t i = 0; i < n_candidates && st == VTA_OK; ++i) {
            if (skip[i]) continue;
            float t = 0.0f;
            st = run(i, round < 0 ? 1 : reps, &t);
            if (st == VTA_ERR_LAUNCH_CONFIG_TOO_LARGE && round < 0) {
                (void)hipStreamSynchronize(ctx->stream);
                skip[i] = 1;
                st = VTA_OK;
                continue;
            }
            if (st == VTA_OK && round >= 0) ms[i] = std::min(ms[i], t / (float)reps);
        }
    }
    if (st != VTA_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->tune[knob] = saved;
        return st;
    }
    size_t bi = n_candidates;
    for (size_t i = 0; i < n_candidates; ++i) {
        if (ms_per_candidate) ms_per_candidate[i] = skip[i] ? -1.0f : ms[i];
        if (!skip[i] && (bi == n_candidates || ms[i] < ms[bi])) bi = i;
    }
    if (bi == n_candidates) {
        ctx->tune[knob] = saved;
        return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "no candidate could be launched");
    }
    ctx->tune[knob] = candidates[bi];
    *best = candidates[bi];
    return VTA_OK;
}

}  // extern "C"
