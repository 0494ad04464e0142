// ALMA host entry points: replaces CudaAlma (src/cuda/moving_averages/
// alma_wrapper.rs).  Grid expansion, input validation, per-combo weights and
// kernel selection run here; the arithmetic runs in kernels/alma.hip.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

// expand_grid (alma_wrapper.rs:1550-1631; same as alma.rs:1198-1289).
bool axis_usize(vta_range_usize r, std::vector<size_t> &v) {
    v.clear();
    if (r.step == 0 || r.start == r.end) {
        v.push_back(r.start);
        return true;
    }
    const size_t span = r.start < r.end ? r.end - r.start : r.start - r.end;
    if (span / r.step >= kMaxGridValues) return false;   // > 2^24 values
    if (r.start < r.end) {
        for (size_t x = r.start; x <= r.end; x += r.step) {
            v.push_back(x);
            if (x > r.end - r.step) break;  // next step would pass end (or overflow)
        }
        return true;
    }
    long long x = (long long)r.start, e = (long long)r.end, st = (long long)r.step;
    while (x >= e) {
        v.push_back((size_t)x);
        x -= st;
    }
    return !v.empty();
}

vta_status expand(const vta_alma_batch_range *range, std::vector<vta_alma_params> &out) {
    if (!range) return fail(VTA_ERR_INVALID_INPUT, "range is null");
    std::vector<size_t> ps;
    std::vector<double> os, ss;
    if (!axis_usize(range->period, ps) || !expand_f64(range->offset, os) ||
        !expand_f64(range->sigma, ss))
        return fail(VTA_ERR_INVALID_INPUT,
                    "Invalid range: period=(%zu, %zu, %zu) offset=(%g, %g, %g) "
                    "sigma=(%g, %g, %g) (empty, infinite or above %zu values)",
                    range->period.start, range->period.end, range->period.step,
                    range->offset.start, range->offset.end, range->offset.step,
                    range->sigma.start, range->sigma.end, range->sigma.step, kMaxGridValues);
    size_t cap;
    if (!mul_ok(ps.size(), os.size(), &cap) || !mul_ok(cap, ss.size(), &cap))
        return fail(VTA_ERR_INVALID_INPUT, "range size overflow");
    if (!grid_fits(cap))
        return fail(VTA_ERR_INVALID_INPUT, "parameter grid too large: %zu combinations", cap);
    out.clear();
    out.reserve(cap);
    for (size_t p : ps)
        for (double o : os)
            for (double s : ss) out.push_back({p, o, s});
    if (out.empty()) return fail(VTA_ERR_INVALID_INPUT, "no parameter combinations");
    return VTA_OK;
}

// prepare_batch_inputs / prepare_batch_inputs_device (alma_wrapper.rs:419-505).
vta_status validate_combos(const std::vector<vta_alma_params> &combos, size_t len,
                           size_t first_valid, size_t *max_period) {
    size_t mp = 0;
    for (auto &c : combos) mp = std::max(mp, c.period);
    if (mp == 0 || first_valid >= len || len - first_valid < mp)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu, valid = %zu)",
                    mp, first_valid < len ? len - first_valid : (size_t)0);
    for (auto &c : combos) {
        if (c.period == 0 || !(c.sigma > 0.0) || !(c.offset >= 0.0 && c.offset <= 1.0))
            return fail(VTA_ERR_INVALID_INPUT, "invalid params: period=%zu, offset=%g, sigma=%g",
                        c.period, c.offset, c.sigma);
    }
    if (mp > (size_t)INT32_MAX || len > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "series or period exceeds int32 range");
    *max_period = mp;
    return VTA_OK;
}

// Normalised Gaussian weights in f64 (the batch form of alma.rs:1426-1436:
// w = exp(-(d*d)/s2), divided by their sum).
void weights_f64(size_t period, double offset, double sigma, std::vector<double> &w) {
    w.resize(period);
    double m = offset * (double)(period - 1);
    double s = (double)period / sigma;
    double s2 = 2.0 * s * s;
    double norm = 0.0;
    for (size_t i = 0; i < period; ++i) {
        double d = (double)i - m;
        w[i] = std::exp(-(d * d) / s2);
        norm += w[i];
    }
    double inv = 1.0 / norm;
    for (auto &x : w) x *= inv;
}

// Flat [C][max_period] f32 table of pre-normalised weights, inv_norms = 1
// (the layout alma_batch_device consumes, alma_wrapper.rs:721-747).
void weights_table_f32(const std::vector<vta_alma_params> &combos, size_t max_period,
                       std::vector<float> &w, std::vector<int32_t> &periods,
                       std::vector<float> &inv) {
    size_t C = combos.size();
    w.assign(C * max_period, 0.0f);
    periods.resize(C);
    inv.assign(C, 1.0f);
    std::vector<double> wd;
    for (size_t c = 0; c < C; ++c) {
        weights_f64(combos[c].period, combos[c].offset, combos[c].sigma, wd);
        for (size_t k = 0; k < combos[c].period; ++k) w[c * max_period + k] = (float)wd[k];
        periods[c] = (int32_t)combos[c].period;
    }
}

constexpr int kFp32TT = 1024;
constexpr int kFp32CB = 8;
constexpr int kTmTT = 256;

vta_status launch_fp32(vta_ctx *ctx, const float *d_prices, const float *d_weights,
                       const int32_t *d_periods, const float *d_inv, int max_period, int T, int C,
                       int fv, float *d_out, bool time_major) {
    size_t lds_limit = ctx->props.sharedMemPerBlock;
    hipFunction_t f;
    if (!time_major) {
        int p4max = (max_period + 3) & ~3;
        size_t smem = (size_t)(p4max + kFp32TT + 8 + kFp32CB * p4max) * sizeof(float);
        if (smem <= lds_limit) {
            VTA_TRY(get_function(ctx, "alma", "alma_batch_f32", &f));
            dim3 grid((T + kFp32TT - 1) / kFp32TT, (C + kFp32CB - 1) / kFp32CB, 1);
            // grid.y chunking (alma_wrapper.rs:357-363), with the per-combo
            // pointers offset by the chunk start (the reference does not, :1105-1128)
            const unsigned max_gy = (unsigned)ctx->props.maxGridSize[1];
            for (unsigned y0 = 0; y0 < grid.y; y0 += max_gy) {
                unsigned gy = std::min(max_gy, grid.y - y0);
                size_t c0 = (size_t)y0 * kFp32CB;
                int cn = (int)std::min((size_t)C - c0, (size_t)gy * kFp32CB);
                VTA_TRY(launch(ctx, f, dim3(grid.x, gy, 1), dim3(256), (unsigned)smem, d_prices,
                               d_weights + c0 * max_period, d_periods + c0, d_inv + c0,
                               max_period, T, cn, fv, d_out + c0 * (size_t)T));
            }
            return VTA_OK;
        }
    } else {
        size_t smem = (size_t)(max_period + kTmTT + 4 + max_period * 64) * sizeof(float);
        if (smem <= lds_limit && (size_t)(C + 63) / 64 <= (size_t)ctx->props.maxGridSize[1]) {
            VTA_TRY(get_function(ctx, "alma", "alma_batch_f32_tm", &f));
            dim3 grid((T + kTmTT - 1) / kTmTT, (C + 63) / 64, 1);
            return launch(ctx, f, grid, dim3(256), (unsigned)smem, d_prices, d_weights,
                          d_periods, d_inv, max_period, T, C, fv, d_out);
        }
    }
    VTA_TRY(get_function(ctx, "alma", "alma_batch_f32_direct", &f));
    size_t total = (size_t)T * (size_t)C;
    size_t blocks = (total + 255) / 256;
    if (blocks > (size_t)ctx->props.maxGridSize[0])
        return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "launch config too large: grid=(%zu,1,1)",
                    blocks);
    int tm = time_major ? 1 : 0;
    return launch(ctx, f, dim3((unsigned)blocks), dim3(256), 0, d_prices, d_weights, d_periods,
                  d_inv, max_period, T, C, fv, d_out, tm);
}

}  // namespace

// Host-side mirror of struct alma_mf_args in kernels/alma.hip (same field
// order and types, so the same layout).
struct alma_mf_args_host {
    const float *prices;
    const void *frags;
    const void *nblocks;
    const int32_t *sorted_period;
    const int32_t *sorted_row;
    const float *w_f32;
    float *out;
    int series_len;
    int first_valid;
    int n_nblocks;
    int tiles_per_wave;
    int n_tiles;
    int kmax;
};

struct alma_mfg_args_host {
    const float *prices;
    const void *frags;
    const void *nblocks;
    const void *groups;
    const int32_t *sorted_period;
    const int32_t *sorted_row;
    const float *w_f32;
    float *out;
    int series_len;
    int first_valid;
    int n_groups;
    int n_tiles;
    int kmax;
    int group_base;
};
static_assert(sizeof(alma_mfg_args_host) == 88, "alma_mfg_args layout");

// Host mirror of struct alma_sw_args (the diagonal-form sweep kernels).
struct alma_sw_args_host {
    const float *prices;
    const void *frags;
    const void *nblocks;
    const void *groups;
    const int32_t *sorted_period;
    const int32_t *sorted_row;
    const float *w_f32;
    const void *cscale;
    float *out;
    int series_len;
    int first_valid;
    int n_groups;
    int n_tiles;
    int kmax;
    int group_base;
};
static_assert(sizeof(alma_sw_args_host) == 96, "alma_sw_args layout");

struct vta_alma_sweep_plan {
    size_t series_len = 0;
    AlmaTables t;   // host copy of the tables and their layout
    DevBuf dev;     // the tables on the device (one buffer, t.blob's layout)
};

namespace {

constexpr int kMfTilesPerWave = 16;

// ALMA_MF_FORM=wave selects the one-wave-per-N-block kernels; the default is
// the shared-staging workgroup form.
bool mf_shared_staging() {
    const char *e = getenv("ALMA_MF_FORM");
    return !(e && strcmp(e, "wave") == 0);
}
constexpr size_t kMfMaxPeriod = 256;

// Diagonal-form sweep kernels (alma_sweep_*): tiles per workgroup run, and the
// variant the F16 form runs by default (direct stores: 299 us against 311 us
// for the transposed stores on the 200,000 x 1,250 sweep at sustained clocks).
constexpr int kSweepRunTiles = 16;
constexpr int kSweepDefaultVariant = 9;

int mfg_stride_host(int npair) { return ((npair + 15) >> 5) * 32 + 16; }

// Dynamic LDS of alma_sweep_* (the layout in alma_sw_entry): prices, sub-run
// anchors and scales, tile flags, block statistics, the staged sub-runs and,
// with the transposed stores, 32 x 36 tiles: one per wave, or rs per wave
// when the stores are deferred.
size_t sweep_lds_bytes(int kmax, int rs, bool xpose, bool defer) {
    const int nsub = kSweepRunTiles / rs;
    const int xsn = 32 * kSweepRunTiles + kmax + 33;
    const int head = (xsn + 3 * nsub + kSweepRunTiles + 96 + 3) & ~3;
    const int stage = nsub * 4 * mfg_stride_host(kmax / 2 + 16 * rs);
    const int tiles = defer ? 4 * rs : xpose ? 4 : 0;
    return 4 * ((size_t)head + (size_t)stage + (size_t)tiles * 32 * 36);
}

bool mf_env_override() {
    return getenv("ALMA_MF_FORM") || getenv("ALMA_MF_TILES") || getenv("ALMA_MF_TJ") ||
           getenv("ALMA_MF_KSPLIT");
}

// Built-in split-BF16 kernel variant (vta_tune_knob numbering), from an
// autotuner sweep over series lengths 2^12..2^20 and grids of 8..800 combos
// (every variant timed in interleaved rounds): with at most 64 combos one wave
// per N-block over 8 tiles (variant 5) is within 4% of the best everywhere;
// up to 2^27 outputs the 8-tile workgroups (2) are 5-35% faster than the
// 16-tile ones, whose staging pays off above that.  Above 2^27 outputs the
// diagonal form with direct stores (9) is the fastest: 250 us against 253
// (8), 257 (11) and 259 (1) on the 200,000 x 1,250 sweep at sustained clocks.
int mf_default_variant(size_t T, size_t C) {
    if (C <= 64) return 5;
    if ((double)T * (double)C <= 134217728.0) return 2;
    return 9;
}

// Appends one table to the blob at the next 256-byte boundary.
template <typename V>
void put_section(AlmaTables &t, int sec, const V &v) {
    const size_t off = (t.blob.size() + 255) & ~(size_t)255;
    const size_t bytes = v.size() * sizeof(typename V::value_type);
    t.blob.resize(off + bytes);
    if (bytes) memcpy(t.blob.data() + off, v.data(), bytes);
    t.off[sec] = off;
}

template <typename P>
const P *section(const AlmaTables &t, const unsigned char *base, int sec) {
    return reinterpret_cast<const P *>(base + t.off[sec]);
}

// IEEE binary16 round-to-nearest-even of a double (|v| < 65520), as bits,
// rounded once from the double (no f32 step): the significand is cut to 11
// bits for a normal result or to the 2^-24 quantum for a subnormal one, ties
// to the even neighbour; a carry out of the significand bumps the exponent.
uint16_t f16_rne(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    const uint16_t sign = (uint16_t)((b >> 48) & 0x8000u);
    const int ex = (int)((b >> 52) & 0x7ff);
    if (ex == 0) return sign;                    // zero or a double subnormal
    const int e = ex - 1023;                     // |v| = 1.m 2^e
    if (e < -25) return sign;                    // below half the smallest subnormal
    const uint64_t sig = (b & ((1ull << 52) - 1)) | (1ull << 52);
    const int shift = e >= -14 ? 42 : 28 - e;    // dropped low bits
    const uint64_t q = sig >> shift;
    const uint64_t rem = sig & ((1ull << shift) - 1);
    const uint64_t half = 1ull << (shift - 1);
    uint64_t r = q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
    if (e >= -14) {
        int E = e + 15;
        if (r == 2048) {                         // rounded up to the next binade
            r = 1024;
            ++E;
        }
        return (uint16_t)(sign | (uint16_t)(E << 10) | (uint16_t)(r - 1024));
    }
    return (uint16_t)(sign | (uint16_t)r);       // subnormal (r == 1024: smallest normal)
}

double f16_value(uint16_t h) {
    const int ex = (h >> 10) & 0x1f, mant = h & 0x3ff;
    const double mag = ex == 0 ? std::ldexp((double)mant, -24)
                               : std::ldexp(1.0 + mant / 1024.0, ex - 15);
    return (h & 0x8000u) ? -mag : mag;
}

// Builds the MFMA operand tables.  Combos are sorted by period (stable) and
// cut into N-blocks of 32; block b contracts over K_b = round_up(max period of
// the block, 16) taps with shorter periods zero padded at the front.  Each
// weight is split hi / lo and stored in v_mfma_f32_32x32x16 B-operand order:
// for K-step s and half h, lane l holds taps 16s + 8(l/32) + j (j < 8) of
// combo l%32.  BF16: hi = bf16(w), lo = bf16(w - hi).  F16: the combo's
// weights are first scaled by 2^sw (largest in [0.5, 1)), hi = f16(w_s), lo =
// f16(w_s - hi); cscale holds {2^-sw, (2^-22 sum w_s + sum |w_s - hi - lo| +
// 2^-11 sum |lo|) 2^-sw} per sorted combo, the factor of the kernel's
// split-error bound (rounded up).
void build_mfma_tables(AlmaTables &t, const std::vector<vta_alma_params> &combos, bool f16) {
    const size_t C = combos.size();
    std::vector<size_t> order(C);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return combos[a].period < combos[b].period; });
    const int nb = (int)((C + 31) / 32);
    const size_t cpad = (size_t)nb * 32;
    std::vector<int32_t> sp(cpad, 1), sr(cpad, -1);
    std::vector<int32_t> nbd((size_t)nb * 4, 0);
    std::vector<uint16_t> frags;
    std::vector<float> csc(2 * cpad, 0.0f);
    for (size_t i = 0; i < cpad; ++i) csc[2 * i] = 1.0f;
    int kmax = 0;
    for (int b = 0; b < nb; ++b) {
        size_t maxp = 0;
        for (int n = 0; n < 32; ++n) {
            size_t i = (size_t)b * 32 + n;
            if (i < C) maxp = std::max(maxp, combos[order[i]].period);
        }
        int K = (int)((maxp + 15) / 16) * 16;
        kmax = std::max(kmax, K);
        nbd[b * 4 + 0] = K / 16;
        nbd[b * 4 + 1] = (int32_t)(frags.size() / 8);  // offset in 16-byte chunks
        nbd[b * 4 + 2] = b * 32;
        // hi / lo of W[k][n], zero padded at the front (F16: scaled by 2^sw per
        // combo), each weight split once
        std::vector<uint16_t> Wh((size_t)K * 32, 0), Wl((size_t)K * 32, 0);
        std::vector<double> wd;
        for (int n = 0; n < 32; ++n) {
            size_t i = (size_t)b * 32 + n;
            if (i >= C) continue;
            const auto &cp = combos[order[i]];
            weights_f64(cp.period, cp.offset, cp.sigma, wd);
            const size_t lead = (size_t)K - cp.period;
            if (f16) {
                const double wmax = *std::max_element(wd.begin(), wd.end());
                int e;
                (void)std::frexp(wmax, &e);            // wmax = f 2^e, f in [0.5, 1)
                const double scale = std::ldexp(1.0, -e);
                double sum_ws = 0.0, eps = 0.0, lsum = 0.0;
                for (size_t k = 0; k < cp.period; ++k) {
                    const double ws = wd[k] * scale;
                    const uint16_t hi = f16_rne(ws);
                    const double hv = f16_value(hi);
                    const uint16_t lo = f16_rne(ws - hv);
                    const double lv = f16_value(lo);
                    sum_ws += ws;
                    eps += std::fabs(ws - hv - lv);
                    lsum += std::fabs(lv);
                    Wh[(lead + k) * 32 + n] = hi;
                    Wl[(lead + k) * 32 + n] = lo;
                }
                const double bound =
                    (std::ldexp(sum_ws, -22) + eps + std::ldexp(lsum, -11)) / scale;
                csc[2 * i + 0] = (float)(1.0 / scale);
                csc[2 * i + 1] = std::nextafter((float)(bound * (1.0 + 1e-6)), INFINITY);
            } else {
                for (size_t k = 0; k < cp.period; ++k) {
                    const double w = wd[k];
                    const uint16_t hi = bf16_rne((float)w);
                    Wh[(lead + k) * 32 + n] = hi;
                    Wl[(lead + k) * 32 + n] = bf16_rne((float)(w - (double)bf16_to_f32(hi)));
                }
            }
            sp[i] = (int32_t)cp.period;
            sr[i] = (int32_t)order[i];
        }
        for (int s = 0; s < K / 16; ++s)
            for (int h = 0; h < 2; ++h) {
                const std::vector<uint16_t> &src = h == 0 ? Wh : Wl;
                for (int l = 0; l < 64; ++l)
                    for (int j = 0; j < 8; ++j)
                        frags.push_back(src[(size_t)(16 * s + 8 * (l / 32) + j) * 32 + (l % 32)]);
            }
    }
    // f32 weight table for the non-finite fallback, [cpad][kmax], taps from
    // the window start.
    std::vector<float> wf(cpad * (size_t)kmax, 0.0f);
    std::vector<double> wd;
    for (size_t i = 0; i < C; ++i) {
        const auto &cp = combos[order[i]];
        weights_f64(cp.period, cp.offset, cp.sigma, wd);
        for (size_t k = 0; k < cp.period; ++k) wf[i * kmax + k] = (float)wd[k];
    }
    // groups of up to 4 consecutive N-blocks (ascending K, so the last one has
    // the group's largest K) for the shared-staging kernel
    std::vector<int32_t> grp;
    for (int b = 0; b < nb; b += 4) {
        const int cnt = std::min(4, nb - b);
        grp.push_back(b);
        grp.push_back(cnt);
        grp.push_back(nbd[(b + cnt - 1) * 4 + 0]);
        grp.push_back(0);
    }
    t.n_groups = (int)(grp.size() / 4);
    t.n_groups_klo = 0;
    while (t.n_groups_klo < t.n_groups && grp[t.n_groups_klo * 4 + 2] <= 8) ++t.n_groups_klo;
    put_section(t, AlmaTables::GROUPS, grp);
    put_section(t, AlmaTables::FRAGS, frags);
    put_section(t, AlmaTables::NBLOCKS, nbd);
    put_section(t, AlmaTables::SORTED_PERIOD, sp);
    put_section(t, AlmaTables::SORTED_ROW, sr);
    put_section(t, AlmaTables::W_F32, wf);
    if (f16) put_section(t, AlmaTables::CSCALE, csc);
    t.n_nblocks = nb;
    t.kmax = kmax;
}

vta_alma_batch_kernel resolve_kernel(vta_alma_batch_kernel policy, size_t n_combos,
                                     size_t max_period) {
    if (policy != VTA_ALMA_BATCH_AUTO) return policy;
    // The matrix-core form needs wide sweeps to fill 32-combo N-blocks; its
    // split-F16 form carries a checked error bound.
    if (n_combos >= 16 && max_period <= kMfMaxPeriod) return VTA_ALMA_BATCH_MFMA_F16X3;
    return VTA_ALMA_BATCH_FP32;
}

vta_status tables_build(const std::vector<vta_alma_params> &combos, size_t max_period,
                        vta_alma_batch_kernel kernel, AlmaTables &t) {
    t = AlmaTables();
    t.kernel = kernel;
    t.n_combos = combos.size();
    t.max_period = max_period;
    if (kernel == VTA_ALMA_BATCH_MFMA_BF16X3 || kernel == VTA_ALMA_BATCH_MFMA_F16X3) {
        if (max_period > kMfMaxPeriod)
            return fail(VTA_ERR_INVALID_POLICY,
                        "invalid policy: MFMA sweep supports periods <= %zu (got %zu)",
                        kMfMaxPeriod, max_period);
        build_mfma_tables(t, combos, kernel == VTA_ALMA_BATCH_MFMA_F16X3);
        return VTA_OK;
    }
    if (kernel != VTA_ALMA_BATCH_FP32) return fail(VTA_ERR_INVALID_POLICY, "invalid policy");
    std::vector<float> w, inv;
    std::vector<int32_t> per;
    weights_table_f32(combos, max_period, w, per, inv);
    put_section(t, AlmaTables::W, w);
    put_section(t, AlmaTables::PERIODS, per);
    put_section(t, AlmaTables::INV, inv);
    return VTA_OK;
}

vta_status plan_create(vta_ctx *ctx, const std::vector<vta_alma_params> &combos, size_t T,
                       size_t max_period, vta_alma_batch_kernel kernel,
                       vta_alma_sweep_plan **out) {
    auto *plan = new vta_alma_sweep_plan();
    plan->series_len = T;
    vta_status st = tables_build(combos, max_period, kernel, plan->t);
    if (st == VTA_OK) st = upload(ctx, plan->dev, plan->t.blob.data(), plan->t.blob.size());
    if (st != VTA_OK) {
        delete plan;
        return st;
    }
    *out = plan;
    return VTA_OK;
}

// Enqueues the sweep over T bars with the tables at device address `base`
// (laid out as t.blob).
vta_status tables_launch(vta_ctx *ctx, const AlmaTables &t, const unsigned char *base,
                         const float *d_prices, size_t fv, float *d_out, size_t series_len) {
    const int T = (int)series_len;
    // kernel variant: the context's VTA_TUNE_ALMA_MFMA_VARIANT knob (1-4 shared
    // staging, 5-7 one wave per N-block), else the ALMA_MF_* environment
    // knobs, else the shared-staging form with 16 tiles
    int variant = (int)ctx->tune[VTA_TUNE_ALMA_MFMA_VARIANT];
    if (variant == 0 && !mf_env_override()) variant = mf_default_variant(T, t.n_combos);
    const bool f16 = t.kernel == VTA_ALMA_BATCH_MFMA_F16X3;
    if (f16 && variant < 8) variant = kSweepDefaultVariant;   // F16 runs the diagonal form only
    // the deferred-store kernels write whole float4s: rows must be 16-byte aligned
    if (variant == 11 && (T & 3) != 0) variant = 8;
    if ((t.kernel == VTA_ALMA_BATCH_MFMA_BF16X3 || f16) && variant >= 8) {
        // diagonal form: 4 N-blocks x 16 tiles per workgroup, sub-runs of RS tiles
        const int rs = variant == 10 ? 4 : 2;
        const bool xpose = variant == 8 || variant == 11;
        const bool defer = variant == 11;
        static const char *const kNames[2][4] = {
            {"alma_sweep_bf16x3", "alma_sweep_bf16x3_d", "alma_sweep_bf16x3_r4",
             "alma_sweep_bf16x3_p"},
            {"alma_sweep_f16x3", "alma_sweep_f16x3_d", "alma_sweep_f16x3_r4", "alma_sweep_f16x3_p"}};
        alma_sw_args_host a{};
        a.prices = d_prices;
        a.frags = section<void>(t, base, AlmaTables::FRAGS);
        a.nblocks = section<void>(t, base, AlmaTables::NBLOCKS);
        a.groups = section<void>(t, base, AlmaTables::GROUPS);
        a.sorted_period = section<int32_t>(t, base, AlmaTables::SORTED_PERIOD);
        a.sorted_row = section<int32_t>(t, base, AlmaTables::SORTED_ROW);
        a.w_f32 = section<float>(t, base, AlmaTables::W_F32);
        a.cscale = f16 ? section<void>(t, base, AlmaTables::CSCALE) : nullptr;
        a.out = d_out;
        a.series_len = T;
        a.first_valid = (int)fv;
        a.n_tiles = (T + 31) / 32;
        a.kmax = t.kmax;
        a.group_base = 0;
        a.n_groups = t.n_groups;
        const size_t smem = sweep_lds_bytes(t.kmax, rs, xpose, defer);
        if (smem > ctx->props.sharedMemPerBlock)
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "MFMA staging needs %zu bytes of LDS",
                        smem);
        hipFunction_t fn;
        VTA_TRY(get_function(ctx, "alma", kNames[f16 ? 1 : 0][variant - 8], &fn));
        ctx->alma_selected = t.kernel;
        const unsigned gx = (unsigned)((a.n_tiles + kSweepRunTiles - 1) / kSweepRunTiles);
        return launch(ctx, fn, dim3(gx, (unsigned)t.n_groups, 1), dim3(256), (unsigned)smem, a);
    }
    const bool shared = variant == 0 ? mf_shared_staging() : variant <= 4;
    if (t.kernel == VTA_ALMA_BATCH_MFMA_BF16X3 && shared) {
        // shared-staging form: 4 N-blocks per workgroup, R tiles per run
        // (ALMA_MF_TILES=8|16, default 16)
        int tiles = 16;
        bool tj2 = false, ks_req = false;
        if (variant == 0) {
            if (const char *e = getenv("ALMA_MF_TILES")) {
                int v = atoi(e);
                if (v == 8 || v == 16) tiles = v;
            }
            // ALMA_MF_TJ=2: two tiles per pass; ALMA_MF_KSPLIT=1: separate
            // launches for K > 128 and K <= 128
            const char *tj_e = getenv("ALMA_MF_TJ");
            const char *ks_e = getenv("ALMA_MF_KSPLIT");
            tj2 = tiles == 16 && tj_e && atoi(tj_e) == 2;
            ks_req = tiles == 16 && ks_e && atoi(ks_e) == 1;
        } else {
            tiles = variant == 2 ? 8 : 16;
            tj2 = variant == 3;
            ks_req = variant == 4;
        }
        const bool ksplit = ks_req && t.n_groups_klo > 0 && t.n_groups_klo < t.n_groups;
        alma_mfg_args_host a{};
        a.prices = d_prices;
        a.frags = section<void>(t, base, AlmaTables::FRAGS);
        a.nblocks = section<void>(t, base, AlmaTables::NBLOCKS);
        a.groups = section<void>(t, base, AlmaTables::GROUPS);
        a.sorted_period = section<int32_t>(t, base, AlmaTables::SORTED_PERIOD);
        a.sorted_row = section<int32_t>(t, base, AlmaTables::SORTED_ROW);
        a.w_f32 = section<float>(t, base, AlmaTables::W_F32);
        a.out = d_out;
        a.series_len = T;
        a.first_valid = (int)fv;
        a.n_tiles = (T + 31) / 32;
        a.kmax = t.kmax;
        const int npair = t.kmax / 2 + 15;
        const int stride = ((npair + 15) >> 5) * 32 + 16;
        // prices, anchors, flags, the staged dwords (rounded to 16 bytes) and
        // four per-wave 32 x 36-float output transpose tiles
        const size_t smem =
            4 * ((size_t)(32 * tiles + t.kmax + 32) + 2 * (size_t)tiles +
                 (((size_t)tiles * 4 * (size_t)stride + 3) & ~(size_t)3) + 4 * 32 * 36);
        if (smem > ctx->props.sharedMemPerBlock)
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "MFMA staging needs %zu bytes of LDS",
                        smem);
        const unsigned gx = (unsigned)((a.n_tiles + tiles - 1) / tiles);
        ctx->alma_selected = VTA_ALMA_BATCH_MFMA_BF16X3;
        struct Part { const char *name; int base, count; };
        Part parts[2];
        int np = 0;
        if (ksplit) {
            parts[np++] = {tj2 ? "alma_batch_mfma_bf16x3_wg16_t2" : "alma_batch_mfma_bf16x3_wg16",
                           t.n_groups_klo, t.n_groups - t.n_groups_klo};
            parts[np++] = {tj2 ? "alma_batch_mfma_bf16x3_wg16_t2_klo" : "alma_batch_mfma_bf16x3_wg16_klo",
                           0, t.n_groups_klo};
        } else {
            parts[np++] = {tiles == 8 ? "alma_batch_mfma_bf16x3_wg8"
                                      : (tj2 ? "alma_batch_mfma_bf16x3_wg16_t2"
                                             : "alma_batch_mfma_bf16x3_wg16"),
                           0, t.n_groups};
        }
        for (int i = 0; i < np; ++i) {
            hipFunction_t f;
            VTA_TRY(get_function(ctx, "alma", parts[i].name, &f));
            a.group_base = parts[i].base;
            a.n_groups = parts[i].count;
            VTA_TRY(launch(ctx, f, dim3(gx, (unsigned)parts[i].count, 1), dim3(256),
                           (unsigned)smem, a));
        }
        return VTA_OK;
    }
    if (t.kernel == VTA_ALMA_BATCH_MFMA_BF16X3) {
        // tiles per wave: 16 by default; ALMA_MF_TILES=8|32 overrides (tuning knob,
        // like the reference's ALMA_TILE, alma_wrapper.rs:314)
        int tiles = kMfTilesPerWave;
        if (variant >= 5) {
            tiles = variant == 5 ? 8 : variant == 6 ? 16 : 32;
        } else if (const char *e = getenv("ALMA_MF_TILES")) {
            int v = atoi(e);
            if (v == 8 || v == 16 || v == 32) tiles = v;
        }
        const char *name = tiles == 8    ? "alma_batch_mfma_bf16x3_r8"
                           : tiles == 32 ? "alma_batch_mfma_bf16x3_r32"
                                         : "alma_batch_mfma_bf16x3";
        hipFunction_t f;
        VTA_TRY(get_function(ctx, "alma", name, &f));
        alma_mf_args_host a;
        a.prices = d_prices;
        a.frags = section<void>(t, base, AlmaTables::FRAGS);
        a.nblocks = section<void>(t, base, AlmaTables::NBLOCKS);
        a.sorted_period = section<int32_t>(t, base, AlmaTables::SORTED_PERIOD);
        a.sorted_row = section<int32_t>(t, base, AlmaTables::SORTED_ROW);
        a.w_f32 = section<float>(t, base, AlmaTables::W_F32);
        a.out = d_out;
        a.series_len = T;
        a.first_valid = (int)fv;
        a.n_nblocks = t.n_nblocks;
        a.tiles_per_wave = tiles;
        a.n_tiles = (T + 31) / 32;
        a.kmax = t.kmax;
        dim3 grid((a.n_tiles + tiles - 1) / tiles, t.n_nblocks, 1);
        ctx->alma_selected = VTA_ALMA_BATCH_MFMA_BF16X3;
        return launch(ctx, f, grid, dim3(64), 0, a);
    }
    ctx->alma_selected = VTA_ALMA_BATCH_FP32;
    return launch_fp32(ctx, d_prices, section<float>(t, base, AlmaTables::W),
                       section<int32_t>(t, base, AlmaTables::PERIODS),
                       section<float>(t, base, AlmaTables::INV), (int)t.max_period, T,
                       (int)t.n_combos, (int)fv, d_out, false);
}

vta_status plan_launch(vta_ctx *ctx, const vta_alma_sweep_plan *plan, const float *d_prices,
                       size_t fv, float *d_out, size_t series_len = 0) {
    // the tables do not depend on the series length; a caller may launch a
    // plan over any length that covers its longest period
    return tables_launch(ctx, plan->t, plan->dev.as<unsigned char>(), d_prices, fv, d_out,
                         series_len ? series_len : plan->series_len);
}
// True when the finite bars of a host series take both signs: the outputs
// then cross zero, where the F16 form's error check would send most tiles to
// the FP32 dot, so AUTO takes the FP32 kernel directly.
bool changes_sign(const float *x, size_t n, size_t fv) {
    bool neg = false, pos = false;
    for (size_t i = fv; i < n; ++i) {
        if (!std::isfinite(x[i])) continue;
        neg = neg || x[i] < 0.0f;
        pos = pos || x[i] > 0.0f;
        if (neg && pos) return true;
    }
    return false;
}

// run_batch_with_prices_device (alma_wrapper.rs:683-781).  host_prices (may
// be NULL) lets AUTO see the series.
vta_status run_batch(vta_ctx *ctx, const float *d_prices, size_t T, size_t fv,
                     const std::vector<vta_alma_params> &combos, size_t max_period,
                     vta_device_array_f32 *out, const float *host_prices = nullptr) {
    size_t n_out;
    if (!mul_ok(combos.size(), T, &n_out) || !mul_ok(n_out, sizeof(float), &n_out))
        return fail(VTA_ERR_INVALID_INPUT, "n_combos*series_len overflow");
    vta_alma_batch_kernel k = resolve_kernel(ctx->alma_policy, combos.size(), max_period);
    if (ctx->alma_policy == VTA_ALMA_BATCH_AUTO && k != VTA_ALMA_BATCH_FP32 && host_prices &&
        changes_sign(host_prices, T, fv))
        k = VTA_ALMA_BATCH_FP32;
    vta_alma_sweep_plan *plan = nullptr;
    VTA_TRY(plan_create(ctx, combos, T, max_period, k, &plan));
    DevBuf dout;
    vta_status st = dout.alloc(n_out);
    if (st == VTA_OK) st = plan_launch(ctx, plan, d_prices, fv, dout.as<float>());
    if (st == VTA_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = fail(VTA_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
    }
    delete plan;
    if (st != VTA_OK) return st;
    out->ptr = static_cast<float *>(dout.release());
    out->rows = combos.size();
    out->cols = T;
    return VTA_OK;
}

vta_status prepare_host(const float *prices, size_t len, const vta_alma_batch_range *range,
                        std::vector<vta_alma_params> &combos, size_t *fv, size_t *max_period) {
    if (!prices || len == 0) return fail(VTA_ERR_INVALID_INPUT, "empty data");
    *fv = first_valid_f32(prices, len);
    if (*fv >= len) return fail(VTA_ERR_INVALID_INPUT, "all values are NaN");
    VTA_TRY(expand(range, combos));
    return validate_combos(combos, len, *fv, max_period);
}

}  // namespace

namespace vta {

vta_status alma_expand_range(const vta_alma_batch_range *range,
                             std::vector<vta_alma_params> &out) {
    return expand(range, out);
}

vta_status alma_grid_plan_create(vta_ctx *ctx, const vta_alma_batch_range *range,
                                 vta_alma_sweep_plan **out, size_t *max_period,
                                 size_t *n_combos) {
    std::vector<vta_alma_params> combos;
    VTA_TRY(expand(range, combos));
    size_t mp = 0;
    for (auto &c : combos) {
        if (c.period == 0 || !(c.sigma > 0.0) || !(c.offset >= 0.0 && c.offset <= 1.0))
            return fail(VTA_ERR_INVALID_INPUT, "invalid params: period=%zu, offset=%g, sigma=%g",
                        c.period, c.offset, c.sigma);
        mp = std::max(mp, c.period);
    }
    if (mp > (size_t)INT32_MAX) return fail(VTA_ERR_INVALID_INPUT, "period too large");
    *max_period = mp;
    *n_combos = combos.size();
    vta_alma_batch_kernel k = resolve_kernel(ctx->alma_policy, combos.size(), mp);
    return plan_create(ctx, combos, mp, mp, k, out);
}

vta_status alma_grid_plan_launch(vta_ctx *ctx, const vta_alma_sweep_plan *plan,
                                 const float *d_prices, size_t series_len, size_t fv,
                                 float *d_out) {
    if (series_len == 0 || series_len > (size_t)INT32_MAX || fv >= series_len ||
        series_len - fv < plan->t.max_period)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)",
                    plan->t.max_period);
    return plan_launch(ctx, plan, d_prices, fv, d_out, series_len);
}

void alma_grid_plan_destroy(vta_alma_sweep_plan *plan) { delete plan; }

vta_status alma_tables_for_range(vta_ctx *ctx, const vta_alma_batch_range *range, size_t len,
                                 size_t fv, AlmaTables &t) {
    std::vector<vta_alma_params> combos;
    size_t mp;
    VTA_TRY(expand(range, combos));
    VTA_TRY(validate_combos(combos, len, fv, &mp));
    return tables_build(combos, mp, resolve_kernel(ctx->alma_policy, combos.size(), mp), t);
}

vta_status alma_tables_launch(vta_ctx *ctx, const AlmaTables &t, const void *d_tables,
                              const float *d_prices, size_t len, size_t fv, float *d_out) {
    return tables_launch(ctx, t, static_cast<const unsigned char *>(d_tables), d_prices, fv,
                         d_out, len);
}

// Host mirror of struct alma_f32_job (the packed FP32 kernel's job record).
struct alma_f32_job_host {
    const float *prices;
    const float *weights;
    const int32_t *periods;
    const float *inv_norms;
    float *out;
    int max_period;
    int series_len;
    int n_combos;
    int first_valid;
};
static_assert(sizeof(alma_f32_job_host) == 56, "alma_f32_job layout");

vta_status alma_pack_build(vta_ctx *ctx, vta_alma_batch_kernel kernel,
                           const std::vector<AlmaPackSrc> &jobs, AlmaPackImage &img) {
    const bool fp32 = kernel == VTA_ALMA_BATCH_FP32;
    const size_t rec = fp32 ? sizeof(alma_f32_job_host) : sizeof(alma_sw_args_host);
    const size_t lds_limit = ctx->props.sharedMemPerBlock;
    img = AlmaPackImage();
    img.kernel = kernel;
    img.records.assign(jobs.size() * rec, 0);
    struct Item {
        int32_t job, x, y, w;
    };
    std::vector<Item> items;
    for (size_t j = 0; j < jobs.size(); ++j) {
        const AlmaTables &t = *jobs[j].t;
        const unsigned char *tb = jobs[j].d_tables;
        const size_t T = jobs[j].len;
        if (t.kernel != kernel)
            return fail(VTA_ERR_INVALID_POLICY, "job %zu: tables of another kernel", j);
        if (T == 0 || T > (size_t)INT32_MAX || jobs[j].fv >= T)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: series_len or first_valid out of range",
                        j);
        size_t need;
        if (fp32) {
            const size_t p4 = (t.max_period + 3) & ~(size_t)3;
            need = (p4 + kFp32TT + 8 + kFp32CB * p4) * sizeof(float);
        } else {
            need = sweep_lds_bytes(t.kmax, 2, false, false);
        }
        if (need > lds_limit)
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE,
                        "job %zu: its windows need %zu bytes of LDS (limit %zu)", j, need,
                        lds_limit);
        img.smem = std::max(img.smem, need);
        if (fp32) {
            alma_f32_job_host r{};
            r.prices = jobs[j].d_prices;
            r.weights = section<float>(t, tb, AlmaTables::W);
            r.periods = section<int32_t>(t, tb, AlmaTables::PERIODS);
            r.inv_norms = section<float>(t, tb, AlmaTables::INV);
            r.out = jobs[j].d_out;
            r.max_period = (int)t.max_period;
            r.series_len = (int)T;
            r.n_combos = (int)t.n_combos;
            r.first_valid = (int)jobs[j].fv;
            memcpy(img.records.data() + j * rec, &r, sizeof r);
            const int gx = (int)((T + kFp32TT - 1) / kFp32TT);
            const int gy = (int)((t.n_combos + kFp32CB - 1) / kFp32CB);
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x)
                    items.push_back({(int32_t)j, x, y, (int)t.max_period});
        } else {
            alma_sw_args_host a{};
            a.prices = jobs[j].d_prices;
            a.frags = section<void>(t, tb, AlmaTables::FRAGS);
            a.nblocks = section<void>(t, tb, AlmaTables::NBLOCKS);
            a.groups = section<void>(t, tb, AlmaTables::GROUPS);
            a.sorted_period = section<int32_t>(t, tb, AlmaTables::SORTED_PERIOD);
            a.sorted_row = section<int32_t>(t, tb, AlmaTables::SORTED_ROW);
            a.w_f32 = section<float>(t, tb, AlmaTables::W_F32);
            a.cscale = kernel == VTA_ALMA_BATCH_MFMA_F16X3
                           ? section<void>(t, tb, AlmaTables::CSCALE)
                           : nullptr;
            a.out = jobs[j].d_out;
            a.series_len = (int)T;
            a.first_valid = (int)jobs[j].fv;
            a.n_groups = t.n_groups;
            a.n_tiles = (int)((T + 31) / 32);
            a.kmax = t.kmax;
            a.group_base = 0;
            memcpy(img.records.data() + j * rec, &a, sizeof a);
            const int32_t *grp =
                reinterpret_cast<const int32_t *>(t.blob.data() + t.off[AlmaTables::GROUPS]);
            const int runs = (a.n_tiles + kSweepRunTiles - 1) / kSweepRunTiles;
            for (int g = 0; g < t.n_groups; ++g)
                for (int r = 0; r < runs; ++r) items.push_back({(int32_t)j, r, g, grp[4 * g + 2]});
        }
        if (items.size() > (size_t)ctx->props.maxGridSize[0])
            return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE, "pack of more than %d work items",
                        ctx->props.maxGridSize[0]);
    }
    // the heaviest work first (FP32: taps per output; MFMA: K steps of the group)
    std::stable_sort(items.begin(), items.end(),
                     [](const Item &a, const Item &b) { return a.w > b.w; });
    img.items.resize(items.size() * 4);
    memcpy(img.items.data(), items.data(), items.size() * sizeof(Item));
    img.n_items = (unsigned)items.size();
    return VTA_OK;
}

vta_status alma_pack_enqueue(vta_ctx *ctx, const AlmaPackImage &img, const void *d_records,
                             const void *d_items) {
    const char *name = img.kernel == VTA_ALMA_BATCH_FP32          ? "alma_batch_f32_pack"
                       : img.kernel == VTA_ALMA_BATCH_MFMA_BF16X3 ? "alma_sweep_pack_bf16x3"
                                                                   : "alma_sweep_pack_f16x3";
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "alma", name, &f));
    ctx->alma_selected = img.kernel;
    return launch(ctx, f, dim3(img.n_items), dim3(256), (unsigned)img.smem, d_records, d_items);
}

const AlmaTables &alma_plan_tables(const vta_alma_sweep_plan *plan) { return plan->t; }
const unsigned char *alma_plan_device(const vta_alma_sweep_plan *plan) {
    return plan->dev.as<unsigned char>();
}

}  // namespace vta

// A pack owns its jobs' tables (one buffer) and its records and items (another).
struct vta_alma_pack {
    DevBuf tables, work;
    AlmaPackImage img;
    size_t items_off = 0;
    std::vector<size_t> combos;
};

extern "C" {

vta_status vta_alma_expand_grid(const vta_alma_batch_range *range, vta_alma_params *out,
                                size_t cap, size_t *n_out) {
    clear_error();
    std::vector<vta_alma_params> v;
    VTA_TRY(expand(range, v));
    if (n_out) *n_out = v.size();
    if (out) memcpy(out, v.data(), std::min(cap, v.size()) * sizeof(vta_alma_params));
    return VTA_OK;
}

vta_status vta_alma_set_batch_policy(vta_ctx *ctx, vta_alma_batch_kernel policy) {
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if (policy < VTA_ALMA_BATCH_AUTO || policy > VTA_ALMA_BATCH_MFMA_F16X3)
        return fail(VTA_ERR_INVALID_POLICY, "invalid policy: %d", (int)policy);
    ctx->alma_policy = policy;
    return VTA_OK;
}

vta_alma_batch_kernel vta_alma_selected_batch_kernel(const vta_ctx *ctx) {
    return ctx ? ctx->alma_selected : VTA_ALMA_BATCH_AUTO;
}

vta_status vta_alma_batch_dev(vta_ctx *ctx, const float *prices, size_t len,
                              const vta_alma_batch_range *range, vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::vector<vta_alma_params> combos;
    size_t fv, mp;
    VTA_TRY(prepare_host(prices, len, range, combos, &fv, &mp));
    size_t req = len * sizeof(float), o;
    if (!mul_ok(combos.size(), len, &o) || !mul_ok(o, sizeof(float), &o) || !add_ok(req, o, &req))
        return fail(VTA_ERR_INVALID_INPUT, "output bytes overflow");
    VTA_TRY(check_fits(req));
    DevBuf dp;
    VTA_TRY(upload(ctx, dp, prices, len * sizeof(float)));
    return run_batch(ctx, dp.as<float>(), len, fv, combos, mp, out, prices);
}

vta_status vta_alma_batch_from_device_prices(vta_ctx *ctx, const float *d_prices, size_t len,
                                             size_t first_valid,
                                             const vta_alma_batch_range *range,
                                             vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !out || !d_prices) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "series_len is zero");
    std::vector<vta_alma_params> combos;
    size_t mp;
    VTA_TRY(expand(range, combos));
    VTA_TRY(validate_combos(combos, len, first_valid, &mp));
    return run_batch(ctx, d_prices, len, first_valid, combos, mp, out);
}

vta_status vta_alma_batch_into_host_f32(vta_ctx *ctx, const float *prices, size_t len,
                                        const vta_alma_batch_range *range, float *out,
                                        size_t out_len) {
    clear_error();
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::vector<vta_alma_params> combos;
    size_t fv, mp;
    VTA_TRY(prepare_host(prices, len, range, combos, &fv, &mp));
    size_t expected;
    if (!mul_ok(combos.size(), len, &expected))
        return fail(VTA_ERR_INVALID_INPUT, "combos*series_len overflow");
    if (out_len != expected)
        return fail(VTA_ERR_INVALID_INPUT, "out slice wrong length: got %zu, expected %zu",
                    out_len, expected);
    DevBuf dp;
    VTA_TRY(upload(ctx, dp, prices, len * sizeof(float)));
    vta_device_array_f32 arr{};
    VTA_TRY(run_batch(ctx, dp.as<float>(), len, fv, combos, mp, &arr, prices));
    vta_status st = vta_device_array_copy_to_host(ctx, &arr, out, out_len);
    vta_device_array_free(ctx, &arr);
    return st;
}

vta_status vta_alma_batch_device(vta_ctx *ctx, const float *d_prices, const float *d_weights,
                                 const int32_t *d_periods, const float *d_inv_norms,
                                 int32_t max_period, int32_t series_len, int32_t n_combos,
                                 int32_t first_valid, float *d_out) {
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if (series_len <= 0 || n_combos <= 0 || max_period <= 0)
        return fail(VTA_ERR_INVALID_INPUT,
                    "series_len, n_combos, and max_period must be positive");
    if (first_valid >= series_len)
        return fail(VTA_ERR_INVALID_INPUT, "first_valid %d >= series_len %d", first_valid,
                    series_len);
    ctx->alma_selected = VTA_ALMA_BATCH_FP32;
    return launch_fp32(ctx, d_prices, d_weights, d_periods, d_inv_norms, max_period, series_len,
                       n_combos, std::max(first_valid, 0), d_out, false);
}

vta_status vta_alma_batch_device_tm(vta_ctx *ctx, const float *d_prices, const float *d_weights,
                                    const int32_t *d_periods, const float *d_inv_norms,
                                    int32_t max_period, int32_t series_len, int32_t n_combos,
                                    int32_t first_valid, float *d_out_tm) {
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if (series_len <= 0 || n_combos <= 0 || max_period <= 0)
        return fail(VTA_ERR_INVALID_INPUT,
                    "series_len, n_combos, and max_period must be positive");
    if (first_valid >= series_len)
        return fail(VTA_ERR_INVALID_INPUT, "first_valid %d >= series_len %d", first_valid,
                    series_len);
    ctx->alma_selected = VTA_ALMA_BATCH_FP32;
    return launch_fp32(ctx, d_prices, d_weights, d_periods, d_inv_norms, max_period, series_len,
                       n_combos, std::max(first_valid, 0), d_out_tm, true);
}

vta_status vta_alma_sweep_plan_create(vta_ctx *ctx, const vta_alma_batch_range *range,
                                      size_t series_len, vta_alma_batch_kernel kernel,
                                      vta_alma_sweep_plan **out) {
    clear_error();
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (series_len == 0 || series_len > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "invalid series_len %zu", series_len);
    std::vector<vta_alma_params> combos;
    VTA_TRY(expand(range, combos));
    size_t mp = 0;
    for (auto &c : combos) {
        if (c.period == 0 || !(c.sigma > 0.0) || !(c.offset >= 0.0 && c.offset <= 1.0))
            return fail(VTA_ERR_INVALID_INPUT, "invalid params: period=%zu, offset=%g, sigma=%g",
                        c.period, c.offset, c.sigma);
        mp = std::max(mp, c.period);
    }
    if (mp > series_len)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu, valid = %zu)",
                    mp, series_len);
    vta_alma_batch_kernel k = resolve_kernel(kernel, combos.size(), mp);
    return plan_create(ctx, combos, series_len, mp, k, out);
}

vta_status vta_alma_sweep_plan_destroy(vta_ctx *ctx, vta_alma_sweep_plan *plan) {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    delete plan;
    return VTA_OK;
}

size_t vta_alma_sweep_plan_combos(const vta_alma_sweep_plan *plan) {
    return plan ? plan->t.n_combos : 0;
}

vta_status vta_alma_sweep_launch(vta_ctx *ctx, const vta_alma_sweep_plan *plan,
                                 const float *d_prices, size_t first_valid, float *d_out) {
    clear_error();
    if (!ctx || !plan || !d_prices || !d_out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (first_valid >= plan->series_len || plan->series_len - first_valid < plan->t.max_period)
        return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu, valid = %zu)",
                    plan->t.max_period,
                    first_valid < plan->series_len ? plan->series_len - first_valid : (size_t)0);
    return plan_launch(ctx, plan, d_prices, first_valid, d_out);
}

vta_status vta_alma_many_series_one_param_time_major_device(
    vta_ctx *ctx, const float *d_prices_tm, const float *d_weights, size_t period,
    float inv_norm, size_t cols, size_t rows, const int32_t *d_first_valids, float *d_out_tm) {
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if (cols == 0 || rows == 0 || period == 0)
        return fail(VTA_ERR_INVALID_INPUT, "cols, rows, and period must be positive");
    if (cols > (size_t)INT32_MAX || rows > (size_t)INT32_MAX || period > rows)
        return fail(VTA_ERR_INVALID_INPUT, "invalid dimensions");
    hipFunction_t f;
    // largest tile height (multiple of 16) whose [tt + period - 1][64] tile fits the LDS
    // (signed: a period above the LDS float count leaves no tile and takes the
    // direct kernel below)
    const long limit = (long)(ctx->props.sharedMemPerBlock / sizeof(float));
    long tt = (limit - (long)period) / 64 - ((long)period - 1);
    tt = (tt / 16) * 16;
    if (tt >= 16 && (size_t)((cols + 63) / 64) <= (size_t)ctx->props.maxGridSize[1]) {
        tt = std::min(tt, 256L);
        VTA_TRY(get_function(ctx, "alma", "alma_ms1p_f32", &f));
        size_t smem = ((size_t)(tt + period - 1) * 64 + period) * sizeof(float);
        dim3 grid((unsigned)((rows + tt - 1) / tt), (unsigned)((cols + 63) / 64), 1);
        return launch(ctx, f, grid, dim3(256), (unsigned)smem, d_prices_tm, d_weights,
                      (int)period, inv_norm, (int)cols, (int)rows, d_first_valids, d_out_tm,
                      (int)tt);
    }
    VTA_TRY(get_function(ctx, "alma", "alma_ms1p_f32_direct", &f));
    size_t blocks = (rows * cols + 255) / 256;
    return launch(ctx, f, dim3((unsigned)blocks), dim3(256), 0, d_prices_tm, d_weights,
                  (int)period, inv_norm, (int)cols, (int)rows, d_first_valids, d_out_tm);
}

vta_status vta_alma_many_series_one_param_time_major_dev(vta_ctx *ctx, const float *data_tm,
                                                         size_t cols, size_t rows,
                                                         const vta_alma_params *params,
                                                         vta_device_array_f32 *out) {
    clear_error();
    if (!ctx || !out || !params || !data_tm) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (cols == 0 || rows == 0) return fail(VTA_ERR_INVALID_INPUT, "cols or rows is zero");
    size_t elems;
    if (!mul_ok(cols, rows, &elems)) return fail(VTA_ERR_INVALID_INPUT, "cols*rows overflow");
    const size_t period = params->period;
    if (period == 0 || !(params->sigma > 0.0) || !(params->offset >= 0.0 && params->offset <= 1.0))
        return fail(VTA_ERR_INVALID_INPUT, "invalid params: period=%zu, offset=%g, sigma=%g",
                    period, params->offset, params->sigma);
    // prepare_many_series_inputs (alma_wrapper.rs:507-563): per-series first_valid
    std::vector<int32_t> fvs(cols);
    for (size_t s = 0; s < cols; ++s) {
        size_t t = 0;
        while (t < rows && std::isnan(data_tm[t * cols + s])) ++t;
        if (t >= rows) return fail(VTA_ERR_INVALID_INPUT, "series %zu all NaN", s);
        if (rows - t < period)
            return fail(VTA_ERR_INVALID_INPUT,
                        "series %zu not enough valid data (needed >= %zu, valid = %zu)", s,
                        period, rows - t);
        fvs[s] = (int32_t)t;
    }
    VTA_TRY(check_fits(2 * elems * sizeof(float) + period * sizeof(float)));
    std::vector<double> wd;
    weights_f64(period, params->offset, params->sigma, wd);
    std::vector<float> wf(wd.begin(), wd.end());
    DevBuf dp, dw, dfv, dout;
    VTA_TRY(upload(ctx, dp, data_tm, elems * sizeof(float)));
    VTA_TRY(upload(ctx, dw, wf.data(), wf.size() * sizeof(float)));
    VTA_TRY(upload(ctx, dfv, fvs.data(), fvs.size() * sizeof(int32_t)));
    VTA_TRY(dout.alloc(elems * sizeof(float)));
    VTA_TRY(vta_alma_many_series_one_param_time_major_device(ctx, dp.as<float>(), dw.as<float>(),
                                                             period, 1.0f, cols, rows,
                                                             dfv.as<int32_t>(),
                                                             dout.as<float>()));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    out->ptr = static_cast<float *>(dout.release());
    out->rows = rows;
    out->cols = cols;
    return VTA_OK;
}

vta_status vta_alma_pack_create(vta_ctx *ctx, vta_alma_batch_kernel kernel,
                                const vta_alma_job *jobs, size_t n_jobs, vta_alma_pack **out) {
    clear_error();
    if (!ctx || !jobs || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = nullptr;
    if (n_jobs == 0 || n_jobs > (size_t)INT32_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "a pack needs 1..2^31-1 jobs (got %zu)", n_jobs);
    if (kernel < VTA_ALMA_BATCH_AUTO || kernel > VTA_ALMA_BATCH_MFMA_F16X3)
        return fail(VTA_ERR_INVALID_POLICY, "invalid policy %d", (int)kernel);
    std::vector<std::vector<vta_alma_params>> cs(n_jobs);
    std::vector<size_t> mps(n_jobs);
    bool all_mfma = true;
    for (size_t j = 0; j < n_jobs; ++j) {
        const vta_alma_job &jb = jobs[j];
        if (!jb.d_prices || !jb.d_out)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: null device pointer", j);
        if (jb.series_len == 0 || jb.series_len > (size_t)INT32_MAX)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: series_len out of range", j);
        if (jb.first_valid >= jb.series_len)
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: first_valid >= series_len", j);
        VTA_TRY(expand(&jb.range, cs[j]));
        VTA_TRY(validate_combos(cs[j], jb.series_len, jb.first_valid, &mps[j]));
        size_t n_out;
        if (!mul_ok(cs[j].size(), jb.series_len, &n_out) || n_out > ((size_t)1 << 60))
            return fail(VTA_ERR_INVALID_INPUT, "job %zu: n_combos*series_len overflow", j);
        all_mfma = all_mfma && cs[j].size() >= 16 && mps[j] <= kMfMaxPeriod;
    }
    const vta_alma_batch_kernel k =
        kernel != VTA_ALMA_BATCH_AUTO ? kernel
                                      : (all_mfma ? VTA_ALMA_BATCH_MFMA_F16X3 : VTA_ALMA_BATCH_FP32);
    std::vector<AlmaTables> ts(n_jobs);
    std::vector<size_t> boff(n_jobs);
    size_t total = 0;
    for (size_t j = 0; j < n_jobs; ++j) {
        VTA_TRY(tables_build(cs[j], mps[j], k, ts[j]));
        boff[j] = (total + 255) & ~(size_t)255;
        total = boff[j] + ts[j].blob.size();
    }
    auto pk = std::make_unique<vta_alma_pack>();
    std::vector<unsigned char> blob(total, 0);
    for (size_t j = 0; j < n_jobs; ++j)
        if (!ts[j].blob.empty()) memcpy(blob.data() + boff[j], ts[j].blob.data(), ts[j].blob.size());
    VTA_TRY(upload(ctx, pk->tables, blob.data(), total));
    std::vector<AlmaPackSrc> src(n_jobs);
    for (size_t j = 0; j < n_jobs; ++j)
        src[j] = {&ts[j], pk->tables.as<unsigned char>() + boff[j], jobs[j].d_prices,
                  jobs[j].d_out, jobs[j].series_len, jobs[j].first_valid};
    VTA_TRY(alma_pack_build(ctx, k, src, pk->img));
    pk->items_off = (pk->img.records.size() + 255) & ~(size_t)255;
    std::vector<unsigned char> work(pk->items_off + pk->img.items.size() * sizeof(int32_t), 0);
    memcpy(work.data(), pk->img.records.data(), pk->img.records.size());
    memcpy(work.data() + pk->items_off, pk->img.items.data(),
           pk->img.items.size() * sizeof(int32_t));
    VTA_TRY(upload(ctx, pk->work, work.data(), work.size()));
    for (auto &c : cs) pk->combos.push_back(c.size());
    *out = pk.release();
    return VTA_OK;
}

vta_status vta_alma_pack_launch(vta_ctx *ctx, const vta_alma_pack *pack) {
    clear_error();
    if (!ctx || !pack) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    const unsigned char *w = pack->work.as<unsigned char>();
    return alma_pack_enqueue(ctx, pack->img, w, w + pack->items_off);
}

vta_alma_batch_kernel vta_alma_pack_kernel(const vta_alma_pack *pack) {
    return pack ? pack->img.kernel : VTA_ALMA_BATCH_AUTO;
}

size_t vta_alma_pack_job_combos(const vta_alma_pack *pack, size_t job) {
    return pack && job < pack->combos.size() ? pack->combos[job] : 0;
}

vta_status vta_alma_pack_destroy(vta_alma_pack *pack) {
    clear_error();
    delete pack;
    return VTA_OK;
}

}  // extern "C"
