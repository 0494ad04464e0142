// Double-ALMA crossover backtest reduction (no reference counterpart: the
// reference names the workload only, README.md:9).  Kernels: kernels/cross.hip.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

struct cross_args_host {
    double *rows;
    const float *prices;
    const double *w64;
    const double *inv64;
    const int32_t *row_period;
    const int32_t *fast_row;
    const int32_t *slow_row;
    long long *part_i;
    double *part_pnl;
    long long *out_i;
    double *out_pnl;
    int T;
    int fv;
    int nfast;
    int nslow;
    int nseg;
    int seg_len;
    int pmax;
    int ntiles_s;
    int row_tile;
    int pad_;
};
static_assert(sizeof(cross_args_host) == 128, "cross_args layout");

constexpr size_t kMaxPeriod = 294;   // LDS staging of the f64 row kernel
constexpr int kRowTile = 2048;       // bars per cross_rows_f64 workgroup (RW_T)
// Workgroups of the pair walk (16 x 16 pairs x one time segment each).
// Tiles of short periods trade more often and walk longer, so small units
// balance the device: 58,300 pairs x 200k bars took 6.73 / 5.76 / 5.28 /
// 5.18 / 5.12 / 5.11 ms at 2,048 / 4,046 / 8,100 / 12,000 / 16,000 / 23,000
// workgroups (at most one segment per 2,048 bars).
constexpr size_t kCrossWorkgroups = 16384;

// f64 batch weights in the oracle's order (alma.rs:1426-1436), not normalised.
double weights_batch_raw(size_t period, double offset, double sigma, double *w) {
    double m = offset * (double)(period - 1);
    double s = (double)period / sigma;
    double s2 = 2.0 * s * s;
    double norm = 0.0;
    for (size_t i = 0; i < period; ++i) {
        double d = (double)i - m;
        double wi = std::exp(-(d * d) / s2);
        w[i] = wi;
        norm += wi;
    }
    return 1.0 / norm;
}

}  // namespace

struct vta_double_alma_plan {
    vta_ctx *ctx = nullptr;
    size_t T = 0, fv = 0;
    int nfast = 0, nslow = 0, nrows = 0, pmax = 0, nseg = 0, seg_len = 0;
    DevBuf prices, rows, w64, inv64, row_period, fast_row, slow_row, part_i, part_pnl, out_i,
        out_pnl;
};

namespace {

vta_status plan_build(vta_ctx *ctx, const float *prices, size_t len,
                      const vta_alma_batch_range *fast_r, const vta_alma_batch_range *slow_r,
                      vta_double_alma_plan **out) {
    if (!ctx || !prices || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "empty data");
    if (len > (size_t)INT32_MAX / 2) return fail(VTA_ERR_INVALID_INPUT, "series too long");
    std::vector<vta_alma_params> fast, slow;
    VTA_TRY(alma_expand_range(fast_r, fast));
    VTA_TRY(alma_expand_range(slow_r, slow));
    if (fast.empty() || slow.empty()) return fail(VTA_ERR_INVALID_INPUT, "empty grid");
    const size_t fv = first_valid_f32(prices, len);
    if (fv >= len) return fail(VTA_ERR_INVALID_INPUT, "all values are NaN");
    for (size_t t = fv; t < len; ++t)
        if (!std::isfinite(prices[t]))
            return fail(VTA_ERR_INVALID_INPUT,
                        "crossover needs finite prices after first_valid (bar %zu)", t);
    // distinct parameter sets -> rows
    std::map<std::tuple<size_t, uint64_t, uint64_t>, int> key_row;
    std::vector<vta_alma_params> rows;
    auto row_of = [&](const vta_alma_params &c) {
        uint64_t ob, sb;
        std::memcpy(&ob, &c.offset, 8);
        std::memcpy(&sb, &c.sigma, 8);
        auto k = std::make_tuple(c.period, ob, sb);
        auto it = key_row.find(k);
        if (it != key_row.end()) return it->second;
        int r = (int)rows.size();
        rows.push_back(c);
        key_row.emplace(k, r);
        return r;
    };
    std::vector<int32_t> frow(fast.size()), srow(slow.size());
    for (size_t i = 0; i < fast.size(); ++i) frow[i] = row_of(fast[i]);
    for (size_t i = 0; i < slow.size(); ++i) srow[i] = row_of(slow[i]);
    size_t pmax = 0;
    for (auto &r : rows) pmax = std::max(pmax, r.period);
    if (pmax > kMaxPeriod)
        return fail(VTA_ERR_INVALID_INPUT, "crossover supports periods <= %zu (got %zu)",
                    kMaxPeriod, pmax);
    const size_t nrows = rows.size();
    size_t npairs, rows_bytes;
    if (!mul_ok(fast.size(), slow.size(), &npairs) || npairs > (size_t)INT32_MAX ||
        !mul_ok(nrows, len * sizeof(double), &rows_bytes))
        return fail(VTA_ERR_INVALID_INPUT, "grid too large");
    // time segments: about kCrossWorkgroups workgroups (VTA_CROSS_WGS overrides)
    const size_t ntiles = ((fast.size() + 15) / 16) * ((slow.size() + 15) / 16);
    size_t wgs = kCrossWorkgroups;
    if (const char *e = getenv("VTA_CROSS_WGS")) {
        const long v = atol(e);
        if (v >= 1) wgs = (size_t)v;
    }
    size_t nseg = std::max<size_t>(1, (wgs + ntiles - 1) / ntiles);
    nseg = std::min(nseg, std::max<size_t>(1, len / 2048));
    const size_t seg_len = (len + nseg - 1) / nseg;
    nseg = (len + seg_len - 1) / seg_len;
    VTA_TRY(check_fits(rows_bytes + nseg * npairs * 40 + npairs * 40 + len * 4));

    auto *plan = new vta_double_alma_plan();
    plan->ctx = ctx;
    plan->T = len;
    plan->fv = fv;
    plan->nfast = (int)fast.size();
    plan->nslow = (int)slow.size();
    plan->nrows = (int)nrows;
    plan->pmax = (int)pmax;
    plan->nseg = (int)nseg;
    plan->seg_len = (int)seg_len;
    std::vector<double> w64(nrows * pmax, 0.0), inv(nrows);
    std::vector<int32_t> per(nrows);
    for (size_t r = 0; r < nrows; ++r) {
        inv[r] = weights_batch_raw(rows[r].period, rows[r].offset, rows[r].sigma,
                                   w64.data() + r * pmax);
        per[r] = (int32_t)rows[r].period;
    }
    for (auto &c : rows)
        if (c.period == 0 || c.period > len - fv || !(c.sigma > 0.0) ||
            !(c.offset >= 0.0 && c.offset <= 1.0)) {
            delete plan;
            return fail(VTA_ERR_INVALID_INPUT, "invalid params: period=%zu, offset=%g, sigma=%g",
                        c.period, c.offset, c.sigma);
        }
    vta_status st = upload(ctx, plan->prices, prices, len * sizeof(float));
    if (st == VTA_OK) st = plan->rows.alloc(rows_bytes);
    if (st == VTA_OK) st = upload(ctx, plan->w64, w64.data(), w64.size() * sizeof(double));
    if (st == VTA_OK) st = upload(ctx, plan->inv64, inv.data(), inv.size() * sizeof(double));
    if (st == VTA_OK) st = upload(ctx, plan->row_period, per.data(), per.size() * 4);
    if (st == VTA_OK) st = upload(ctx, plan->fast_row, frow.data(), frow.size() * 4);
    if (st == VTA_OK) st = upload(ctx, plan->slow_row, srow.data(), srow.size() * 4);
    if (st == VTA_OK) st = plan->part_i.alloc(nseg * npairs * 4 * sizeof(long long));
    if (st == VTA_OK) st = plan->part_pnl.alloc(nseg * npairs * sizeof(double));
    if (st == VTA_OK) st = plan->out_i.alloc(npairs * 4 * sizeof(long long));
    if (st == VTA_OK) st = plan->out_pnl.alloc(npairs * sizeof(double));
    if (st != VTA_OK) {
        delete plan;
        return st;
    }
    *out = plan;
    return VTA_OK;
}

vta_status plan_enqueue(vta_double_alma_plan *p) {
    vta_ctx *ctx = p->ctx;
    cross_args_host a{};
    a.rows = p->rows.as<double>();
    a.prices = p->prices.as<float>();
    a.w64 = p->w64.as<double>();
    a.inv64 = p->inv64.as<double>();
    a.row_period = p->row_period.as<int32_t>();
    a.fast_row = p->fast_row.as<int32_t>();
    a.slow_row = p->slow_row.as<int32_t>();
    a.part_i = p->part_i.as<long long>();
    a.part_pnl = p->part_pnl.as<double>();
    a.out_i = p->out_i.as<long long>();
    a.out_pnl = p->out_pnl.as<double>();
    a.T = (int)p->T;
    a.fv = (int)p->fv;
    a.nfast = p->nfast;
    a.nslow = p->nslow;
    a.nseg = p->nseg;
    a.seg_len = p->seg_len;
    a.pmax = p->pmax;
    a.ntiles_s = (p->nslow + 15) / 16;
    a.row_tile = kRowTile;
    const unsigned ntiles = (unsigned)(((p->nfast + 15) / 16) * a.ntiles_s);
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "cross", "cross_rows_f64", &f));
    VTA_TRY(launch(ctx, f, dim3((unsigned)((p->T + kRowTile - 1) / kRowTile), (unsigned)p->nrows, 1),
                   dim3(256), 0, a));
    VTA_TRY(get_function(ctx, "cross", "cross_pairs_f64", &f));
    VTA_TRY(launch(ctx, f, dim3(ntiles, (unsigned)p->nseg, 1), dim3(256), 0, a));
    const size_t npairs = (size_t)p->nfast * p->nslow;
    VTA_TRY(get_function(ctx, "cross", "cross_reduce", &f));
    return launch(ctx, f, dim3((unsigned)((npairs + 255) / 256), 1, 1), dim3(256), 0, a);
}

vta_status plan_copy(vta_double_alma_plan *p, int64_t *stats4, double *pnl, size_t cap) {
    const size_t npairs = (size_t)p->nfast * p->nslow;
    if ((stats4 || pnl) && cap < npairs)
        return fail(VTA_ERR_INVALID_INPUT, "output capacity %zu < %zu pairs", cap, npairs);
    VTA_HIP(hipStreamSynchronize(p->ctx->stream));
    if (stats4)
        VTA_HIP(hipMemcpy(stats4, p->out_i.p, npairs * 4 * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (pnl) VTA_HIP(hipMemcpy(pnl, p->out_pnl.p, npairs * sizeof(double), hipMemcpyDeviceToHost));
    return VTA_OK;
}

}  // namespace

extern "C" {

vta_status vta_double_alma_plan_create(vta_ctx *ctx, const float *prices, size_t len,
                                       const vta_alma_batch_range *fast,
                                       const vta_alma_batch_range *slow,
                                       vta_double_alma_plan **out) {
    clear_error();
    return plan_build(ctx, prices, len, fast, slow, out);
}

size_t vta_double_alma_plan_pairs(const vta_double_alma_plan *plan) {
    return plan ? (size_t)plan->nfast * plan->nslow : 0;
}

vta_status vta_double_alma_plan_launch(vta_double_alma_plan *plan) {
    clear_error();
    if (!plan) return fail(VTA_ERR_INVALID_INPUT, "plan is null");
    return plan_enqueue(plan);
}

vta_status vta_double_alma_plan_results(vta_double_alma_plan *plan, int64_t *stats4,
                                        double *pnl, size_t cap) {
    clear_error();
    if (!plan) return fail(VTA_ERR_INVALID_INPUT, "plan is null");
    return plan_copy(plan, stats4, pnl, cap);
}

vta_status vta_double_alma_plan_destroy(vta_double_alma_plan *plan) {
    clear_error();
    if (plan) {
        (void)hipStreamSynchronize(plan->ctx->stream);
        delete plan;
    }
    return VTA_OK;
}

vta_status vta_double_alma_crossover(vta_ctx *ctx, const float *prices, size_t len,
                                     const vta_alma_batch_range *fast,
                                     const vta_alma_batch_range *slow, int64_t *stats4,
                                     double *pnl, size_t cap, size_t *n_pairs) {
    clear_error();
    if (!stats4 && !pnl) {
        // size query: the grids alone
        std::vector<vta_alma_params> f, s;
        VTA_TRY(alma_expand_range(fast, f));
        VTA_TRY(alma_expand_range(slow, s));
        if (n_pairs) *n_pairs = f.size() * s.size();
        return VTA_OK;
    }
    vta_double_alma_plan *p = nullptr;
    VTA_TRY(plan_build(ctx, prices, len, fast, slow, &p));
    if (n_pairs) *n_pairs = (size_t)p->nfast * p->nslow;
    vta_status st = plan_enqueue(p);
    if (st == VTA_OK) st = plan_copy(p, stats4, pnl, cap);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    return st;
}

}  // extern "C"
