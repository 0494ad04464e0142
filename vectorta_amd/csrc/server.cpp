// Request server: continuous batching of indicator requests on one device.
//
// No reference counterpart (the reference serves one call at a time through
// its wrappers).  Requests (ALMA, EMA, RSI, Bollinger Bands, ADX one-series
// batches) enter a FIFO queue.  Each of n_slots worker threads owns a vta_ctx
// (its own stream), so admitted requests run concurrently on the device.  A
// request is admitted when a slot is free and the HBM budget holds its
// outputs plus, on a cache miss, its series upload; outputs stay reserved
// until the client releases them.  Series live in a device cache keyed by the
// client's 64-bit key; unreferenced entries are evicted least-recently-used
// when admission needs room.  The kernels are the library's public entry
// points, so a result is bit-identical whether the request runs alone or in
// a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime_internal.h"

using namespace vta;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Host-side metadata of a series (computed at submit time).
struct SeriesMeta {
    size_t len = 0;
    bool hlc = false;
    int32_t fv_close = 0;   // first non-NaN close
    int32_t z_close = 0;    // first non-finite close at or after fv_close (len if none)
    int32_t fv_adx = 0;     // max of the three first non-NaN indices (adx.rs:190-198)
};

struct CacheEntry {
    uint64_t key = 0;
    SeriesMeta meta;
    float *d_close = nullptr, *d_high = nullptr, *d_low = nullptr;
    size_t bytes = 0;
    int refs = 0;
    bool ready = false;
    bool cached = false;               // false: freed when its request completes
    std::list<uint64_t>::iterator lru;
};

struct Req {
    uint64_t id = 0;
    vta_request r{};
    std::vector<float> close, high, low;   // host copies
    SeriesMeta meta;
    size_t out_bytes = 0;                  // reserved for outputs
    size_t series_bytes = 0;               // device bytes of the series
    size_t series_reserved = 0;            // reserved at admission (misses only)
    bool uploader = false;                 // this request uploads the series
    std::shared_ptr<CacheEntry> entry;
    vta_result res{};
    bool done = false;
};

size_t first_nan_end(const float *x, size_t n) {
    size_t i = 0;
    while (i < n && std::isnan(x[i])) ++i;
    return i;
}

// Best-fit allocator over one device allocation; coalesces on free.
class Arena {
  public:
    vta_status init(size_t cap) {
        cap = (cap + 255) & ~(size_t)255;
        VTA_HIP(hipMalloc(reinterpret_cast<void **>(&base_), cap));
        cap_ = cap;
        free_.emplace(0, cap);
        bylen_.emplace(cap, 0);
        return VTA_OK;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    void *alloc(size_t n) {
        n = std::max<size_t>(256, (n + 255) & ~(size_t)255);
        std::lock_guard<std::mutex> lk(mu_);
        auto it = bylen_.lower_bound(n);
        if (it == bylen_.end()) return nullptr;
        const size_t len = it->first, off = it->second;
        bylen_.erase(it);
        free_.erase(off);
        if (len > n) {
            free_.emplace(off + n, len - n);
            bylen_.emplace(len - n, off + n);
        }
        used_.emplace(off, n);
        return base_ + off;
    }
    void release(void *p) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        size_t off = static_cast<size_t>(static_cast<char *>(p) - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t n = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + n) {
            erase_len(next->second, next->first);
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                erase_len(prev->second, prev->first);
                off = prev->first;
                n += prev->second;
                free_.erase(prev);
            }
        }
        free_.emplace(off, n);
        bylen_.emplace(n, off);
    }

  private:
    void erase_len(size_t len, size_t off) {
        auto r = bylen_.equal_range(len);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == off) {
                bylen_.erase(it);
                return;
            }
    }
    char *base_ = nullptr;
    size_t cap_ = 0;
    std::map<size_t, size_t> free_;          // offset -> length
    std::multimap<size_t, size_t> bylen_;    // length -> offset
    std::unordered_map<size_t, size_t> used_;
    std::mutex mu_;
};

}  // namespace

struct vta_server {
    int device = 0;
    size_t budget = 0;
    size_t in_use = 0;          // outputs of live results + cache + pending uploads
    size_t peak_in_use = 0;
    int in_flight = 0, peak_in_flight = 0;
    size_t packs = 0, packed_requests = 0;
    bool stop = false;
    uint64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_admit;   // queue / budget / shutdown changes
    std::condition_variable cv_done;    // request completions, entry readiness
    std::deque<std::shared_ptr<Req>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> reqs;
    std::unordered_map<uint64_t, std::shared_ptr<CacheEntry>> cache;
    std::list<uint64_t> lru;            // front = least recently used
    std::vector<vta_ctx *> ctxs;
    std::vector<std::thread> workers;
    Arena arena;                        // all served-path device memory
    std::vector<void *> pending_free;   // collected under mu, returned after unlocking
    std::vector<float *> pinned;        // per-slot pinned staging for series uploads
    std::vector<size_t> pinned_cap;
    double t0 = 0.0;
    // ALMA grid -> sweep tables (guarded by plan_mu; entries live until destroy)
    std::mutex plan_mu;
    std::unordered_map<std::string, vta_alma_sweep_plan *> alma_plans;
    std::unordered_map<std::string, std::pair<size_t, size_t>> alma_plan_dims;
};

namespace {

size_t grid_outputs(const vta_request &r, size_t *n_combos, int *n_out) {
    *n_out = 1;
    std::vector<size_t> ps;
    switch (r.kind) {
        case VTA_REQ_ALMA: {
            size_t n = 0;
            if (vta_alma_expand_grid(&r.alma, nullptr, 0, &n) != VTA_OK) return 0;
            *n_combos = n;
            return n;
        }
        case VTA_REQ_EMA:
        case VTA_REQ_RSI:
        case VTA_REQ_ADX: {
            size_t n = 0;
            vta_status st = r.kind == VTA_REQ_EMA   ? vta_ema_expand_grid(&r.period, nullptr, 0, &n)
                            : r.kind == VTA_REQ_RSI ? vta_rsi_expand_grid(&r.period, nullptr, 0, &n)
                                                    : vta_adx_expand_grid(&r.period, nullptr, 0, &n);
            if (st != VTA_OK) return 0;
            *n_combos = n;
            return n;
        }
        case VTA_REQ_BB: {
            size_t n = 0;
            if (vta_bb_expand_grid(&r.bb, nullptr, 0, &n) != VTA_OK) return 0;
            *n_combos = n;
            *n_out = 3;
            return 3 * n;
        }
        default:
            return 0;
    }
}

void evict_lru(vta_server *s, size_t need) {
    auto it = s->lru.begin();
    while (s->in_use + need > s->budget && it != s->lru.end()) {
        auto e = s->cache.find(*it);
        if (e == s->cache.end() || e->second->refs > 0 || !e->second->ready) {
            ++it;
            continue;
        }
        std::shared_ptr<CacheEntry> ent = e->second;
        s->pending_free.push_back(ent->d_close);
        if (ent->d_high) s->pending_free.push_back(ent->d_high);
        if (ent->d_low) s->pending_free.push_back(ent->d_low);
        s->in_use -= ent->bytes;
        s->cache.erase(e);
        it = s->lru.erase(it);
    }
}

// Called with s->mu held: admits the queued request at `it` if the HBM budget
// holds its outputs (and its series, on a cache miss), evicting cached series
// if needed.  Returns it (removed from the queue) or null.
std::shared_ptr<Req> admit_at(vta_server *s, std::deque<std::shared_ptr<Req>>::iterator it) {
    std::shared_ptr<Req> q = *it;
    auto hit = q->r.series_key ? s->cache.find(q->r.series_key) : s->cache.end();
    const bool is_hit = hit != s->cache.end();
    const size_t series = is_hit ? 0 : q->series_bytes;
    const size_t need = q->out_bytes + series;
    if (s->in_use + need > s->budget) evict_lru(s, need);
    if (s->in_use + need > s->budget) return nullptr;
    s->queue.erase(it);
    s->in_use += need;
    s->peak_in_use = std::max(s->peak_in_use, s->in_use);
    q->series_reserved = series;
    q->uploader = !is_hit;
    if (is_hit) {
        q->entry = hit->second;
        q->entry->refs += 1;
        s->lru.splice(s->lru.end(), s->lru, q->entry->lru);   // most recently used
    } else {
        auto e = std::make_shared<CacheEntry>();
        e->key = q->r.series_key;
        e->meta = q->meta;
        e->bytes = q->series_bytes;
        e->refs = 1;
        e->cached = q->r.cache && q->r.series_key != 0;
        if (e->cached) {
            s->lru.push_back(e->key);
            e->lru = std::prev(s->lru.end());
            s->cache.emplace(e->key, e);
        }
        q->entry = e;
    }
    return q;
}

// Called with s->mu held.  Returns the admitted request or null.
std::shared_ptr<Req> try_admit(vta_server *s) {
    while (!s->queue.empty()) {
        std::shared_ptr<Req> q = s->queue.front();
        const bool is_hit = q->r.series_key && s->cache.count(q->r.series_key);
        if (q->out_bytes + (is_hit ? 0 : q->series_bytes) > s->budget) {
            // can never fit: fail it now (nothing was reserved for it)
            s->queue.pop_front();
            q->out_bytes = 0;
            q->res.status = VTA_ERR_OUT_OF_MEMORY;
            q->res.t_start = q->res.t_done = now_s() - s->t0;
            q->done = true;
            s->cv_done.notify_all();
            continue;
        }
        return admit_at(s, s->queue.begin());
    }
    return nullptr;
}

// Small one-series requests (ALMA, EMA, RSI, ADX, Bollinger Bands) are served together:
// a slot that admits one also admits queued requests of the same kind, from
// the first kPackScan in the queue, up to kPackMax, and runs them as one
// packed batch (one launch per scan pass for all of them).  Small = at most
// kPackOutputs outputs, where launch overhead and per-request latency chains
// dominate; outputs are bit-identical to the request served alone.
constexpr size_t kPackOutputs = 4u << 20;
constexpr size_t kPackMax = 32;
constexpr size_t kPackScan = 64;

bool packable(const Req &q) {
    return (q.r.kind == VTA_REQ_EMA || q.r.kind == VTA_REQ_RSI || q.r.kind == VTA_REQ_ADX ||
            q.r.kind == VTA_REQ_BB || q.r.kind == VTA_REQ_ALMA) &&
           q.out_bytes <= kPackOutputs * sizeof(float);
}

// Called with s->mu held: the next packable request of `kind` among the
// first kPackScan queued, admitted, or null.
std::shared_ptr<Req> try_admit_kind(vta_server *s, int kind) {
    size_t seen = 0;
    for (auto it = s->queue.begin(); it != s->queue.end() && seen < kPackScan; ++it, ++seen) {
        const Req &q = **it;
        if (q.r.kind != kind || !packable(q)) continue;
        return admit_at(s, it);
    }
    return nullptr;
}

// Copies the series through the slot's pinned staging buffer (grown on
// demand), so the DMA engine reads page-locked memory directly.
vta_status upload_series(vta_server *s, int slot, vta_ctx *ctx, Req &q) {
    CacheEntry &e = *q.entry;
    const size_t n = q.meta.len * sizeof(float);
    const int arrays = q.meta.hlc ? 3 : 1;
    if (s->pinned_cap[slot] < arrays * n) {
        if (s->pinned[slot]) VTA_HIP(hipHostFree(s->pinned[slot]));
        s->pinned[slot] = nullptr;
        s->pinned_cap[slot] = 0;
        void *p = nullptr;
        VTA_HIP(hipHostMalloc(&p, arrays * n, hipHostMallocDefault));
        s->pinned[slot] = static_cast<float *>(p);
        s->pinned_cap[slot] = arrays * n;
    }
    float *stage = s->pinned[slot];
    std::memcpy(stage, q.close.data(), n);
    if (q.meta.hlc) {
        std::memcpy(stage + q.meta.len, q.high.data(), n);
        std::memcpy(stage + 2 * q.meta.len, q.low.data(), n);
    }
    e.d_close = static_cast<float *>(s->arena.alloc(n));
    if (!e.d_close) return fail(VTA_ERR_OUT_OF_MEMORY, "device arena exhausted");
    VTA_HIP(hipMemcpyAsync(e.d_close, stage, n, hipMemcpyHostToDevice, ctx->stream));
    if (q.meta.hlc) {
        e.d_high = static_cast<float *>(s->arena.alloc(n));
        e.d_low = static_cast<float *>(s->arena.alloc(n));
        if (!e.d_high || !e.d_low) {
            (void)hipStreamSynchronize(ctx->stream);
            return fail(VTA_ERR_OUT_OF_MEMORY, "device arena exhausted");
        }
        VTA_HIP(hipMemcpyAsync(e.d_high, stage + q.meta.len, n, hipMemcpyHostToDevice,
                               ctx->stream));
        VTA_HIP(hipMemcpyAsync(e.d_low, stage + 2 * q.meta.len, n, hipMemcpyHostToDevice,
                               ctx->stream));
    }
    VTA_HIP(hipStreamSynchronize(ctx->stream));   // the staging buffer is reused next request
    return VTA_OK;
}

// A small parameter table in the arena, returned when it goes out of scope.
// An early error return can leave kernels that read it queued on the slot's
// stream, so the stream is drained first (a no-op after run_request's own
// synchronisation): another slot may be handed the memory at once.
struct ArenaTable {
    Arena *ar = nullptr;
    vta_ctx *ctx = nullptr;
    void *p = nullptr;
    ~ArenaTable() {
        if (p) {
            (void)hipStreamSynchronize(ctx->stream);
            ar->release(p);
        }
    }
    template <typename T>
    vta_status put(vta_server *s, vta_ctx *c, const std::vector<T> &v) {
        ar = &s->arena;
        ctx = c;
        p = ar->alloc(v.size() * sizeof(T));
        if (!p) return fail(VTA_ERR_OUT_OF_MEMORY, "device arena exhausted");
        VTA_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                               c->stream));
        return VTA_OK;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

vta_status alloc_out(vta_server *s, vta_device_array_f32 &o, size_t rows, size_t cols) {
    void *p = s->arena.alloc(rows * cols * sizeof(float));
    if (!p) return fail(VTA_ERR_OUT_OF_MEMORY, "device arena exhausted");
    o.ptr = static_cast<float *>(p);
    o.rows = rows;
    o.cols = cols;
    return VTA_OK;
}

// The grid's sweep plan, built on a miss outside plan_mu (the host table build
// of a large grid takes milliseconds and must not hold up the other slots);
// of two slots that race on one new grid, the first insert wins and the
// other's plan is dropped.
vta_status alma_plan_for(vta_server *s, vta_ctx *ctx, const vta_alma_batch_range &r,
                         vta_alma_sweep_plan **plan, size_t *max_period, size_t *n_combos) {
    std::string key(reinterpret_cast<const char *>(&r), sizeof(r));
    {
        std::lock_guard<std::mutex> lk(s->plan_mu);
        auto it = s->alma_plans.find(key);
        if (it != s->alma_plans.end()) {
            *plan = it->second;
            *max_period = s->alma_plan_dims[key].first;
            *n_combos = s->alma_plan_dims[key].second;
            return VTA_OK;
        }
    }
    vta_alma_sweep_plan *p = nullptr;
    size_t mp = 0, nc = 0;
    VTA_TRY(alma_grid_plan_create(ctx, &r, &p, &mp, &nc));
    std::lock_guard<std::mutex> lk(s->plan_mu);
    auto ins = s->alma_plans.emplace(key, p);
    if (ins.second) {
        s->alma_plan_dims.emplace(key, std::make_pair(mp, nc));
    } else {
        alma_grid_plan_destroy(p);   // another slot built the same grid first
    }
    *plan = ins.first->second;
    *max_period = s->alma_plan_dims[key].first;
    *n_combos = s->alma_plan_dims[key].second;
    return VTA_OK;
}

vta_status run_request(vta_server *s, vta_ctx *ctx, Req &q) {
    CacheEntry &e = *q.entry;
    const double t0 = s->t0;
    const SeriesMeta &m = e.meta;
    const vta_request &r = q.r;
    vta_result &res = q.res;
    const size_t T = m.len;
    switch (r.kind) {
        case VTA_REQ_ALMA: {
            vta_alma_sweep_plan *plan = nullptr;
            size_t mp = 0, nc = 0;
            VTA_TRY(alma_plan_for(s, ctx, r.alma, &plan, &mp, &nc));
            if ((size_t)m.fv_close >= T || T - (size_t)m.fv_close < mp)
                return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", mp);
            res.n_outputs = 1;
            res.n_combos = nc;
            VTA_TRY(alloc_out(s, res.out[0], nc, T));
            VTA_TRY(alma_grid_plan_launch(ctx, plan, e.d_close, T, (size_t)m.fv_close,
                                          res.out[0].ptr));
            res.t_launch = now_s() - t0;
            VTA_HIP(hipStreamSynchronize(ctx->stream));
            return VTA_OK;
        }
        case VTA_REQ_EMA:
        case VTA_REQ_RSI:
        case VTA_REQ_ADX: {
            size_t n = 0;
            vta_status (*grid)(const vta_range_usize *, size_t *, size_t, size_t *) =
                r.kind == VTA_REQ_EMA ? vta_ema_expand_grid
                : r.kind == VTA_REQ_RSI ? vta_rsi_expand_grid
                                        : vta_adx_expand_grid;
            VTA_TRY(grid(&r.period, nullptr, 0, &n));
            std::vector<size_t> ps(n);
            VTA_TRY(grid(&r.period, ps.data(), n, &n));
            if (n == 0 || T > (size_t)INT32_MAX || n > (size_t)INT32_MAX)
                return fail(VTA_ERR_INVALID_INPUT, "invalid request sizes");
            const size_t maxp = *std::max_element(ps.begin(), ps.end());
            const int32_t fv = r.kind == VTA_REQ_ADX ? m.fv_adx : m.fv_close;
            const size_t need = r.kind == VTA_REQ_EMA   ? maxp
                                : r.kind == VTA_REQ_RSI ? maxp + 1
                                                        : 2 * maxp;
            if (maxp == 0 || (size_t)fv >= T || T - (size_t)fv < need)
                return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", need);
            if (r.kind == VTA_REQ_ADX && !m.hlc)
                return fail(VTA_ERR_INVALID_INPUT, "ADX needs high, low and close");
            std::vector<int32_t> per(ps.begin(), ps.end());
            ArenaTable dp;
            VTA_TRY(dp.put(s, ctx, per));
            res.n_outputs = 1;
            res.n_combos = n;
            VTA_TRY(alloc_out(s, res.out[0], n, T));
            if (r.kind == VTA_REQ_EMA)
                VTA_TRY(vta_ema_batch_device(ctx, e.d_close, dp.as<int32_t>(), nullptr, T,
                                             (size_t)fv, n, res.out[0].ptr));
            else if (r.kind == VTA_REQ_RSI)
                VTA_TRY(vta_rsi_batch_device(ctx, e.d_close, dp.as<int32_t>(), (int32_t)maxp,
                                             (int32_t)T, fv, (int32_t)n, res.out[0].ptr));
            else
                VTA_TRY(vta_adx_batch_device(ctx, e.d_high, e.d_low, e.d_close, dp.as<int32_t>(),
                                             (int32_t)maxp, (int32_t)T, fv, (int32_t)n,
                                             res.out[0].ptr));
            res.t_launch = now_s() - t0;
            VTA_HIP(hipStreamSynchronize(ctx->stream));
            return VTA_OK;
        }
        case VTA_REQ_BB: {
            size_t n = 0;
            VTA_TRY(vta_bb_expand_grid(&r.bb, nullptr, 0, &n));
            std::vector<vta_bb_params> cs(n);
            VTA_TRY(vta_bb_expand_grid(&r.bb, cs.data(), n, &n));
            if (n == 0 || T > (size_t)INT32_MAX || n > (size_t)INT32_MAX)
                return fail(VTA_ERR_INVALID_INPUT, "invalid request sizes");
            std::vector<int32_t> per(n);
            std::vector<double> up(n), dn(n);
            size_t maxp = 0;
            for (size_t i = 0; i < n; ++i) {
                if (cs[i].period == 0 || cs[i].period > T)
                    return fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", cs[i].period);
                per[i] = (int32_t)cs[i].period;
                up[i] = cs[i].devup;
                dn[i] = cs[i].devdn;
                maxp = std::max(maxp, cs[i].period);
            }
            if ((size_t)m.fv_close >= T || T - (size_t)m.fv_close < maxp)
                return fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", maxp);
            ArenaTable dp, du, dd;
            VTA_TRY(dp.put(s, ctx, per));
            VTA_TRY(du.put(s, ctx, up));
            VTA_TRY(dd.put(s, ctx, dn));
            res.n_outputs = 3;
            res.n_combos = n;
            for (int i = 0; i < 3; ++i) VTA_TRY(alloc_out(s, res.out[i], n, T));
            VTA_TRY(vta_bollinger_bands_batch_device(
                ctx, e.d_close, (int32_t)T, m.fv_close, m.z_close, dp.as<int32_t>(),
                du.as<double>(), dd.as<double>(), (int32_t)n, (int32_t)maxp, res.out[0].ptr,
                res.out[1].ptr, res.out[2].ptr));
            res.t_launch = now_s() - t0;
            VTA_HIP(hipStreamSynchronize(ctx->stream));
            return VTA_OK;
        }
        default:
            return fail(VTA_ERR_INVALID_INPUT, "unknown request kind %d", r.kind);
    }
}

// Moves the result's device pointers to `out` (freed later, outside mu).
void take_outputs(vta_result &res, std::vector<void *> &out) {
    for (int i = 0; i < 3; ++i)
        if (res.out[i].ptr) {
            out.push_back(res.out[i].ptr);
            res.out[i].ptr = nullptr;
        }
}

// Called without mu held: stream-ordered frees of everything collected.
void flush_frees(vta_server *s, std::vector<void *> &mine) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        mine.insert(mine.end(), s->pending_free.begin(), s->pending_free.end());
        s->pending_free.clear();
    }
    for (void *p : mine) s->arena.release(p);
    mine.clear();
}

// Uploads the request's series (its uploader) or waits until another slot
// has (a cache hit or a concurrent miss on the same key).
vta_status prepare_series(vta_server *s, int slot, vta_ctx *ctx, Req &q) {
    CacheEntry &e = *q.entry;
    vta_status st = VTA_OK;
    if (q.uploader) {
        st = upload_series(s, slot, ctx, q);
        std::lock_guard<std::mutex> lk(s->mu);
        if (st != VTA_OK) {
            (void)hipStreamSynchronize(ctx->stream);
            if (e.d_close) s->pending_free.push_back(e.d_close);
            if (e.d_high) s->pending_free.push_back(e.d_high);
            if (e.d_low) s->pending_free.push_back(e.d_low);
            e.d_close = e.d_high = e.d_low = nullptr;
            if (e.cached) {
                // drop the broken entry; its bytes go back with this request
                s->cache.erase(e.key);
                s->lru.erase(e.lru);
                e.cached = false;
            }
        }
        e.ready = true;
        s->cv_done.notify_all();
    } else {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return e.ready; });
    }
    q.res.t_series = now_s() - s->t0;
    if (st == VTA_OK && !e.d_close) st = fail(VTA_ERR_HIP, "series upload failed");
    return st;
}

// Records the request's outcome and returns what it holds.
void finish_request(vta_server *s, vta_ctx *ctx, Req &q, vta_status st) {
    CacheEntry &e = *q.entry;
    std::lock_guard<std::mutex> lk(s->mu);
    q.res.status = st;
    q.res.t_done = now_s() - s->t0;
    if (st != VTA_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        take_outputs(q.res, s->pending_free);
        s->in_use -= q.out_bytes;
        q.out_bytes = 0;
    }
    e.refs -= 1;
    if (!e.cached) {
        // a one-off series, or an entry dropped after a failed upload
        if (e.refs == 0) {
            if (e.d_close) s->pending_free.push_back(e.d_close);
            if (e.d_high) s->pending_free.push_back(e.d_high);
            if (e.d_low) s->pending_free.push_back(e.d_low);
            e.d_close = e.d_high = e.d_low = nullptr;
        }
        s->in_use -= q.series_reserved;
        q.series_reserved = 0;
    }
    s->in_flight -= 1;
    q.done = true;
    s->cv_done.notify_all();
    s->cv_admit.notify_all();
}

// The packed form of run_request for a group of Bollinger Bands requests
// (see run_packed).
bool run_packed_bb(vta_server *s, vta_ctx *ctx, std::vector<std::shared_ptr<Req>> &g,
                   std::vector<vta_status> &st) {
    std::vector<vta_bb_job> jobs;
    std::vector<size_t> member;
    std::vector<std::unique_ptr<ArenaTable>> tables;
    std::vector<std::vector<int32_t>> pers(g.size());
    std::vector<std::vector<double>> ups(g.size()), dns(g.size());
    for (size_t i = 0; i < g.size(); ++i) {
        Req &q = *g[i];
        const CacheEntry &e = *q.entry;
        const SeriesMeta &m = e.meta;
        const size_t T = m.len;
        size_t n = 0;
        vta_status v = vta_bb_expand_grid(&q.r.bb, nullptr, 0, &n);
        std::vector<vta_bb_params> cs(n);
        if (v == VTA_OK) v = vta_bb_expand_grid(&q.r.bb, cs.data(), n, &n);
        if (v == VTA_OK && (n == 0 || T > (size_t)INT32_MAX || n > (size_t)INT32_MAX))
            v = fail(VTA_ERR_INVALID_INPUT, "invalid request sizes");
        size_t maxp = 0;
        for (size_t c = 0; v == VTA_OK && c < n; ++c) {
            if (cs[c].period == 0 || cs[c].period > T)
                v = fail(VTA_ERR_INVALID_INPUT, "invalid period %zu", cs[c].period);
            maxp = std::max(maxp, cs[c].period);
        }
        if (v == VTA_OK && ((size_t)m.fv_close >= T || T - (size_t)m.fv_close < maxp))
            v = fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", maxp);
        if (v == VTA_OK) {
            for (size_t c = 0; c < n; ++c) {
                pers[i].push_back((int32_t)cs[c].period);
                ups[i].push_back(cs[c].devup);
                dns[i].push_back(cs[c].devdn);
            }
            for (int t = 0; t < 3 && v == VTA_OK; ++t) {
                tables.push_back(std::make_unique<ArenaTable>());
                v = t == 0 ? tables.back()->put(s, ctx, pers[i])
                           : tables.back()->put(s, ctx, t == 1 ? ups[i] : dns[i]);
            }
        }
        for (int o = 0; o < 3 && v == VTA_OK; ++o) v = alloc_out(s, q.res.out[o], n, T);
        if (v != VTA_OK) {
            st[i] = v;
            continue;
        }
        q.res.n_outputs = 3;
        q.res.n_combos = n;
        const size_t tb = tables.size();
        vta_bb_job j{};
        j.d_prices = e.d_close;
        j.d_periods = tables[tb - 3]->as<int32_t>();
        j.d_devup = tables[tb - 2]->as<double>();
        j.d_devdn = tables[tb - 1]->as<double>();
        j.d_upper = q.res.out[0].ptr;
        j.d_middle = q.res.out[1].ptr;
        j.d_lower = q.res.out[2].ptr;
        j.series_len = (int32_t)T;
        j.first_valid = m.fv_close;
        j.first_nonfinite = m.z_close;
        j.n_combos = (int32_t)n;
        j.max_period = (int32_t)maxp;
        jobs.push_back(j);
        member.push_back(i);
    }
    if (jobs.empty()) return true;
    std::vector<BbPackJob> tab;
    unsigned blocks = 0;
    vta_status v = bb_pack_layout(ctx, jobs.data(), jobs.size(), tab, &blocks);
    void *d_tab = nullptr;
    if (v == VTA_OK) {
        d_tab = s->arena.alloc(tab.size() * sizeof(BbPackJob));
        if (!d_tab) {
            (void)hipStreamSynchronize(ctx->stream);   // the parameter tables' copies
            for (auto &q : g)
                for (int o = 0; o < 3; ++o)
                    if (q->res.out[o].ptr) {
                        s->arena.release(q->res.out[o].ptr);
                        q->res.out[o].ptr = nullptr;
                    }
            return false;
        }
        hipError_t e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(BbPackJob),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) v = fail(VTA_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
    }
    if (v == VTA_OK) v = bb_pack_launch(ctx, d_tab, (int)tab.size(), blocks);
    const double t_launch = now_s() - s->t0;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (v == VTA_OK && e != hipSuccess)
        v = fail(VTA_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
    if (d_tab) s->arena.release(d_tab);
    for (size_t i : member) {
        st[i] = v;
        g[i]->res.t_launch = t_launch;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->packs += 1;
    s->packed_requests += member.size();
    return true;
}

// The packed form of run_request for a group of EMA / RSI / ADX requests of
// one kind whose series are on the device.  st[i] receives each member's
// status; a member that fails validation is left out of the pack.  Returns
// false (nothing launched, st untouched) when the arena cannot hold the
// pack's table and carry, and the members are then run one by one.
// The packed form of run_request for a group of ALMA requests: the requests
// whose grid's cached sweep plan runs the first one's kernel (the split-F16
// form for 16+ combos under the servers' AUTO policy, else FP32) share one
// launch over their plans' tables (vta_alma_pack's kernels, bit-identical to
// each plan's own launch); the others are served alone.
bool run_packed_alma(vta_server *s, vta_ctx *ctx, std::vector<std::shared_ptr<Req>> &g,
                     std::vector<vta_status> &st) {
    std::vector<AlmaPackSrc> src;
    std::vector<size_t> member, alone;
    int kernel = -1;
    for (size_t i = 0; i < g.size(); ++i) {
        Req &q = *g[i];
        const CacheEntry &e = *q.entry;
        const SeriesMeta &m = e.meta;
        const size_t T = m.len;
        vta_alma_sweep_plan *plan = nullptr;
        size_t mp = 0, nc = 0;
        vta_status v = alma_plan_for(s, ctx, q.r.alma, &plan, &mp, &nc);
        if (v == VTA_OK && ((size_t)m.fv_close >= T || T - (size_t)m.fv_close < mp))
            v = fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", mp);
        if (v != VTA_OK) {
            st[i] = v;
            continue;
        }
        const AlmaTables &t = alma_plan_tables(plan);
        if (kernel < 0) kernel = (int)t.kernel;
        if ((int)t.kernel != kernel) {
            alone.push_back(i);
            continue;
        }
        v = alloc_out(s, q.res.out[0], nc, T);
        if (v != VTA_OK) {
            st[i] = v;
            continue;
        }
        q.res.n_outputs = 1;
        q.res.n_combos = nc;
        src.push_back({&t, alma_plan_device(plan), e.d_close, q.res.out[0].ptr, T,
                       (size_t)m.fv_close});
        member.push_back(i);
    }
    vta_status v = VTA_OK;
    AlmaPackImage img;
    void *d_rec = nullptr, *d_items = nullptr;
    if (!src.empty()) {
        v = alma_pack_build(ctx, (vta_alma_batch_kernel)kernel, src, img);
        if (v == VTA_OK) {
            d_rec = s->arena.alloc(img.records.size());
            d_items = s->arena.alloc(img.items.size() * sizeof(int32_t));
            if (!d_rec || !d_items) v = fail(VTA_ERR_OUT_OF_MEMORY, "device arena exhausted");
        }
        if (v != VTA_OK) {
            // served one by one instead
            for (void *p : {d_rec, d_items})
                if (p) s->arena.release(p);
            d_rec = d_items = nullptr;
            for (size_t i : member) {
                s->arena.release(g[i]->res.out[0].ptr);
                g[i]->res.out[0].ptr = nullptr;
                alone.push_back(i);
            }
            member.clear();
            v = VTA_OK;
        }
    }
    if (!member.empty()) {
        hipError_t e = hipMemcpyAsync(d_rec, img.records.data(), img.records.size(),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_items, img.items.data(), img.items.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) v = fail(VTA_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
        if (v == VTA_OK) v = alma_pack_enqueue(ctx, img, d_rec, d_items);
        const double t_launch = now_s() - s->t0;
        e = hipStreamSynchronize(ctx->stream);
        if (v == VTA_OK && e != hipSuccess)
            v = fail(VTA_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
        for (size_t i : member) {
            st[i] = v;
            g[i]->res.t_launch = t_launch;
        }
    }
    for (void *p : {d_rec, d_items})
        if (p) s->arena.release(p);
    for (size_t i : alone) {
        clear_error();
        st[i] = run_request(s, ctx, *g[i]);
    }
    if (!member.empty()) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->packs += 1;
        s->packed_requests += member.size();
    }
    return true;
}

bool run_packed(vta_server *s, vta_ctx *ctx, std::vector<std::shared_ptr<Req>> &g,
                std::vector<vta_status> &st) {
    const int kind = g[0]->r.kind;
    if (kind == VTA_REQ_BB) return run_packed_bb(s, ctx, g, st);
    if (kind == VTA_REQ_ALMA) return run_packed_alma(s, ctx, g, st);
    const int rk = kind == VTA_REQ_EMA ? VTA_REC_EMA : kind == VTA_REQ_RSI ? VTA_REC_RSI
                                                                             : VTA_REC_ADX;
    std::vector<vta_rec_job> jobs;
    std::vector<size_t> member;                     // job -> index in g
    std::vector<std::unique_ptr<ArenaTable>> tables;
    std::vector<std::vector<int32_t>> pers(g.size());
    for (size_t i = 0; i < g.size(); ++i) {
        Req &q = *g[i];
        const CacheEntry &e = *q.entry;
        const SeriesMeta &m = e.meta;
        const size_t T = m.len;
        size_t n = 0;
        vta_status (*grid)(const vta_range_usize *, size_t *, size_t, size_t *) =
            kind == VTA_REQ_EMA ? vta_ema_expand_grid
            : kind == VTA_REQ_RSI ? vta_rsi_expand_grid : vta_adx_expand_grid;
        vta_status v = grid(&q.r.period, nullptr, 0, &n);
        std::vector<size_t> ps(n);
        if (v == VTA_OK) v = grid(&q.r.period, ps.data(), n, &n);
        if (v == VTA_OK && (n == 0 || T > (size_t)INT32_MAX || n > (size_t)INT32_MAX))
            v = fail(VTA_ERR_INVALID_INPUT, "invalid request sizes");
        size_t maxp = 0;
        for (size_t p : ps) maxp = std::max(maxp, p);
        const int32_t fv = kind == VTA_REQ_ADX ? m.fv_adx : m.fv_close;
        const size_t need = kind == VTA_REQ_EMA ? maxp : kind == VTA_REQ_RSI ? maxp + 1 : 2 * maxp;
        if (v == VTA_OK && (maxp == 0 || (size_t)fv >= T || T - (size_t)fv < need))
            v = fail(VTA_ERR_INVALID_INPUT, "not enough valid data (needed >= %zu)", need);
        if (v == VTA_OK && kind == VTA_REQ_ADX && !m.hlc)
            v = fail(VTA_ERR_INVALID_INPUT, "ADX needs high, low and close");
        if (v == VTA_OK) {
            pers[i].assign(ps.begin(), ps.end());
            tables.push_back(std::make_unique<ArenaTable>());
            v = tables.back()->put(s, ctx, pers[i]);
        }
        if (v == VTA_OK) v = alloc_out(s, q.res.out[0], n, T);
        if (v != VTA_OK) {
            st[i] = v;
            continue;
        }
        q.res.n_outputs = 1;
        q.res.n_combos = n;
        vta_rec_job j{};
        j.d_close = e.d_close;
        j.d_high = e.d_high;
        j.d_low = e.d_low;
        j.d_periods = tables.back()->as<int32_t>();
        j.d_out = q.res.out[0].ptr;
        j.series_len = (int32_t)T;
        j.first_valid = fv;
        j.n_combos = (int32_t)n;
        j.max_period = (int32_t)maxp;
        jobs.push_back(j);
        member.push_back(i);
    }
    if (jobs.empty()) return true;
    std::vector<PackJob> tab;
    std::vector<size_t> off;
    size_t elems = 0;
    PackDims dims;
    vta_status v = pack_layout(ctx, rk, jobs.data(), jobs.size(), tab, off, &elems, &dims);
    void *d_tab = nullptr, *d_flags = nullptr, *d_carry = nullptr;
    if (v == VTA_OK) {
        d_tab = s->arena.alloc(tab.size() * sizeof(PackJob));
        d_flags = s->arena.alloc(tab.size() * sizeof(int32_t));
        d_carry = s->arena.alloc(std::max<size_t>(elems, 1) * sizeof(double));
        if (!d_tab || !d_flags || !d_carry) {
            for (void *p : {d_tab, d_flags, d_carry})
                if (p) s->arena.release(p);
            (void)hipStreamSynchronize(ctx->stream);   // the period tables' copies
            for (auto &q : g)
                if (q->res.out[0].ptr) {
                    s->arena.release(q->res.out[0].ptr);
                    q->res.out[0].ptr = nullptr;
                }
            return false;
        }
        for (size_t j = 0; j < tab.size(); ++j)
            tab[j].carry = static_cast<double *>(d_carry) + off[j];
        hipError_t e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(PackJob),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) v = fail(VTA_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
    }
    if (v == VTA_OK) v = pack_launch(ctx, dims, d_tab, static_cast<int32_t *>(d_flags));
    const double t_launch = now_s() - s->t0;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (v == VTA_OK && e != hipSuccess)
        v = fail(VTA_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
    for (void *p : {d_tab, d_flags, d_carry})
        if (p) s->arena.release(p);
    for (size_t i : member) {
        st[i] = v;
        g[i]->res.t_launch = t_launch;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->packs += 1;
    s->packed_requests += member.size();
    return true;
}

void worker(vta_server *s, int slot) {
    vta_ctx *ctx = s->ctxs[slot];
    (void)hipSetDevice(s->device);
    std::vector<void *> local_frees;
    for (;;) {
        std::vector<std::shared_ptr<Req>> g;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                std::shared_ptr<Req> q = try_admit(s);
                if (q) {
                    g.push_back(q);
                    break;
                }
                s->cv_admit.wait(lk);
            }
            if (packable(*g[0]))
                while (g.size() < kPackMax) {
                    std::shared_ptr<Req> q = try_admit_kind(s, g[0]->r.kind);
                    if (!q) break;
                    g.push_back(q);
                }
            const double t = now_s() - s->t0;
            for (auto &q : g) q->res.t_start = t;
            s->in_flight += (int)g.size();
            s->peak_in_flight = std::max(s->peak_in_flight, s->in_flight);
        }
        flush_frees(s, local_frees);   // evictions made room for these requests
        std::vector<vta_status> st(g.size(), VTA_OK);
        for (size_t i = 0; i < g.size(); ++i) st[i] = prepare_series(s, slot, ctx, *g[i]);
        bool done = false;
        if (g.size() > 1) {
            std::vector<std::shared_ptr<Req>> ready;
            std::vector<size_t> idx;
            for (size_t i = 0; i < g.size(); ++i)
                if (st[i] == VTA_OK) {
                    ready.push_back(g[i]);
                    idx.push_back(i);
                }
            if (ready.size() > 1) {
                clear_error();
                std::vector<vta_status> rs(ready.size(), VTA_OK);
                done = run_packed(s, ctx, ready, rs);
                if (done)
                    for (size_t k = 0; k < idx.size(); ++k) st[idx[k]] = rs[k];
            }
        }
        if (!done)
            for (size_t i = 0; i < g.size(); ++i)
                if (st[i] == VTA_OK) {
                    clear_error();
                    st[i] = run_request(s, ctx, *g[i]);
                }
        for (size_t i = 0; i < g.size(); ++i) finish_request(s, ctx, *g[i], st[i]);
        flush_frees(s, local_frees);
    }
}

}  // namespace

extern "C" {

vta_status vta_server_create(int device_id, int n_slots, size_t hbm_budget, vta_server **out) {
    clear_error();
    if (!out || n_slots <= 0 || n_slots > 64 || hbm_budget == 0)
        return fail(VTA_ERR_INVALID_INPUT, "invalid server arguments");
    auto *s = new vta_server();
    s->device = device_id;
    s->budget = hbm_budget;
    s->t0 = now_s();
    {
        // the budget plus slack for parameter tables and fragmentation
        vta_status st = VTA_OK;
        if (hipSetDevice(device_id) != hipSuccess)
            st = fail(VTA_ERR_HIP, "hipSetDevice(%d) failed", device_id);
        if (st == VTA_OK) st = s->arena.init(hbm_budget + hbm_budget / 16 + (64u << 20));
        if (st != VTA_OK) {
            delete s;
            return st;
        }
    }
    for (int i = 0; i < n_slots; ++i) {
        vta_ctx *c = nullptr;
        vta_status st = vta_ctx_create(device_id, &c);
        if (st == VTA_OK) st = vta_ctx_preload(c);
        if (st != VTA_OK) {
            if (c) vta_ctx_destroy(c);
            for (auto *x : s->ctxs) vta_ctx_destroy(x);
            delete s;
            return st;
        }
        s->ctxs.push_back(c);
    }
    s->pinned.assign(n_slots, nullptr);
    s->pinned_cap.assign(n_slots, 0);
    for (int i = 0; i < n_slots; ++i) s->workers.emplace_back(worker, s, i);
    *out = s;
    return VTA_OK;
}

vta_status vta_server_destroy(vta_server *s) {
    clear_error();
    if (!s) return VTA_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_admit.notify_all();
        s->cv_done.notify_all();
    }
    for (auto &t : s->workers) t.join();
    (void)hipSetDevice(s->device);
    std::vector<void *> frees;
    for (auto &kv : s->reqs)
        if (kv.second->done) take_outputs(kv.second->res, frees);
    for (auto &kv : s->cache) {
        frees.push_back(kv.second->d_close);
        if (kv.second->d_high) frees.push_back(kv.second->d_high);
        if (kv.second->d_low) frees.push_back(kv.second->d_low);
    }
    flush_frees(s, frees);
    (void)hipDeviceSynchronize();
    for (auto &kv : s->alma_plans) alma_grid_plan_destroy(kv.second);
    for (float *p : s->pinned)
        if (p) (void)hipHostFree(p);
    for (auto *c : s->ctxs) vta_ctx_destroy(c);
    delete s;
    return VTA_OK;
}

vta_status vta_server_submit(vta_server *s, const vta_request *r, uint64_t *req_id) {
    clear_error();
    if (!s || !r || !req_id || !r->close || r->len == 0)
        return fail(VTA_ERR_INVALID_INPUT, "invalid request");
    if (r->kind == VTA_REQ_ADX && (!r->high || !r->low))
        return fail(VTA_ERR_INVALID_INPUT, "ADX needs high, low and close");
    auto q = std::make_shared<Req>();
    q->r = *r;
    const size_t n = r->len;
    q->meta.len = n;
    q->meta.hlc = r->high && r->low;
    q->close.assign(r->close, r->close + n);
    if (q->meta.hlc) {
        q->high.assign(r->high, r->high + n);
        q->low.assign(r->low, r->low + n);
    }
    q->r.close = q->r.high = q->r.low = nullptr;
    const size_t fvc = first_nan_end(q->close.data(), n);
    q->meta.fv_close = (int32_t)std::min(fvc, n);
    size_t z = fvc;
    while (z < n && std::isfinite(q->close[z])) ++z;
    q->meta.z_close = (int32_t)z;
    if (q->meta.hlc) {
        size_t a = first_nan_end(q->high.data(), n), b = first_nan_end(q->low.data(), n);
        q->meta.fv_adx = (int32_t)std::max({a, b, fvc});
    }
    size_t n_combos = 0;
    int n_out = 1;
    const size_t cells = grid_outputs(*r, &n_combos, &n_out);
    if (cells == 0) return fail(VTA_ERR_INVALID_INPUT, "invalid parameter grid");
    if (!mul_ok(cells, n, &q->out_bytes) || !mul_ok(q->out_bytes, sizeof(float), &q->out_bytes))
        return fail(VTA_ERR_INVALID_INPUT, "output size overflow");
    q->series_bytes = (q->meta.hlc ? 3 : 1) * n * sizeof(float);
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    q->res.t_submit = now_s() - s->t0;
    s->reqs.emplace(q->id, q);
    s->queue.push_back(q);
    *req_id = q->id;
    s->cv_admit.notify_one();
    return VTA_OK;
}

vta_status vta_server_wait(vta_server *s, uint64_t req_id, int timeout_ms, vta_result *res) {
    clear_error();
    if (!s || !res) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return fail(VTA_ERR_INVALID_INPUT, "unknown request %llu",
                                         (unsigned long long)req_id);
    std::shared_ptr<Req> q = it->second;
    auto ready = [&] { return q->done; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return fail(VTA_ERR_TIMEOUT, "request %llu not finished", (unsigned long long)req_id);
    }
    *res = q->res;
    return VTA_OK;
}

vta_status vta_server_release(vta_server *s, uint64_t req_id) {
    clear_error();
    if (!s) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::vector<void *> frees;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req_id);
        if (it == s->reqs.end()) return fail(VTA_ERR_INVALID_INPUT, "unknown request");
        std::shared_ptr<Req> q = it->second;
        if (!q->done) return fail(VTA_ERR_INVALID_INPUT, "request still running");
        take_outputs(q->res, frees);
        s->in_use -= q->out_bytes;
        q->out_bytes = 0;
        s->reqs.erase(it);
        s->cv_admit.notify_all();
    }
    (void)hipSetDevice(s->device);
    flush_frees(s, frees);
    return VTA_OK;
}

vta_status vta_server_get_stats(vta_server *s, vta_server_stats *out) {
    clear_error();
    if (!s || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->budget = s->budget;
    out->in_use = s->in_use;
    out->peak_in_use = s->peak_in_use;
    out->queued = s->queue.size();
    out->in_flight = (size_t)s->in_flight;
    out->peak_in_flight = (size_t)s->peak_in_flight;
    out->cached_series = s->cache.size();
    out->packs = s->packs;
    out->packed_requests = s->packed_requests;
    return VTA_OK;
}

double vta_server_now(const vta_server *s) { return s ? now_s() - s->t0 : 0.0; }

}  // extern "C"
