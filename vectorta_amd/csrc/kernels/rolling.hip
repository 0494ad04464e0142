// Rolling-window indicators for gfx950: Bollinger Bands (SMA, population sd),
// percentile_nearest_rank and the rolling median absolute deviation.
//
// Semantics (CPU oracle):
//   BB   src/indicators/bollinger_bands.rs:566-672 (running f64 sums; a NaN
//        after first_valid poisons every later output)
//   PNR  src/indicators/percentile_nearest_rank.rs:190-266 (sorted window of the
//        non-NaN values, k = round(p/100 * wl) - 1 clamped)
//   MAD  src/indicators/medium_ad.rs:293-402 (median, then median of |x - med|;
//        even windows average the two middle order statistics; any NaN -> NaN)
//
// Time is cut into chunks; every chunk rebuilds its window state exactly from
// the W bars before it (windowed indicators need no burn-in approximation).
// BB re-anchors its f64 sums at every chunk start; its "poison forever" rule
// is reproduced from the series' first non-finite bar (see bb_run).
//
// PNR and MAD keep each lane's window sorted in LDS (column layout, stride 64
// lanes, so every lane hits its own bank): one binary search for the outgoing
// value, one for the incoming value (run interleaved), and a single shift of
// the span between them (four moves per batch).  Each step is a chain of
// dependent LDS round trips and ~1,000 instructions at 1.5 waves per SIMD
// (the windows' LDS caps residency), so the kernels are latency-bound: LDS is
// busy ~40% of the cycles and waves wait about half their time.  The
// many-series forms run time wave-uniform with their bars loaded a block
// ahead.  NaN bars are kept as +Inf keys and counted, so every step is one
// replace.  PNR reads its rank among the first (window - NaNs) keys; MAD takes
// the median and selects from the two monotone deviation sequences either side
// of it by a two-array order-statistic search.  All comparisons are on the
// f32 inputs and all arithmetic is f64 in the oracle's order, so selected values
// and MAD are f32(oracle) exactly.
#include "vta_common.h"

struct roll_args {
    const float *x;
    const int32_t *p_len;    // batch: per combo
    const double *p_a;       // batch: BB devup / PNR percentage per combo
    const double *p_b;       // batch: BB devdn per combo
    const int32_t *fvs;      // many-series: per series; batch: [0]
    const int32_t *poison;   // BB: first non-finite bar >= first_valid, per series (ms)
                             // or [0] (batch); null means none known (one chunk)
    float *out0;
    float *out1;
    float *out2;
    double a1;               // many-series: devup / percentage
    double b1;               // many-series: devdn
    int T;
    int nlanes;
    int chunk;
    int wcap;                // LDS window capacity per lane (PNR, MAD); many-series BB:
                             // ring length (0: outgoing bars read from HBM)
    int p1;                  // many-series: period / length
    int pad_;
};

#define RL_NAN_D __builtin_bit_cast(double, 0x7ff8000000000000ull)

__device__ __forceinline__ int rl_wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// ---------------------------------------------------------------- BB engine

struct BbEng {
    double inv_n, sum, sum_sq, up, dn;
    int p, fv;
    __device__ void setup(int p_, int fv_, double up_, double dn_) {
        p = p_; fv = fv_; up = up_; dn = dn_;
        inv_n = 1.0 / (double)p;
    }
    __device__ void out(double m, double var, float *o) const {
        const double sd = sqrt(fmax(var, 0.0));
        o[0] = (float)(m + up * sd);
        o[1] = (float)m;
        o[2] = (float)(m - dn * sd);
    }
    __device__ void step(double nw, double old) {
        sum += nw - old;                                    // bollinger_bands.rs:640-641
        sum_sq = (sum_sq - old * old) + nw * nw;
    }
    __device__ void emit(float *o) const {
        const double m = sum * inv_n;
        const double var = (sum_sq * inv_n) - (m * m);
        out(m, var, o);
    }
};

// One-series batch outputs are rows [combo][t] and a lane (combo) produces
// one bar per step, so a wave stages RB_W bars of its 64 rows in LDS and then
// stores row segments: each store instruction covers two rows x 32
// consecutive bars instead of 64 scattered dwords.
#define RB_W 32
#define RB_LD 33

__device__ __forceinline__ void rb_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void rb_flush(const float *ob, float *out, int row0, int nrows, int T,
                                         int t0, int nb) {
    const int lane = threadIdx.x & 63;
    const int k = lane & 31;
#pragma unroll 4
    for (int i = 0; i < 32; ++i) {
        const int r = 2 * i + (lane >> 5);
        const int row = row0 + r;
        if (row < nrows && k < nb) out[(size_t)row * T + t0 + k] = ob[r * RB_LD + k];
    }
}

// Window sums at bar t from scratch: sum over [t-p+1, t] in order (the
// oracle's first-window loop, bollinger_bands.rs:613-620).
template <class LD>
__device__ __forceinline__ void bb_anchor(BbEng &e, int t, LD ld) {
    double s = 0.0, q = 0.0;
    for (int k = 0; k < e.p; ++k) {
        const double v = ld(t - e.p + 1 + k);
        s += v;
        q += v * v;
    }
    e.sum = s;
    e.sum_sq = q;
}

// Outputs of the chunk [c0, c1).  z is the first non-finite bar at or after
// first_valid (T if none).  Windows ending before z are finite, so any start
// bar before z may re-anchor.  The running sums turn non-finite at z and NaN
// for good once bar z leaves the window (x - Inf = -Inf, Inf + -Inf = NaN; a
// NaN bar poisons at once), so every output from z + p on is NaN.  Between z
// and z + p - 1 the sums' class (+-Inf or NaN) depends only on the non-finite
// bars in the window, which an anchor at max(warm, z) sees too: a chunk
// starting after z replays at most p bars from there.
//
// RING (many-series, windows up to 256): the lane's last p bars also sit in an
// LDS ring (slot u mod p of the lane's column `ring`, stride 64): bar u's
// outgoing bar u - p is read from the slot bar u then overwrites.  Read from
// HBM instead, the outgoing row was fetched again (the output stream had
// pushed it out of L2 since the same wave read it p bars before).
template <bool RING, class LD, class ST>
__device__ __forceinline__ void bb_run(BbEng &e, int c0, int c1, int z, bool active, LD ld,
                                       ST st, float *ring) {
    if (!active) return;
    const int warm = e.fv + e.p - 1;
    const long zp = (long)z + e.p;
    const float q = VTA_QNAN_F32;
    const float nan3[3] = {q, q, q};
    for (int u = c0; u < min(c1, warm); ++u) st(u, nan3);
    const int cs = max(c0, warm);
    const int ce = (int)min((long)c1, zp);
    if (cs < ce) {
        int t = cs < z ? cs : max(warm, z);
        int slot = 0;                                 // ring slot of bar t + 1
        if (RING) {
            // the anchor's bars t - p + 1 .. t into their slots
            auto ldr = [&](int u) -> double {
                const double v = ld(u);
                ring[64 * (u % e.p)] = (float)v;
                return v;
            };
            bb_anchor(e, t, ldr);
            slot = (t + 1) % e.p;
        } else {
            bb_anchor(e, t, ld);
        }
        auto emit = [&](int tt) {
            const double m = e.sum * e.inv_n;
            const double var = (e.sum_sq * e.inv_n) - (m * m);
            float o[3];
            e.out(m, var, o);
            if (tt >= cs) st(tt, o);
        };
        emit(t);
        // blocks of 8 bars: the 16 loads of a block are issued before its
        // sequential updates (indices clamped into [t+1, ce-1])
        while (t + 1 < ce) {
            double nw[8], od[8];
            const int nk = min(8, ce - 1 - t);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int u = min(t + 1 + k, ce - 1);
                nw[k] = ld(u);
                if (!RING) od[k] = ld(u - e.p);
            }
            if (RING) {
                // outgoing bars: ring slots (or, p < 8, this block's own new
                // bars), all read before the block's bars go in
                int sl = slot;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    od[k] = k < e.p ? (double)ring[64 * sl] : nw[k - e.p < 0 ? 0 : k - e.p];
                    sl = sl + 1 == e.p ? 0 : sl + 1;
                }
                sl = slot;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (k < nk) {
                        ring[64 * sl] = (float)nw[k];
                        sl = sl + 1 == e.p ? 0 : sl + 1;
                    }
                }
                slot = sl;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k < nk) {
                    e.sum += nw[k] - od[k];                     // bollinger_bands.rs:640-641
                    e.sum_sq = (e.sum_sq - od[k] * od[k]) + nw[k] * nw[k];
                    emit(t + 1 + k);
                }
            }
            t += nk;
        }
    }
    for (int u = (int)max((long)c0, min(zp, (long)c1)); u < c1; ++u) st(u, nan3);
}

// many series: grid.x = 64-series groups, grid.y = chunks; dynamic LDS
// 64 * period floats when the period fits the ring (a.wcap > 0), else none
extern "C" __global__ __launch_bounds__(64) void bb_ms1p_f32(roll_args a) {
    extern __shared__ float bb_ring[];
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    const int sc = min(s, N - 1);
    BbEng e;
    e.setup(a.p1, a.fvs[sc], a.a1, a.b1);
    const int z = a.poison ? a.poison[sc] : a.T;
    auto ld = [&](int t) -> double { return (double)a.x[(size_t)t * N + sc]; };
    auto st = [&](int t, const float *o) {
        const size_t off = (size_t)t * N + s;
        a.out0[off] = o[0];
        a.out1[off] = o[1];
        a.out2[off] = o[2];
    };
    if (a.wcap > 0) bb_run<true>(e, c0, c1, z, active, ld, st, bb_ring + lane);
    else bb_run<false>(e, c0, c1, z, active, ld, st, nullptr);
}

// one series x many (period, devup, devdn): grid.x = combos / 64, grid.y = chunks.
// Lane = combo; all lanes step through the chunk's bars together (the same
// output rule as bb_run: re-anchor at the first output bar of the chunk,
// replaying at most p bars after the first non-finite bar z, NaN from z + p).
// What one block of a one-series BB batch works on: the series, the combos'
// periods / deviations and outputs, the chunking, first_valid, the first
// non-finite bar z (T if none), and which combo group and chunk it is.
struct bb_view {
    const float *x;
    const int32_t *p_len;
    const double *p_a, *p_b;
    float *out0, *out1, *out2;
    int T, nlanes, chunk, fv, z, group, k;
};

__device__ __forceinline__ void bb_batch_body(const bb_view &a) {
    __shared__ float ob[3][64 * RB_LD];
    const int lane = threadIdx.x & 63;
    const int c = a.group * 64 + lane;
    const bool active = c < a.nlanes;
    const int T = a.T;
    const int c0 = a.k * a.chunk;
    const int c1 = min(c0 + a.chunk, T);
    const int cc = min(c, a.nlanes - 1);
    BbEng e;
    e.setup(a.p_len[cc], a.fv, a.p_a[cc], a.p_b[cc]);
    const int z = a.z;
    auto ld = [&](int t) -> double { return (double)a.x[t]; };
    const int warm = e.fv + e.p - 1;
    const long zp = (long)z + e.p;
    const int cs = max(c0, warm);
    const int ce = (int)min((long)c1, zp);
    const float q = VTA_QNAN_F32;
    bool have = false;
    for (int t0 = c0; t0 < c1; t0 += RB_W) {
        const int nb = min(RB_W, c1 - t0);
        for (int k8 = 0; k8 < nb; k8 += 8) {
            // the 16 loads of 8 bars first (indices clamped into the chunk)
            double nw[8], od[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int u = min(t0 + k8 + j, c1 - 1);
                nw[j] = ld(u);
                od[j] = ld(max(u - e.p, 0));
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k8 + j;
                const int t = t0 + k;
                float o[3] = {q, q, q};
                if (k < nb && active && t >= cs && t < ce) {
                    if (!have) {
                        const int ta = cs < z ? cs : max(warm, z);
                        bb_anchor(e, ta, ld);
                        for (int u = ta + 1; u <= t; ++u) e.step(ld(u), ld(u - e.p));
                        have = true;
                    } else {
                        e.step(nw[j], od[j]);
                    }
                    e.emit(o);
                }
                if (k < nb) {
                    ob[0][lane * RB_LD + k] = o[0];
                    ob[1][lane * RB_LD + k] = o[1];
                    ob[2][lane * RB_LD + k] = o[2];
                }
            }
        }
        rb_sync();
        rb_flush(ob[0], a.out0, a.group * 64, a.nlanes, T, t0, nb);
        rb_flush(ob[1], a.out1, a.group * 64, a.nlanes, T, t0, nb);
        rb_flush(ob[2], a.out2, a.group * 64, a.nlanes, T, t0, nb);
        rb_sync();
    }
}

extern "C" __global__ __launch_bounds__(64) void bb_batch_f32(roll_args a) {
    bb_view v;
    v.x = a.x;
    v.p_len = a.p_len;
    v.p_a = a.p_a;
    v.p_b = a.p_b;
    v.out0 = a.out0;
    v.out1 = a.out1;
    v.out2 = a.out2;
    v.T = a.T;
    v.nlanes = a.nlanes;
    v.chunk = a.chunk;
    v.fv = a.fvs[0];
    v.z = a.poison ? a.poison[0] : a.T;
    v.group = blockIdx.x;
    v.k = blockIdx.y;
    bb_batch_body(v);
}

// Packed one-series BB batches: job j = one series with its own combos;
// a 1-D grid over every job's (combo group, chunk) blocks, job j owning
// [block0, block0 + groups * ceil(T / chunk)); each block finds its job by
// binary search and runs exactly the block of the job's direct launch.
struct bb_job {
    const float *x;
    const int32_t *p_len;
    const double *p_a, *p_b;
    float *out0, *out1, *out2;
    int T, nlanes, fv, z, chunk, groups, block0, pad_;
};

extern "C" __global__ __launch_bounds__(64) void bb_packed_f32(const bb_job *jobs, int n_jobs) {
    const int flat = blockIdx.x;
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].block0 <= flat) lo = mid; else hi = mid - 1;
    }
    const bb_job &j = jobs[lo];
    const int local = flat - j.block0;
    bb_view v;
    v.x = j.x;
    v.p_len = j.p_len;
    v.p_a = j.p_a;
    v.p_b = j.p_b;
    v.out0 = j.out0;
    v.out1 = j.out1;
    v.out2 = j.out2;
    v.T = j.T;
    v.nlanes = j.nlanes;
    v.chunk = j.chunk;
    v.fv = j.fv;
    v.z = j.z;
    v.group = local % j.groups;
    v.k = local / j.groups;
    bb_batch_body(v);
}

// ---------------------------------------------------------------- sorted window
//
// A lane's sorted multiset lives in LDS column `col` (element i at col[64*i]).

struct SortedWin {
    float *col;
    int n;
    __device__ float at(int i) const { return col[64 * i]; }
    // number of elements < v (lower bound); `steps` binary-search rounds cover n
    __device__ int lower(float v, int steps) const {
        int lo = 0, hi = n;
        for (int r = 0; r < steps; ++r) {
            if (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (at(mid) < v) lo = mid + 1; else hi = mid;
            }
        }
        return lo;
    }
    __device__ void insert(float v, int steps) {
        const int pos = lower(v, steps);
        int i = n;
        for (; i - 4 >= pos; i -= 4) {
            const float a = col[64 * (i - 1)], b = col[64 * (i - 2)], c = col[64 * (i - 3)],
                        d = col[64 * (i - 4)];
            col[64 * i] = a;
            col[64 * (i - 1)] = b;
            col[64 * (i - 2)] = c;
            col[64 * (i - 3)] = d;
        }
        for (; i > pos; --i) col[64 * i] = col[64 * (i - 1)];
        col[64 * pos] = v;
        ++n;
    }
    __device__ void set(int i, float v) { col[64 * i] = v; }
    // Replaces one element equal to `vo` (present) by `vi`: the elements
    // strictly between the two positions move one slot towards vo's slot.
    // The step is bound by dependent LDS round trips, not LDS bandwidth, so
    // the two binary searches run interleaved (their reads in flight
    // together) and the moves go in batches of four whose reads are all
    // issued before their writes (a batch's reads lie past the previous
    // batch's writes), then at most three single moves.
    __device__ void replace(float vo, float vi, int steps) {
        int lo_o = 0, hi_o = n, lo_i = 0, hi_i = n;
        for (int r = 0; r < steps; ++r) {
            const int mo = (lo_o + hi_o) >> 1, mi = (lo_i + hi_i) >> 1;
            const float ao = at(min(mo, n - 1)), ai = at(min(mi, n - 1));
            if (lo_o < hi_o) {
                if (ao < vo) lo_o = mo + 1; else hi_o = mo;
            }
            if (lo_i < hi_i) {
                if (ai < vi) lo_i = mi + 1; else hi_i = mi;
            }
        }
        const int po = lo_o, pi = lo_i;
        int j = po;
        if (pi > po) {
            // col[j] = col[j + 1] for j in [po, pi - 1)
            int span = pi - 1 - po;
            for (; span >= 4; span -= 4, j += 4) {
                const float a = at(j + 1), b = at(j + 2), c = at(j + 3), d = at(j + 4);
                set(j, a);
                set(j + 1, b);
                set(j + 2, c);
                set(j + 3, d);
            }
            for (; span > 0; --span, ++j) set(j, at(j + 1));
        } else {
            // col[j] = col[j - 1] for j from po down to pi + 1
            int span = po - pi;
            for (; span >= 4; span -= 4, j -= 4) {
                const float a = at(j - 1), b = at(j - 2), c = at(j - 3), d = at(j - 4);
                set(j, a);
                set(j - 1, b);
                set(j - 2, c);
                set(j - 3, d);
            }
            for (; span > 0; --span, --j) set(j, at(j - 1));
        }
        set(j, vi);
    }
};

__device__ __forceinline__ int bsearch_steps(int cap) {
    int s = 1;
    while ((1 << s) <= cap) ++s;
    return s;
}

// NaN bars enter the window as +Inf keys (counted separately).
__device__ __forceinline__ float rl_key(float v) { return isnan(v) ? INFINITY : v; }

// percentile_nearest_rank.rs:218-230: k = round(fma(p_frac, wl, 0)) - 1 (Rust
// round = half away from zero, as C round), clamped to [0, wl - 1].
__device__ __forceinline__ int pnr_rank(double p_frac, int wl) {
    const long raw = (long)round(fma(p_frac, (double)wl, 0.0)) - 1;
    int k = raw <= 0 ? 0 : (int)raw;
    if (k >= wl) k = wl - 1;
    return k;
}

// Fills the window with the `len` bars ending at t (NaN as +Inf keys).
template <class LD>
__device__ __forceinline__ int fill_window(SortedWin &w, int t, int len, int steps, LD ld) {
    int nan_cnt = 0;
    w.n = 0;
    for (int i = t - len + 1; i <= t; ++i) {
        const float v = (float)ld(i);
        nan_cnt += isnan(v) ? 1 : 0;
        w.insert(rl_key(v), steps);
    }
    return nan_cnt;
}

// One lane's sorted window over the output bars [c0, c1) of a many-series
// chunk: first output bar `first` (the window is filled there), window `len`.
// Time runs wave-uniform, so every bar's loads and stores are one row of the
// time-major matrix, and bars arrive in blocks of RW_B whose loads (incoming
// x[t], outgoing x[t - len]) are issued one block ahead: each step's LDS
// work then never waits on a global load.  out(w, nan_cnt) gives the output
// of the current window; st(t, v) stores it (NaN before `first`).
#define RW_B 8
template <class LD, class ST, class OUT>
__device__ __forceinline__ void window_run(SortedWin &w, int len, int first, int c0, int c1,
                                           bool active, int steps, LD ld, ST st, OUT out) {
    float ci[RW_B], co[RW_B], ni[RW_B], no[RW_B];
    auto load = [&](int t0, float *bi, float *bo) {
#pragma unroll
        for (int k = 0; k < RW_B; ++k) {
            const int t = min(t0 + k, c1 - 1);
            bi[k] = (float)ld(t);
            bo[k] = (float)ld(max(t - len, 0));
        }
    };
    int nan_cnt = 0;
    load(c0, ci, co);
    for (int t0 = c0; t0 < c1; t0 += RW_B) {
        if (t0 + RW_B < c1) load(t0 + RW_B, ni, no);
#pragma unroll
        for (int k = 0; k < RW_B; ++k) {
            const int t = t0 + k;
            if (t >= c1) break;
            float o = VTA_QNAN_F32;
            if (active && t >= first) {
                if (t == first) {
                    nan_cnt = fill_window(w, t, len, steps, ld);
                } else {
                    nan_cnt += (isnan(ci[k]) ? 1 : 0) - (isnan(co[k]) ? 1 : 0);
                    w.replace(rl_key(co[k]), rl_key(ci[k]), steps);
                }
                o = out(w, nan_cnt);
            }
            if (active) st(t, o);
        }
#pragma unroll
        for (int k = 0; k < RW_B; ++k) {
            ci[k] = ni[k];
            co[k] = no[k];
        }
    }
}

// ---------------------------------------------------------------- PNR

// The order statistics of the non-NaN bars of a window equal the first
// (len - nan_cnt) order statistics of the keyed window, since every NaN key
// (+Inf) is >= every real value.
template <class LD, class ST>
__device__ __forceinline__ void pnr_run(SortedWin &w, int len, double pct, int fv, int c0, int c1,
                                        bool active, int steps, LD ld, ST st) {
    const int start_i = fv + len - 1;                 // first output bar
    int t = max(c0, start_i);
    if (!active) return;
    for (int u = c0; u < min(t, c1); ++u) st(u, VTA_QNAN_F32);
    if (t >= c1) return;
    const double p_frac = pct * 0.01;
    int nan_cnt = fill_window(w, t, len, steps, ld);
    for (;;) {
        const int wl = len - nan_cnt;
        st(t, wl == 0 ? VTA_QNAN_F32 : w.at(pnr_rank(p_frac, wl)));
        if (t + 1 >= c1) break;
        const float vo = (float)ld(t + 1 - len);
        const float vi = (float)ld(t + 1);
        nan_cnt += (isnan(vi) ? 1 : 0) - (isnan(vo) ? 1 : 0);
        w.replace(rl_key(vo), rl_key(vi), steps);
        ++t;
    }
}

extern "C" __global__ __launch_bounds__(64) void pnr_ms1p_f32(roll_args a) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    SortedWin w{lds + lane, 0};
    const int sc = min(s, N - 1);
    auto ld = [&](int t) -> double { return (double)a.x[(size_t)t * N + sc]; };
    auto st = [&](int t, float v) { a.out0[(size_t)t * N + s] = v; };
    const int len = a.p1;
    const double p_frac = a.a1 * 0.01;
    const int fv = active ? a.fvs[s] : a.T;
    const int first = max(c0, fv + len - 1);
    auto out = [&](const SortedWin &win, int nan_cnt) -> float {
        const int wl = len - nan_cnt;
        return wl == 0 ? VTA_QNAN_F32 : win.at(pnr_rank(p_frac, wl));
    };
    window_run(w, len, first, c0, c1, active, bsearch_steps(a.wcap), ld, st, out);
}

// Lockstep one-series batch: lane = (length, percentage) combo; dynamic LDS =
// 64 (wcap + 1) window slots followed by the 64 x RB_LD output staging.
extern "C" __global__ __launch_bounds__(64) void pnr_batch_f32(roll_args a) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 64 + lane;
    const bool active = c < a.nlanes;
    const int T = a.T;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, T);
    const int cc = min(c, a.nlanes - 1);
    SortedWin w{lds + lane, 0};
    float *ob = lds + 64 * (a.wcap + 1);
    const int steps = bsearch_steps(a.wcap);
    const int len = a.p_len[cc];
    const double p_frac = a.p_a[cc] * 0.01;
    const int cs = max(c0, a.fvs[0] + len - 1);      // first output bar of the chunk
    auto ld = [&](int t) -> double { return (double)a.x[t]; };
    int nan_cnt = 0;
    bool have = false;
    for (int t0 = c0; t0 < c1; t0 += RB_W) {
        const int nb = min(RB_W, c1 - t0);
        for (int k = 0; k < nb; ++k) {
            const int t = t0 + k;
            float o = VTA_QNAN_F32;
            if (active && t >= cs) {
                if (!have) {
                    nan_cnt = fill_window(w, t, len, steps, ld);
                    have = true;
                } else {
                    const float vo = a.x[t - len];
                    const float vi = a.x[t];
                    nan_cnt += (isnan(vi) ? 1 : 0) - (isnan(vo) ? 1 : 0);
                    w.replace(rl_key(vo), rl_key(vi), steps);
                }
                const int wl = len - nan_cnt;
                o = wl == 0 ? VTA_QNAN_F32 : w.at(pnr_rank(p_frac, wl));
            }
            ob[lane * RB_LD + k] = o;
        }
        rb_sync();
        rb_flush(ob, a.out0, blockIdx.x * 64, a.nlanes, T, t0, nb);
        rb_sync();
    }
}

// ---------------------------------------------------------------- MAD

// k-th smallest (0-based) deviation |a_i - med| over the sorted window and the
// (k+1)-th in *next.  Left of `split` the deviations ascend right-to-left
// (L_i = |a[split-1-i] - med|), from `split` on they ascend (R_j = |a[split+j]
// - med|).  Binary search for i = how many of the k+1 smallest come from L:
// the least i with L_i >= R_{k-i}.
__device__ __forceinline__ double mad_kth(const SortedWin &w, double med, int split, int k,
                                          int steps, double *next) {
    const int nl = split, nr = w.n - split;
    auto L = [&](int i) { return fabs((double)w.at(split - 1 - i) - med); };
    auto R = [&](int j) { return fabs((double)w.at(split + j) - med); };
    int lo = max(0, k + 1 - nr), hi = min(k + 1, nl);
    for (int r = 0; r < steps; ++r) {
        if (lo < hi) {
            const int i = (lo + hi) >> 1;
            if (L(i) < R(k - i)) lo = i + 1; else hi = i;
        }
    }
    const int i = lo, j = k + 1 - lo;
    const double a = i > 0 ? L(i - 1) : -INFINITY;
    const double b = j > 0 ? R(j - 1) : -INFINITY;
    const double c = i < nl ? L(i) : INFINITY;
    const double d = j < nr ? R(j) : INFINITY;
    *next = fmin(c, d);
    return fmax(a, b);
}

// MAD of the current (NaN-free) sorted window of p values.
__device__ __forceinline__ float mad_value(const SortedWin &w, int p, int steps) {
    const int mid = p >> 1;
    double med;
    if (p & 1) med = (double)w.at(mid);
    else med = 0.5 * ((double)w.at(mid - 1) + (double)w.at(mid));
    // elements < med go left; ties with med (deviation 0) are on the right
    int sp = w.lower((float)med, steps);
    while (sp > 0 && (double)w.at(sp - 1) >= med) --sp;
    while (sp < w.n && (double)w.at(sp) < med) ++sp;
    double nxt;
    const double kth = mad_kth(w, med, sp, (p & 1) ? mid : mid - 1, steps, &nxt);
    return (float)((p & 1) ? kth : 0.5 * (kth + nxt));
}

template <class LD, class ST>
__device__ __forceinline__ void mad_run(SortedWin &w, int p, int fv, int c0, int c1, bool active,
                                        int steps, LD ld, ST st) {
    const int warm = fv + p - 1;
    int t = max(c0, warm);
    if (!active) return;
    for (int u = c0; u < min(t, c1); ++u) st(u, VTA_QNAN_F32);
    if (t >= c1) return;
    if (p == 1) {                                     // medium_ad.rs:326-333
        for (; t < c1; ++t) st(t, isnan((float)ld(t)) ? VTA_QNAN_F32 : 0.0f);
        return;
    }
    // any NaN in the window makes the output NaN (medium_ad.rs:343-348)
    int nan_cnt = fill_window(w, t, p, steps, ld);
    for (;;) {
        st(t, nan_cnt == 0 ? mad_value(w, p, steps) : VTA_QNAN_F32);
        if (t + 1 >= c1) break;
        const float vo = (float)ld(t + 1 - p);
        const float vi = (float)ld(t + 1);
        nan_cnt += (isnan(vi) ? 1 : 0) - (isnan(vo) ? 1 : 0);
        w.replace(rl_key(vo), rl_key(vi), steps);
        ++t;
    }
}

extern "C" __global__ __launch_bounds__(64) void mad_ms1p_f32(roll_args a) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    SortedWin w{lds + lane, 0};
    const int sc = min(s, N - 1);
    auto ld = [&](int t) -> double { return (double)a.x[(size_t)t * N + sc]; };
    auto st = [&](int t, float v) { a.out0[(size_t)t * N + s] = v; };
    const int p = a.p1;
    const int steps = bsearch_steps(a.wcap);
    const int fv = active ? a.fvs[s] : a.T;
    if (p == 1) {
        mad_run(w, p, fv, c0, c1, active, steps, ld, st);
        return;
    }
    // any NaN in the window makes the output NaN (medium_ad.rs:343-348)
    auto out = [&](const SortedWin &win, int nan_cnt) -> float {
        return nan_cnt == 0 ? mad_value(win, p, steps) : VTA_QNAN_F32;
    };
    window_run(w, p, max(c0, fv + p - 1), c0, c1, active, steps, ld, st, out);
}

// Lockstep one-series batch: lane = period; dynamic LDS as pnr_batch_f32.
extern "C" __global__ __launch_bounds__(64) void mad_batch_f32(roll_args a) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 64 + lane;
    const bool active = c < a.nlanes;
    const int T = a.T;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, T);
    const int cc = min(c, a.nlanes - 1);
    SortedWin w{lds + lane, 0};
    float *ob = lds + 64 * (a.wcap + 1);
    const int steps = bsearch_steps(a.wcap);
    const int p = a.p_len[cc];
    const int cs = max(c0, a.fvs[0] + p - 1);
    auto ld = [&](int t) -> double { return (double)a.x[t]; };
    int nan_cnt = 0;
    bool have = false;
    for (int t0 = c0; t0 < c1; t0 += RB_W) {
        const int nb = min(RB_W, c1 - t0);
        for (int k = 0; k < nb; ++k) {
            const int t = t0 + k;
            float o = VTA_QNAN_F32;
            if (active && t >= cs) {
                if (p == 1) {                                 // medium_ad.rs:326-333
                    o = isnan(a.x[t]) ? VTA_QNAN_F32 : 0.0f;
                } else {
                    if (!have) {
                        nan_cnt = fill_window(w, t, p, steps, ld);
                        have = true;
                    } else {
                        const float vo = a.x[t - p];
                        const float vi = a.x[t];
                        nan_cnt += (isnan(vi) ? 1 : 0) - (isnan(vo) ? 1 : 0);
                        w.replace(rl_key(vo), rl_key(vi), steps);
                    }
                    if (nan_cnt == 0) o = mad_value(w, p, steps);
                }
            }
            ob[lane * RB_LD + k] = o;
        }
        rb_sync();
        rb_flush(ob, a.out0, blockIdx.x * 64, a.nlanes, T, t0, nb);
        rb_sync();
    }
}

// ---------------------------------------------------------------- register windows
//
// Many-series PNR and MAD with each lane's sorted window held in VGPRs: CAP
// slots, CAP a compile-time multiple of 16 chosen as the window length rounded
// up (one kernel per CAP), slots past the window holding +Inf.  Time runs
// wave-uniform and every slot takes the same branch-free update each bar, so
// a step is a fixed run of VALU instructions: no LDS round trips, no per-lane
// shift counts, and no LDS to cap residency.  With the window keyed as in the
// LDS form (NaN bars as +Inf, counted):
//   removal of vo (present):  a[p] = w[p] < vo ? w[p] : w[p + 1]
//     drops the first slot not below vo, the oracle's lower_bound erase;
//   insertion of vi:          w'[p] = a[p] < vi ? a[p] : (a[p - 1] < vi ? vi : a[p - 1])
//     puts vi at its lower_bound slot and moves the slots above it up one.
// Zeros enter the window as +0.0 keys, so keys that compare equal are bitwise
// equal and the insertion is clamp(vi, a[p - 1], a[p]), one v_med3_f32.  The
// zero signs live beside the window: the oracle inserts a zero before every
// zero already in its window and erases the first one, so its zero block,
// front to back, is a stack of signs (last pushed first), kept here as a bit
// stack per lane.  The k-th slot is read by a select tree over the bits of k.

constexpr int rw_pow2(int n) { return n <= 1 ? 1 : 2 * rw_pow2((n + 1) >> 1); }
constexpr int rw_log2(int n) { return n <= 1 ? 0 : 1 + rw_log2(n >> 1); }

template <int N> struct RwVecF { typedef float v __attribute__((ext_vector_type(N))); };
template <int N> struct RwVecD { typedef double v __attribute__((ext_vector_type(N))); };

// The window: even slots in one vector, odd slots in another.  Every select
// between neighbouring slots (the removal, the tree's first level) then reads
// two different vectors; a select between two elements of one vector would be
// folded by the compiler into an indexed read, lowered as a compare-and-select
// over every slot.
template <int CAP>
struct RegWin {
    static_assert(CAP % 16 == 0, "register window capacity");
    typename RwVecF<CAP / 2>::v e, o;
    __device__ __forceinline__ float get(int p) const {
        return p >= CAP ? INFINITY : (p & 1) ? o[p >> 1] : e[p >> 1];
    }
    __device__ __forceinline__ void set(int p, float v) {
        if (p & 1) o[p >> 1] = v;
        else e[p >> 1] = v;
    }
};

// Insertion of vi (no removal): the fill of a chunk's first window.  Slots go
// from the top down, so each med3 reads the slot it overwrites and the one
// below it, still old (bottom up, every old slot would need a copy).
template <int CAP>
__device__ __forceinline__ void rw_insert(RegWin<CAP> &w, float vi) {
#pragma unroll
    for (int p = CAP - 1; p >= 0; --p)
        w.set(p, __builtin_amdgcn_fmed3f(p > 0 ? w.get(p - 1) : -INFINITY, w.get(p), vi));
}

// Removal of vo followed by insertion of vi, in one pass over the slots (slot
// p's removal reads the old slot p + 1, which the pass has not yet written).
template <int CAP>
__device__ __forceinline__ void rw_replace(RegWin<CAP> &w, float vo, float vi) {
    float pa = -INFINITY;
#pragma unroll
    for (int p = 0; p < CAP; ++p) {
        const float cur = w.get(p);
        const float ap = cur < vo ? cur : w.get(p + 1);
        w.set(p, __builtin_amdgcn_fmed3f(pa, ap, vi));
        pa = ap;
    }
}

// Signs of the window's zero keys, last inserted at bit 0 (NW 32-bit words,
// enough for a window of CAP keys).
template <int NW>
struct ZeroSigns {
    unsigned b[NW];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int j = 0; j < NW; ++j) b[j] = 0;
    }
    __device__ __forceinline__ void push(unsigned neg) {
#pragma unroll
        for (int j = NW - 1; j > 0; --j) b[j] = (b[j] << 1) | (b[j - 1] >> 31);
        b[0] = (b[0] << 1) | neg;
    }
    __device__ __forceinline__ void pop() {
#pragma unroll
        for (int j = 0; j + 1 < NW; ++j) b[j] = (b[j] >> 1) | (b[j + 1] << 31);
        b[NW - 1] >>= 1;
    }
    __device__ __forceinline__ bool neg_at(int d) const {   // d-th from the top
        unsigned wd = b[0];
#pragma unroll
        for (int j = 1; j < NW; ++j) wd = (d >> 5) == j ? b[j] : wd;
        return (wd >> (d & 31)) & 1u;
    }
};

// Window key of a bar: NaN as +Inf, either zero as +0.0.
__device__ __forceinline__ float rw_key(float v) {
    return isnan(v) ? INFINITY : v == 0.0f ? 0.0f : v;
}

// Slot k (per lane, k < CAP) by a select tree over the bits of k, bit 0 at
// the leaves (even/odd neighbours); slots from CAP on read as +Inf.
template <int CAP, int LO, int N>
struct RwTree {
    static __device__ __forceinline__ float pick(const RegWin<CAP> &w, int k) {
        if (LO >= CAP) return INFINITY;
        const float a = RwTree<CAP, LO, N / 2>::pick(w, k);
        const float b = RwTree<CAP, LO + N / 2, N / 2>::pick(w, k);
        return (k & (N / 2)) ? b : a;
    }
};
template <int CAP, int LO>
struct RwTree<CAP, LO, 1> {
    static __device__ __forceinline__ float pick(const RegWin<CAP> &w, int) { return w.get(LO); }
};

template <int CAP>
__device__ __forceinline__ float rw_pick(const RegWin<CAP> &w, int k) {
    return RwTree<CAP, 0, rw_pow2(CAP)>::pick(w, k);
}

// Slot k for a wave-uniform k: a binary search of scalar branches down to one
// register copy.  The copy is an asm statement, so the compiler can neither
// speculate the branches into selects nor fold them into an indexed read.
template <int CAP, int LO, int N>
struct RwSwitch {
    static __device__ __forceinline__ float pick(const RegWin<CAP> &w, int k) {
        if (LO >= CAP) return INFINITY;
        if (k < LO + N / 2) return RwSwitch<CAP, LO, N / 2>::pick(w, k);
        return RwSwitch<CAP, LO + N / 2, N / 2>::pick(w, k);
    }
};
template <int CAP, int LO>
struct RwSwitch<CAP, LO, 1> {
    static __device__ __forceinline__ float pick(const RegWin<CAP> &w, int) {
        float v = w.get(LO);
        asm volatile("" : "+v"(v));
        return v;
    }
};

// Slot k: the scalar search when every lane of the wave wants the same slot
// (every lane's window holds the same number of NaN bars), else the tree.
template <int CAP>
__device__ __forceinline__ float rw_pick_any(const RegWin<CAP> &w, int k) {
    const int k0 = __builtin_amdgcn_readfirstlane(k);
    if (__all(k == k0)) return RwSwitch<CAP, 0, rw_pow2(CAP)>::pick(w, k0);
    return rw_pick<CAP>(w, k);
}

// Bars [t0, t1) in blocks of RW_B whose loads (x[t] and, for lag > 0,
// x[t - lag]) are issued one block ahead; bars before 0 read as NaN.
// f(t, vi, vo) runs per bar, wave-uniform.  The per-bar loop stays rolled (a
// step is hundreds of instructions): the block's values rotate down one
// register per bar, so the step always reads slot 0.
#define RW_RB 4   // a step is ~2,000 cycles: loads four bars ahead are in time
template <class LD, class F>
__device__ __forceinline__ void rw_bars(int t0, int t1, int lag, LD ld, F f) {
    float ci[RW_RB], co[RW_RB], ni[RW_RB], no[RW_RB];
    auto load = [&](int b0, float *bi, float *bo) {
#pragma unroll
        for (int k = 0; k < RW_RB; ++k) {
            const int t = min(b0 + k, t1 - 1);
            bi[k] = t >= 0 ? ld(t) : VTA_QNAN_F32;
            bo[k] = (lag > 0 && t - lag >= 0) ? ld(t - lag) : VTA_QNAN_F32;
        }
    };
    if (t0 >= t1) return;
    load(t0, ci, co);
    for (int b0 = t0; b0 < t1; b0 += RW_RB) {
        if (b0 + RW_RB < t1) load(b0 + RW_RB, ni, no);
        const int nb = min(RW_RB, t1 - b0);
#pragma unroll 1
        for (int k = 0; k < nb; ++k) {
            f(b0 + k, ci[0], co[0]);
#pragma unroll
            for (int j = 0; j + 1 < RW_RB; ++j) {
                ci[j] = ci[j + 1];
                co[j] = co[j + 1];
            }
        }
#pragma unroll
        for (int k = 0; k < RW_RB; ++k) {
            ci[k] = ni[k];
            co[k] = no[k];
        }
    }
}

// One chunk [c0, c1) of a many-series PNR / MAD over register windows.  The
// chunk's first window (bars c0 - len + 1 .. c0, bars before 0 as NaN) is
// built by insertions, so every lane runs the same steps whatever its first
// valid bar; outputs before a lane's first output bar are NaN.  out(w, st)
// gives a window's output from the keys and the counts in st.  ZS: keep the
// zero signs (PNR returns a key itself; MAD's value does not depend on a
// zero's sign).
struct RwCounts {
    int nan_cnt;   // NaN bars in the window
    int neg_cnt;   // keys below zero
};

template <int CAP, bool ZS, class OUT>
__device__ __forceinline__ void rw_chunk(const roll_args &a, OUT out) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const int N = a.nlanes;
    const bool active = s < N;
    const int sc = min(s, N - 1);
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    const int len = a.p1;
    const int first = (active ? a.fvs[s] : a.T) + len - 1;
    const float *xs = a.x + sc;
    float *os = a.out0 + s;
    auto ld = [&](int t) -> float { return xs[(size_t)t * N]; };
    auto st = [&](int t, float v) {
        if (active) os[(size_t)t * N] = v;
    };
    if (rl_wave_min(first) >= c1) {                   // no output in this chunk
        for (int t = c0; t < c1; ++t) st(t, VTA_QNAN_F32);
        return;
    }
    RegWin<CAP> w;
#pragma unroll
    for (int p = 0; p < CAP; ++p) w.set(p, INFINITY);
    ZeroSigns<(CAP + 31) / 32> zs;
    zs.clear();
    RwCounts c{0, 0};
    auto emit = [&](int t) {
        float o = t >= first ? out(w, c) : VTA_QNAN_F32;
        if (ZS && __any(o == 0.0f)) {
            // the zero at slot k is the (k - neg_cnt)-th of the zero block
            if (o == 0.0f && zs.neg_at(out.slot(c) - c.neg_cnt)) o = -0.0f;
        }
        st(t, o);
    };
    rw_bars(c0 - len + 1, c0 + 1, 0, ld, [&](int, float vi, float) {
        c.nan_cnt += isnan(vi) ? 1 : 0;
        c.neg_cnt += vi < 0.0f ? 1 : 0;
        if (ZS && __any(vi == 0.0f)) {
            if (vi == 0.0f) zs.push(signbit(vi) ? 1u : 0u);
        }
        rw_insert<CAP>(w, rw_key(vi));
    });
    emit(c0);
    rw_bars(c0 + 1, c1, len, ld, [&](int t, float vi, float vo) {
        c.nan_cnt += (isnan(vi) ? 1 : 0) - (isnan(vo) ? 1 : 0);
        c.neg_cnt += (vi < 0.0f ? 1 : 0) - (vo < 0.0f ? 1 : 0);
        if (ZS && __any(vi == 0.0f || vo == 0.0f)) {   // erase, then insert
            if (vo == 0.0f) zs.pop();
            if (vi == 0.0f) zs.push(signbit(vi) ? 1u : 0u);
        }
        rw_replace<CAP>(w, rw_key(vo), rw_key(vi));
        emit(t);
    });
}

template <int CAP>
struct PnrRegOut {
    int len;
    double p_frac;
    __device__ __forceinline__ int slot(const RwCounts &c) const {
        return pnr_rank(p_frac, len - c.nan_cnt);
    }
    __device__ __forceinline__ float operator()(const RegWin<CAP> &w, const RwCounts &c) const {
        return len == c.nan_cnt ? VTA_QNAN_F32 : rw_pick_any<CAP>(w, slot(c));
    }
};

template <int CAP>
__device__ __forceinline__ void pnr_reg_body(const roll_args &a) {
    rw_chunk<CAP, true>(a, PnrRegOut<CAP>{a.p1, a.a1 * 0.01});
}

// MAD of a NaN-free register window of len keys (len wave-uniform).  med as
// the oracle's median (f64 from the middle key or keys); the deviations
// |w[i] - med| over the sorted slots fall then rise (a bitonic sequence, +Inf
// slots last, padded with +Inf to a power of two NP), so the k-th smallest is
// found by half-cleaners: at half size h, min(d[i], d[i + h]) holds the h
// smallest and max(...) the h largest, each again bitonic; the path keeps the
// half holding rank k.  Ranks mid - 1 and mid (even len) share the path until
// their halves part.  The ranks follow from len alone, so every path choice
// is a scalar branch; the first two levels run together (each path holds
// NP / 4 values).  A non-finite median gives NaN, as the oracle's sort of NaN
// deviations does.
template <bool MX>
__device__ __forceinline__ double rw_op(double x, double y) {
    return MX ? fmax(x, y) : fmin(x, y);
}

template <int CAP, bool A1, bool A2, bool B1, bool B2>
__device__ __forceinline__ void rw_mad_l12(const RegWin<CAP> &w, double med,
                                           typename RwVecD<rw_pow2(CAP) / 4>::v &A,
                                           typename RwVecD<rw_pow2(CAP) / 4>::v &B) {
    constexpr int NP = rw_pow2(CAP), h1 = NP / 2, h2 = NP / 4;
    constexpr bool SPLIT = A1 != B1 || A2 != B2;
    // an opaque copy of med keeps the deviations inside this case of the
    // caller's switch (hoisted above it, all NP of them would be live at once)
    double m = med;
    asm volatile("" : "+v"(m));
    auto dev = [&](int i) -> double { return i < CAP ? fabs((double)w.get(i) - m) : INFINITY; };
#pragma unroll
    for (int i = 0; i < h2; ++i) {
        const double d0 = dev(i), d1 = dev(i + h2), d2 = dev(i + h1), d3 = dev(i + h1 + h2);
        A[i] = rw_op<A2>(rw_op<A1>(d0, d2), rw_op<A1>(d1, d3));
        if (SPLIT) B[i] = rw_op<B2>(rw_op<B1>(d0, d2), rw_op<B1>(d1, d3));
    }
}

// one half-cleaner level on the first 2h values of v, keeping the max or min half
template <int H, int N>
__device__ __forceinline__ void rw_mad_level(typename RwVecD<N>::v &v, bool mx) {
    if (mx) {
#pragma unroll
        for (int i = 0; i < H; ++i) v[i] = fmax(v[i], v[i + H]);
    } else {
#pragma unroll
        for (int i = 0; i < H; ++i) v[i] = fmin(v[i], v[i + H]);
    }
}

template <int CAP>
__device__ __forceinline__ float rw_mad(const RegWin<CAP> &w, int len) {
    constexpr int NP = rw_pow2(CAP);
    static_assert(NP >= 16, "register MAD window");
    constexpr int Q = NP / 4;
    typedef typename RwVecD<Q>::v dvec;
    len = __builtin_amdgcn_readfirstlane(len);
    const int mid = len >> 1;
    const bool odd = len & 1;
    const float hi = RwSwitch<CAP, 0, NP>::pick(w, mid);
    const double med =
        odd ? (double)hi : 0.5 * ((double)RwSwitch<CAP, 0, NP>::pick(w, mid - 1) + (double)hi);
    if (!isfinite(med)) return VTA_QNAN_F32;
    const int k1 = odd ? mid : mid - 1, k2 = mid;
    dvec A, B;
    const int code = ((k1 & (NP / 2)) ? 8 : 0) | ((k1 & Q) ? 4 : 0) | ((k2 & (NP / 2)) ? 2 : 0) |
                     ((k2 & Q) ? 1 : 0);
    switch (code) {
#define RW_L12(C)                                                                          \
    case C:                                                                                \
        rw_mad_l12<CAP, (C & 8) != 0, (C & 4) != 0, (C & 2) != 0, (C & 1) != 0>(w, med, A, B); \
        break;
        RW_L12(0) RW_L12(1) RW_L12(2) RW_L12(3) RW_L12(4) RW_L12(5) RW_L12(6) RW_L12(7)
        RW_L12(8) RW_L12(9) RW_L12(10) RW_L12(11) RW_L12(12) RW_L12(13) RW_L12(14) RW_L12(15)
#undef RW_L12
    }
    bool split = (code >> 2) != (code & 3);
#pragma unroll
    for (int lv = 0; lv < rw_log2(NP) - 2; ++lv) {
        const int h = Q / 2 >> lv;
        const bool ba = k1 & h, bb = k2 & h;
        if (!split && ba != bb) {            // the paths part: B continues from A
#pragma unroll
            for (int i = 0; i < 2 * h; ++i) B[i] = A[i];
            split = true;
        }
        switch (h) {
#define RW_LV(HH)                                                                          \
    case HH:                                                                               \
        if (HH < Q) {                                                                      \
            rw_mad_level<(HH < Q ? HH : 1), Q>(A, ba);                                     \
            if (split) rw_mad_level<(HH < Q ? HH : 1), Q>(B, bb);                          \
        }                                                                                  \
        break;
            RW_LV(1) RW_LV(2) RW_LV(4) RW_LV(8) RW_LV(16)
#undef RW_LV
        }
    }
    const double d1 = A[0], d2 = split ? B[0] : A[0];
    return (float)(odd ? d1 : 0.5 * (d1 + d2));
}

// MAD for a window length fixed at compile time (LEN >= 4): the ranks, the
// median slots and every half-cleaner path choice are constants, so the
// network is straight-line code (no path switch) over NP = LEN rounded up to
// a power of two; entries past LEN are +Inf.  Same operations and order of
// selection as rw_mad.
template <int LEN>
struct MadLen {
    static constexpr int CAP = (LEN + 15) / 16 * 16;
    static constexpr int NP = rw_pow2(LEN);
    static constexpr int MID = LEN >> 1;
    static constexpr bool ODD = (LEN & 1) != 0;
    static constexpr int K1 = ODD ? MID : MID - 1;
    static constexpr int K2 = MID;
};

template <int LEN>
__device__ __forceinline__ float rw_mad_len(const RegWin<MadLen<LEN>::CAP> &w) {
    typedef MadLen<LEN> M;
    static_assert(M::NP >= 4, "MAD window length");
    constexpr int NP = M::NP, h1 = NP / 2, h2 = NP / 4;
    const float hi = w.get(M::MID);
    const double med = M::ODD ? (double)hi : 0.5 * ((double)w.get(M::MID - 1) + (double)hi);
    if (!isfinite(med)) return VTA_QNAN_F32;
    auto dev = [&](int i) -> double { return i < LEN ? fabs((double)w.get(i) - med) : INFINITY; };
    constexpr bool a1 = (M::K1 & h1) != 0, a2 = (M::K1 & h2) != 0;
    constexpr bool b1 = (M::K2 & h1) != 0, b2 = (M::K2 & h2) != 0;
    constexpr bool split12 = a1 != b1 || a2 != b2;
    double A[h2], B[h2];
#pragma unroll
    for (int i = 0; i < h2; ++i) {
        const double d0 = dev(i), d1 = dev(i + h2), d2 = dev(i + h1), d3 = dev(i + h1 + h2);
        const double lo0 = fmin(d0, d2), hi0 = fmax(d0, d2), lo1 = fmin(d1, d3), hi1 = fmax(d1, d3);
        const double x0 = a1 ? hi0 : lo0, x1 = a1 ? hi1 : lo1;
        A[i] = a2 ? fmax(x0, x1) : fmin(x0, x1);
        if (split12) {
            const double y0 = b1 ? hi0 : lo0, y1 = b1 ? hi1 : lo1;
            B[i] = b2 ? fmax(y0, y1) : fmin(y0, y1);
        }
    }
    bool split = split12;
#pragma unroll
    for (int lv = 0; lv < rw_log2(NP) - 2; ++lv) {
        const int h = h2 / 2 >> lv;
        const bool ba = (M::K1 & h) != 0, bb = (M::K2 & h) != 0;
        if (!split && ba != bb) {
#pragma unroll
            for (int i = 0; i < h2; ++i) B[i] = A[i];
            split = true;
        }
#pragma unroll
        for (int i = 0; i < h; ++i) {
            A[i] = ba ? fmax(A[i], A[i + h]) : fmin(A[i], A[i + h]);
            if (split) B[i] = bb ? fmax(B[i], B[i + h]) : fmin(B[i], B[i + h]);
        }
    }
    const double d1 = A[0], d2 = split ? B[0] : A[0];
    return (float)(M::ODD ? d1 : 0.5 * (d1 + d2));
}

template <int LEN>
struct MadLenOut {
    __device__ __forceinline__ int slot(const RwCounts &) const { return 0; }
    __device__ __forceinline__ float operator()(const RegWin<MadLen<LEN>::CAP> &w,
                                                const RwCounts &c) const {
        return c.nan_cnt != 0 ? VTA_QNAN_F32 : rw_mad_len<LEN>(w);
    }
};

// period 1: 0 for every non-NaN bar, +-Inf included (medium_ad.rs:326-333)
template <int CAP>
struct MadRegOut {
    int len;
    __device__ __forceinline__ int slot(const RwCounts &) const { return 0; }
    __device__ __forceinline__ float operator()(const RegWin<CAP> &w, const RwCounts &c) const {
        return c.nan_cnt != 0 ? VTA_QNAN_F32 : len == 1 ? 0.0f : rw_mad<CAP>(w, len);
    }
};

template <int CAP>
__device__ __forceinline__ void mad_reg_body(const roll_args &a) {
    rw_chunk<CAP, false>(a, MadRegOut<CAP>{a.p1});
}

// Waves per SIMD hinted per kernel (2 at least, 3 for PNR windows up to 112
// slots), the most that fit the VGPRs without spilling.
#define RW_PNR(CAP)                                                                    \
    extern "C" __global__ __launch_bounds__(64, CAP <= 112 ? 3 : 2) void pnr_ms1p_reg##CAP( \
        roll_args a) {                                                                     \
        pnr_reg_body<CAP>(a);                                                              \
    }
#define RW_MAD(CAP)                                                                        \
    extern "C" __global__ __launch_bounds__(64, 2) void mad_ms1p_reg##CAP(roll_args a) {   \
        mad_reg_body<CAP>(a);                                                              \
    }
RW_PNR(16) RW_PNR(32) RW_PNR(48) RW_PNR(64) RW_PNR(80) RW_PNR(96) RW_PNR(112) RW_PNR(128)
// periods 1-3 (the length-templated kernels below start at 4)
RW_MAD(16)
#define RW_MAD_LEN(L)                                                                      \
    extern "C" __global__ __launch_bounds__(64, 2) void mad_ms1p_len##L(roll_args a) {     \
        rw_chunk<MadLen<L>::CAP, false>(a, MadLenOut<L>{});                                \
    }
RW_MAD_LEN(4) RW_MAD_LEN(5) RW_MAD_LEN(6) RW_MAD_LEN(7) RW_MAD_LEN(8) RW_MAD_LEN(9) RW_MAD_LEN(10) RW_MAD_LEN(11)
RW_MAD_LEN(12) RW_MAD_LEN(13) RW_MAD_LEN(14) RW_MAD_LEN(15) RW_MAD_LEN(16) RW_MAD_LEN(17) RW_MAD_LEN(18) RW_MAD_LEN(19)
RW_MAD_LEN(20) RW_MAD_LEN(21) RW_MAD_LEN(22) RW_MAD_LEN(23) RW_MAD_LEN(24) RW_MAD_LEN(25) RW_MAD_LEN(26) RW_MAD_LEN(27)
RW_MAD_LEN(28) RW_MAD_LEN(29) RW_MAD_LEN(30) RW_MAD_LEN(31) RW_MAD_LEN(32) RW_MAD_LEN(33) RW_MAD_LEN(34) RW_MAD_LEN(35)
RW_MAD_LEN(36) RW_MAD_LEN(37) RW_MAD_LEN(38) RW_MAD_LEN(39) RW_MAD_LEN(40) RW_MAD_LEN(41) RW_MAD_LEN(42) RW_MAD_LEN(43)
RW_MAD_LEN(44) RW_MAD_LEN(45) RW_MAD_LEN(46) RW_MAD_LEN(47) RW_MAD_LEN(48) RW_MAD_LEN(49) RW_MAD_LEN(50) RW_MAD_LEN(51)
RW_MAD_LEN(52) RW_MAD_LEN(53) RW_MAD_LEN(54) RW_MAD_LEN(55) RW_MAD_LEN(56) RW_MAD_LEN(57) RW_MAD_LEN(58) RW_MAD_LEN(59)
RW_MAD_LEN(60) RW_MAD_LEN(61) RW_MAD_LEN(62) RW_MAD_LEN(63) RW_MAD_LEN(64)
