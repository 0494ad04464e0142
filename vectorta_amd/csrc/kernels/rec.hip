// Wilder / exponential recurrences for gfx950: EMA, RSI, ATR, ADX and the fused
// EMA+RSI+ATR many-series launch.
//
// Each lane runs one stream (a combo of a one-series batch, or one series of
// a time-major matrix) through an f64 engine that restates the reference's
// scalar arithmetic in the same order:
//   EMA  src/indicators/moving_averages/ema.rs:430-468
//   RSI  src/indicators/rsi.rs:327-414
//   ATR  src/indicators/atr.rs:329-460
//   ADX  src/indicators/adx.rs:299-422
//   SMMA src/indicators/moving_averages/smma.rs:243-271
//   Wilders src/indicators/moving_averages/wilders.rs:262-320
// so a lane that starts exactly at first_valid reproduces the oracle's f64
// values and only rounds them to f32 on output.
//
// Time parallelism: the series is cut into chunks of L bars, one wave per
// (64 streams, chunk).  A wave whose chunk starts far enough after the
// warm-up begins K bars early with an approximate state and discards those
// outputs ("burn-in").  All four recurrences contract the state error by
// beta = 1 - 1/p (EMA: 1 - 2/(p+1)) per bar, so after K >= 20 / -ln(beta)
// bars (ADX, two coupled stages: 32 / -ln(beta)) the error is below
// e^-20 ~ 2.1e-9 of the state: ~30x under f32 rounding of the output and four
// orders under the 1e-5 parity bound.  This holds only when
// every bar after first_valid is finite (non-finite bars stop EMA decay and
// poison ATR forever), so each stream carries a "clean" flag; an unclean
// stream always starts exactly at first_valid.
//
// One-series batches (EMA, RSI, ATR, ADX, SMMA, Wilders) use an exact chunked
// scan instead when the series is clean: every step after first_valid, the
// warm-up included, is affine in the state, s_t = a_t s_{t-1} + f_t (a seed
// sum has a_t = 1, the step that takes its mean 1/p, the smoothing beta), so
// a chunk maps its start state s to g s + B, where g is the product of its
// a_t and B is the chunk run from a zero state.  Pass 1 (phase 1) computes B
// for every (chunk, combo) -- the chunk holding first_valid runs exactly from
// there --, a per-combo pass (*_scan_combine) folds the chunks in order into
// the exact state at every chunk end, and pass 3 (phase 3) runs every chunk
// again from its exact start state and writes the outputs.  No lane replays
// bars outside its chunk, so the chunk length is chosen for waves in flight
// alone (Wilders, whose seed is summed in groups of four, starts carried
// chunks only past its warm-up).
#include "vta_common.h"

struct rec_args {
    const float *x;          // close / prices
    const float *h;          // high (ATR, ADX)
    const float *l;          // low  (ATR, ADX)
    const int32_t *periods;  // batch: per combo; many-series: [0] only
    const int32_t *fvs;      // many-series: per series; batch: [0] only
    const int32_t *clean;    // many-series: per series; batch: [0] only
    const int32_t *fvs2;     // fused kernel: per-series first_valid of (high, low, close)
    float *out0;
    float *out1;
    float *out2;
    double *carry;           // batch scan: [combo][chunk][states] chunk-end states
    const float *alphas;     // EMA / Wilders: caller's smoothing factor (batch: per combo,
                             // many-series: [0]); NULL derives it from the period
    double *seeds;           // Wilders batch: per-combo value at bar first_valid + period - 1
                             // (wilders_seed), or NULL (exact lanes replay the warm-up)
    int T;                   // bars
    int nlanes;              // combos (batch) or series (many-series)
    int chunk;               // bars per chunk
    int p1, p2, p3;          // fused kernel periods (EMA, RSI, ATR)
    int phase;               // batch: 0 burn-in/exact, 1 scan pass, 3 output pass of the scan
    int clean_epoch;         // batch: 0, or clean = (clean[0] != clean_epoch), the word
                             // batch_prepare stores the call's epoch in on a non-finite bar
};

#define REC_NAN __builtin_bit_cast(double, 0x7ff8000000000000ull)

__device__ __forceinline__ bool fin(double v) { return vta_isfinite_d(v); }

// Burn-in length in nats of state decay: the carried error of a burnt-in
// start is below e^-20 ~ 2.1e-9 of the state, ~30x under the f32 rounding of
// an output and four orders under the 1e-5 parity bound (ADX, whose two
// coupled stages compound the error, uses 32).
#define REC_NATS 20.0

__device__ __forceinline__ int burn_bars(double beta, double nats) {
    if (!(beta > 0.0)) return 1;
    double k = nats / -log(beta);
    if (k > 1.0e8) return 100000000;
    return max(1, (int)ceil(k));
}

// Bars t of [c0, c1) with t >= from.
__device__ __forceinline__ int bars_from(int c0, int c1, int from) {
    return max(0, c1 - max(c0, from));
}

// ---------------------------------------------------------------- engines
//
// Chunked-scan interface (one-series batches): gain(c0, c1) is the product of
// the per-bar multipliers of the state over bars [c0, c1) (every step, warm-up
// included, is affine in the state: s_t = a_t s_{t-1} + f_t), begin_carry(t0,
// ...) restarts the engine after bar t0 from a carried state and
// get_state(t1, ...) reads it after bar t1; the state is whichever quantity
// the step carries at that bar (a seed sum during the warm-up, the average
// after).  kWarmScan engines accept a carried start anywhere after
// first_valid; the others only past the warm-up.

struct EmaEng {
    static constexpr bool kSeedScan = false;
    double alpha, beta, prev;
    int p, fv, vc;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        alpha = 2.0 / ((double)p + 1.0);
        beta = 1.0 - alpha;
    }
    __device__ void set_alpha(double a_) { alpha = a_; beta = 1.0 - a_; }
    __device__ int kb() const { return burn_bars(beta, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p; }   // b >= this may burn in
    // exact start at bar fv
    __device__ double begin_exact(double x) { prev = x; vc = 1; return prev; }
    __device__ void begin_burn(double x) { prev = fin(x) ? x : 0.0; vc = p; }
    __device__ double step(int t, double x) {
        if (t < fv + p) {                       // running-mean warm-up (ema.rs:441-452)
            if (fin(x)) {
                vc += 1;
                double v = (double)vc;
                prev = ((v - 1.0) * prev + x) / v;
            }
        } else if (fin(x)) {
            prev = fma(beta, prev, alpha * x);  // ema.rs:458-466
        }
        return prev;
    }
    __device__ double out_before_start(int t) const { return REC_NAN; }
    static constexpr int kScan = 1;                  // states carried by the chunked scan
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = false;
    static constexpr int kPre = 0;
    static constexpr int kStageW = 64;
    static constexpr bool kSeeded = false;
    // steps t < seed_end() after the start are the warm-up accumulation only
    __device__ int seed_end() const { return fv + p; }
    __device__ double seed_step(int t, double x) {
        if (fin(x)) {
            vc += 1;
            double v = (double)vc;
            prev = ((v - 1.0) * prev + x) / v;
        }
        return prev;
    }
    // the step for t >= burn_ok_from() (no warm-up branches)
    __device__ double steady(double x) {
        if (fin(x)) prev = fma(beta, prev, alpha * x);
        return prev;
    }
    // running mean: (v-1)/v per bar with v = t - fv + 1 (the product
    // telescopes), then beta
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        const int m = min(c1, fv + p);
        const double g = c0 < m ? (double)(c0 - fv) / (double)(m - fv) : 1.0;
        const int n = bars_from(c0, c1, fv + p);
        return n > 0 ? g * pow(beta, (double)n) : g;
    }
    __device__ void begin_carry(int t0, double x, const double *st) {
        prev = st[0];
        vc = t0 - fv + 1;
    }
    __device__ void get_state(int t1, double *st) const { st[0] = prev; }
};

// num / den for den > 0: an f32 reciprocal estimate refined by one f64 Newton
// step (relative error ~1e-14 instead of the IEEE division sequence; the
// outputs are f32).
__device__ __forceinline__ double rec_ratio(double num, double den) {
    if (!(den > 1e-30 && den < 1e30)) return num / den;   // outside the f32 estimate's range
    double r = (double)__builtin_amdgcn_rcpf((float)den);
    r = fma(fma(-den, r, 1.0), r, r);
    return num * r;
}

// RSI output 100 ag / (ag + al) (50 when the denominator is 0), formed in
// f32 from the f64 averages: both rounded to f32, their f32 sum, a
// reciprocal estimate refined by one Newton step and two f32 products --
// within ~6 ulp of f32 (~4e-7 relative; the output is f32, the parity bound
// 1e-5).  A sum outside [1e-30, 1e30] takes the f64 quotient (and 50 for a
// zero denominator).  The ratio is an output only; the state stays f64.
__device__ __forceinline__ float rsi_ratio_f32(double ag, double al) {
    const float a = (float)ag, b = (float)al;
    const float den = a + b;
    if (!(den > 1e-30f && den < 1e30f)) {
        const double dd = ag + al;
        return dd == 0.0 ? 50.0f : (float)(100.0 * ag / dd);
    }
    float r = __builtin_amdgcn_rcpf(den);
    r = fmaf(fmaf(-den, r, 1.0f), r, r);
    return (100.0f * a) * r;
}

// DX = 100 |+DM - -DM| / (+DM + -DM) for den > 0, divided in f32 (reciprocal
// estimate and one Newton step, <= ~2 ulp of f32); outside the estimate's
// range the f64 quotient.  Each DX enters the ADX as a weighted mean, so its
// ~1e-7 relative error is not amplified.
__device__ __forceinline__ double dx_ratio(double num, double den) {
    if (!(den > 1e-30 && den < 1e30)) return num / den;
    const float d = (float)den;
    float r = __builtin_amdgcn_rcpf(d);
    r = fmaf(fmaf(-d, r, 1.0f), r, r);
    return (double)((float)num * r);
}

struct RsiEng {
    static constexpr bool kSeedScan = true;
    double inv_p, beta, ag, al, px;
    int p, fv;
    bool poisoned;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        inv_p = 1.0 / (double)p;
        beta = 1.0 - inv_p;
        poisoned = false;
    }
    __device__ int kb() const { return burn_bars(beta, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p + 1; }
    __device__ double begin_exact(double x) { px = x; ag = 0.0; al = 0.0; return REC_NAN; }
    __device__ void begin_burn(double x) { px = x; ag = 0.0; al = 0.0; }
    // the seed window (fv, fv+p] has a non-finite delta -> NaN forever (rsi.rs:340-361)
    __device__ void seed_scan(int t, double x) {
        if (t > fv && t <= fv + p && !fin(x - px)) poisoned = true;
        px = x;
    }
    __device__ double step(int t, double x) {
        double d = x - px;
        px = x;
        if (t <= fv + p) {
            if (!poisoned) {
                if (!fin(d)) poisoned = true;
                else if (d > 0.0) ag += d;
                else if (d < 0.0) al -= d;
            }
            if (t < fv + p) return REC_NAN;
            if (poisoned) { ag = REC_NAN; al = REC_NAN; return REC_NAN; }
            ag *= inv_p;
            al *= inv_p;
        } else {
            double g = d > 0.0 ? d : 0.0;
            double lo = d < 0.0 ? -d : 0.0;
            ag = fma(ag, beta, inv_p * g);
            al = fma(al, beta, inv_p * lo);
        }
        if (poisoned) return REC_NAN;
        return (double)rsi_ratio_f32(ag, al);
    }
    static constexpr int kScan = 2;
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = true;
    // the gain and loss of a bar are two ops: reading them from LDS
    // (see drive_batch) measured slower (0.55 vs 0.47 ms, 1,250 x 200k)
    static constexpr int kPre = 0;
    static constexpr int kStageW = 32;
    static constexpr bool kSeeded = false;
    __device__ int seed_end() const { return fv + p; }
    __device__ double seed_step(int t, double x) {
        double d = x - px;
        px = x;
        if (!poisoned) {
            if (!fin(d)) poisoned = true;
            else if (d > 0.0) ag += d;
            else if (d < 0.0) al -= d;
        }
        return REC_NAN;
    }
    __device__ int advance_from() const { return burn_ok_from(); }
    // pass 1 of the chunked scan: the state update without the output ratio
    __device__ void advance(double x) {
        double d = x - px;
        px = x;
        double g = d > 0.0 ? d : 0.0;
        double lo = d < 0.0 ? -d : 0.0;
        ag = fma(ag, beta, inv_p * g);
        al = fma(al, beta, inv_p * lo);
    }
    __device__ double steady(double x) {
        double d = x - px;
        px = x;
        double g = d > 0.0 ? d : 0.0;
        double lo = d < 0.0 ? -d : 0.0;
        ag = fma(ag, beta, inv_p * g);
        al = fma(al, beta, inv_p * lo);
        if (poisoned) return REC_NAN;
        return (double)rsi_ratio_f32(ag, al);
    }
    // seed sums (multiplier 1), times 1/p at fv+p, then beta
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        const int ws = fv + p;
        const double g = (c0 <= ws && ws < c1) ? inv_p : 1.0;
        const int n = bars_from(c0, c1, ws + 1);
        return n > 0 ? g * pow(beta, (double)n) : g;
    }
    __device__ void begin_carry(int t0, double x, const double *st) {
        px = x;
        ag = st[0];
        al = st[1];
    }
    __device__ void get_state(int t1, double *st) const { st[0] = ag; st[1] = al; }
};

__device__ __forceinline__ double true_range(double hi, double lo, double pc) {
    double tr = hi - lo;
    double hc = fabs(hi - pc);
    if (hc > tr) tr = hc;
    double lc = fabs(lo - pc);
    if (lc > tr) tr = lc;
    return tr;
}

struct AtrEng {
    static constexpr bool kSeedScan = false;
    double alpha, sum, rma, pc;
    int p, fv;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        alpha = 1.0 / (double)p;
    }
    __device__ int kb() const { return burn_bars(1.0 - alpha, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p; }
    __device__ double begin_exact(double h, double l, double c) {
        sum = h - l;
        pc = c;
        if (p == 1) { rma = sum / (double)p; return rma; }
        return REC_NAN;
    }
    __device__ void begin_burn(double h, double l, double c) { rma = h - l; pc = c; }
    __device__ double step(int t, double h, double l, double c) {
        double tr = true_range(h, l, pc);
        pc = c;
        if (t < fv + p - 1) { sum += tr; return REC_NAN; }
        if (t == fv + p - 1) { sum += tr; rma = sum / (double)p; return rma; }
        rma = fma(-alpha, rma, rma) + alpha * tr;  // atr.rs:393
        return rma;
    }
    static constexpr int kScan = 1;
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = false;
    static constexpr int kPre = 1;
    static constexpr int kStageW = 64;
    static constexpr bool kSeeded = false;
    __device__ int seed_end() const { return fv + p - 1; }
    __device__ double seed_step(int t, double h, double l, double c) {
        sum += true_range(h, l, pc);
        pc = c;
        return REC_NAN;
    }
    __device__ double steady(double h, double l, double c) {
        double tr = true_range(h, l, pc);
        pc = c;
        rma = fma(-alpha, rma, rma) + alpha * tr;
        return rma;
    }
    // per-bar term shared by every combo: the true range of bar t
    static __device__ void pre(double h, double l, double, double, double, double pc, double *o) {
        o[0] = true_range(h, l, pc);
    }
    __device__ void set_prev(double, double, double c) { pc = c; }
    __device__ void advance_pre(const double *) {}
    __device__ double steady_pre(const double *o) {
        rma = fma(-alpha, rma, rma) + alpha * o[0];
        return rma;
    }
    // seed sum up to fv+p-1, where the mean is taken (1/p), then 1 - alpha
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        const int ws = fv + p - 1;
        const double g = (c0 <= ws && ws < c1) ? 1.0 / (double)p : 1.0;
        const int n = bars_from(c0, c1, ws + 1);
        return n > 0 ? g * pow(1.0 - alpha, (double)n) : g;
    }
    __device__ void begin_carry(int t0, double h, double l, double c, const double *st) {
        if (t0 < fv + p - 1) sum = st[0];
        else rma = st[0];
        pc = c;
    }
    __device__ void get_state(int t1, double *st) const { st[0] = t1 < fv + p - 1 ? sum : rma; }
};

// ATR fed by a precomputed true-range series (atr_wrapper.rs:626-700,
// atr_batch_device_with_tr): seed = mean of tr[fv .. fv+p) at fv+p-1, then
// the Wilder update of atr.rs:393 on tr.
struct AtrTrEng {
    static constexpr bool kSeedScan = false;
    double alpha, sum, rma;
    int p, fv;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        alpha = 1.0 / (double)p;
    }
    __device__ int kb() const { return burn_bars(1.0 - alpha, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p; }
    __device__ double begin_exact(double tr) {
        sum = tr;
        if (p == 1) { rma = sum; return rma; }
        return REC_NAN;
    }
    __device__ void begin_burn(double tr) { rma = tr; }
    __device__ double step(int t, double tr) {
        if (t < fv + p - 1) { sum += tr; return REC_NAN; }
        if (t == fv + p - 1) { sum += tr; rma = sum / (double)p; return rma; }
        rma = fma(-alpha, rma, rma) + alpha * tr;
        return rma;
    }
    __device__ double out_before_start(int t) const { return REC_NAN; }
    static constexpr int kScan = 1;
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = false;
    static constexpr int kPre = 0;
    static constexpr int kStageW = 64;
    static constexpr bool kSeeded = false;
    __device__ int seed_end() const { return fv + p - 1; }
    __device__ double seed_step(int t, double tr) {
        sum += tr;
        return REC_NAN;
    }
    __device__ double steady(double tr) {
        rma = fma(-alpha, rma, rma) + alpha * tr;
        return rma;
    }
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        const int ws = fv + p - 1;
        const double g = (c0 <= ws && ws < c1) ? 1.0 / (double)p : 1.0;
        const int n = bars_from(c0, c1, ws + 1);
        return n > 0 ? g * pow(1.0 - alpha, (double)n) : g;
    }
    __device__ void begin_carry(int t0, double tr, const double *st) {
        if (t0 < fv + p - 1) sum = st[0];
        else rma = st[0];
    }
    __device__ void get_state(int t1, double *st) const { st[0] = t1 < fv + p - 1 ? sum : rma; }
};

struct AdxEng {
    static constexpr bool kSeedScan = false;
    double rp, omr, pm1, atr, pds, mds, dx_sum, adx, ph, pl, pc;
    int p, fv, dx_count;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        double pf = (double)p;
        rp = 1.0 / pf;
        omr = 1.0 - rp;
        pm1 = pf - 1.0;
    }
    __device__ int kb() const { return burn_bars(omr, 32.0); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + 2 * p; }
    __device__ double begin_exact(double h, double l, double c) {
        atr = 0.0; pds = 0.0; mds = 0.0; dx_sum = 0.0; adx = 0.0; dx_count = 0;
        ph = h; pl = l; pc = c;
        return REC_NAN;
    }
    __device__ void begin_burn(double h, double l, double c) {
        atr = 0.0; pds = 0.0; mds = 0.0; adx = 0.0; dx_sum = 0.0; dx_count = p;
        ph = h; pl = l; pc = c;
    }
    __device__ double step(int t, double ch, double cl, double cc) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        ph = ch; pl = cl; pc = cc;
        int i = t - fv;   // index within the slice starting at first_valid
        if (i <= p) {     // adx.rs:327-348 seed sums
            if (up > down && up > 0.0) pds += up;
            if (down > up && down > 0.0) mds += down;
            atr += tr;
            if (i < p) return REC_NAN;
            double pdi = 0.0, mdi = 0.0;
            if (atr != 0.0) { pdi = (pds / atr) * 100.0; mdi = (mds / atr) * 100.0; }
            double sdi = pdi + mdi;
            dx_sum = sdi != 0.0 ? (fabs(pdi - mdi) / sdi) * 100.0 : 0.0;
            dx_count = 1;
            return REC_NAN;
        }
        double pdm = (up > down && up > 0.0) ? up : 0.0;
        double mdm = (down > up && down > 0.0) ? down : 0.0;
        atr = atr * omr + tr;
        pds = pds * omr + pdm;
        mds = mds * omr + mdm;
        // DX = 100 |DI+ - DI-| / (DI+ + DI-) with DI = 100 DM / ATR: the ATR
        // cancels, leaving one ratio (zero when ATR or DM+ + DM- is zero, as
        // adx.rs:370-390 gives)
        const double sdm = pds + mds;
        double dx = (atr != 0.0 && sdm != 0.0) ? dx_ratio(100.0 * fabs(pds - mds), sdm) : 0.0;
        if (dx_count < p) {
            dx_sum += dx;
            dx_count += 1;
            if (dx_count == p) { adx = dx_sum * rp; return adx; }
            return REC_NAN;
        }
        adx = (adx * pm1 + dx) * rp;
        return adx;
    }
    // two stages: (atr, +DM, -DM) decay by omr; adx decays by pm1 * rp once
    // the DX sequence is known, which needs the first stage's exact states
    static constexpr int kScan = 4;
    static constexpr bool kTwoStage = true;
    static constexpr bool kAdvance = true;
    static constexpr int kPre = 3;
    static constexpr int kStageW = 32;
    static constexpr bool kSeeded = false;
    __device__ int seed_end() const { return fv + p; }
    __device__ double seed_step(int t, double ch, double cl, double cc) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        ph = ch; pl = cl; pc = cc;
        if (up > down && up > 0.0) pds += up;
        if (down > up && down > 0.0) mds += down;
        atr += tr;
        return REC_NAN;
    }
    // pass 1 needs the first stage only, which is a plain smoothing from
    // i = p + 1 on
    __device__ int advance_from() const { return fv + p + 1; }
    // DX seed (adx.rs:369-409): bars fv+p < t < fv+2p-1 smooth the first
    // stage and add DX to the seed sum; the ADX itself is not yet defined
    __device__ int dx_seed_end() const { return fv + 2 * p - 1; }
    __device__ void dx_seed_step(double ch, double cl, double cc) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        ph = ch; pl = cl; pc = cc;
        double pdm = (up > down && up > 0.0) ? up : 0.0;
        double mdm = (down > up && down > 0.0) ? down : 0.0;
        atr = atr * omr + tr;
        pds = pds * omr + pdm;
        mds = mds * omr + mdm;
        const double sdm = pds + mds;
        double dx = (atr != 0.0 && sdm != 0.0) ? dx_ratio(100.0 * fabs(pds - mds), sdm) : 0.0;
        dx_sum += dx;
    }
    // first stage: seed sums (multiplier 1) up to i = p, then omr; second
    // stage: restarted at i = p (first DX), DX summed (1) up to i = 2p-2,
    // the mean at i = 2p-1 (rp), then pm1 * rp
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        const int n = bars_from(c0, c1, fv + p + 1);
        return n > 0 ? pow(omr, (double)n) : 1.0;
    }
    __device__ double gain2(int c0, int c1) const {
        if (c0 <= fv + p) return 0.0;
        const int w = fv + 2 * p - 1;
        const double g = (c0 <= w && w < c1) ? rp : 1.0;
        const int n = bars_from(c0, c1, w + 1);
        return n > 0 ? g * pow(pm1 * rp, (double)n) : g;
    }
    // second-stage state after bar t: undefined before i = p, the DX sum up
    // to i = 2p-2, the ADX from i = 2p-1 on
    __device__ void begin_carry(int t0, double h, double l, double c, const double *st) {
        atr = st[0]; pds = st[1]; mds = st[2];
        const int i0 = t0 - fv;
        if (i0 < p) {
            dx_sum = 0.0; adx = 0.0; dx_count = 0;
        } else if (i0 < 2 * p - 1) {
            dx_sum = st[3]; adx = 0.0; dx_count = i0 - p + 1;
        } else {
            dx_sum = 0.0; adx = st[3]; dx_count = p;
        }
        ph = h; pl = l; pc = c;
    }
    __device__ void get_state(int t1, double *st) const {
        st[0] = atr; st[1] = pds; st[2] = mds;
        const int i1 = t1 - fv;
        st[3] = i1 < p ? 0.0 : (i1 < 2 * p - 1 ? dx_sum : adx);
    }
    // pass 1 of the chunked scan: the first stage only (the ADX stage is
    // folded after pass 2)
    __device__ void advance(double ch, double cl, double cc) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        ph = ch; pl = cl; pc = cc;
        double pdm = (up > down && up > 0.0) ? up : 0.0;
        double mdm = (down > up && down > 0.0) ? down : 0.0;
        atr = atr * omr + tr;
        pds = pds * omr + pdm;
        mds = mds * omr + mdm;
    }
    __device__ double steady(double ch, double cl, double cc) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        ph = ch; pl = cl; pc = cc;
        double pdm = (up > down && up > 0.0) ? up : 0.0;
        double mdm = (down > up && down > 0.0) ? down : 0.0;
        atr = atr * omr + tr;
        pds = pds * omr + pdm;
        mds = mds * omr + mdm;
        const double sdm = pds + mds;
        double dx = (atr != 0.0 && sdm != 0.0) ? dx_ratio(100.0 * fabs(pds - mds), sdm) : 0.0;
        adx = (adx * pm1 + dx) * rp;
        return adx;
    }
    // per-bar terms shared by every combo: the true range, +DM and -DM of bar t
    static __device__ void pre(double ch, double cl, double, double ph_, double pl_, double pc_,
                               double *o) {
        double hl = ch - cl;
        double hpc = fabs(ch - pc_);
        double lpc = fabs(cl - pc_);
        o[0] = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph_;
        double down = pl_ - cl;
        o[1] = (up > down && up > 0.0) ? up : 0.0;
        o[2] = (down > up && down > 0.0) ? down : 0.0;
    }
    __device__ void set_prev(double h, double l, double c) { ph = h; pl = l; pc = c; }
    __device__ void advance_pre(const double *o) {
        atr = atr * omr + o[0];
        pds = pds * omr + o[1];
        mds = mds * omr + o[2];
    }
    __device__ double steady_pre(const double *o) {
        atr = atr * omr + o[0];
        pds = pds * omr + o[1];
        mds = mds * omr + o[2];
        const double sdm = pds + mds;
        double dx = (atr != 0.0 && sdm != 0.0) ? dx_ratio(100.0 * fabs(pds - mds), sdm) : 0.0;
        adx = (adx * pm1 + dx) * rp;
        return adx;
    }
};

// SMMA (smma.rs:243-271): the mean of the first p bars at fv+p-1, then
// prev = (prev*(p-1) + x) / p (here times 1/p).  p = 1 copies the input.
struct SmmaEng {
    static constexpr bool kSeedScan = false;
    double pm1, rp, prev, sum;
    int p, fv;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        pm1 = (double)p - 1.0;
        rp = 1.0 / (double)p;
    }
    __device__ int kb() const { return burn_bars(1.0 - rp, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p; }
    __device__ double begin_exact(double x) {
        sum = x;
        prev = x;
        if (p == 1) return x;
        return REC_NAN;
    }
    __device__ void begin_burn(double x) { prev = x; }
    __device__ double step(int t, double x) {
        if (p == 1) { prev = x; return x; }
        const int ws = fv + p - 1;
        if (t < ws) { sum += x; return REC_NAN; }
        if (t == ws) { sum += x; prev = sum / (double)p; return prev; }
        prev = (prev * pm1 + x) * rp;
        return prev;
    }
    __device__ double out_before_start(int t) const { return REC_NAN; }
    static constexpr int kScan = 1;
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = false;
    static constexpr int kPre = 0;
    static constexpr int kStageW = 64;
    static constexpr bool kSeeded = false;
    __device__ int seed_end() const { return p == 1 ? fv : fv + p - 1; }
    __device__ double seed_step(int t, double x) {
        sum += x;
        return REC_NAN;
    }
    __device__ double steady(double x) {
        if (p == 1) { prev = x; return x; }
        prev = (prev * pm1 + x) * rp;
        return prev;
    }
    // seed sum up to fv+p-1 (mean: 1/p), then pm1 * rp; p = 1 copies (0)
    static constexpr bool kWarmScan = true;
    __device__ double gain(int c0, int c1) const {
        if (p == 1) return 0.0;
        const int ws = fv + p - 1;
        const double g = (c0 <= ws && ws < c1) ? 1.0 / (double)p : 1.0;
        const int n = bars_from(c0, c1, ws + 1);
        return n > 0 ? g * pow(pm1 * rp, (double)n) : g;
    }
    __device__ void begin_carry(int t0, double x, const double *st) {
        if (p > 1 && t0 < fv + p - 1) sum = st[0];
        else prev = st[0];
    }
    __device__ void get_state(int t1, double *st) const {
        st[0] = (p > 1 && t1 < fv + p - 1) ? sum : prev;
    }
};

// Wilders (wilders.rs:262-320): seed summed in groups of four,
// sum += ((x0 + x1) + x2) + x3, y = sum * (1/p) at fv+p-1, then
// y = fma(x - y, 1/p, y).  p = 1 copies after the seed bar.
struct WildersEng {
    static constexpr bool kSeedScan = false;
    double al, y, sum, g;
    int p, fv;
    __device__ void setup(int p_, int fv_) {
        p = p_; fv = fv_;
        al = 1.0 / (double)p;
    }
    __device__ void set_alpha(double a_) { al = a_; }
    __device__ int kb() const { return burn_bars(1.0 - al, REC_NATS); }
    __device__ int exact_start() const { return fv; }
    __device__ int burn_ok_from() const { return fv + p; }
    __device__ double seed_bar(int k, double x) {       // k = t - fv < p
        if ((k & 3) == 0) g = x; else g += x;
        if ((k & 3) == 3 || k == p - 1) sum += g;
        if (k == p - 1) { y = sum * al; return y; }
        return REC_NAN;
    }
    __device__ double begin_exact(double x) { sum = 0.0; return seed_bar(0, x); }
    __device__ void begin_burn(double x) { y = x; }
    __device__ double step(int t, double x) {
        const int k = t - fv;
        if (k < p) return seed_bar(k, x);
        if (p == 1) { y = x; return x; }
        y = fma(x - y, al, y);
        return y;
    }
    __device__ double out_before_start(int t) const { return REC_NAN; }
    static constexpr int kScan = 1;
    static constexpr bool kTwoStage = false;
    static constexpr bool kAdvance = false;
    static constexpr int kPre = 0;
    static constexpr int kStageW = 64;
    // a scan lane that would replay its warm-up from first_valid starts at
    // its last warm-up bar from the precomputed seed instead (wilders_seed)
    static constexpr bool kSeeded = true;
    __device__ double begin_seeded(double seed) { y = seed; return y; }
    __device__ int seed_end() const { return fv + p - 1; }
    __device__ double seed_step(int t, double x) { return seed_bar(t - fv, x); }
    __device__ double steady(double x) {
        if (p == 1) { y = x; return x; }
        y = fma(x - y, al, y);
        return y;
    }
    // the seed is summed in groups of four (a two-word state), so a carried
    // start is taken only past the warm-up: fma(x - y, al, y), i.e. 1 - al
    static constexpr bool kWarmScan = false;
    __device__ double gain(int c0, int c1) const { return pow(1.0 - al, (double)(c1 - c0)); }
    __device__ void begin_carry(int t0, double x, const double *st) { y = st[0]; }
    __device__ void get_state(int t1, double *st) const { st[0] = y; }
};

// ---------------------------------------------------------------- chunk planning

// A caller-supplied smoothing factor (the reference's d_alphas / alpha
// arguments of the EMA and Wilders device forms) replaces the one the engine
// derives from the period; the other engines have none.
template <class E>
__device__ __forceinline__ void apply_alpha(E &e, const float *alphas, int i) {
    if constexpr (__is_same(E, EmaEng) || __is_same(E, WildersEng)) {
        if (alphas) e.set_alpha((double)alphas[i]);
    }
}

// Bar at which a lane starts (exact or burn-in) for chunk [c0, c1).
template <class E>
__device__ __forceinline__ int lane_start(const E &e, int c0, bool clean, bool *burn) {
    int es = e.exact_start();
    if (clean) {
        int b = c0 - e.kb();
        if (b >= e.burn_ok_from() && b > es) { *burn = true; return b; }
    }
    *burn = false;
    return es;
}

// Wave-wide minimum, returned as a scalar (readfirstlane) so loops bounded by
// it keep their induction variables and row addresses in SGPRs.
__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// ---------------------------------------------------------------- many series (time-major)
//
// grid.x: groups of 64 series, grid.y: chunks.  Lane = series; every load and
// store is one coalesced row segment of the [T][N] matrix.
template <class E, int NIN>
__device__ __forceinline__ void drive_ms(const rec_args &a, int p, float *out) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    const int fv = active ? a.fvs[s] : a.T;
    const bool clean = active ? (a.clean != nullptr && a.clean[s] != 0) : true;
    E e;
    e.setup(p, fv);
    apply_alpha(e, a.alphas, 0);
    bool burn = false;
    int start = active ? lane_start(e, c0, clean, &burn) : a.T;
    if (start > a.T) start = a.T;
    const int t_begin = wave_min(min(start, c0));
    const int lo = active ? min(start, c0) : a.T - 1;
    const int sc = min(s, N - 1);
    // blocks of 8 bars: the block's loads (rows clamped into [lo, T-1]) are
    // issued before its sequential updates; a block in which every lane is
    // past its start and its warm-up takes the engine's steady update only
    for (int tb = t_begin; tb < c1; tb += 8) {
        double xv[8], hv[8], lv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long loff = (long)min(max(tb + k, lo), a.T - 1) * N + sc;
            xv[k] = (double)a.x[loff];
            if (NIN == 3) {
                hv[k] = (double)a.h[loff];
                lv[k] = (double)a.l[loff];
            }
        }
        const bool lean = tb + 8 <= c1 && __all(!active || (start < tb && tb >= e.burn_ok_from()));
        if (lean) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                double y;
                if constexpr (NIN == 1) y = e.steady(xv[k]);
                else y = e.steady(hv[k], lv[k], xv[k]);
                if (active && tb + k >= c0) out[(size_t)(tb + k) * N + s] = (float)y;
            }
            continue;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = tb + k;
            if (t >= c1) break;
            double y = REC_NAN;
            if (active) {
                if (t == start) {
                    if constexpr (NIN == 1) {
                        if (burn) e.begin_burn(xv[k]); else y = e.begin_exact(xv[k]);
                    } else {
                        if (burn) e.begin_burn(hv[k], lv[k], xv[k]);
                        else y = e.begin_exact(hv[k], lv[k], xv[k]);
                    }
                } else if (t > start) {
                    if constexpr (NIN == 1) y = e.step(t, xv[k]); else y = e.step(t, hv[k], lv[k], xv[k]);
                }
                if (t >= c0) out[(size_t)t * N + s] = (float)y;
            }
        }
    }
}

// RSI burn-in needs the seed-window NaN check, done by a short scan first.
__device__ __forceinline__ void drive_ms_rsi(const rec_args &a, int p, float *out) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    const int fv = active ? a.fvs[s] : a.T;
    const bool clean = active ? (a.clean != nullptr && a.clean[s] != 0) : true;
    RsiEng e;
    e.setup(p, fv);
    bool burn = false;
    int start = active ? lane_start(e, c0, clean, &burn) : a.T;
    if (burn) {
        e.px = (double)a.x[(size_t)fv * N + s];
        for (int t = fv + 1; t <= min(fv + p, a.T - 1); ++t) e.seed_scan(t, (double)a.x[(size_t)t * N + s]);
    }
    const int t_begin = wave_min(min(start, c0));
    for (int t = t_begin; t < c1; ++t) {
        const size_t off = (size_t)t * N + s;
        const double x = (double)a.x[(long)min(max(t, active ? min(start, c0) : a.T - 1),
                                               a.T - 1) * N + min(s, N - 1)];
        double y = REC_NAN;
        if (active) {
            if (t == start) {
                if (burn) e.begin_burn(x); else y = e.begin_exact(x);
            } else if (t > start) {
                y = e.step(t, x);
            }
            if (t >= c0) out[off] = (float)y;
        }
    }
}

extern "C" __global__ __launch_bounds__(64) void ema_ms1p_f32(rec_args a) {
    drive_ms<EmaEng, 1>(a, a.periods[0], a.out0);
}
extern "C" __global__ __launch_bounds__(64) void rsi_ms1p_f32(rec_args a) {
    drive_ms_rsi(a, a.periods[0], a.out0);
}
extern "C" __global__ __launch_bounds__(64) void atr_ms1p_f32(rec_args a) {
    drive_ms<AtrEng, 3>(a, a.periods[0], a.out0);
}
extern "C" __global__ __launch_bounds__(64) void smma_ms1p_f32(rec_args a) {
    drive_ms<SmmaEng, 1>(a, a.periods[0], a.out0);
}
extern "C" __global__ __launch_bounds__(64) void wilders_ms1p_f32(rec_args a) {
    drive_ms<WildersEng, 1>(a, a.periods[0], a.out0);
}
extern "C" __global__ __launch_bounds__(64) void adx_ms1p_f32(rec_args a) {
    drive_ms<AdxEng, 3>(a, a.periods[0], a.out0);
}

// ---------------------------------------------------------------- fused EMA + RSI + ATR
//
// One pass over close/high/low [T][N]: EMA(p1) and RSI(p2) of close and
// ATR(p3) of (high, low, close), three time-major outputs.
//
// Every lane has a bar `steady` from which its three engines are in pure
// recurrence.  Each bar, lanes at or past it run the lean update (three
// recurrences, no phase checks, the RSI ratio divided in f32 via rcp + one
// Newton step: <= ~2 ulp of f32, far inside the 1e-5 bound); lanes still in a
// warm-up run the general engines.  When every lane of the wave is steady the
// general path is skipped entirely, so a series start costs its own lane ~p
// general bars.  Loads of the next U rows are issued before the current U
// rows are consumed (register double buffering); addresses are clamped so
// the loads are unconditional and stay in flight.  U = 8 keeps the kernel at
// 167 VGPRs (3 waves per SIMD); U = 16 (185 VGPRs, 2 waves) was 1.5% slower
// on 4,096 x 50,000, U = 32 (286 VGPRs) holds one wave.
#define FUSED_U 8


extern "C" __global__ __launch_bounds__(64) void ms_ema_rsi_atr_f32(rec_args a) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, a.T);
    const int fv = active ? a.fvs[s] : a.T;          // first non-NaN close
    const int fv3 = active ? a.fvs2[s] : a.T;        // first bar with high, low, close
    const bool clean = active ? (a.clean != nullptr && a.clean[s] != 0) : true;
    const bool live = active && fv < c1;             // has an output in this chunk
    EmaEng ee; RsiEng re; AtrEng ae;
    ee.setup(a.p1, fv); re.setup(a.p2, fv); ae.setup(a.p3, fv3);
    bool be = false, br = false, ba = false;
    int se = live ? lane_start(ee, c0, clean, &be) : a.T;
    int sr = live ? lane_start(re, c0, clean, &br) : a.T;
    int sa = live ? lane_start(ae, c0, clean, &ba) : a.T;
    if (br) {
        re.px = (double)a.x[(size_t)fv * N + s];
        for (int t = fv + 1; t <= min(fv + a.p2, a.T - 1); ++t)
            re.seed_scan(t, (double)a.x[(size_t)t * N + s]);
    }
    int steady = a.T;
    if (live && clean) {
        steady = max(be ? se + 1 : fv + a.p1, br ? sr + 1 : fv + a.p2 + 1);
        steady = max(steady, ba ? sa + 1 : fv3 + a.p3);
    }
    const int smin = live ? min(min(se, sr), min(sa, c0)) : c0;
    const int t_begin = wave_min(smin);
    // bars before a lane's first engine start need no work at all
    const int lane_first = live ? min(min(se, sr), sa) : a.T;
    const bool rsi_nan_fixed = re.poisoned;
    const int sc = min(s, N - 1);
    // Row addresses are wave-uniform (scalar base + 32-bit lane offset).

    float bc[FUSED_U], bh[FUSED_U], bl[FUSED_U];
    float nc[FUSED_U], nh[FUSED_U], nl[FUSED_U];
#pragma unroll
    for (int u = 0; u < FUSED_U; ++u) {
        const size_t rb = (size_t)min(t_begin + u, a.T - 1) * (size_t)N;
        bc[u] = (a.x + rb)[sc];
        bh[u] = (a.h + rb)[sc];
        bl[u] = (a.l + rb)[sc];
    }
    for (int t0 = t_begin; t0 < c1; t0 += FUSED_U) {
#pragma unroll
        for (int u = 0; u < FUSED_U; ++u) {
            const size_t rb = (size_t)min(t0 + FUSED_U + u, a.T - 1) * (size_t)N;
            nc[u] = (a.x + rb)[sc];
            nh[u] = (a.h + rb)[sc];
            nl[u] = (a.l + rb)[sc];
        }
        // wave-uniform: does any lane have warm-up bars inside this block?
        const bool mixed = __any(live && lane_first < t0 + FUSED_U && steady > t0);
        if (!mixed) {
            // every live lane is steady for the whole block, or has not started
            // (its state is rebuilt when it starts, so updating it is harmless)
#pragma unroll
            for (int u = 0; u < FUSED_U; ++u) {
                const int t = t0 + u;
                const double x = bc[u], hh = bh[u], ll = bl[u];
                ee.prev = fma(ee.beta, ee.prev, ee.alpha * x);          // ema.rs:462
                const double d = x - re.px;
                re.px = x;
                re.ag = fma(re.ag, re.beta, re.inv_p * (d > 0.0 ? d : 0.0));  // rsi.rs:377
                re.al = fma(re.al, re.beta, re.inv_p * (d < 0.0 ? -d : 0.0));
                const double tr = true_range(hh, ll, ae.pc);
                ae.pc = x;
                ae.rma = fma(-ae.alpha, ae.rma, ae.rma) + ae.alpha * tr;      // atr.rs:393
                const bool ok = t >= steady;
                const float o0 = ok ? (float)ee.prev : VTA_QNAN_F32;
                const float o1 = (ok && !rsi_nan_fixed) ? rsi_ratio_f32(re.ag, re.al)
                                                        : VTA_QNAN_F32;
                const float o2 = ok ? (float)ae.rma : VTA_QNAN_F32;
                if (t >= c0 && t < c1 && active) {
                    const size_t rb = (size_t)t * (size_t)N;
                    (a.out0 + rb)[s] = o0;
                    (a.out1 + rb)[s] = o1;
                    (a.out2 + rb)[s] = o2;
                }
            }
        } else {
#pragma unroll
        for (int u = 0; u < FUSED_U; ++u) {
            const int t = t0 + u;
            if (t >= c1) break;
            const double x = bc[u], hh = bh[u], ll = bl[u];
            float o0 = VTA_QNAN_F32, o1 = VTA_QNAN_F32, o2 = VTA_QNAN_F32;
            if (t >= steady) {
                // lean steady-state update (clean, started lanes only)
                ee.prev = fma(ee.beta, ee.prev, ee.alpha * x);          // ema.rs:462
                const double d = x - re.px;
                re.px = x;
                re.ag = fma(re.ag, re.beta, re.inv_p * (d > 0.0 ? d : 0.0));  // rsi.rs:377
                re.al = fma(re.al, re.beta, re.inv_p * (d < 0.0 ? -d : 0.0));
                const double tr = true_range(hh, ll, ae.pc);
                ae.pc = x;
                ae.rma = fma(-ae.alpha, ae.rma, ae.rma) + ae.alpha * tr;      // atr.rs:393
                o0 = (float)ee.prev;
                o1 = rsi_nan_fixed ? VTA_QNAN_F32 : rsi_ratio_f32(re.ag, re.al);
                o2 = (float)ae.rma;
            } else if (t >= lane_first) {
                double y0 = REC_NAN, y1 = REC_NAN, y2 = REC_NAN;
                if (t == se) { if (be) ee.begin_burn(x); else y0 = ee.begin_exact(x); }
                else if (t > se) y0 = ee.step(t, x);
                if (t == sr) { if (br) re.begin_burn(x); else y1 = re.begin_exact(x); }
                else if (t > sr) y1 = re.step(t, x);
                if (t == sa) { if (ba) ae.begin_burn(hh, ll, x); else y2 = ae.begin_exact(hh, ll, x); }
                else if (t > sa) y2 = ae.step(t, hh, ll, x);
                o0 = (float)y0;
                o1 = (float)y1;
                o2 = (float)y2;
            }
            if (active && t >= c0) {
                const size_t off = (size_t)t * N + s;
                a.out0[off] = o0;
                a.out1[off] = o1;
                a.out2[off] = o2;
            }
        }
        }
#pragma unroll
        for (int u = 0; u < FUSED_U; ++u) { bc[u] = nc[u]; bh[u] = nh[u]; bl[u] = nl[u]; }
    }
}

// ---------------------------------------------------------------- one series, many params
//
// grid.x: groups of 64 combos, grid.y: chunks.  Lane = combo.  The wave
// stages 64 bars of the series in LDS (one coalesced load, then broadcast
// reads), and buffers 64 bars x 64 combos of output in LDS so each output row
// segment is written as one coalesced 256-byte store.
// What one block of a one-series batch pass (or one wave of a fold) works
// on: the series, the combos' periods and outputs, the chunking, and which
// combo group and chunk this block is (for a fold: `group` = the combo).
struct batch_view {
    const float *x, *h, *l;
    const int32_t *periods;
    double *carry;           // [combo][chunk][states]
    const float *alphas;     // per combo, or NULL (see apply_alpha)
    const double *seeds;     // Wilders: per-combo seed, or NULL
    int T, nlanes, chunk, nch, phase, fv, group, k;
    bool clean;
};

// The view of a block of a direct launch: grid.x = combo groups (a fold: the
// combo), grid.y = chunks.
__device__ __forceinline__ batch_view view_of(const rec_args &a) {
    batch_view v;
    v.x = a.x; v.h = a.h; v.l = a.l;
    v.periods = a.periods;
    v.carry = a.carry;
    v.alphas = a.alphas;
    v.seeds = a.seeds;
    v.T = a.T;
    v.nlanes = a.nlanes;
    v.chunk = a.chunk;
    v.nch = (a.T + a.chunk - 1) / a.chunk;
    v.phase = a.phase;
    v.fv = a.fvs[0];
    v.group = blockIdx.x;
    v.k = blockIdx.y;
    v.clean = a.clean != nullptr &&
              (a.clean_epoch != 0 ? a.clean[0] != a.clean_epoch : a.clean[0] != 0);
    return v;
}

template <class E, int NIN>
__device__ __forceinline__ void drive_batch(const batch_view &a, float *out) {
    __shared__ __align__(16) float xs[3][64];
    // per-bar terms every combo shares (E::kPre of them, f64): lane j computes
    // them for bar tb + j, so the steady and pass-1 loops read them instead of
    // recomputing them in every lane
    constexpr int NPRE = E::kPre;
    __shared__ __align__(16) double pv[NPRE > 0 ? NPRE : 1][64];
    // 64 x 65 output staging in dynamic LDS, launched only with the passes
    // that write rows (a pass without stores keeps more waves per CU)
    // 64 rows x SW bars (+1 pad), flushed every SW bars: engines whose steps
    // are long (RSI, ADX) flush every 32 bars, which leaves LDS for ~3 waves
    // per SIMD where 64-bar staging allows 2; the others keep 64-bar row
    // segments
    constexpr int SW = E::kStageW;    // the host launches 64 x (SW + 1) floats
                                      // of dynamic LDS (batch_stage_bytes, rec.cpp)
    extern __shared__ float ob_dyn[];
    float (*ob)[SW + 1] = reinterpret_cast<float (*)[SW + 1]>(ob_dyn);
    const int lane = threadIdx.x & 63;
    const int c = a.group * 64 + lane;
    const bool active = c < a.nlanes;
    const int T = a.T;
    const int k = a.k;
    const bool clean = a.clean;
    // a scan pass over an unclean series: its lanes run exactly from
    // first_valid, so the output pass runs the whole series in chunk 0 and
    // every other block of the scan has nothing to do
    if (a.phase != 0 && !clean && (a.phase != 3 || k != 0)) return;
    const int c0 = (a.phase != 0 && !clean) ? 0 : k * a.chunk;
    const int c1 = (a.phase != 0 && !clean) ? T : min(c0 + a.chunk, T);
    const int fv = a.fv;
    const int p = active ? a.periods[c] : 1;
    E e;
    e.setup(p, fv);
    if (active) apply_alpha(e, a.alphas, c);
    bool burn = false, carried = false, seeded = false;
    int start;
    if (a.phase != 0) {
        // chunked scan: chunks after first_valid (past the warm-up for engines
        // without kWarmScan) start from a carried state at c0 - 1; an unclean
        // series runs every chunk exactly from first_valid
        carried = active && clean &&
                  (E::kWarmScan ? c0 > e.exact_start() : c0 >= e.burn_ok_from());
        start = active ? (carried ? c0 - 1 : e.exact_start()) : T;
    } else {
        start = active ? lane_start(e, c0, clean, &burn) : T;
    }
    double seedv = 0.0;
    if constexpr (E::kSeeded) {
        // an exact start whose seed is known begins at the last warm-up bar
        if (active && !carried && !burn && a.seeds) {
            seeded = true;
            start = e.fv + e.p - 1;
            seedv = a.seeds[c];
        }
    }
    if constexpr (E::kSeedScan) {
        // RSI burn-in: the seed window's NaN rule still has to be seen
        if (burn) {
            e.px = (double)a.x[fv];
            for (int t = fv + 1; t <= min(fv + p, T - 1); ++t) e.seed_scan(t, (double)a.x[t]);
        }
    }
    // carried start state: zeros (pass 1), the exact first-stage states with a
    // zero second stage (pass 2 of a two-stage engine), or all exact (pass 3)
    constexpr int S = E::kScan > 0 ? E::kScan : 1;
    double st0[S];
#pragma unroll
    for (int i = 0; i < S; ++i) st0[i] = 0.0;
    if (carried && a.phase >= 2) {
        const double *cs = a.carry + ((size_t)c * a.nch + (k - 1)) * S;
        const int n = (E::kTwoStage && a.phase == 2) ? S - 1 : S;
#pragma unroll
        for (int i = 0; i < S; ++i)
            if (i < n) st0[i] = cs[i];
    }
    const bool emit = a.phase == 0 || a.phase == 3;
    // 32-bar staging: rows of bars tb + 32 hh .. + 31, each store two rows x
    // 32 consecutive bars (lane & 31 = bar, lane >> 5 = row parity)
    const int nrow = min(64, a.nlanes - a.group * 64);
    auto flush_half = [&](int tb, int hh) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int t = tb + 32 * hh + (lane & 31);
        if (tb + 32 * hh + 32 > c0) {
            const bool ok = t >= c0 && t < c1;
#pragma unroll 4
            for (int i = 0; i < 32; ++i) {
                const int r = 2 * i + (lane >> 5);
                if (ok && r < nrow) out[(size_t)(a.group * 64 + r) * T + t] = ob[r][lane & (SW - 1)];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // a carried lane restarts after bar c0 - 1 here, from that one bar read
    // directly, so the wave need not run the 64-bar block holding it
    if (carried) {
        const int tp = c0 - 1;
        if constexpr (NIN == 1) e.begin_carry(tp, (double)a.x[tp], st0);
        else e.begin_carry(tp, (double)a.h[tp], (double)a.l[tp], (double)a.x[tp], st0);
    }
    const int t_begin = wave_min(carried ? c0 : min(start, c0)) & ~63;
    // bars of the next block are loaded into registers while the current one
    // is processed (with the bar before each, for the per-bar terms)
    float nx0, nx1 = 0.0f, nx2 = 0.0f, np0 = 0.0f, np1 = 0.0f, np2 = 0.0f;
    auto load_next = [&](int tb0) {
        const long t = min(tb0 + lane, T - 1);
        nx0 = a.x[t];
        if (NIN == 3) {
            nx1 = a.h[t];
            nx2 = a.l[t];
        }
        if constexpr (NPRE > 0) {
            const long tp = max(t - 1, 0L);
            np0 = a.x[tp];
            if (NIN == 3) {
                np1 = a.h[tp];
                np2 = a.l[tp];
            }
        }
    };
    // a whole block that seeded lanes can run without the per-bar start logic
    auto seeded_block = [&](int tb) {
        if constexpr (E::kSeeded && NIN == 1)
            return tb + 64 <= c1 &&
                   __all(!active || seeded || (start < tb && tb >= e.burn_ok_from()));
        else
            return false;
    };
    load_next(t_begin);
    for (int tb = t_begin; tb < c1; tb += 64) {
        xs[0][lane] = nx0;
        if (NIN == 3) {
            xs[1][lane] = nx1;
            xs[2][lane] = nx2;
        }
        if constexpr (NPRE > 0) {
            double o[NPRE];
            E::pre((double)nx1, (double)nx2, (double)nx0, (double)np1, (double)np2, (double)np0, o);
#pragma unroll
            for (int i = 0; i < NPRE; ++i) pv[i][lane] = o[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        load_next(tb + 64);
        // lean path: a full block in which every lane is past its start and
        // its warm-up -- only the engine's steady update, bars read four at a
        // time
        const bool lean = tb + 64 <= c1 && __all(!active || (start < tb && tb >= e.burn_ok_from()));
        bool advanced = false;
        if constexpr (E::kAdvance) {
            if (a.phase == 1 && tb + 64 <= c1 &&
                __all(!active || (start < tb && tb >= e.advance_from()))) {
                // pass 1: the state update only, no outputs
                if constexpr (NPRE > 0) {
#pragma unroll 4
                    for (int j = 0; j < 64; ++j) {
                        double o[NPRE];
#pragma unroll
                        for (int i = 0; i < NPRE; ++i) o[i] = pv[i][j];
                        e.advance_pre(o);
                    }
                    e.set_prev(xs[1][63], xs[2][63], xs[0][63]);
                } else
#pragma unroll 2
                for (int j = 0; j < 64; j += 4) {
                    const float4 x4 = *reinterpret_cast<const float4 *>(&xs[0][j]);
                    if constexpr (NIN == 1) {
                        e.advance(x4.x);
                        e.advance(x4.y);
                        e.advance(x4.z);
                        e.advance(x4.w);
                    } else {
                        const float4 h4 = *reinterpret_cast<const float4 *>(&xs[1][j]);
                        const float4 l4 = *reinterpret_cast<const float4 *>(&xs[2][j]);
                        e.advance(h4.x, l4.x, x4.x);
                        e.advance(h4.y, l4.y, x4.y);
                        e.advance(h4.z, l4.z, x4.z);
                        e.advance(h4.w, l4.w, x4.w);
                    }
                }
                advanced = true;
            }
        }
        if constexpr (E::kTwoStage) {
            // ADX's DX seed: every lane between its first-stage seed and its
            // first ADX for the whole block (outputs NaN)
            if (!advanced && tb + 64 <= c1 &&
                __all(!active || (start < tb && tb > e.fv + e.p && tb + 64 <= e.dx_seed_end()))) {
#pragma unroll 2
                for (int j = 0; j < 64; j += 4) {
                    const float4 x4 = *reinterpret_cast<const float4 *>(&xs[0][j]);
                    const float4 h4 = *reinterpret_cast<const float4 *>(&xs[1][j]);
                    const float4 l4 = *reinterpret_cast<const float4 *>(&xs[2][j]);
                    e.dx_seed_step(h4.x, l4.x, x4.x);
                    e.dx_seed_step(h4.y, l4.y, x4.y);
                    e.dx_seed_step(h4.z, l4.z, x4.z);
                    e.dx_seed_step(h4.w, l4.w, x4.w);
                    if (emit) {
                        ob[lane][j & (SW - 1)] = (float)REC_NAN;
                        ob[lane][(j + 1) & (SW - 1)] = (float)REC_NAN;
                        ob[lane][(j + 2) & (SW - 1)] = (float)REC_NAN;
                        ob[lane][(j + 3) & (SW - 1)] = (float)REC_NAN;
                        if (SW == 32 && ((j + 4) & 31) == 0) flush_half(tb, j >> 5);
                    }
                }
                e.dx_count += 64;
                advanced = true;
            }
        }
        if (advanced) {
            // block done
        } else if (lean) {
          if constexpr (NPRE > 0) {
#pragma unroll 4
            for (int j = 0; j < 64; ++j) {
                double o[NPRE];
#pragma unroll
                for (int i = 0; i < NPRE; ++i) o[i] = pv[i][j];
                const double y = e.steady_pre(o);
                if (emit) {
                    ob[lane][j & (SW - 1)] = (float)y;
                    if (SW == 32 && (j & 31) == 31) flush_half(tb, j >> 5);
                }
            }
            e.set_prev(xs[1][63], xs[2][63], xs[0][63]);
          } else {
#pragma unroll 2
            for (int j = 0; j < 64; j += 4) {
                const float4 x4 = *reinterpret_cast<const float4 *>(&xs[0][j]);
                double y0, y1, y2, y3;
                if constexpr (NIN == 1) {
                    y0 = e.steady(x4.x);
                    y1 = e.steady(x4.y);
                    y2 = e.steady(x4.z);
                    y3 = e.steady(x4.w);
                } else {
                    const float4 h4 = *reinterpret_cast<const float4 *>(&xs[1][j]);
                    const float4 l4 = *reinterpret_cast<const float4 *>(&xs[2][j]);
                    y0 = e.steady(h4.x, l4.x, x4.x);
                    y1 = e.steady(h4.y, l4.y, x4.y);
                    y2 = e.steady(h4.z, l4.z, x4.z);
                    y3 = e.steady(h4.w, l4.w, x4.w);
                }
                if (emit) {
                    ob[lane][j & (SW - 1)] = (float)y0;
                    ob[lane][(j + 1) & (SW - 1)] = (float)y1;
                    ob[lane][(j + 2) & (SW - 1)] = (float)y2;
                    ob[lane][(j + 3) & (SW - 1)] = (float)y3;
                    if (SW == 32 && ((j + 4) & 31) == 0) flush_half(tb, j >> 5);
                }
            }
          }
        } else if (seeded_block(tb)) {
            // lanes that start from a seed in this block or a later one run the
            // steady update on every bar, take the seed at their start bar and
            // write NaN before it; the others are past their warm-ups
            if constexpr (E::kSeeded) {
                auto sbar = [&](int t, float x) {
                    double y = e.steady(x);
                    if (t == start) y = e.begin_seeded(seedv);
                    return t < start ? REC_NAN : y;
                };
#pragma unroll 2
                for (int j = 0; j < 64; j += 4) {
                    const float4 x4 = *reinterpret_cast<const float4 *>(&xs[0][j]);
                    const double y0 = sbar(tb + j, x4.x);
                    const double y1 = sbar(tb + j + 1, x4.y);
                    const double y2 = sbar(tb + j + 2, x4.z);
                    const double y3 = sbar(tb + j + 3, x4.w);
                    if (emit) {
                        ob[lane][j & (SW - 1)] = (float)y0;
                        ob[lane][(j + 1) & (SW - 1)] = (float)y1;
                        ob[lane][(j + 2) & (SW - 1)] = (float)y2;
                        ob[lane][(j + 3) & (SW - 1)] = (float)y3;
                        if (SW == 32 && ((j + 4) & 31) == 0) flush_half(tb, j >> 5);
                    }
                }
            }
        } else if (tb + 64 <= c1 && __all(!active || (start < tb && tb + 64 <= e.seed_end()))) {
            // every lane still accumulating its warm-up seed for the whole block
#pragma unroll 2
            for (int j = 0; j < 64; j += 4) {
                const float4 x4 = *reinterpret_cast<const float4 *>(&xs[0][j]);
                double y0, y1, y2, y3;
                if constexpr (NIN == 1) {
                    y0 = e.seed_step(tb + j, x4.x);
                    y1 = e.seed_step(tb + j + 1, x4.y);
                    y2 = e.seed_step(tb + j + 2, x4.z);
                    y3 = e.seed_step(tb + j + 3, x4.w);
                } else {
                    const float4 h4 = *reinterpret_cast<const float4 *>(&xs[1][j]);
                    const float4 l4 = *reinterpret_cast<const float4 *>(&xs[2][j]);
                    y0 = e.seed_step(tb + j, h4.x, l4.x, x4.x);
                    y1 = e.seed_step(tb + j + 1, h4.y, l4.y, x4.y);
                    y2 = e.seed_step(tb + j + 2, h4.z, l4.z, x4.z);
                    y3 = e.seed_step(tb + j + 3, h4.w, l4.w, x4.w);
                }
                if (emit) {
                    ob[lane][j & (SW - 1)] = (float)y0;
                    ob[lane][(j + 1) & (SW - 1)] = (float)y1;
                    ob[lane][(j + 2) & (SW - 1)] = (float)y2;
                    ob[lane][(j + 3) & (SW - 1)] = (float)y3;
                    if (SW == 32 && ((j + 4) & 31) == 0) flush_half(tb, j >> 5);
                }
            }
        } else
        for (int j = 0; j < 64; ++j) {
            const int t = tb + j;
            double y = REC_NAN;
            if (active && t < c1) {
                const double x = xs[0][j];
                if (t == start) {
                    // (a carried lane began before the loop)
                    if constexpr (NIN == 1) {
                        if (carried) {
                        } else if (seeded) {
                            if constexpr (E::kSeeded) y = e.begin_seeded(seedv);
                        } else if (burn) e.begin_burn(x);
                        else y = e.begin_exact(x);
                    } else {
                        if (carried) {
                        } else if (burn) e.begin_burn(xs[1][j], xs[2][j], x);
                        else y = e.begin_exact(xs[1][j], xs[2][j], x);
                    }
                } else if (t > start) {
                    if constexpr (NIN == 1) y = e.step(t, x);
                    else y = e.step(t, xs[1][j], xs[2][j], x);
                }
            }
            if (emit) {
                ob[lane][j & (SW - 1)] = (float)y;
                if (SW == 32 && (j & 31) == 31) flush_half(tb, j >> 5);
            }
        }
        // 64-bar staging: one flush per block, after whichever path ran; each
        // store writes one row x 64 bars (lane = bar)
        if (SW == 64 && emit) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (tb + 64 > c0) {
                const int t = tb + lane;
                const bool ok = t >= c0 && t < c1;
                for (int r = 0; r < nrow; ++r)
                    if (ok) out[(size_t)(a.group * 64 + r) * T + t] = ob[r][lane & (SW - 1)];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if constexpr (E::kScan > 0) {
        // chunk-end states: all of them after pass 1; after pass 2 (two-stage
        // engines) only the second stage's, next to the folded first stage
        if ((a.phase == 1 || a.phase == 2) && active && k + 1 < a.nch) {
            double st[S];
            e.get_state(c1 - 1, st);
            double *cs = a.carry + ((size_t)c * a.nch + k) * S;
            const int lo = a.phase == 2 ? S - 1 : 0;
#pragma unroll
            for (int i = 0; i < S; ++i)
                if (i >= lo) cs[i] = st[i];
        }
    }
}

// Folds the chunk-end states of pass 1 in chunk order, in place: a carried
// chunk k maps the state at its start to s_k = g_k s_{k-1} + b_k, with b_k its
// zero-start run and g_k = gain(c0, c1); a chunk that ran exactly from
// first_valid has g_k = 0.  One wave per combo; the carry holds [combo]
// [chunk][state].  64 chunks at a time: an inclusive scan of the affine maps
// across the lanes, (g, b) o (g', b') = (g g', g b' + b), then the previous
// row's last state applied.  Two-stage engines fold their first-stage states
// after pass 1 (a.phase = 1) and their second-stage state after pass 2.
// A chunk that ends before first_valid stores an undefined state; the chunk
// after it always has g = 0, so a g = 0 step selects b instead of forming
// 0 * s (which is NaN for an undefined s).
template <class E>
__device__ __forceinline__ void scan_combine(const batch_view &a) {
    const int c = a.group;                  // the combo
    const int lane = threadIdx.x & 63;
    if (c >= a.nlanes || !a.clean) return;
    E e;
    e.setup(a.periods[c], a.fv);
    apply_alpha(e, a.alphas, c);
    constexpr int S = E::kScan > 0 ? E::kScan : 1;
    int lo = 0, hi = S;
    bool second = false;
    if constexpr (E::kTwoStage) {
        if (a.phase == 2) {
            lo = S - 1;
            second = true;
        } else {
            hi = S - 1;
        }
    }
    const int nch = (a.T + a.chunk - 1) / a.chunk;
    const int nst = nch - 1;                 // chunk-end states read by later chunks
    double *base = a.carry + (size_t)c * nch * S;
    double cin[S];
#pragma unroll
    for (int i = 0; i < S; ++i) cin[i] = 0.0;
    for (int r0 = 0; r0 < nst; r0 += 64) {
        const int k = r0 + lane;
        const bool ok = k < nst;
        const int c0 = k * a.chunk;
        const int c1 = min(c0 + a.chunk, a.T);
        const bool carried =
            ok && (E::kWarmScan ? c0 > e.exact_start() : c0 >= e.burn_ok_from());
        double g = 0.0;
        if (carried) {
            if constexpr (E::kTwoStage) g = second ? e.gain2(c0, c1) : e.gain(c0, c1);
            else g = e.gain(c0, c1);
        }
        double b[S];
#pragma unroll
        for (int i = 0; i < S; ++i) b[i] = (ok && i >= lo && i < hi) ? base[(size_t)k * S + i] : 0.0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double gu = __shfl_up(g, o);
            double bu[S];
#pragma unroll
            for (int i = 0; i < S; ++i) bu[i] = __shfl_up(b[i], o);
            if (lane >= o) {
#pragma unroll
                for (int i = 0; i < S; ++i) b[i] = g != 0.0 ? fma(g, bu[i], b[i]) : b[i];
                g *= gu;
            }
        }
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const double y = g != 0.0 ? fma(g, cin[i], b[i]) : b[i];
            if (ok && i >= lo && i < hi) base[(size_t)k * S + i] = y;
            cin[i] = __shfl(y, 63);
        }
    }
}

extern "C" __global__ __launch_bounds__(64) void ema_batch_f32(rec_args a) {
    drive_batch<EmaEng, 1>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void rsi_batch_f32(rec_args a) {
    drive_batch<RsiEng, 1>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void atr_batch_f32(rec_args a) {
    drive_batch<AtrEng, 3>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void smma_batch_f32(rec_args a) {
    drive_batch<SmmaEng, 1>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void wilders_batch_f32(rec_args a) {
    drive_batch<WildersEng, 1>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void adx_batch_f32(rec_args a) {
    drive_batch<AdxEng, 3>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void adx_scan_combine(rec_args a) {
    scan_combine<AdxEng>(view_of(a));
}
extern "C" __global__ __launch_bounds__(64) void atr_tr_batch_f32(rec_args a) {
    drive_batch<AtrTrEng, 1>(view_of(a), a.out0);
}
extern "C" __global__ __launch_bounds__(64) void atr_tr_scan_combine(rec_args a) {
    scan_combine<AtrTrEng>(view_of(a));
}

// True range of one series (tr_from_hlc_f32 in the reference): 0 before
// first_valid, high - low at first_valid, then max(h - l, |h - c[t-1]|,
// |l - c[t-1]|), evaluated in f64 (exact for f32 inputs) and rounded once.
// ---------------------------------------------------------------- packed batches
//
// Many one-series batches of one indicator in one launch per scan pass: job j
// is a series (its own length, first_valid and output) with its own period
// table.  A pass is a 1-D grid over every job's (combo group, chunk) blocks,
// job j owning blocks [block0, block0 + groups * nch); a fold is a 1-D grid
// over every job's combos, job j owning [combo0, combo0 + nlanes).  A block
// finds its job by binary search over block0 (combo0), then runs exactly the
// block of that job's direct launch, so a job's outputs are bit-identical
// alone and packed.
struct rec_job {
    const float *x, *h, *l;
    const int32_t *periods;
    float *out;
    double *carry;           // this job's [combo][chunk][states]
    int T, nlanes, fv, chunk, nch, groups, block0, combo0;
};

struct rec_pack_args {
    const rec_job *jobs;
    const int32_t *flags;    // per job: 1 if a bar at or after first_valid is non-finite
    int n_jobs;
    int phase;
};

// The last job whose first block (combo) is <= flat; jobs are in order.
__device__ __forceinline__ int pack_find(const rec_pack_args &p, int flat, bool combos) {
    int lo = 0, hi = p.n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int first = combos ? p.jobs[mid].combo0 : p.jobs[mid].block0;
        if (first <= flat) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ batch_view pack_view(const rec_pack_args &p, int j) {
    const rec_job &jb = p.jobs[j];
    batch_view v;
    v.x = jb.x; v.h = jb.h; v.l = jb.l;
    v.periods = jb.periods;
    v.carry = jb.carry;
    v.alphas = nullptr;
    v.seeds = nullptr;
    v.T = jb.T;
    v.nlanes = jb.nlanes;
    v.chunk = jb.chunk;
    v.nch = jb.nch;
    v.phase = p.phase;
    v.fv = jb.fv;
    v.clean = p.flags[j] == 0;
    v.group = 0;
    v.k = 0;
    return v;
}

template <class E, int NIN>
__device__ __forceinline__ void packed_pass(const rec_pack_args &p) {
    const int flat = blockIdx.x;
    const int j = pack_find(p, flat, false);
    batch_view v = pack_view(p, j);
    const int local = flat - p.jobs[j].block0;
    const int groups = p.jobs[j].groups;
    v.group = local % groups;
    v.k = local / groups;
    drive_batch<E, NIN>(v, p.jobs[j].out);
}

template <class E>
__device__ __forceinline__ void packed_fold(const rec_pack_args &p) {
    const int flat = blockIdx.x;
    const int j = pack_find(p, flat, true);
    batch_view v = pack_view(p, j);
    v.group = flat - p.jobs[j].combo0;
    scan_combine<E>(v);
}

#define REC_PACKED(NAME, ENG, NIN)                                                           \
    extern "C" __global__ __launch_bounds__(64) void NAME##_packed_f32(rec_pack_args p) {    \
        packed_pass<ENG, NIN>(p);                                                            \
    }                                                                                        \
    extern "C" __global__ __launch_bounds__(64) void NAME##_packed_fold(rec_pack_args p) {   \
        packed_fold<ENG>(p);                                                                 \
    }
REC_PACKED(ema, EmaEng, 1)
REC_PACKED(rsi, RsiEng, 1)
REC_PACKED(atr, AtrEng, 3)
REC_PACKED(adx, AdxEng, 3)
REC_PACKED(smma, SmmaEng, 1)
REC_PACKED(wilders, WildersEng, 1)

// Per-job non-finite flags: grid (blocks, jobs); block (b, j) checks bars
// first_valid + b*256 + tid, stepping by gridDim.x * 256.  flags[j] must be 0
// on entry; every failing lane stores 1.
extern "C" __global__ __launch_bounds__(256) void packed_nonfinite_flag(const rec_job *jobs,
                                                                        int32_t *flags,
                                                                        int hlc) {
    const rec_job &jb = jobs[blockIdx.y];
    bool bad = false;
    for (long t = (long)jb.fv + (long)blockIdx.x * 256 + threadIdx.x; t < jb.T;
         t += (long)gridDim.x * 256) {
        bad = bad || !vta_isfinite(jb.x[t]);
        if (hlc) bad = bad || !vta_isfinite(jb.h[t]) || !vta_isfinite(jb.l[t]);
    }
    if (bad) flags[blockIdx.y] = 1;
}

extern "C" __global__ __launch_bounds__(256) void tr_from_hlc_f32(
    const float *__restrict__ h, const float *__restrict__ l, const float *__restrict__ c, int T,
    int first_valid, float *__restrict__ tr) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= T) return;
    float v = 0.0f;
    if (t == first_valid) v = (float)((double)h[t] - (double)l[t]);
    else if (t > first_valid) v = (float)true_range(h[t], l[t], c[t - 1]);
    tr[t] = v;
}

extern "C" __global__ __launch_bounds__(64) void ema_scan_combine(rec_args a) {
    scan_combine<EmaEng>(view_of(a));
}
extern "C" __global__ __launch_bounds__(64) void rsi_scan_combine(rec_args a) {
    scan_combine<RsiEng>(view_of(a));
}
extern "C" __global__ __launch_bounds__(64) void atr_scan_combine(rec_args a) {
    scan_combine<AtrEng>(view_of(a));
}
extern "C" __global__ __launch_bounds__(64) void smma_scan_combine(rec_args a) {
    scan_combine<SmmaEng>(view_of(a));
}
extern "C" __global__ __launch_bounds__(64) void wilders_scan_combine(rec_args a) {
    scan_combine<WildersEng>(view_of(a));
}

// Per-combo Wilders seed of a one-series batch (wave = combo): the value the
// lane reaches at bar fv + p - 1 by replaying its warm-up (seed_bar): groups
// of four bars g = ((x0 + x1) + x2) + x3, each group's own bars added in
// order, summed from 0.0 one group after another, times the smoothing
// factor.  Lanes form 64 groups (256 bars) at a time from one coalesced
// read; the wave then adds them in group order (each read out of its lane),
// so the sum is the same sequence of f64 operations as the replay.  Combos
// whose warm-up passes the series end get no seed (their lanes have no
// output).
extern "C" __global__ __launch_bounds__(256) void wilders_seed(rec_args a) {
    const int lane = threadIdx.x & 63;
    const int c = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (c >= a.nlanes) return;
    const int fv = a.fvs[0];
    WildersEng e;
    e.setup(__builtin_amdgcn_readfirstlane(a.periods[c]), fv);
    apply_alpha(e, a.alphas, c);
    const int p = e.p;
    if (p < 1 || (long)fv + p - 1 >= a.T) return;
    const float *x = a.x + fv;
    auto group = [&](int b) {
        const int k0 = b + 4 * lane;
        double g = 0.0;
        if (k0 < p) {
            g = (double)x[k0];
            if (k0 + 1 < p) g += (double)x[k0 + 1];
            if (k0 + 2 < p) g += (double)x[k0 + 2];
            if (k0 + 3 < p) g += (double)x[k0 + 3];
        }
        return g;
    };
    double sum = 0.0;
    double g = group(0);
    for (int b = 0; b < p; b += 256) {
        const double gn = b + 256 < p ? group(b + 256) : 0.0;   // next 256 bars in flight
        const int ng = min(64, (p - b + 3) >> 2);
        const unsigned long long gb = __builtin_bit_cast(unsigned long long, g);
        const int lo = (int)(unsigned)gb, hi = (int)(unsigned)(gb >> 32);
        for (int j = 0; j < ng; ++j) {
            const unsigned long long v =
                ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, j) << 32) |
                (unsigned)__builtin_amdgcn_readlane(lo, j);
            sum += __builtin_bit_cast(double, v);
        }
        g = gn;
    }
    if (lane == 0) a.seeds[c] = sum * e.al;
}

// ---------------------------------------------------------------- preparation
//
// Per time-major series: first_valid and the clean flag (every input finite
// at every bar from first_valid on).  One lane per series walks its column;
// each row is one coalesced load across lanes.
//   mode 0: x only (EMA, RSI)
//   mode 1: first bar with high, low and x all non-NaN (ATR, atr.rs:197-207)
//   mode 2: max of the per-input first non-NaN bars (ADX, adx.rs:190-198)
//   mode 3: fv = first non-NaN x, fv2 = mode-1 bar, clean over all inputs
//           from fv on (fused EMA/RSI/ATR)
//   mode 4: fv = first non-NaN x; clean[s] receives the first bar at or after
//           fv whose x is not finite (T if none), for the windowed averages
extern "C" __global__ __launch_bounds__(64) void ms_prepare_meta(
    const float *__restrict__ x, const float *__restrict__ h, const float *__restrict__ l,
    int T, int N, int mode, int32_t *__restrict__ fvs, int32_t *__restrict__ fvs2,
    int32_t *__restrict__ clean) {
    const int s = blockIdx.x * 64 + (threadIdx.x & 63);
    if (s >= N) return;
    const bool hl = mode == 1 || mode == 2 || mode == 3;
    int fx = T, fh = T, fl = T, f3 = T;
    for (int t = 0; t < T; ++t) {
        const size_t off = (size_t)t * N + s;
        const bool nx = !isnan(x[off]);
        const bool nh = hl ? !isnan(h[off]) : true;
        const bool nl = hl ? !isnan(l[off]) : true;
        if (fx == T && nx) fx = t;
        if (fh == T && nh) fh = t;
        if (fl == T && nl) fl = t;
        if (f3 == T && nx && nh && nl) f3 = t;
        if (fx < T && f3 < T && fh < T && fl < T) break;
    }
    int f = mode == 0 ? fx : mode == 1 ? f3 : mode == 2 ? max(fx, max(fh, fl)) : fx;
    int zf = T;
    for (int t = f; t < T; ++t) {
        const size_t off = (size_t)t * N + s;
        const bool ok = vta_isfinite(x[off]) &&
                        (!hl || (vta_isfinite(h[off]) && vta_isfinite(l[off])));
        if (!ok) {
            zf = t;
            break;
        }
    }
    fvs[s] = f;
    if (fvs2) fvs2[s] = f3;
    clean[s] = mode == 4 ? zf : (zf == T ? 1 : 0);
}

// Per series s of a time-major [T][N] matrix, from the caller's first_valid
// (clamped into [0, T]): clean[s] = 1 when every bar from it on is finite
// (the recurrences' device forms), or, with want_index, the first non-finite
// bar (T if none; the finite-window averages' device forms).
__device__ __forceinline__ void ms_scan_from_fv(const float *__restrict__ x, int T, int N,
                                                const int32_t *__restrict__ fvs,
                                                int32_t *__restrict__ out, bool want_index) {
    const int s = blockIdx.x * 64 + (threadIdx.x & 63);
    if (s >= N) return;
    const int f = min(max((int)fvs[s], 0), T);
    int zf = T;
    for (int t = f; t < T; ++t) {
        if (!vta_isfinite(x[(size_t)t * N + s])) {
            zf = t;
            break;
        }
    }
    out[s] = want_index ? zf : (zf == T ? 1 : 0);
}
extern "C" __global__ __launch_bounds__(64) void ms_clean_from_first_valids(
    const float *__restrict__ x, int T, int N, const int32_t *__restrict__ fvs,
    int32_t *__restrict__ clean) {
    ms_scan_from_fv(x, T, N, fvs, clean, false);
}
extern "C" __global__ __launch_bounds__(64) void ms_first_nonfinite_from_first_valids(
    const float *__restrict__ x, int T, int N, const int32_t *__restrict__ fvs,
    int32_t *__restrict__ z) {
    ms_scan_from_fv(x, T, N, fvs, z, true);
}

// Single series: clean flag = no non-finite input in [first_valid, T).
// The one launch before a device-resident one-series batch: meta[0] =
// first_valid (the batch kernels' fvs[0]); meta[2] = epoch when a bar of x
// (and h, l) at or after first_valid is not finite, so the batch reads its
// clean flag as meta[2] != epoch and the word needs no clearing (each call
// brings a new epoch); and, with warms, vta_check_warms' check of the
// reference's warm indices against first_valid + period - 1 (a mismatch
// stores 1 in *err).  Grid: enough threads for max(T - first_valid, n).
extern "C" __global__ __launch_bounds__(256) void batch_prepare(
    const float *__restrict__ x, const float *__restrict__ h, const float *__restrict__ l,
    int T, int first_valid, int epoch, int32_t *__restrict__ meta,
    const int32_t *__restrict__ periods, const int32_t *__restrict__ warms, int n,
    int32_t *__restrict__ err) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) meta[0] = first_valid;
    const int t = i + first_valid;
    if (t < T) {
        bool bad = !vta_isfinite(x[t]);
        if (h) bad = bad || !vta_isfinite(h[t]);
        if (l) bad = bad || !vta_isfinite(l[t]);
        if (bad) meta[2] = epoch;
    }
    if (warms && i < n && (long long)warms[i] != (long long)first_valid + periods[i] - 1)
        err[0] = 1;
}

// flag[0] must be 0 on entry (see vta_fill_i32); every failing lane stores 1.
extern "C" __global__ __launch_bounds__(256) void series_nonfinite_flag(
    const float *__restrict__ x, const float *__restrict__ h, const float *__restrict__ l,
    int T, int first_valid, int32_t *__restrict__ flag) {
    const int t = blockIdx.x * 256 + threadIdx.x + first_valid;
    if (t >= T) return;
    bool bad = !vta_isfinite(x[t]);
    if (h) bad = bad || !vta_isfinite(h[t]);
    if (l) bad = bad || !vta_isfinite(l[t]);
    if (bad) flag[0] = 1;
}

// clean[0] = (flag[0] == 0): turns the non-finite flag into the engines' clean word.
extern "C" __global__ __launch_bounds__(64) void flag_to_clean(const int32_t *__restrict__ flag,
                                                               int32_t *__restrict__ clean) {
    if (threadIdx.x == 0) clean[0] = flag[0] == 0 ? 1 : 0;
}

// The reference's device forms of SMMA, SWMA and Wilders take per-combo warm
// indices that its host code sets to first_valid + period - 1; the kernels
// here derive the warm-up from the period, so the indices are checked: a
// mismatch stores 1 in *err (the context reports it at its next synchronize).
// first_valid < 0 derives first_valid from combo 0 (SWMA passes none) and
// stores it in *fv_out; a derived value outside [0, T - max_period] is an
// error too, and T (no output bar) is stored instead.
extern "C" __global__ __launch_bounds__(256) void vta_check_warms(
    const int32_t *__restrict__ periods, const int32_t *__restrict__ warms, int n, int T,
    int first_valid, int max_period, int32_t *__restrict__ fv_out, int32_t *__restrict__ err) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    long long fv = first_valid;
    bool bad = false;
    if (first_valid < 0) {
        fv = (long long)warms[0] - (long long)periods[0] + 1;
        const bool ok = fv >= 0 && fv <= (long long)T - (long long)max_period;
        if (c == 0 && fv_out) *fv_out = ok ? (int32_t)fv : T;
        bad = !ok;
    }
    if (c < n && (long long)warms[c] != fv + (long long)periods[c] - 1) bad = true;
    if (bad) err[0] = 1;
}

extern "C" __global__ __launch_bounds__(256) void vta_fill_i32(int32_t *__restrict__ p, int n,
                                                                int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
