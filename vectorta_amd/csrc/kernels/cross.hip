// Double-ALMA crossover backtest reduction for gfx950.
//
// For every (fast, slow) pair of ALMA parameter sets the position is
// pos_t = sign(ALMA_fast[t] - ALMA_slow[t]) from the first bar w where both
// are valid, and the per-pair statistics are
//   trades   = #{t > w : pos_t != pos_{t-1}}
//   n_long   = #{t >= w : pos_t = +1},  n_short = #{t >= w : pos_t = -1}
//   checksum = sum_t pos_t * (t + 1)                  (exact, int64)
//   pnl      = sum_{t <= T-2} pos_t * (x_{t+1} - x_t) (f64)
// The 58,300 x 200,000 position matrix is never materialised.
//
// The rows of the distinct ALMA parameter sets are computed here in f64 with
// the oracle's exact operation order (cross_rows_f64: 4-wide grouped dot,
// then * inv_norm, no contraction, weights from the host with glibc exp), so
// they are bit-identical to the f64 rows and every sign -- exact ties
// included -- equals the oracle's without any guard band.

// Positions change rarely, so all statistics are accumulated per run of
// constant position when the run closes: the per-bar work is one subtract and
// one compare.  A run [a, b) adds (b - a) to n_long or n_short, pos * (sum of
// t + 1 over the run) to the checksum and pos * (x[min(b, T-1)] - x[a]) to the
// pnl; the differences of f32 prices are exact in f64 and so are their sums
// (all are multiples of the smallest price's ulp), which makes the pnl equal
// to the bar-by-bar f64 sum.
//
// Work split: a workgroup takes 16 fast x 16 slow rows (256 pairs, one per
// lane) over one time segment; 128-bar chunks of the 32 f64 rows are staged in
// LDS (row stride 129 doubles: the 16 rows a wave reads at one bar fall on
// distinct bank pairs).  Each segment looks back one bar for the previous position, so the
// segments are independent; cross_reduce sums them in segment order.
#include "vta_common.h"

struct cross_args {
    double *rows;               // [nrows][T] f64 ALMA rows
    const float *prices;        // [T]
    const double *w64;          // [nrows][pmax] f64 batch weights, not normalised
    const double *inv64;        // [nrows] 1 / sum(w)
    const int32_t *row_period;  // [nrows]
    const int32_t *fast_row;    // [nfast] -> row
    const int32_t *slow_row;    // [nslow] -> row
    long long *part_i;          // [nseg][npairs][4]
    double *part_pnl;           // [nseg][npairs]
    long long *out_i;           // [npairs][4]
    double *out_pnl;            // [npairs]
    int T;
    int fv;
    int nfast;
    int nslow;
    int nseg;
    int seg_len;
    int pmax;
    int ntiles_s;
    int row_tile;               // bars per cross_rows_f64 workgroup
    int pad_;
};

#define CR_CH 64
#define CR_LD (CR_CH + 1)
#define CR_RPP (256 / CR_CH)
#define RW_NB 8               // consecutive bars per lane of the row kernel
#define RW_T (256 * RW_NB)    // bars per row workgroup
#define RW_PMAX 296

// f64 ALMA rows in the oracle's order (alma.rs:1505-1527): for output t,
// sum over 4-tap groups g = (x0 w0 + x1 w1) + x2 w2 + x3 w3, sum += g, then
// the tail taps one at a time, out = sum * inv_norm.  NaN before fv + p - 1.
// grid.x = bar tiles of RW_T, grid.y = rows; lane l computes bars
// t0 + RW_NB l .. t0 + RW_NB (l + 1) - 1.
extern "C" __global__ __launch_bounds__(256) void cross_rows_f64(cross_args a) {
#pragma clang fp contract(off)
    __shared__ double xs[RW_T + RW_PMAX];
    __shared__ double ws[RW_PMAX];
    const int tid = threadIdx.x;
    const int r = blockIdx.y;
    const int p = a.row_period[r];
    const int T = a.T;
    const int t0 = blockIdx.x * RW_T;
    const int xlo = t0 - p + 1;
    for (int i = tid; i < RW_T + p - 1; i += 256) {
        const int idx = xlo + i;
        xs[i] = (idx >= 0 && idx < T) ? (double)a.prices[idx] : 0.0;
    }
    for (int i = tid; i < p; i += 256) ws[i] = a.w64[(size_t)r * a.pmax + i];
    __syncthreads();
    const double inv = a.inv64[r];
    const int p4 = p & ~3;
    const int warm = a.fv + p - 1;
    double *row = a.rows + (size_t)r * T;
    // lane = RW_NB consecutive bars tb .. tb + RW_NB - 1: a 4-tap group of
    // all of them reads the RW_NB + 3 prices x[tb + k ..] once and 4 weights,
    // each output keeping its own operation order
    const int tb = RW_NB * tid;
    const double *x = xs + tb;              // x[i + k] = price[t0 + tb + i - p + 1 + k]
    double acc[RW_NB];
#pragma unroll
    for (int i = 0; i < RW_NB; ++i) acc[i] = 0.0;
    for (int k = 0; k < p4; k += 4) {
        double v[RW_NB + 3];
#pragma unroll
        for (int j = 0; j < RW_NB + 3; ++j) v[j] = x[k + j];
        const double w0 = ws[k], w1 = ws[k + 1], w2 = ws[k + 2], w3 = ws[k + 3];
#pragma unroll
        for (int i = 0; i < RW_NB; ++i) {
            double g = v[i] * w0 + v[i + 1] * w1;
            g = g + v[i + 2] * w2;
            g = g + v[i + 3] * w3;
            acc[i] += g;
        }
    }
    for (int k = p4; k < p; ++k) {
        const double w = ws[k];
#pragma unroll
        for (int i = 0; i < RW_NB; ++i) acc[i] += x[k + i] * w;
    }
#pragma unroll
    for (int i = 0; i < RW_NB; ++i) {
        const int t = t0 + tb + i;
        if (t < T) row[t] = t >= warm ? acc[i] * inv : __builtin_bit_cast(double, 0x7ff8000000000000ull);
    }
}

struct CrossRun {
    int tr = 0, nl = 0, ns = 0;                  // per segment: < 2^31
    long long ck = 0;
    double pnl = 0.0;
    // run [r0, r1) at position p (r1 >= r0); xa = x[r0], xe = x[min(r1, T - 1)],
    // t0 = tri(r0), t1 = tri(r1).  Branch-free: p = 0 or an empty run adds
    // zeros (an empty run has xe = xa).
    __device__ void close(int p, int r0, int r1, long long t0, long long t1, float xa, float xe) {
        const int n = r1 - r0;
        nl += p > 0 ? n : 0;
        ns += p < 0 ? n : 0;
        const long long s = t1 - t0;             // sum of t + 1 over [r0, r1)
        ck += p > 0 ? s : (p < 0 ? -s : 0LL);
        pnl += (double)p * ((double)xe - (double)xa);
    }
};

// r (r + 1) / 2: the sum of t + 1 over [0, r)
__device__ __forceinline__ long long tri(int r) { return ((long long)r * ((long long)r + 1)) >> 1; }

extern "C" __global__ __launch_bounds__(256) void cross_pairs_f64(cross_args a) {
    __shared__ double lf[16 * CR_LD];
    __shared__ double ls[16 * CR_LD];
    __shared__ float lx[CR_CH];                  // the chunk's prices
    __shared__ int rowid[32];
    const int tid = threadIdx.x;
    const int tf = blockIdx.x / a.ntiles_s;
    const int ts = blockIdx.x % a.ntiles_s;
    const int fi = tid >> 4, sj = tid & 15;
    const int f = tf * 16 + fi, s = ts * 16 + sj;
    const bool active = f < a.nfast && s < a.nslow;
    const int T = a.T;
    if (tid < 32) {
        rowid[tid] = tid < 16 ? a.fast_row[min(tf * 16 + tid, a.nfast - 1)]
                              : a.slow_row[min(ts * 16 + tid - 16, a.nslow - 1)];
    }
    const int rf = a.fast_row[min(f, a.nfast - 1)];
    const int rs = a.slow_row[min(s, a.nslow - 1)];
    const int w = a.fv + max(a.row_period[rf], a.row_period[rs]) - 1;
    const int seg0 = blockIdx.y * a.seg_len;
    const int seg1 = min(T, seg0 + a.seg_len);
    const int b0 = max(seg0, w);                 // first bar counted by this segment
    const int tstart = b0 > w ? b0 - 1 : b0;     // one bar of look-back for the position
    const bool walk = active && rf != rs && b0 < seg1;   // identical rows: always 0
    CrossRun acc;
    // the position at tstart, read from the rows directly, opens the first run
    int prev = 0, run_a = b0;
    double prevd = 0.0;                          // prev as a double
    float xa = 0.0f;                             // price at run_a
    long long ta = 0;                            // tri(run_a)
    if (walk) {
        const double d = a.rows[(size_t)rf * T + tstart] - a.rows[(size_t)rs * T + tstart];
        prev = d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
        prevd = (double)prev;
        xa = a.prices[b0];
        ta = tri(b0);
    }
    __syncthreads();
    for (int c0 = max(seg0 - 1, 0); c0 < seg1; c0 += CR_CH) {
        // stage the tile's 32 rows over [c0, c0 + CR_CH), two rows per pass
        const int rh = tid / CR_CH;              // CR_RPP rows per staging pass
        const int tc = tid % CR_CH;
        const int t = min(c0 + tc, T - 1);
        if (rh == 0) lx[tc] = a.prices[t];
#pragma unroll 4
        for (int r2 = 0; r2 < 32; r2 += CR_RPP) {
            const int r = r2 + rh;
            const double v = a.rows[(size_t)rowid[r] * T + t];
            if (r < 16) lf[r * CR_LD + tc] = v;
            else ls[(r - 16) * CR_LD + tc] = v;
        }
        __syncthreads();
        if (walk) {
            const int lo = max(tstart + 1 - c0, 0);
            const int hi = min(seg1 - c0, CR_CH);
            const double *pf = lf + fi * CR_LD;
            const double *ps = ls + sj * CR_LD;
            auto bar = [&](int tl, double d) {   // d = the oracle's f_t - s_t
                if (prevd * d > 0.0) return;   // position unchanged
                const int pos = d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
                if (pos != prev) {
                    // tt < seg1 <= T: its price is this chunk's lx[tl]
                    const int tt = c0 + tl;
                    const float xt = lx[tl];
                    const long long tt1 = tri(tt);
                    acc.close(prev, run_a, tt, ta, tt1, xa, xt);
                    acc.tr += 1;
                    prev = pos;
                    prevd = (double)pos;
                    run_a = tt;
                    xa = xt;
                    ta = tt1;
                }
            };
            // four bars' differences are read (one LDS wait) before their
            // sequential position logic; reads past hi are clamped into the chunk
            for (int tl0 = lo; tl0 < hi; tl0 += 4) {
                double d[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int tl = min(tl0 + k, hi - 1);
                    d[k] = pf[tl] - ps[tl];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (tl0 + k < hi) bar(tl0 + k, d[k]);
            }
        }
        __syncthreads();
    }
    if (walk) acc.close(prev, run_a, seg1, ta, tri(seg1), xa, a.prices[min(seg1, T - 1)]);
    if (active) {
        const size_t pair = (size_t)f * a.nslow + s;
        const size_t npairs = (size_t)a.nfast * a.nslow;
        const size_t o = (size_t)blockIdx.y * npairs + pair;
        a.part_i[o * 4 + 0] = acc.tr;
        a.part_i[o * 4 + 1] = acc.nl;
        a.part_i[o * 4 + 2] = acc.ns;
        a.part_i[o * 4 + 3] = acc.ck;
        a.part_pnl[o] = acc.pnl;
    }
}

// Sum of the segment partials in segment order: one lane per pair.
extern "C" __global__ __launch_bounds__(256) void cross_reduce(cross_args a) {
    const size_t npairs = (size_t)a.nfast * a.nslow;
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npairs) return;
    long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    double pn = 0.0;
    for (int g = 0; g < a.nseg; ++g) {
        const size_t o = (size_t)g * npairs + p;
        s0 += a.part_i[o * 4 + 0];
        s1 += a.part_i[o * 4 + 1];
        s2 += a.part_i[o * 4 + 2];
        s3 += a.part_i[o * 4 + 3];
        pn += a.part_pnl[o];
    }
    a.out_i[p * 4 + 0] = s0;
    a.out_i[p * 4 + 1] = s1;
    a.out_i[p * 4 + 2] = s2;
    a.out_i[p * 4 + 3] = s3;
    a.out_pnl[p] = pn;
}
