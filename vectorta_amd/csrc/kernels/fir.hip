// Finite-window moving averages for gfx950: SMA, WMA, SWMA.
//
// One series x many periods (fir_batch_f32).  A workgroup owns a tile of
// consecutive bars [t0, t0 + tile) and a group of FIR_GC combos.  It stages
// the tile plus the group's look-back once and turns it into two local f64
// prefix sums in LDS (block scan: wave scan by shuffles, wave totals through
// LDS), with the local index i of bar s0 + i:
//   A[i+1] = sum_{i' <= i} x[s0 + i'],   B[i+1] = sum_{i' <= i} i' * x[s0 + i']
// (bars before first_valid count as 0, like the reference's batch prefixes).
// Every output is then O(1) differences of prefixes (Lt = A index of bar t):
//   SMA_p  = (A[Lt] - A[Lt-p]) / p
//   WMA_p  = ((B[Lt] - B[Lt-p]) - q (A[Lt] - A[Lt-p])) / (p(p+1)/2),
//            q = local index of bar t-p (weights 1..p, newest p)
//   SWMA_p = the triangle as two boxes of lengths a and b (a + b = p + 1):
//            sum_{i<b} box_a(t-i) = (C[Lt] - C[Lt-b]) - (C[Lt-a] - C[Lt-a-b]),
//            C[k] = k A[k] - B[k] (the prefix of the prefix), times 1/(a b).
// Local prefixes keep the magnitudes at (tile + period) bars instead of the
// whole series, so the differences lose ~eps * (tile + p)^k / p^k, far below
// f32 rounding; the kernel is bound by the coalesced row stores.
//
// Many series x one period, time-major (fir_ms1p_f32): one lane per series,
// one wave per (64 series, time chunk).  Each chunk replays the reference's
// scalar running sums exactly from p bars before the chunk (the windows have
// finite memory, so the state at the chunk start is the exact one), reading
// the lagged bars again through the cache.
//
// Non-finite input after first_valid (z = first such bar), as the reference's
// CPU paths give it:
//   SMA   (sma.rs:302-330, batch prefix :909-1021): window sums up to z+p-1
//         (inf or NaN), NaN from z+p on.  p = 1 in the scalar form copies.
//   WMA   batch prefix (wma.rs:656-760): NaN from z on; scalar (many-series,
//         wma.rs:290-320): the window sum at z, NaN after.
//   SWMA  (swma.rs:406-476, batch row :1155-1217): the running sums at
//         [z, z+a), NaN from z+a on; p = 1 copies and p = 2 averages pairs.
#include "vta_common.h"

struct fir_args {
    const float *x;           // batch: [T]; many-series: [T][N] time-major
    const int32_t *periods;   // batch: per combo
    const int32_t *meta;      // batch: [0] first_valid, [1] first non-finite bar (T if none)
    const int32_t *fvs;       // many-series: per series first_valid
    const int32_t *zs;        // many-series: per series first non-finite bar, or null
    float *out;
    int T;
    int nlanes;               // combos (batch) or series (many-series)
    int kind;                 // 0 SMA, 1 WMA, 2 SWMA
    int tile;                 // batch: bars per workgroup
    int span_cap;             // batch: LDS doubles per prefix array, minus one
    int chunk;                // many-series: bars per chunk
    int period;               // many-series
    int pad_;
};

#define FIR_GC 16
#define FIR_NAN_F __builtin_bit_cast(float, 0x7fc00000u)

__device__ __forceinline__ void swma_ab(int p, int *a, int *b) {
    if (p & 1) {
        *a = (p + 1) >> 1;
        *b = *a;
    } else {
        *a = p >> 1;
        *b = *a + 1;
    }
}

extern "C" __global__ __launch_bounds__(256) void fir_batch_f32(fir_args a) {
    extern __shared__ double lds[];
    __shared__ double wtot[2][4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int T = a.T;
    const int t0 = blockIdx.x * a.tile;
    const int c0 = blockIdx.y * FIR_GC;
    const int nc = min(FIR_GC, a.nlanes - c0);
    const int fv = a.meta[0];
    const int z = a.meta[1];
    int pg = 1;
    bool bad = false;
    for (int j = 0; j < nc; ++j) {
        const int p = a.periods[c0 + j];
        pg = max(pg, p);
        bad = bad || p < 1;
    }
    const int tend = min(t0 + a.tile, T);
    if (bad || pg > a.span_cap - a.tile - 1) {
        // a period outside [1, max_period] (device-side input of the enqueue
        // form): the window would not fit the LDS, so the rows stay NaN
        for (int j = 0; j < nc; ++j)
            for (int t = t0 + tid; t < tend; t += 256) a.out[(size_t)(c0 + j) * T + t] = FIR_NAN_F;
        return;
    }
    const int s0 = t0 - pg - 1;            // first staged bar (may be negative)
    const int span = tend - s0;            // staged bars [s0, tend), <= span_cap
    double *A = lds;
    double *B = lds + a.span_cap + 1;
    const bool wantB = a.kind != 0;

    // ---- local prefix sums A, B over the staged bars
    double carA = 0.0, carB = 0.0;
    if (tid == 0) {
        A[0] = 0.0;
        if (wantB) B[0] = 0.0;
    }
    for (int base = 0; base < span; base += 256) {
        const int i = base + tid;
        const int t = s0 + i;
        double xv = 0.0;
        if (i < span && t >= fv && t < T) xv = (double)a.x[t];
        double sa = xv;
        double sb = (double)i * xv;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double ya = __shfl_up(sa, o);
            const double yb = __shfl_up(sb, o);
            if (lane >= o) {
                sa += ya;
                sb += yb;
            }
        }
        if (lane == 63) {
            wtot[0][w] = sa;
            wtot[1][w] = sb;
        }
        __syncthreads();
        double oa = carA, ob = carB;
        for (int k = 0; k < w; ++k) {
            oa += wtot[0][k];
            ob += wtot[1][k];
        }
        if (i < span) {
            A[i + 1] = oa + sa;
            if (wantB) B[i + 1] = ob + sb;
        }
        for (int k = 0; k < 4; ++k) {
            carA += wtot[0][k];
            carB += wtot[1][k];
        }
        __syncthreads();
    }

    // ---- outputs, one coalesced row segment per combo
    for (int j = 0; j < nc; ++j) {
        const int c = c0 + j;
        const int p = a.periods[c];
        float *row = a.out + (size_t)c * T;
        if (a.kind == 0) {
            const int warm = fv + p - 1;
            const double inv = 1.0 / (double)p;
            for (int t = t0 + tid; t < tend; t += 256) {
                float v = FIR_NAN_F;
                if (t >= warm && t - p < z) {
                    const int Lt = t - s0 + 1;
                    v = (float)((A[Lt] - A[Lt - p]) * inv);
                }
                row[t] = v;
            }
        } else if (a.kind == 1) {
            const int warm = fv + p - 1;
            const double inv = 1.0 / ((double)p * ((double)p + 1.0) * 0.5);
            for (int t = t0 + tid; t < tend; t += 256) {
                float v = FIR_NAN_F;
                if (t >= warm && t < z) {
                    const int Lt = t - s0 + 1;
                    const int Lq = Lt - p;
                    const double sa = A[Lt] - A[Lq];
                    const double sb = B[Lt] - B[Lq];
                    v = (float)((sb - (double)(Lq - 1) * sa) * inv);
                }
                row[t] = v;
            }
        } else {
            if (p <= 2) {
                // window-only forms (swma.rs:429-445): copy, or pair average
                const int warm = fv + p - 1;
                for (int t = t0 + tid; t < tend; t += 256) {
                    float v = FIR_NAN_F;
                    if (t >= warm) {
                        if (p == 1) v = a.x[t];
                        else v = (float)(((double)a.x[t - 1] + (double)a.x[t]) * 0.5);
                    }
                    row[t] = v;
                }
                continue;
            }
            int sa_, sb_;
            swma_ab(p, &sa_, &sb_);
            const int wa = sa_, wb = sb_;
            const int warm = fv + p - 1;
            const double inv = 1.0 / ((double)wa * (double)wb);
            for (int t = t0 + tid; t < tend; t += 256) {
                float v = FIR_NAN_F;
                if (t >= warm && t - wa < z) {
                    const int Lt = t - s0 + 1;
                    if (t >= z) {
                        // running sums hold the sum of the non-finite bars in [z, t]
                        v = (float)(A[Lt] * inv);
                    } else {
                        const double c_t = (double)Lt * A[Lt] - B[Lt];
                        const int k1 = Lt - wb, k2 = Lt - wa, k3 = Lt - wa - wb;
                        const double c_1 = (double)k1 * A[k1] - B[k1];
                        const double c_2 = (double)k2 * A[k2] - B[k2];
                        const double c_3 = (double)k3 * A[k3] - B[k3];
                        v = (float)(((c_t - c_1) - (c_2 - c_3)) * inv);
                    }
                }
                row[t] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- many series
//
// grid.x: groups of 64 series, grid.y: chunks of a.chunk bars.  Lane = series;
// each load/store is one coalesced row segment of the [T][N] matrix.
extern "C" __global__ __launch_bounds__(64) void fir_ms1p_f32(fir_args a) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int T = a.T;
    const int p = a.period;
    const int sc = min(s, N - 1);
    const int c0 = blockIdx.y * a.chunk;
    const int c1 = min(c0 + a.chunk, T);
    const int fv = active ? min(max(a.fvs[s], 0), T) : T;
    const int z = active ? (a.zs ? a.zs[s] : T) : T;
    const float *xc = a.x + sc;
    float *oc = a.out + sc;
    const int warm = fv + p - 1;
    // replay start: p bars before the chunk (never before first_valid)
    const int r0 = active ? max(fv, c0 - p + 1) : T;
    int tb = r0;
    for (int o = 32; o > 0; o >>= 1) tb = min(tb, __shfl_xor(tb, o));
    tb = __builtin_amdgcn_readfirstlane(tb);
    tb = min(tb, c0);

    if (a.kind == 0 || (a.kind == 2 && p <= 2)) {
        const bool copy = p == 1;
        const bool pair = a.kind == 2 && p == 2;
        const double inv = 1.0 / (double)p;
        double sum = 0.0;
        for (int t = tb; t < c1; ++t) {
            const size_t off = (size_t)t * N;
            const double xv = (double)xc[off];
            float v = FIR_NAN_F;
            if (copy) {
                v = (float)xv;
            } else if (pair) {
                if (t >= 1) v = (float)(((double)xc[off - N] + xv) * 0.5);
            } else if (t >= r0) {
                // sma.rs:318-328: the first p bars summed, then sum += x[t] - x[t-p]
                if (t < r0 + p) sum += xv;
                else sum += xv - (double)xc[off - (size_t)p * N];
                v = (float)(sum * inv);
                if (t - p >= z) v = FIR_NAN_F;
            }
            if (active && t >= c0) oc[off] = t >= warm ? v : FIR_NAN_F;
        }
    } else if (a.kind == 1) {
        // wma.rs:290-320: weights 1..p, running weighted sum and plain sum
        const double pf = (double)p;
        const double inv = 1.0 / (pf * (pf + 1.0) * 0.5);
        double s1 = 0.0, s2 = 0.0;
        for (int t = tb; t < c1; ++t) {
            const size_t off = (size_t)t * N;
            const double xv = (double)xc[off];
            float v = FIR_NAN_F;
            if (t >= r0) {
                const int k = t - r0;
                if (k < p - 1) {
                    s2 += xv * ((double)k + 1.0);
                    s1 += xv;
                } else {
                    s2 += xv * pf;
                    s1 += xv;
                    v = (float)(s2 * inv);
                    s2 -= s1;
                    s1 -= (double)xc[off - (size_t)(p - 1) * N];
                }
                if (t > z) v = FIR_NAN_F;
            }
            if (active && t >= c0) oc[off] = t >= warm ? v : FIR_NAN_F;
        }
    } else {
        // swma.rs:406-476 with the ring of the last b box sums replaced by a
        // second running box sum lagged by b bars (same operations, same bits)
        int wa, wb;
        swma_ab(p, &wa, &wb);
        const double inv = 1.0 / ((double)wa * (double)wb);
        double s1 = 0.0, s1b = 0.0, s2 = 0.0;
        for (int t = tb; t < c1; ++t) {
            const size_t off = (size_t)t * N;
            float v = FIR_NAN_F;
            if (t >= r0) {
                const int sa0 = r0 + wa - 1;         // first full box
                s1 += (double)xc[off];
                double old = 0.0;
                const int tl = t - wb;                // bar whose box leaves the sum
                if (tl >= r0) s1b += (double)xc[off - (size_t)wb * N];
                if (tl >= sa0) old = s1b;
                if (t >= sa0) {
                    s2 = s2 + (s1 - old);
                    v = (float)(s2 * inv);
                    s1 -= (double)xc[off - (size_t)(wa - 1) * N];
                }
                if (tl >= sa0) s1b -= (double)xc[off - (size_t)(wb + wa - 1) * N];
                if (t - wa >= z) v = FIR_NAN_F;
            }
            if (active && t >= c0) oc[off] = t >= warm ? v : FIR_NAN_F;
        }
    }
}

// Single series: the first bar at or after first_valid whose value is not
// finite.  z[0] must hold T on entry (vta_fill_i32).
extern "C" __global__ __launch_bounds__(256) void series_first_nonfinite(
    const float *__restrict__ x, int T, int first_valid, int32_t *__restrict__ z) {
    const int t = blockIdx.x * 256 + threadIdx.x + first_valid;
    if (t >= T) return;
    if (!vta_isfinite(x[t])) atomicMin(z, t);
}
