// Finite-window moving averages for gfx950: SMA, WMA, SWMA.
//
// One series x many periods (fir_batch_f32).  A workgroup owns a tile of
// consecutive bars [t0, t0 + tile) and a group of consecutive combos.  It stages
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
// Many series x one period, time-major (fir_ms1p_{sma,wma,swma}_f32): one lane per series,
// one wave per (64 series, time chunk).  Each chunk replays the reference's
// scalar running sums exactly from p bars before the chunk (the windows have
// finite memory, so the state at the chunk start is the exact one).  The
// lagged bars (at most p back) come either from a per-wave LDS ring of the
// last ring rows (each lane reads only its own column, so no barrier) or,
// with no ring, again from global memory through the cache.
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
    int group;                // batch: combos per workgroup (<= FIR_GMAX)
    int ring;                 // many-series: rows of the LDS ring (>= longest lag + 8), 0 = none
    int pad_;
};

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

// Combos per workgroup at most (the host's group size).
#define FIR_GMAX 256

extern "C" __global__ __launch_bounds__(256) void fir_batch_f32(fir_args a) {
    extern __shared__ double lds[];
    __shared__ double wtot[2][4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int T = a.T;
    const int t0 = blockIdx.x * a.tile;
    const int c0 = blockIdx.y * a.group;
    const int nc = min(a.group, a.nlanes - c0);
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
    // per-combo constants of the group (period, 1/normaliser, SWMA boxes)
    __shared__ int g_p[FIR_GMAX], g_a[FIR_GMAX], g_b[FIR_GMAX];
    __shared__ double g_inv[FIR_GMAX];
    if (tid < nc) {
        const int p = a.periods[c0 + tid];
        g_p[tid] = p;
        int wa = 1, wb = 1;
        swma_ab(p, &wa, &wb);
        g_a[tid] = wa;
        g_b[tid] = wb;
        g_inv[tid] = a.kind == 0 ? 1.0 / (double)p
                     : a.kind == 1 ? 1.0 / ((double)p * ((double)p + 1.0) * 0.5)
                                   : 1.0 / ((double)wa * (double)wb);
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

    // SWMA reads only C[k] = k A[k] - B[k] (the prefix of the prefix) and A:
    // fold C into the B array once per tile
    if (a.kind == 2) {
        for (int k = tid; k <= span; k += 256) B[k] = (double)k * A[k] - B[k];
        __syncthreads();
    }
    const double *C = B;

    // ---- outputs: lane l owns bars t0 + l + 256 m (m < 4); the prefix values
    // at its bars are read once into registers and shared by every combo of
    // the group, so a combo costs one LDS read (SMA), two (WMA) or three
    // (SWMA) per output plus its constants once per four bars.  Each store
    // covers 64 consecutive bars of one row.
    int Lt[4];
    bool in[4];
    double At[4], Pt[4], Xt[4], Xp[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int t = t0 + tid + 256 * m;
        in[m] = t < tend;
        Lt[m] = in[m] ? t - s0 + 1 : span;
        At[m] = A[Lt[m]];
        Pt[m] = a.kind == 0 ? 0.0 : B[Lt[m]];          // B (WMA) or C (SWMA)
        Xt[m] = (a.kind == 2 && in[m]) ? (double)a.x[t] : 0.0;
        Xp[m] = (a.kind == 2 && in[m] && t >= 1) ? (double)a.x[t - 1] : 0.0;
    }
    for (int j = 0; j < nc; ++j) {
        const int p = g_p[j];
        const double inv = g_inv[j];
        const int warm = fv + p - 1;
        float *row = a.out + (size_t)(c0 + j) * T;
        if (a.kind == 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int t = t0 + tid + 256 * m;
                float v = FIR_NAN_F;
                if (t >= warm && t - p < z) v = (float)((At[m] - A[Lt[m] - p]) * inv);
                if (in[m]) row[t] = v;
            }
        } else if (a.kind == 1) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int t = t0 + tid + 256 * m;
                float v = FIR_NAN_F;
                if (t >= warm && t < z) {
                    const int Lq = Lt[m] - p;
                    const double sa = At[m] - A[Lq];
                    const double sb = Pt[m] - B[Lq];
                    v = (float)((sb - (double)(Lq - 1) * sa) * inv);
                }
                if (in[m]) row[t] = v;
            }
        } else if (p <= 2) {
            // window-only forms (swma.rs:429-445): copy, or pair average
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int t = t0 + tid + 256 * m;
                float v = FIR_NAN_F;
                if (t >= warm) v = p == 1 ? (float)Xt[m] : (float)((Xp[m] + Xt[m]) * 0.5);
                if (in[m]) row[t] = v;
            }
        } else {
            const int wa = g_a[j], wb = g_b[j];
            const int ab = wa + wb;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int t = t0 + tid + 256 * m;
                const int L = Lt[m];
                const double c1 = C[L - wb], c2 = C[L - wa], c3 = C[L - ab];
                float v = (float)(((Pt[m] - c1) - (c2 - c3)) * inv);
                // inside [z, z+a) the running sums hold the sum of the
                // non-finite bars in [z, t]; NaN from z+a on and in the warm-up
                if (t >= z) v = t - wa < z ? (float)(At[m] * inv) : FIR_NAN_F;
                if (t < warm) v = FIR_NAN_F;
                if (in[m]) row[t] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- many series
//
// grid.x: groups of 64 series, grid.y: chunks of a.chunk bars.  Lane = series;
// each load/store is one coalesced row segment of the [T][N] matrix.  Bars go
// in blocks of 8: the block's loads (the new bars and the lagged bars that
// leave the running sums, indices clamped into the matrix) are issued before
// its sequential updates.
// FMS_B = bars per block (loads in flight per wave).
template <int KIND, int FMS_B>
__device__ __forceinline__ void fir_ms1p(const fir_args &a) {
    extern __shared__ float ring[];       // [a.ring rows][64 lanes]
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * 64 + lane;
    const bool active = s < a.nlanes;
    const int N = a.nlanes;
    const int T = a.T;
    const int p = a.period;
    const int R = a.ring;
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
    auto ldf = [&](int t) -> float { return xc[(size_t)min(max(t, 0), T - 1) * N]; };
    auto put = [&](int t, float v) {
        if (active && t >= c0) oc[(size_t)t * N] = t >= warm ? v : FIR_NAN_F;
    };

    // Blocks of FMS_B bars: stage() loads bars [t, t+FMS_B) into cur and the
    // ring; lagged(o, lag) fills o[k] with bar t+k-lag.  Every lagged bar the
    // sums use lies at or after r0 >= tb (the replay loaded it) and at most
    // p <= R-FMS_B bars back, so the ring still holds it.
    int slot = 0;                         // ring slot of bar t
    int t = tb;
    float cur[FMS_B];
    auto stage = [&]() {
#pragma unroll
        for (int k = 0; k < FMS_B; ++k) cur[k] = ldf(t + k);
        if (R) {
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) {
                int sl = slot + k;
                sl -= sl >= R ? R : 0;
                ring[sl * 64 + lane] = cur[k];
            }
        }
    };
    auto lagged = [&](float (&o)[FMS_B], int lag) {
        if (R) {
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) {
                int sl = slot + k - lag;
                sl += sl < 0 ? R : 0;
                sl -= sl >= R ? R : 0;
                o[k] = ring[sl * 64 + lane];
            }
        } else {
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) o[k] = ldf(t + k - lag);
        }
    };
    auto next = [&]() {
        slot += FMS_B;
        slot -= slot >= R ? R : 0;
        t += FMS_B;
    };

    if (KIND == 0 || (KIND == 2 && p <= 2)) {
        const bool copy = p == 1;
        const bool pair = KIND == 2 && p == 2;
        const int lag = pair ? 1 : p;
        const double inv = 1.0 / (double)p;
        double sum = 0.0;
        while (t < c1) {
            stage();
            float xo[FMS_B];
            if (!copy) lagged(xo, lag);
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) {
                const int u = t + k;
                if (u >= c1) break;
                const double xn = (double)cur[k];
                float v = FIR_NAN_F;
                if (copy) {
                    v = cur[k];
                } else if (pair) {
                    // stored only from warm = fv + 1 and c0 on, where u-1 >= r0
                    if (u > r0) v = (float)(((double)xo[k] + xn) * 0.5);
                } else if (u >= r0) {
                    // sma.rs:318-328: the first p bars summed, then sum += x[t] - x[t-p]
                    if (u < r0 + p) sum += xn;
                    else sum += xn - (double)xo[k];
                    v = (float)(sum * inv);
                    if (u - p >= z) v = FIR_NAN_F;
                }
                put(u, v);
            }
            next();
        }
    } else if constexpr (KIND == 1) {
        // wma.rs:290-320: weights 1..p, running weighted sum and plain sum
        const double pf = (double)p;
        const double inv = 1.0 / (pf * (pf + 1.0) * 0.5);
        double s1 = 0.0, s2 = 0.0;
        while (t < c1) {
            stage();
            float xo[FMS_B];
            lagged(xo, p - 1);
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) {
                const int u = t + k;
                if (u >= c1) break;
                const double xn = (double)cur[k];
                float v = FIR_NAN_F;
                if (u >= r0) {
                    const int q = u - r0;
                    if (q < p - 1) {
                        s2 += xn * ((double)q + 1.0);
                        s1 += xn;
                    } else {
                        s2 += xn * pf;
                        s1 += xn;
                        v = (float)(s2 * inv);
                        s2 -= s1;
                        s1 -= (double)xo[k];
                    }
                    if (u > z) v = FIR_NAN_F;
                }
                put(u, v);
            }
            next();
        }
    } else if constexpr (KIND == 2) {
        // swma.rs:406-476 with the ring of the last b box sums replaced by a
        // second running box sum lagged by b bars (same operations, same bits)
        int wa, wb;
        swma_ab(p, &wa, &wb);
        const double inv = 1.0 / ((double)wa * (double)wb);
        const int sa0 = r0 + wa - 1;                  // first full box
        double s1 = 0.0, s1b = 0.0, s2 = 0.0;
        while (t < c1) {
            stage();
            float xa[FMS_B], xb[FMS_B], xab[FMS_B];
            lagged(xa, wa - 1);
            lagged(xb, wb);
            lagged(xab, wb + wa - 1);
#pragma unroll
            for (int k = 0; k < FMS_B; ++k) {
                const int u = t + k;
                if (u >= c1) break;
                float v = FIR_NAN_F;
                if (u >= r0) {
                    s1 += (double)cur[k];
                    double old = 0.0;
                    const int tl = u - wb;            // bar whose box leaves the sum
                    if (tl >= r0) s1b += (double)xb[k];
                    if (tl >= sa0) old = s1b;
                    if (u >= sa0) {
                        s2 = s2 + (s1 - old);
                        v = (float)(s2 * inv);
                        s1 -= (double)xa[k];
                    }
                    if (tl >= sa0) s1b -= (double)xab[k];
                    if (u - wa >= z) v = FIR_NAN_F;
                }
                put(u, v);
            }
            next();
        }
    }
}

// One kernel per kind, so each gets its own register allocation.
// SMA / WMA read their one lagged stream through the cache, 8-bar blocks;
// SWMA takes its three lagged streams from the ring, 16-bar blocks.
extern "C" __global__ __launch_bounds__(64) void fir_ms1p_sma_f32(fir_args a) { fir_ms1p<0, 8>(a); }
extern "C" __global__ __launch_bounds__(64) void fir_ms1p_wma_f32(fir_args a) { fir_ms1p<1, 8>(a); }
extern "C" __global__ __launch_bounds__(64) void fir_ms1p_swma_f32(fir_args a) { fir_ms1p<2, 16>(a); }

// Single series: the first bar at or after first_valid whose value is not
// finite.  z[0] must hold T on entry (vta_fill_i32).
// The same with first_valid read from the device (*fvp, in [0, T]): grid over
// the whole series.
extern "C" __global__ __launch_bounds__(256) void series_first_nonfinite_dfv(
    const float *__restrict__ x, int T, const int32_t *__restrict__ fvp,
    int32_t *__restrict__ z) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= T || t < fvp[0]) return;
    if (!vta_isfinite(x[t])) atomicMin(z, t);
}

extern "C" __global__ __launch_bounds__(256) void series_first_nonfinite(
    const float *__restrict__ x, int T, int first_valid, int32_t *__restrict__ z) {
    const int t = blockIdx.x * 256 + threadIdx.x + first_valid;
    if (t >= T) return;
    if (!vta_isfinite(x[t])) atomicMin(z, t);
}
