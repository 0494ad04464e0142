/*
 * vta_oracle.c -- CPU f64 restatement of VectorTA's scalar indicator kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * batch-indicator path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU baseline).  The product library never links or calls it.
 *
 * Every function restates the reference's scalar f64 arithmetic in the same
 * operation order.  The build compiles this file with -ffp-contract=off so no
 * multiply-add is fused except where the reference calls f64::mul_add (written
 * here as fma()).  Rust's f64::exp / sqrt / abs / round / max map onto glibc
 * exp / sqrt / fabs / round / fmax (fmax has the same NaN rule as f64::max:
 * a NaN argument yields the other argument).
 *
 * References (paths relative to the VectorTA checkout):
 *   ALMA  src/indicators/moving_averages/alma.rs:279-340, 426-463, 1198-1289,
 *         1368-1527; the AVX2 row :1529-1573 (a timing baseline only)
 *   EMA   src/indicators/moving_averages/ema.rs:287-324, 430-468, 858-893
 *   RSI   src/indicators/rsi.rs:189-227, 327-414 (batch row :903-990 identical)
 *   ATR   src/indicators/atr.rs:197-230, 329-460; batch-prefix :1101-1213,
 *         TR :1267-1295
 *   BB    src/indicators/bollinger_bands.rs:253-295, 566-672 (SMA, devtype 0)
 *   PNR   src/indicators/percentile_nearest_rank.rs:148-266
 *   MAD   src/indicators/medium_ad.rs:175-230, 293-402
 *   ADX   src/indicators/adx.rs:190-298, 299-422
 *   SMA   src/indicators/moving_averages/sma.rs:245-330 (scalar), 909-1021 (batch
 *         prefix rows)
 *   WMA   src/indicators/moving_averages/wma.rs:243-320 (scalar), 656-760 (batch
 *         prefix rows)
 *   SWMA  src/indicators/moving_averages/swma.rs:203-231, 406-476 (scalar; the
 *         batch row swma_row_scalar :1155-1217 is the same algorithm)
 *   SMMA  src/indicators/moving_averages/smma.rs:243-271, 369-400 (batch rows
 *         call the scalar kernel, :775-862)
 *   Wilders src/indicators/moving_averages/wilders.rs:195-231, 262-320
 *   NaN prefixes: src/utilities/helpers.rs:102-164 (quiet NaN 0x7ff8...)
 *   Double-ALMA crossover: not in the reference (README.md:9 names the
 *         workload only); position = sign(ALMA_fast - ALMA_slow) of the f64
 *         batch rows above, reduced to per-pair statistics (or_cross_pairs).
 *
 * Status codes: 0 ok, 1 empty input, 2 all NaN, 3 invalid period,
 * 4 not enough valid data, 5 invalid parameter (sigma/offset/percentage).
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_OK 0
#define OR_EMPTY 1
#define OR_ALLNAN 2
#define OR_PERIOD 3
#define OR_NOTENOUGH 4
#define OR_PARAM 5

static double qnan(void) {
    uint64_t b = 0x7ff8000000000000ULL;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

static void fill_nan(double *out, size_t n) {
    double q = qnan();
    for (size_t i = 0; i < n; ++i) out[i] = q;
}

/* first index whose value is not NaN; len if none */
size_t or_first_valid(const double *x, size_t len) {
    size_t i = 0;
    while (i < len && isnan(x[i])) ++i;
    return i;
}

/* ------------------------------------------------------------------ ALMA */

/* alma.rs:315-332 single-series weights: w = exp((-d*d) * inv_s2),
 * inv_norm = 1/sum(w).  Returns inv_norm, writes period weights. */
double or_alma_weights_single(size_t period, double offset, double sigma, double *w) {
    double m = offset * (double)(period - 1);
    double s = (double)period / sigma;
    double s2 = 2.0 * s * s;
    double inv_s2 = 1.0 / s2;
    double norm = 0.0;
    for (size_t i = 0; i < period; ++i) {
        double diff = (double)i - m;
        double wi = exp(-diff * diff * inv_s2);
        w[i] = wi;
        norm += wi;
    }
    return 1.0 / norm;
}

/* alma.rs:1426-1436 batch weights: w = exp(-(d.powi(2)) / s2). */
double or_alma_weights_batch(size_t period, double offset, double sigma, double *w) {
    double m = offset * (double)(period - 1);
    double s = (double)period / sigma;
    double s2 = 2.0 * s * s;
    double norm = 0.0;
    for (size_t i = 0; i < period; ++i) {
        double d = (double)i - m;
        double wi = exp(-(d * d) / s2);
        w[i] = wi;
        norm += wi;
    }
    return 1.0 / norm;
}

/* alma.rs:426-463 (alma_scalar) and :1505-1527 (alma_row_scalar): the same
 * 4-wide grouped dot, ((d0w0 + d1w1) + d2w2) + d3w3 added to the running
 * sum, then the tail one term at a time; out = sum * inv_norm. */
static void alma_dot_rows(const double *data, size_t len, size_t first, size_t period,
                          const double *w, double inv_n, double *out) {
    size_t p4 = period & ~(size_t)3;
    for (size_t i = first + period - 1; i < len; ++i) {
        size_t start = i + 1 - period;
        const double *d = data + start;
        double sum = 0.0;
        for (size_t k = 0; k < p4; k += 4) {
            double g = d[k] * w[k] + d[k + 1] * w[k + 1];
            g = g + d[k + 2] * w[k + 2];
            g = g + d[k + 3] * w[k + 3];
            sum += g;
        }
        for (size_t k = p4; k < period; ++k) sum += d[k] * w[k];
        out[i] = sum * inv_n;
    }
}

/* alma.rs:279-350 alma_prepare + alma_with_kernel(Scalar). */
int or_alma_single(const double *data, size_t len, size_t period, double offset,
                   double sigma, double *out) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (period == 0 || period > len) return OR_PERIOD;
    if (len - first < period) return OR_NOTENOUGH;
    if (!(sigma > 0.0)) return OR_PARAM;
    if (!(offset >= 0.0 && offset <= 1.0)) return OR_PARAM;
    double *w = (double *)malloc(sizeof(double) * period);
    double inv_n = or_alma_weights_single(period, offset, sigma, w);
    fill_nan(out, first + period - 1);
    alma_dot_rows(data, len, first, period, w, inv_n, out);
    free(w);
    return OR_OK;
}

/* alma.rs:1368-1502 alma_batch_inner_into with Kernel::ScalarBatch for an
 * explicit combo list (the grid expansion is restated in the product's host
 * runtime and in tests).  out is [ncombos][len] row-major.  The batch path
 * needs len - first >= round_up8(max period) (alma.rs:1388-1398). */
int or_alma_batch(const double *data, size_t len, const size_t *periods, const double *offsets,
                  const double *sigmas, size_t ncombos, double *out, int nthreads) {
    if (len == 0) return OR_ALLNAN;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    size_t max_p8 = 0;
    for (size_t c = 0; c < ncombos; ++c) {
        size_t r8 = (periods[c] + 7) & ~(size_t)7;
        if (r8 > max_p8) max_p8 = r8;
    }
    if (len - first < max_p8) return OR_NOTENOUGH;
    for (size_t c = 0; c < ncombos; ++c) {
        if (!(sigmas[c] > 0.0)) return OR_PARAM;
        if (!(offsets[c] >= 0.0 && offsets[c] <= 1.0)) return OR_PARAM;
        if (periods[c] == 0) return OR_PERIOD;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long c = 0; c < (long)ncombos; ++c) {
        size_t p = periods[c];
        double *w = (double *)malloc(sizeof(double) * p);
        double inv_n = or_alma_weights_batch(p, offsets[c], sigmas[c], w);
        double *row = out + (size_t)c * len;
        fill_nan(row, first + p - 1);
        alma_dot_rows(data, len, first, p, w, inv_n, row);
        free(w);
    }
    return OR_OK;
}

/* alma.rs:1529-1573 alma_row_avx2 (the reference's "nightly-avx" build; its
 * default build runs alma_row_scalar, restated above): four f64 lanes of FMA
 * over the window in blocks of four, a masked tail block, then
 * ((l0 + l2) + (l1 + l3)) * inv_n.  A CPU timing baseline only: its summation
 * order differs from alma_row_scalar, so it is checked against or_alma_batch
 * to a relative 1e-12, not used as the parity reference. */
__attribute__((target("avx2,fma"))) static void alma_dot_rows_avx2(
    const double *data, size_t len, size_t first, size_t period, const double *w, double inv_n,
    double *out) {
    const size_t blocks = period / 4, tail = period % 4;
    const __m256i mask = _mm256_setr_epi64x(tail > 0 ? -1 : 0, tail > 1 ? -1 : 0,
                                            tail > 2 ? -1 : 0, 0);
    for (size_t i = first + period - 1; i < len; ++i) {
        const double *d = data + (i + 1 - period);
        __m256d acc = _mm256_setzero_pd();
        for (size_t b = 0; b < blocks; ++b)
            acc = _mm256_fmadd_pd(_mm256_loadu_pd(d + 4 * b), _mm256_loadu_pd(w + 4 * b), acc);
        if (tail)
            acc = _mm256_fmadd_pd(_mm256_maskload_pd(d + 4 * blocks, mask),
                                  _mm256_maskload_pd(w + 4 * blocks, mask), acc);
        const __m128d s2 = _mm_add_pd(_mm256_castpd256_pd128(acc), _mm256_extractf128_pd(acc, 1));
        const __m128d s1 = _mm_add_pd(s2, _mm_unpackhi_pd(s2, s2));
        out[i] = _mm_cvtsd_f64(s1) * inv_n;
    }
}

/* or_alma_batch with the AVX2 row; -1 when the CPU lacks AVX2 or FMA. */
int or_alma_batch_avx2(const double *data, size_t len, const size_t *periods,
                       const double *offsets, const double *sigmas, size_t ncombos, double *out,
                       int nthreads) {
    __builtin_cpu_init();
    if (!__builtin_cpu_supports("avx2") || !__builtin_cpu_supports("fma")) return -1;
    if (len == 0) return OR_ALLNAN;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    size_t max_p8 = 0;
    for (size_t c = 0; c < ncombos; ++c) {
        size_t r8 = (periods[c] + 7) & ~(size_t)7;
        if (r8 > max_p8) max_p8 = r8;
    }
    if (len - first < max_p8) return OR_NOTENOUGH;
    for (size_t c = 0; c < ncombos; ++c) {
        if (!(sigmas[c] > 0.0)) return OR_PARAM;
        if (!(offsets[c] >= 0.0 && offsets[c] <= 1.0)) return OR_PARAM;
        if (periods[c] == 0) return OR_PERIOD;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long c = 0; c < (long)ncombos; ++c) {
        size_t p = periods[c];
        double *w = (double *)malloc(sizeof(double) * p);
        double inv_n = or_alma_weights_batch(p, offsets[c], sigmas[c], w);
        double *row = out + (size_t)c * len;
        fill_nan(row, first + p - 1);
        alma_dot_rows_avx2(data, len, first, p, w, inv_n, row);
        free(w);
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ EMA */

static int is_finite_bits(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    return (b & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL;
}

/* ema.rs:430-468 ema_scalar_into (batch row ema.rs:858-893 is identical).
 * NaN prefix = first (ema.rs:362). */
int or_ema(const double *data, size_t len, size_t period, double *out) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (period == 0 || period > len) return OR_PERIOD;
    if (len - first < period) return OR_NOTENOUGH;
    double alpha = 2.0 / ((double)period + 1.0);
    double beta = 1.0 - alpha;
    fill_nan(out, first);
    double mean = data[first];
    out[first] = mean;
    size_t vc = 1;
    size_t warm_end = first + period < len ? first + period : len;
    for (size_t i = first + 1; i < warm_end; ++i) {
        double x = data[i];
        if (is_finite_bits(x)) {
            vc += 1;
            double v = (double)vc;
            mean = ((v - 1.0) * mean + x) / v;
        }
        out[i] = mean;
    }
    double prev = mean;
    for (size_t i = warm_end; i < len; ++i) {
        double x = data[i];
        if (is_finite_bits(x)) prev = fma(beta, prev, alpha * x);
        out[i] = prev;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ RSI */

/* rsi.rs:189-227 + rsi_compute_into_scalar :327-414.  NaN prefix first+period. */
int or_rsi(const double *data, size_t len, size_t period, double *out) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (period == 0 || period > len) return OR_PERIOD;
    if (len - first < period) return OR_NOTENOUGH;
    fill_nan(out, first + period < len ? first + period : len);
    double inv_p = 1.0 / (double)period;
    double beta = 1.0 - inv_p;
    double ag = 0.0, al = 0.0;
    int has_nan = 0;
    size_t warm_last = first + period;
    if (len - 1 < warm_last) warm_last = len - 1;
    for (size_t i = first + 1; i <= warm_last; ++i) {
        double d = data[i] - data[i - 1];
        if (!isfinite(d)) {
            has_nan = 1;
            break;
        }
        if (d > 0.0)
            ag += d;
        else if (d < 0.0)
            al -= d;
    }
    size_t idx0 = first + period;
    if (has_nan) {
        ag = qnan();
        al = qnan();
        if (idx0 < len) out[idx0] = qnan();
    } else {
        ag *= inv_p;
        al *= inv_p;
        if (idx0 < len) {
            double den = ag + al;
            out[idx0] = den == 0.0 ? 50.0 : 100.0 * ag / den;
        }
    }
    for (size_t j = idx0 + 1; j < len; ++j) {
        double d = data[j] - data[j - 1];
        double g = d > 0.0 ? d : 0.0;
        double l = d < 0.0 ? -d : 0.0;
        ag = fma(ag, beta, inv_p * g);
        al = fma(al, beta, inv_p * l);
        double den = ag + al;
        out[j] = den == 0.0 ? 50.0 : 100.0 * ag / den;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ ATR */

/* atr.rs:197-207 */
size_t or_first_valid_hlc(const double *h, const double *l, const double *c, size_t len) {
    size_t i = 0;
    while (i < len && (isnan(h[i]) || isnan(l[i]) || isnan(c[i]))) ++i;
    return i;
}

static double true_range(double hi, double lo, double pc) {
    double tr = hi - lo;
    double hc = fabs(hi - pc);
    if (hc > tr) tr = hc;
    double lc = fabs(lo - pc);
    if (lc > tr) tr = lc;
    return tr;
}

/* atr.rs:236-273 + atr_compute_into_scalar :329-460 (the 4-way unrolled loop
 * is the same sequential recurrence).  NaN prefix first+length-1. */
int or_atr(const double *h, const double *l, const double *c, size_t len, size_t length,
           double *out) {
    if (length == 0) return OR_PERIOD;
    if (len == 0) return OR_EMPTY;
    if (length > len) return OR_NOTENOUGH;
    size_t first = or_first_valid_hlc(h, l, c, len);
    if (first >= len) return OR_ALLNAN;
    if (len - first < length) return OR_NOTENOUGH;
    size_t warm = first + length - 1;
    double alpha = 1.0 / (double)length;
    fill_nan(out, warm);
    double sum_tr = h[first] - l[first];
    for (size_t i = first + 1; i <= warm; ++i) sum_tr += true_range(h[i], l[i], c[i - 1]);
    double rma = sum_tr / (double)length;
    out[warm] = rma;
    for (size_t i = warm + 1; i < len; ++i) {
        double tr = true_range(h[i], l[i], c[i - 1]);
        rma = fma(-alpha, rma, rma) + alpha * tr;
        out[i] = rma;
    }
    return OR_OK;
}

/* atr.rs:1101-1213 batch row: TR precomputed (:1267-1295), seed from the
 * prefix-sum difference ps[warm+1]-ps[first]. */
int or_atr_batch_row(const double *h, const double *l, const double *c, size_t len,
                     size_t length, double *out) {
    size_t first = or_first_valid_hlc(h, l, c, len);
    if (first >= len) return OR_ALLNAN;
    if (length == 0) return OR_PERIOD;
    if (len - first < length) return OR_NOTENOUGH;
    double *tr = (double *)calloc(len, sizeof(double));
    double *ps = (double *)malloc(sizeof(double) * (len + 1));
    tr[first] = h[first] - l[first];
    for (size_t i = first + 1; i < len; ++i) tr[i] = true_range(h[i], l[i], c[i - 1]);
    ps[0] = 0.0;
    for (size_t i = 0; i < len; ++i) ps[i + 1] = ps[i] + tr[i];
    size_t warm = first + length - 1;
    fill_nan(out, warm);
    double rma = (ps[warm + 1] - ps[first]) / (double)length;
    out[warm] = rma;
    double alpha = 1.0 / (double)length;
    for (size_t i = warm + 1; i < len; ++i) {
        rma = fma(-alpha, rma, rma) + alpha * tr[i];
        out[i] = rma;
    }
    free(tr);
    free(ps);
    return OR_OK;
}

/* ------------------------------------------------------------------ BB */

/* bollinger_bands.rs:253-295 + bb_row_scalar_classic_sma :566-672 for
 * matype "sma", devtype 0 (population variance). */
int or_bb(const double *data, size_t len, size_t period, double devup, double devdn,
          double *out_u, double *out_m, double *out_l) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (period == 0 || period > len) return OR_PERIOD;
    if (len - first < period) return OR_NOTENOUGH;
    size_t warm = first + period - 1;
    fill_nan(out_u, warm);
    fill_nan(out_m, warm);
    fill_nan(out_l, warm);
    double inv_n = 1.0 / (double)period;
    double sum = 0.0, sum_sq = 0.0;
    for (size_t k = 0; k < period; ++k) {
        double v = data[first + k];
        sum += v;
        sum_sq += v * v;
    }
    double mean = sum * inv_n;
    double var0 = (sum_sq * inv_n) - (mean * mean);
    double sd0 = sqrt(fmax(var0, 0.0));
    out_m[warm] = mean;
    out_u[warm] = mean + devup * sd0;
    out_l[warm] = mean - devdn * sd0;
    for (size_t i = warm + 1; i < len; ++i) {
        double nw = data[i];
        double old = data[i - period];
        sum += nw - old;
        sum_sq = (sum_sq - old * old) + nw * nw;
        double m = sum * inv_n;
        double var = (sum_sq * inv_n) - (m * m);
        double sd = sqrt(fmax(var, 0.0));
        out_m[i] = m;
        out_u[i] = m + devup * sd;
        out_l[i] = m - devdn * sd;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ PNR */

/* percentile_nearest_rank.rs:218-230: k = round(fma(p%/100, wl, 0)) - 1,
 * clamped to [0, wl-1]. */
static size_t pnr_rank(double p_frac, size_t wl) {
    double raw_f = round(fma(p_frac, (double)wl, 0.0));
    long raw = (long)raw_f - 1;
    size_t k = raw <= 0 ? 0 : (size_t)raw;
    if (k >= wl) k = wl - 1;
    return k;
}

static size_t lower_bound_d(const double *a, size_t n, double v) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* percentile_nearest_rank.rs:148-266: sorted window of the non-NaN values,
 * binary erase of the outgoing value and binary insert of the incoming one.
 * Output is an order statistic of the window, so any exact selection method
 * gives the same value.  NaN prefix first+length-1. */
int or_pnr(const double *data, size_t len, size_t length, double percentage, double *out) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (length == 0 || length > len) return OR_PERIOD;
    if (!(percentage >= 0.0 && percentage <= 100.0)) return OR_PARAM;
    if (len - first < length) return OR_NOTENOUGH;
    size_t start_i = first + length - 1;
    fill_nan(out, start_i);
    double *sorted = (double *)malloc(sizeof(double) * (length + 1));
    size_t wl = 0;
    for (size_t idx = start_i + 1 - length; idx <= start_i; ++idx) {
        double v = data[idx];
        if (!isnan(v)) {
            size_t pos = lower_bound_d(sorted, wl, v);
            memmove(sorted + pos + 1, sorted + pos, (wl - pos) * sizeof(double));
            sorted[pos] = v;
            ++wl;
        }
    }
    double p_frac = percentage * 0.01;
    for (size_t i = start_i;; ++i) {
        out[i] = wl == 0 ? qnan() : sorted[pnr_rank(p_frac, wl)];
        if (i + 1 >= len) break;
        double vo = data[i + 1 - length];
        if (!isnan(vo)) {
            size_t pos = lower_bound_d(sorted, wl, vo);
            if (pos < wl && sorted[pos] == vo) {
                memmove(sorted + pos, sorted + pos + 1, (wl - pos - 1) * sizeof(double));
                --wl;
            }
        }
        double vi = data[i + 1];
        if (!isnan(vi)) {
            size_t pos = lower_bound_d(sorted, wl, vi);
            memmove(sorted + pos + 1, sorted + pos, (wl - pos) * sizeof(double));
            sorted[pos] = vi;
            ++wl;
        }
    }
    free(sorted);
    return OR_OK;
}

/* ------------------------------------------------------------------ MAD */

static void insertion_sort(double *a, size_t n) {
    for (size_t i = 1; i < n; ++i) {
        double v = a[i];
        size_t j = i;
        while (j > 0 && a[j - 1] > v) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = v;
    }
}

/* medium_ad.rs:302-324 median_from after select_nth_unstable(mid): odd n ->
 * element mid; even n -> 0.5 * (max of the lower mid elements + element mid),
 * i.e. 0.5 * (sorted[mid-1] + sorted[mid]). */
static double median_sorted(const double *s, size_t n) {
    size_t mid = n >> 1;
    if (n & 1) return s[mid];
    return 0.5 * (s[mid - 1] + s[mid]);
}

/* medium_ad.rs:175-230 + medium_ad_scalar :293-402.  NaN prefix
 * first+period-1; any NaN in the window -> NaN; period 1 -> 0. */
int or_mad(const double *data, size_t len, size_t period, double *out) {
    if (len == 0) return OR_EMPTY;
    size_t first = or_first_valid(data, len);
    if (first >= len) return OR_ALLNAN;
    if (period == 0 || period > len) return OR_PERIOD;
    if (len - first < period) return OR_NOTENOUGH;
    fill_nan(out, first + period - 1);
    if (period == 1) {
        for (size_t i = first; i < len; ++i) out[i] = isnan(data[i]) ? qnan() : 0.0;
        return OR_OK;
    }
    double *buf = (double *)malloc(sizeof(double) * period);
    for (size_t i = first + period - 1; i < len; ++i) {
        size_t start = i + 1 - period;
        int has_nan = 0;
        for (size_t k = 0; k < period; ++k) {
            buf[k] = data[start + k];
            if (isnan(buf[k])) has_nan = 1;
        }
        if (has_nan) {
            out[i] = qnan();
            continue;
        }
        insertion_sort(buf, period);
        double med = median_sorted(buf, period);
        for (size_t k = 0; k < period; ++k) buf[k] = fabs(buf[k] - med);
        insertion_sort(buf, period);
        out[i] = median_sorted(buf, period);
    }
    free(buf);
    return OR_OK;
}

/* ------------------------------------------------------------------ ADX */

/* adx.rs:190-198: the max of the three per-array first non-NaN indices. */
size_t or_first_valid_triple(const double *h, const double *l, const double *c, size_t len) {
    size_t fh = or_first_valid(h, len), fl = or_first_valid(l, len), fc = or_first_valid(c, len);
    size_t m = fh > fl ? fh : fl;
    return m > fc ? m : fc;
}

/* adx.rs:299-422 adx_scalar on the slices starting at first. */
static void adx_scalar(const double *high, const double *low, const double *close, size_t len,
                       size_t period, double *out) {
    if (len <= period) return;
    double pf = (double)period;
    double rp = 1.0 / pf;
    double omr = 1.0 - rp;
    double pm1 = pf - 1.0;
    double tr_sum = 0.0, pdm_sum = 0.0, mdm_sum = 0.0;
    double ph = high[0], pl = low[0], pc = close[0];
    for (size_t i = 1; i <= period; ++i) {
        double ch = high[i], cl = low[i];
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        if (up > down && up > 0.0) pdm_sum += up;
        if (down > up && down > 0.0) mdm_sum += down;
        tr_sum += tr;
        ph = ch;
        pl = cl;
        pc = close[i];
    }
    double atr = tr_sum, pds = pdm_sum, mds = mdm_sum;
    double pdi0 = 0.0, mdi0 = 0.0;
    if (atr != 0.0) {
        pdi0 = (pds / atr) * 100.0;
        mdi0 = (mds / atr) * 100.0;
    }
    double sdi0 = pdi0 + mdi0;
    double dx_sum = sdi0 != 0.0 ? (fabs(pdi0 - mdi0) / sdi0) * 100.0 : 0.0;
    size_t dx_count = 1;
    double last_adx = 0.0;
    ph = high[period];
    pl = low[period];
    pc = close[period];
    for (size_t i = period + 1; i < len; ++i) {
        double ch = high[i], cl = low[i];
        double hl = ch - cl;
        double hpc = fabs(ch - pc);
        double lpc = fabs(cl - pc);
        double tr = fmax(fmax(hl, hpc), lpc);
        double up = ch - ph;
        double down = pl - cl;
        double pdm = (up > down && up > 0.0) ? up : 0.0;
        double mdm = (down > up && down > 0.0) ? down : 0.0;
        atr = atr * omr + tr;
        pds = pds * omr + pdm;
        mds = mds * omr + mdm;
        double pdi = 0.0, mdi = 0.0;
        if (atr != 0.0) {
            pdi = (pds / atr) * 100.0;
            mdi = (mds / atr) * 100.0;
        }
        double sdi = pdi + mdi;
        double dx = sdi != 0.0 ? (fabs(pdi - mdi) / sdi) * 100.0 : 0.0;
        if (dx_count < period) {
            dx_sum += dx;
            dx_count += 1;
            if (dx_count == period) {
                last_adx = dx_sum * rp;
                out[i] = last_adx;
            }
        } else {
            last_adx = (last_adx * pm1 + dx) * rp;
            out[i] = last_adx;
        }
        ph = ch;
        pl = cl;
        pc = close[i];
    }
}

/* adx.rs:211-298 adx_with_kernel(Scalar).  NaN prefix first+2*period-1. */
int or_adx(const double *h, const double *l, const double *c, size_t len, size_t period,
           double *out) {
    if (len == 0) return OR_EMPTY;
    if (period == 0 || period > len) return OR_PERIOD;
    if (or_first_valid(h, len) >= len || or_first_valid(l, len) >= len ||
        or_first_valid(c, len) >= len)
        return OR_ALLNAN;
    size_t first = or_first_valid_triple(h, l, c, len);
    if (len - first < period + 1) return OR_NOTENOUGH;
    size_t warm_end = first + 2 * period - 1;
    fill_nan(out, warm_end < len ? warm_end : len);
    adx_scalar(h + first, l + first, c + first, len - first, period, out + first);
    return OR_OK;
}

/* -------------------------------------------------- many-series helpers */

/* Runs a single-series oracle over each column of a time-major [rows][cols]
 * matrix (the reference's GPU tests build the expected many-series output
 * this way, e.g. tests/alma_cuda.rs:118-140).  kind: 0 alma, 1 ema, 2 rsi,
 * 3 pnr, 4 mad.  p0/p1/p2 carry the indicator parameters.  Per-column status
 * codes are written to status[cols]. */
void or_many_series_1in(int kind, const double *tm, size_t cols, size_t rows, double p0,
                        double p1, double p2, double *out_tm, int *status, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long s = 0; s < (long)cols; ++s) {
        double *x = (double *)malloc(sizeof(double) * rows);
        double *y = (double *)malloc(sizeof(double) * rows);
        for (size_t t = 0; t < rows; ++t) x[t] = tm[t * cols + s];
        fill_nan(y, rows);
        int st = OR_PARAM;
        switch (kind) {
            case 0: st = or_alma_single(x, rows, (size_t)p0, p1, p2, y); break;
            case 1: st = or_ema(x, rows, (size_t)p0, y); break;
            case 2: st = or_rsi(x, rows, (size_t)p0, y); break;
            case 3: st = or_pnr(x, rows, (size_t)p0, p1, y); break;
            case 4: st = or_mad(x, rows, (size_t)p0, y); break;
            default: break;
        }
        status[s] = st;
        for (size_t t = 0; t < rows; ++t) out_tm[t * cols + s] = y[t];
        free(x);
        free(y);
    }
}

/* Three-input (high, low, close) single-series oracle over each column of
 * time-major matrices.  kind: 0 atr, 1 adx, 2 atr batch-prefix row. */
void or_many_series_3in(int kind, const double *h_tm, const double *l_tm, const double *c_tm,
                        size_t cols, size_t rows, size_t period, double *out_tm, int *status,
                        int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long s = 0; s < (long)cols; ++s) {
        double *h = (double *)malloc(sizeof(double) * rows);
        double *l = (double *)malloc(sizeof(double) * rows);
        double *c = (double *)malloc(sizeof(double) * rows);
        double *y = (double *)malloc(sizeof(double) * rows);
        for (size_t t = 0; t < rows; ++t) {
            h[t] = h_tm[t * cols + s];
            l[t] = l_tm[t * cols + s];
            c[t] = c_tm[t * cols + s];
        }
        fill_nan(y, rows);
        int st = OR_PARAM;
        if (kind == 0) st = or_atr(h, l, c, rows, period, y);
        else if (kind == 1) st = or_adx(h, l, c, rows, period, y);
        else if (kind == 2) st = or_atr_batch_row(h, l, c, rows, period, y);
        status[s] = st;
        for (size_t t = 0; t < rows; ++t) out_tm[t * cols + s] = y[t];
        free(h);
        free(l);
        free(c);
        free(y);
    }
}

/* Series-major driver (the CPU baselines' layout): n_series contiguous series
 * of len bars, series s at a + s * len (b, c: low and close for ATR), outputs
 * at out + s * len (BB: upper, middle, lower in out, out2, out3), OpenMP over
 * series.  kind: 0 ema(p0), 1 rsi(p0), 2 pnr(p0, percentage p1), 3 mad(p0),
 * 4 bb(p0, devup p1, devdn p2), 5 atr(p0) over (high a, low b, close c). */
void or_series_major(int kind, const double *a, const double *b, const double *c,
                     size_t n_series, size_t len, double p0, double p1, double p2, double *out,
                     double *out2, double *out3, int *status, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) if (nthreads != 1)
#endif
    for (long s = 0; s < (long)n_series; ++s) {
        const size_t o = (size_t)s * len;
        double *y = out + o;
        fill_nan(y, len);
        int st = OR_PARAM;
        switch (kind) {
            case 0: st = or_ema(a + o, len, (size_t)p0, y); break;
            case 1: st = or_rsi(a + o, len, (size_t)p0, y); break;
            case 2: st = or_pnr(a + o, len, (size_t)p0, p1, y); break;
            case 3: st = or_mad(a + o, len, (size_t)p0, y); break;
            case 4:
                fill_nan(out2 + o, len);
                fill_nan(out3 + o, len);
                st = or_bb(a + o, len, (size_t)p0, p1, p2, y, out2 + o, out3 + o);
                break;
            case 5: st = or_atr(a + o, b + o, c + o, len, (size_t)p0, y); break;
            default: break;
        }
        status[s] = st;
    }
}

/* One series x many periods (the recurrence batches' CPU baseline): row r
 * of out is the single-series oracle at periods[r], OpenMP over rows.
 * kind: 0 ema, 1 rsi (series c), 2 atr (the batch path's row, atr_batch_row),
 * 3 adx (high h, low l, close c). */
void or_batch_periods(int kind, const double *h, const double *l, const double *c, size_t len,
                      const int *periods, size_t n, double *out, int *status, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) if (nthreads != 1)
#endif
    for (long r = 0; r < (long)n; ++r) {
        double *y = out + (size_t)r * len;
        fill_nan(y, len);
        const size_t p = (size_t)periods[r];
        int st = OR_PARAM;
        switch (kind) {
            case 0: st = or_ema(c, len, p, y); break;
            case 1: st = or_rsi(c, len, p, y); break;
            case 2: st = or_atr_batch_row(h, l, c, len, p, y); break;
            case 3: st = or_adx(h, l, c, len, p, y); break;
            default: break;
        }
        status[r] = st;
    }
}

/* ------------------------------------------------------------ double-ALMA crossover */

/* Per pair (fast row f, slow row s of the f64 ALMA batch rows, [nrows][len]):
 * from w = first bar where both rows are non-NaN, pos_t = sign(f_t - s_t)
 * (+1, -1, or 0 when equal), and
 *   stats[0] trades   = #{t > w : pos_t != pos_{t-1}}
 *   stats[1] n_long   = #{t >= w : pos_t = +1}
 *   stats[2] n_short  = #{t >= w : pos_t = -1}
 *   stats[3] checksum = sum_{t >= w} pos_t * (t + 1)
 *   pnl              = sum_{t = w}^{len-2} pos_t * (x_{t+1} - x_t)
 * summed bar by bar in time order. */
void or_cross_pairs(const double *rows, size_t nrows, size_t len, const double *x,
                    const int *pair_fast, const int *pair_slow, size_t npairs,
                    long long *stats, double *pnl, int nthreads) {
    (void)nrows;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
    for (long p = 0; p < (long)npairs; ++p) {
        const double *f = rows + (size_t)pair_fast[p] * len;
        const double *s = rows + (size_t)pair_slow[p] * len;
        size_t w = 0;
        while (w < len && (isnan(f[w]) || isnan(s[w]))) ++w;
        long long tr = 0, nl = 0, ns = 0, ck = 0;
        double pn = 0.0;
        int prev = 0;
        for (size_t t = w; t < len; ++t) {
            double d = f[t] - s[t];
            int pos = d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
            if (t > w && pos != prev) ++tr;
            if (pos > 0) ++nl;
            if (pos < 0) ++ns;
            ck += (long long)pos * (long long)(t + 1);
            if (t + 1 < len) pn += (double)pos * (x[t + 1] - x[t]);
            prev = pos;
        }
        stats[p * 4 + 0] = tr;
        stats[p * 4 + 1] = nl;
        stats[p * 4 + 2] = ns;
        stats[p * 4 + 3] = ck;
        pnl[p] = pn;
    }
}

/* ------------------------------------------------------------------ FIR moving averages */

/* Validation shared by the one-period moving averages (sma.rs:245-280,
 * swma.rs:203-231, smma.rs:369-400, wilders.rs:195-231): empty input, all
 * NaN, period outside [min_p, len], fewer than period bars after first. */
static int ma_prepare(const double *data, size_t len, size_t period, size_t min_p,
                      size_t *first) {
    if (len == 0) return OR_EMPTY;
    *first = or_first_valid(data, len);
    if (*first >= len) return OR_ALLNAN;
    if (period < min_p || period > len) return OR_PERIOD;
    if (len - *first < period) return OR_NOTENOUGH;
    return OR_OK;
}

/* sma.rs:302-330 sma_scalar: running sum, sum += x[i] - x[i-p]. */
int or_sma(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 1, &first);
    if (st) return st;
    fill_nan(out, first + period - 1);
    if (period == 1) {
        for (size_t i = first; i < len; ++i) out[i] = data[i];
        return OR_OK;
    }
    double sum = 0.0;
    for (size_t k = 0; k < period; ++k) sum += data[first + k];
    double inv = 1.0 / (double)period;
    out[first + period - 1] = sum * inv;
    for (size_t i = first + period; i < len; ++i) {
        sum += data[i] - data[i - period];
        out[i] = sum * inv;
    }
    return OR_OK;
}

/* sma.rs:909-1021 sma_batch_inner_into, one row: inclusive prefix ps[i]
 * (0 before first), out[i] = (ps[i] - ps[i-p]) * (1/p). */
int or_sma_batch_row(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 1, &first);
    if (st) return st;
    double *ps = (double *)calloc(len, sizeof(double));
    ps[first] = data[first];
    for (size_t i = first + 1; i < len; ++i) ps[i] = ps[i - 1] + data[i];
    size_t warm = first + period - 1;
    fill_nan(out, warm);
    double inv = 1.0 / (double)period;
    double lo = warm >= period ? ps[warm - period] : 0.0;
    out[warm] = (ps[warm] - lo) * inv;
    for (size_t i = warm + 1; i < len; ++i) out[i] = (ps[i] - ps[i - period]) * inv;
    free(ps);
    return OR_OK;
}

/* wma.rs:243-320 wma_prepare + wma_scalar: weights 1..p (newest p), divided
 * by p(p+1)/2; running weight_sum and sum. */
int or_wma(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 2, &first);
    if (st) return st;
    fill_nan(out, first + period - 1);
    size_t lookback = period - 1;
    double pf = (double)period;
    double weights = pf * (pf + 1.0) * 0.5;
    const double *base = data + first;
    double sum = 0.0, wsum = 0.0;
    for (size_t k = 0; k < lookback; ++k) {
        double v = base[k];
        wsum += v * ((double)k + 1.0);
        sum += v;
    }
    for (size_t i = first + lookback, j = 0; i < len; ++i, ++j) {
        double v = data[i];
        wsum += v * pf;
        sum += v;
        out[i] = wsum / weights;
        wsum -= sum;
        sum -= base[j];
    }
    return OR_OK;
}

/* wma.rs:656-760 wma_batch_inner_into, one row: prefixes pa (sum x) and pb
 * (sum i*x), x = 0 before first; wsum = s_b - ((i+1-p) - 1) * s_a. */
int or_wma_batch_row(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 2, &first);
    if (st) return st;
    double *pa = (double *)malloc(sizeof(double) * (len + 1));
    double *pb = (double *)malloc(sizeof(double) * (len + 1));
    pa[0] = 0.0;
    pb[0] = 0.0;
    for (size_t i = 0; i < len; ++i) {
        double x = i < first ? 0.0 : data[i];
        pa[i + 1] = pa[i] + x;
        pb[i + 1] = pb[i] + (double)i * x;
    }
    size_t warm = first + period - 1;
    fill_nan(out, warm);
    double denom = (double)(period * (period + 1)) / 2.0;
    double inv_div = 1.0 / denom;
    for (size_t i = warm; i < len; ++i) {
        double s_a = pa[i + 1] - pa[i + 1 - period];
        double s_b = pb[i + 1] - pb[i + 1 - period];
        double w = s_b - ((double)(i + 1 - period) - 1.0) * s_a;
        out[i] = w * inv_div;
    }
    free(pa);
    free(pb);
    return OR_OK;
}

/* swma.rs:406-476 swma_scalar (batch row swma_row_scalar :1155-1217 is the
 * same): the symmetric triangle as two running box sums of lengths a and b
 * (odd p: a = b = (p+1)/2; even p: a = p/2, b = a+1), a ring of the last b
 * first-stage sums, out = s2 / (a*b).  p = 1 copies, p = 2 averages pairs. */
int or_swma(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 1, &first);
    if (st) return st;
    fill_nan(out, first + period - 1);
    if (period == 1) {
        for (size_t i = first; i < len; ++i) out[i] = data[i];
        return OR_OK;
    }
    if (period == 2) {
        for (size_t i = first + 1; i < len; ++i) out[i] = (data[i - 1] + data[i]) * 0.5;
        return OR_OK;
    }
    size_t a, b;
    if (period & 1) {
        a = (period + 1) >> 1;
        b = a;
    } else {
        a = period >> 1;
        b = a + 1;
    }
    double inv_ab = 1.0 / ((double)a * (double)b);
    size_t start_a = first + a - 1, start_ab = first + period - 1;
    double *ring = (double *)calloc(b, sizeof(double));
    size_t rb = 0;
    double s1 = 0.0, s2 = 0.0;
    for (size_t i = first; i < len; ++i) {
        s1 += data[i];
        if (i >= start_a) {
            double old = ring[rb];
            s2 = s2 + (s1 - old);
            ring[rb] = s1;
            if (++rb == b) rb = 0;
            if (i >= start_ab) out[i] = s2 * inv_ab;
            s1 -= data[i + 1 - a];
        }
    }
    free(ring);
    return OR_OK;
}

/* smma.rs:243-271 smma_scalar: seed = mean of the first p bars at first+p-1,
 * then prev = (prev*(p-1) + x) / p.  p = 1 copies. */
int or_smma(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 1, &first);
    if (st) return st;
    size_t end = first + period;
    fill_nan(out, end - 1);
    if (period == 1) {
        for (size_t i = first; i < len; ++i) out[i] = data[i];
        return OR_OK;
    }
    double sum = 0.0;
    for (size_t i = first; i < end; ++i) sum += data[i];
    double pf = (double)period, pm1 = pf - 1.0;
    double prev = sum / pf;
    out[end - 1] = prev;
    for (size_t i = end; i < len; ++i) {
        prev = (prev * pm1 + data[i]) / pf;
        out[i] = prev;
    }
    return OR_OK;
}

/* wilders.rs:195-231 (the first p bars after first must be finite) and
 * wilders_scalar :262-320: seed summed in groups of four, y = sum * (1/p),
 * then y = fma(x - y, 1/p, y).  p = 1 copies after the seed bar. */
int or_wilders(const double *data, size_t len, size_t period, double *out) {
    size_t first;
    int st = ma_prepare(data, len, period, 1, &first);
    if (st) return st;
    for (size_t i = 0; i < period; ++i)
        if (!is_finite_bits(data[first + i])) return OR_NOTENOUGH;
    size_t ws = first + period - 1;
    fill_nan(out, ws);
    const double *p = data + first;
    double sum = 0.0;
    size_t c4 = period / 4;
    for (size_t g = 0; g < c4; ++g, p += 4) sum += p[0] + p[1] + p[2] + p[3];
    switch (period - c4 * 4) {
        case 3: sum += p[0] + p[1] + p[2]; break;
        case 2: sum += p[0] + p[1]; break;
        case 1: sum += p[0]; break;
        default: break;
    }
    double inv_n = 1.0 / (double)period;
    double y = sum * inv_n;
    out[ws] = y;
    if (period == 1) {
        for (size_t i = ws + 1; i < len; ++i) out[i] = data[i];
        return OR_OK;
    }
    for (size_t i = ws + 1; i < len; ++i) {
        y = fma(data[i] - y, inv_n, y);
        out[i] = y;
    }
    return OR_OK;
}

/* One-period moving average of every column of a time-major [rows][cols]
 * matrix.  kind: 0 sma, 1 wma, 2 swma, 3 smma, 4 wilders (scalar kernels),
 * 5 sma batch row, 6 wma batch row (prefix forms). */
void or_ma_many(int kind, const double *tm, size_t cols, size_t rows, size_t period,
                double *out_tm, int *status, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) if (nthreads != 1)
#endif
    for (long s = 0; s < (long)cols; ++s) {
        double *x = (double *)malloc(sizeof(double) * rows);
        double *y = (double *)malloc(sizeof(double) * rows);
        for (size_t t = 0; t < rows; ++t) x[t] = tm[t * cols + s];
        fill_nan(y, rows);
        int st = OR_PARAM;
        switch (kind) {
            case 0: st = or_sma(x, rows, period, y); break;
            case 1: st = or_wma(x, rows, period, y); break;
            case 2: st = or_swma(x, rows, period, y); break;
            case 3: st = or_smma(x, rows, period, y); break;
            case 4: st = or_wilders(x, rows, period, y); break;
            case 5: st = or_sma_batch_row(x, rows, period, y); break;
            case 6: st = or_wma_batch_row(x, rows, period, y); break;
            default: break;
        }
        status[s] = st;
        for (size_t t = 0; t < rows; ++t) out_tm[t * cols + s] = y[t];
        free(x);
        free(y);
    }
}
