// ALMA kernels for gfx950 (MI355X).
//
// Semantics follow the reference GPU path (kernels/cuda/moving_averages/
// alma_kernel.cu, launched from src/cuda/moving_averages/alma_wrapper.rs):
// out[c][t] = sum_k w_c[k] * x[t - p_c + 1 + k] * inv_norm_c for
// t >= first_valid + p_c - 1, NaN before.  The CPU oracle is alma.rs:426-463 /
// 1505-1527 (f64).
//
//   alma_batch_f32         one series x many combos, row-major [combo][t];
//                          LDS-staged price tile, 4 outputs per lane, FP32 FMA.
//   alma_batch_f32_tm      same, time-major output [t][combo].
//   alma_ms1p_f32          many series (time-major) x one param.
//   alma_batch_mfma_bf16x3 one series x many combos on the matrix cores: a
//                          Hankel tile of anchored prices (M = 32 bars) times a
//                          per-combo weight matrix (N = 32 combos), split-BF16
//                          hi/lo with three products and FP32 accumulation.
#include "vta_common.h"

// ---------------------------------------------------------------- FP32 sliding dot
//
// Block: 256 lanes, time tile of 1024 bars (4 consecutive outputs per lane),
// ALMA_CB combos per block sharing the staged price tile.  Weights are front
// padded with zeros to a multiple of 4 taps so every price read is an aligned
// 16-byte LDS load; each 4-tap step is 16 FMAs against two float4 of prices.
#define ALMA_TT 1024
#define ALMA_CB 8

// One block's tile: bars [bx * ALMA_TT, +ALMA_TT) of combos [by * ALMA_CB, +ALMA_CB).
__device__ __forceinline__ void alma_f32_tile(const float *__restrict__ prices,
                                              const float *__restrict__ weights,
                                              const int32_t *__restrict__ periods,
                                              const float *__restrict__ inv_norms, int max_period,
                                              int series_len, int n_combos, int first_valid,
                                              float *__restrict__ out, int bx, int by) {
    extern __shared__ float4 f32_tile_lds[];
    float *smem = reinterpret_cast<float *>(f32_tile_lds);
    const int p4max = (max_period + 3) & ~3;
    const int xs_len = p4max + ALMA_TT + 8;
    float *xs = smem;                 // xs[e] = x[t0 - p4max + e]
    float *ws = smem + xs_len;        // [ALMA_CB][p4max]
    const int t0 = bx * ALMA_TT;
    const int c0 = by * ALMA_CB;
    const int tid = threadIdx.x;

    // Staged prices are sanitised (0 outside [first_valid, T) and for
    // non-finite values) because zero-padded taps read bars outside each
    // combo's window; a block whose tile holds a non-finite input at or after
    // first_valid takes the exact per-window path below instead.
    int bad = 0;
    for (int e = tid; e < xs_len; e += 256) {
        long idx = (long)t0 - p4max + e;
        const float x = vta_ld_clamped(prices, idx, first_valid, series_len);
        const bool in = idx >= first_valid && idx < series_len;
        float v = 0.0f;
        if (in) {
            if (vta_isfinite(x)) v = x; else bad = 1;
        }
        xs[e] = v;
    }
    for (int cc = 0; cc < ALMA_CB; ++cc) {
        int c = c0 + cc;
        if (c >= n_combos) break;
        int p = periods[c];
        int p4 = (p + 3) & ~3;
        float inv = inv_norms[c];
        for (int k = tid; k < p4max; k += 256) {
            int src = k - (p4 - p);
            ws[cc * p4max + k] =
                (k < p4 && src >= 0) ? weights[(size_t)c * max_period + src] * inv : 0.0f;
        }
    }
    const int block_bad = __syncthreads_or(bad);

    const int tl = tid * 4;
    const bool vec_ok = ((series_len & 3) == 0) && (t0 + tl + 3 < series_len);
    for (int cc = 0; cc < ALMA_CB; ++cc) {
        const int c = c0 + cc;
        if (c >= n_combos) break;
        const int p = periods[c];
        const int p4 = (p + 3) & ~3;
        const float *w = ws + cc * p4max;
        // output t = t0+tl+j uses x[t - p4 + 1 + k], k < p4 -> xs[tl + j + p4max - p4 + 1 + k]
        const float *xb = xs + tl + (p4max - p4);
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        float4 A = *reinterpret_cast<const float4 *>(xb);
        // xb is 16-byte aligned; the window starts one element in (offset +1).
        for (int k = 0; k < p4 && !block_bad; k += 4) {
            const float4 B = *reinterpret_cast<const float4 *>(xb + k + 4);
            const float4 W = *reinterpret_cast<const float4 *>(w + k);
            // window element q of output j is x at offset (k + 1 + j + q)
            a0 = fmaf(W.x, A.y, a0); a0 = fmaf(W.y, A.z, a0); a0 = fmaf(W.z, A.w, a0); a0 = fmaf(W.w, B.x, a0);
            a1 = fmaf(W.x, A.z, a1); a1 = fmaf(W.y, A.w, a1); a1 = fmaf(W.z, B.x, a1); a1 = fmaf(W.w, B.y, a1);
            a2 = fmaf(W.x, A.w, a2); a2 = fmaf(W.y, B.x, a2); a2 = fmaf(W.z, B.y, a2); a2 = fmaf(W.w, B.z, a2);
            a3 = fmaf(W.x, B.x, a3); a3 = fmaf(W.y, B.y, a3); a3 = fmaf(W.z, B.z, a3); a3 = fmaf(W.w, B.w, a3);
            A = B;
        }
        const int warm = first_valid + p - 1;
        const int t = t0 + tl;
        if (block_bad) {
            // exact window, taps in order, NaN/Inf propagate as in the reference
            const float inv = inv_norms[c];
            const float *wr = weights + (size_t)c * max_period;
            float acc[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                const int tj = t + j;
                if (tj >= warm && tj < series_len)
                    for (int k = 0; k < p; ++k) s = fmaf(wr[k] * inv, prices[tj - p + 1 + k], s);
                acc[j] = s;
            }
            a0 = acc[0]; a1 = acc[1]; a2 = acc[2]; a3 = acc[3];
        }
        float r0 = (t + 0 >= warm) ? a0 : VTA_QNAN_F32;
        float r1 = (t + 1 >= warm) ? a1 : VTA_QNAN_F32;
        float r2 = (t + 2 >= warm) ? a2 : VTA_QNAN_F32;
        float r3 = (t + 3 >= warm) ? a3 : VTA_QNAN_F32;
        float *row = out + (size_t)c * series_len;
        if (vec_ok) {
            *reinterpret_cast<float4 *>(row + t) = make_float4(r0, r1, r2, r3);
        } else {
            if (t + 0 < series_len) row[t + 0] = r0;
            if (t + 1 < series_len) row[t + 1] = r1;
            if (t + 2 < series_len) row[t + 2] = r2;
            if (t + 3 < series_len) row[t + 3] = r3;
        }
    }
}

extern "C" __global__ __launch_bounds__(256) void alma_batch_f32(
    const float *__restrict__ prices, const float *__restrict__ weights,
    const int32_t *__restrict__ periods, const float *__restrict__ inv_norms, int max_period,
    int series_len, int n_combos, int first_valid, float *__restrict__ out) {
    alma_f32_tile(prices, weights, periods, inv_norms, max_period, series_len, n_combos,
                  first_valid, out, blockIdx.x, blockIdx.y);
}

// Packed FP32 sweeps: several one-series batches (uneven lengths and grids) in
// one launch.  Block b runs tile (items[b].y, items[b].z) of job items[b].x, so
// every job's outputs are those of its own alma_batch_f32 launch.
struct alma_f32_job {
    const float *prices;
    const float *weights;        // [n_combos][max_period]
    const int32_t *periods;
    const float *inv_norms;
    float *out;                  // [n_combos][series_len]
    int max_period;
    int series_len;
    int n_combos;
    int first_valid;
};

extern "C" __global__ __launch_bounds__(256) void alma_batch_f32_pack(
    const alma_f32_job *__restrict__ jobs, const int4 *__restrict__ items) {
    const int4 it = items[blockIdx.x];
    const alma_f32_job j = jobs[it.x];
    alma_f32_tile(j.prices, j.weights, j.periods, j.inv_norms, j.max_period, j.series_len,
                  j.n_combos, j.first_valid, j.out, it.y, it.z);
}

// ---------------------------------------------------------------- FP32, time-major out
//
// Lanes map to combos so each output row [t][c0..c0+63] is one coalesced
// store.  Block: 64 combos x 4 waves; each wave walks 64 bars of the tile with
// four outputs in flight.  Weights are stored tap-major in LDS ([k][64]) so a
// wave reads 64 consecutive floats per tap; prices are wave-uniform broadcasts.
#define ALMA_TM_TT 256
extern "C" __global__ __launch_bounds__(256) void alma_batch_f32_tm(
    const float *__restrict__ prices, const float *__restrict__ weights,
    const int32_t *__restrict__ periods, const float *__restrict__ inv_norms, int max_period,
    int series_len, int n_combos, int first_valid, float *__restrict__ out_tm) {
    extern __shared__ float4 smem4[];
    float *smem = reinterpret_cast<float *>(smem4);
    const int xs_len = max_period + ALMA_TM_TT + 4;
    float *xs = smem;               // xs[e] = x[t0 - max_period + 1 + e]
    float *ws = smem + xs_len;      // ws[k*64 + lane] = w_c[k - (max_period - p_c)]
    const int t0 = blockIdx.x * ALMA_TM_TT;
    const int c0 = blockIdx.y * 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    int bad = 0;
    for (int e = tid; e < xs_len; e += 256) {
        long idx = (long)t0 - max_period + 1 + e;
        float v = 0.0f;
        if (idx >= first_valid && idx < series_len) {
            float x = prices[idx];
            if (vta_isfinite(x)) v = x; else bad = 1;
        }
        xs[e] = v;
    }
    for (int i = tid; i < max_period * 64; i += 256) {
        int k = i >> 6, cl = i & 63, c = c0 + cl;
        float v = 0.0f;
        if (c < n_combos) {
            int p = periods[c];
            int src = k - (max_period - p);
            if (src >= 0) v = weights[(size_t)c * max_period + src] * inv_norms[c];
        }
        ws[i] = v;
    }
    const int block_bad = __syncthreads_or(bad);
    const int c = c0 + lane;
    const int p = (c < n_combos) ? periods[c] : 1;
    const int warm = first_valid + p - 1;
    for (int tb = wv * 4; tb < ALMA_TM_TT; tb += 16) {
        // outputs t0+tb .. t0+tb+3, window xs[tb + j + k], k < max_period
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        float x0 = xs[tb], x1 = xs[tb + 1], x2 = xs[tb + 2];
        for (int k = 0; k < max_period; ++k) {
            float x3 = xs[tb + k + 3];
            float w = ws[k * 64 + lane];
            a0 = fmaf(w, x0, a0);
            a1 = fmaf(w, x1, a1);
            a2 = fmaf(w, x2, a2);
            a3 = fmaf(w, x3, a3);
            x0 = x1; x1 = x2; x2 = x3;
        }
        if (c < n_combos) {
            float acc[4] = {a0, a1, a2, a3};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int t = t0 + tb + j;
                if (block_bad && t >= warm && t < series_len) {
                    const float *wr = weights + (size_t)c * max_period;
                    const float inv = inv_norms[c];
                    float s = 0.0f;
                    for (int k = 0; k < p; ++k) s = fmaf(wr[k] * inv, prices[t - p + 1 + k], s);
                    acc[j] = s;
                }
                if (t < series_len)
                    out_tm[(size_t)t * n_combos + c] = (t >= warm) ? acc[j] : VTA_QNAN_F32;
            }
        }
    }
}

// ---------------------------------------------------------------- many series, one param
//
// Time-major [T][N] input: lanes map to series (coalesced rows).  A block
// stages a [MS_TT + period - 1][64] tile in LDS (MS_TT, a multiple of 16, is
// chosen by the host to fit the LDS budget); each lane produces 4 consecutive
// bars of its series per step.  Weights are wave-uniform.
extern "C" __global__ __launch_bounds__(256) void alma_ms1p_f32(
    const float *__restrict__ prices_tm, const float *__restrict__ weights, int period,
    float inv_norm, int num_series, int series_len, const int32_t *__restrict__ first_valids,
    float *__restrict__ out_tm, int MS_TT) {
    extern __shared__ float4 smem4[];
    float *tile = reinterpret_cast<float *>(smem4);   // [(MS_TT + period - 1)][64]
    float *ws = tile + (size_t)(MS_TT + period - 1) * 64;
    const int t0 = blockIdx.x * MS_TT;
    const int s0 = blockIdx.y * 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int rows = MS_TT + period - 1;
    for (int i = tid; i < rows * 64; i += 256) {
        int r = i >> 6, sl = i & 63;
        long t = (long)t0 - period + 1 + r;
        int s = s0 + sl;
        tile[i] = (t >= 0 && t < series_len && s < num_series)
                      ? prices_tm[(size_t)t * num_series + s] : 0.0f;
    }
    for (int k = tid; k < period; k += 256) ws[k] = weights[k] * inv_norm;
    __syncthreads();
    const int s = s0 + lane;
    if (s >= num_series) return;
    const int warm = first_valids[s] + period - 1;
    for (int tb = wv * 4; tb < MS_TT; tb += 16) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        float x0 = tile[(tb + 0) * 64 + lane];
        float x1 = tile[(tb + 1) * 64 + lane];
        float x2 = tile[(tb + 2) * 64 + lane];
        for (int k = 0; k < period; ++k) {
            float x3 = (tb + k + 3 < rows) ? tile[(tb + k + 3) * 64 + lane] : 0.0f;
            float w = ws[k];
            a0 = fmaf(w, x0, a0);
            a1 = fmaf(w, x1, a1);
            a2 = fmaf(w, x2, a2);
            a3 = fmaf(w, x3, a3);
            x0 = x1; x1 = x2; x2 = x3;
        }
        float acc[4] = {a0, a1, a2, a3};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int t = t0 + tb + j;
            if (t < series_len)
                out_tm[(size_t)t * num_series + s] = (t >= warm) ? acc[j] : VTA_QNAN_F32;
        }
    }
}

// ---------------------------------------------------------------- MFMA split-BF16
//
// Problem: out[row(n)][t] = sum_{k<K} W[k][n] * x[t - K + 1 + k] for the 32
// combos n of an N-block, where K (a multiple of 16) covers the block's longest
// period and shorter periods are zero padded at the front.  Each wave owns one
// N-block and a run of MF_TILES_PER_WAVE 32-bar M-tiles.  It loads the prices
// of its whole run into LDS once, then per tile
//   1. stages x' = x - a (a = one anchor price per tile) as BF16 hi and
//      lo = bf16(x' - hi) in two copies: copy 0 holds x'[0..], copy 1 holds
//      x'[1..], so every lane's 8-element A-fragment starts on a dword
//      boundary in one of them whatever the lane's row parity;
//   2. runs K/16 steps of v_mfma_f32_32x32x16_bf16 with the Hankel tile as
//      the A operand and the weights as B, so D is [bar][combo]: hi*hi into
//      one accumulator, hi*lo + lo*hi into a second;
//   3. writes out = (acc_hi + acc_corr) + a (weights sum to 1), NaN before the
//      warm-up of each combo, as 16-byte stores of 4 consecutive bars.
// (The transposed [combo][bar] D layout makes every store a full 128-byte row
// segment but needs four times the store instructions; it measured ~8% slower.)
// Split error: |x' - hi - lo| <= 2^-18 |x'|, same for the weights, dropped
// lo*lo <= 2^-18 |w x'|; anchoring makes |x'| a few percent of the price, so
// the contraction carries FP32-level relative accuracy.  Tiles that contain a
// non-finite input at or after first_valid are recomputed with a plain FP32
// dot from the f32 weight table so NaN/Inf behave like the FP32 kernel.
//
// Operand layouts (v_mfma_f32_32x32x16_bf16, wave64):
//   A (32x16): lane l holds A[l%32][8*(l/32) + j], j = 0..7   (Hankel x')
//   B (16x32): lane l holds B[8*(l/32) + j][l%32]             (weights W)
//   D (32x32): lane l, element i holds D[8*(i/4) + 4*(l/32) + i%4][l%32]
typedef __attribute__((ext_vector_type(8))) __bf16 vta_bf16x8;
typedef __attribute__((ext_vector_type(16))) float vta_f32x16;

// R = 32-bar M-tiles per wave; one kernel per R (8, 16, 32), picked at launch.
#define MF_XS_MAX(R) ((((R) * 32 + 256 + 63) / 64) * 64)
// One staged array (a copy of hi or lo) holds K/2 + 15 dwords; 160 dwords
// covers K <= 256.  The four arrays sit 168 dwords apart so the two copies a
// wave reads in one instruction fall on different LDS banks.
#define MF_ARR_STRIDE 168
#define MF_ARR_WORDS (4 * MF_ARR_STRIDE)

struct alma_mf_args {
    const float *prices;
    const uint4 *frags;          // [sum_b ks_b*2*64] 16-byte lane chunks
    const int4 *nblocks;         // {ks, frag_off, combo_base, 0}
    const int32_t *sorted_period;
    const int32_t *sorted_row;   // output row, -1 for padding
    const float *w_f32;          // [ncombos_sorted][kmax] f32 weights (fallback path)
    float *out;
    int series_len;
    int first_valid;
    int n_nblocks;
    int tiles_per_wave;
    int n_tiles;
    int kmax;
};

__device__ __forceinline__ uint32_t mf_pack(uint16_t lo_elem, uint16_t hi_elem) {
    return (uint32_t)lo_elem | ((uint32_t)hi_elem << 16);
}

template <int KS, int R>
__device__ __forceinline__ void alma_mf_wave(const alma_mf_args &a, int nb, int tile_begin,
                                             uint32_t *arr, float *xs) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;    // K-half for A/B, row-quad selector for D
    const int m = lane & 31;    // A row (bar in tile) / B,D column (combo)
    const int4 nbd = a.nblocks[nb];
    const int frag_off = nbd.y;
    const int combo_base = nbd.z;
    constexpr int K = KS * 16;
    constexpr int NPAIR = K / 2 + 15;          // dwords staged per array
    uint32_t *hi0 = arr;                       // copy 0 (x'[2i], x'[2i+1])
    uint32_t *hi1 = arr + MF_ARR_STRIDE;       // copy 1 (x'[2i+1], x'[2i+2])
    uint32_t *lo0 = arr + 2 * MF_ARR_STRIDE;
    uint32_t *lo1 = arr + 3 * MF_ARR_STRIDE;

    // B fragments stay in registers for all tiles of this wave.
    uint4 bhi[KS], blo[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bhi[s] = a.frags[frag_off + (2 * s + 0) * 64 + lane];
        blo[s] = a.frags[frag_off + (2 * s + 1) * 64 + lane];
    }
    const int T = a.series_len;
    const int fv = a.first_valid;
    const int my_period = a.sorted_period[combo_base + m];
    const int my_row = a.sorted_row[combo_base + m];
    const int my_warm = fv + my_period - 1;

    const int tile_end = min(tile_begin + R, a.n_tiles);
    // ---- prologue: the wave's whole price range x[t_lo, t_lo + NX) into LDS
    // (0 outside [first_valid, T), non-finite values kept), one batch of loads.
    const int t_lo = tile_begin * 32 - K + 1;
    constexpr int NXI = (R * 32 + K + 63) / 64;
    {
        float v[NXI];
#pragma unroll
        for (int i = 0; i < NXI; ++i) {
            const int idx = t_lo + lane + 64 * i;
            const float x = vta_ld_clamped(a.prices, idx, fv, T);
            v[i] = (idx >= fv && idx < T) ? x : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < NXI; ++i) xs[lane + 64 * i] = v[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // A-fragment dword offset of this lane inside a copy, K-step 0
    const int r = m & 1;
    const int d0 = (m >> 1) + 4 * g;
    const uint32_t *ahi_base = (r ? hi1 : hi0) + d0;
    const uint32_t *alo_base = (r ? lo1 : lo0) + d0;

    for (int tile = tile_begin; tile < tile_end; ++tile) {
        const int t_tile = tile * 32;
        const int xoff = t_tile - K + 1 - t_lo;   // xs index of x'[0]
        const int ai = min(t_tile + 31, T - 1);   // anchor: the tile's last in-range bar
        float anchor = xs[ai - t_lo];
        if (!vta_isfinite(anchor)) anchor = 0.0f;

        // ---- stage pairs i: copy0[i] = (x'[2i], x'[2i+1]), copy1[i] = (x'[2i+1], x'[2i+2])
        bool bad = false;
#pragma unroll
        for (int it = 0; it < (NPAIR + 63) / 64; ++it) {
            const int i = lane + 64 * it;
            if (i >= NPAIR) break;
            const float x0 = xs[xoff + 2 * i];
            const float x1 = xs[xoff + 2 * i + 1];
            const float x2 = xs[xoff + 2 * i + 2];
            float v0 = x0 - anchor, v1 = x1 - anchor, v2 = x2 - anchor;
            if (!vta_isfinite(x0)) { bad = true; v0 = 0.0f; }
            if (!vta_isfinite(x1)) { bad = true; v1 = 0.0f; }
            if (!vta_isfinite(x2)) { bad = true; v2 = 0.0f; }
            const uint16_t h0 = vta_bf16_rne(v0), h1 = vta_bf16_rne(v1), h2 = vta_bf16_rne(v2);
            const uint16_t l0 = vta_bf16_rne(v0 - vta_bf16_to_f32(h0));
            const uint16_t l1 = vta_bf16_rne(v1 - vta_bf16_to_f32(h1));
            const uint16_t l2 = vta_bf16_rne(v2 - vta_bf16_to_f32(h2));
            hi0[i] = mf_pack(h0, h1);
            hi1[i] = mf_pack(h1, h2);
            lo0[i] = mf_pack(l0, l1);
            lo1[i] = mf_pack(l1, l2);
        }
        const bool tile_bad = __any(bad);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        // ---- contraction: lane chunk at K-step s starts at x'[m + 16s + 8g],
        // i.e. dword d0 + 8s of copy (m & 1)
        vta_f32x16 acc_h = {};
        vta_f32x16 acc_c = {};
        __builtin_amdgcn_s_setprio(1);   // favour the wave that is feeding the matrix cores
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            uint4 ahi_u, alo_u;
            ahi_u.x = ahi_base[8 * s + 0]; ahi_u.y = ahi_base[8 * s + 1];
            ahi_u.z = ahi_base[8 * s + 2]; ahi_u.w = ahi_base[8 * s + 3];
            alo_u.x = alo_base[8 * s + 0]; alo_u.y = alo_base[8 * s + 1];
            alo_u.z = alo_base[8 * s + 2]; alo_u.w = alo_base[8 * s + 3];
            const vta_bf16x8 ahi = __builtin_bit_cast(vta_bf16x8, ahi_u);
            const vta_bf16x8 alo = __builtin_bit_cast(vta_bf16x8, alo_u);
            const vta_bf16x8 bh = __builtin_bit_cast(vta_bf16x8, bhi[s]);
            const vta_bf16x8 bl = __builtin_bit_cast(vta_bf16x8, blo[s]);
            acc_h = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bh, acc_h, 0, 0, 0);
            acc_c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bl, acc_c, 0, 0, 0);
            acc_c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bh, acc_c, 0, 0, 0);
        }
        __builtin_amdgcn_s_setprio(0);
        // LDS is rewritten by the next tile: make sure every lane's reads are done.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        // ---- epilogue: lane owns combo m, bars 8q + 4g + (0..3)
        if (my_row >= 0) {
            float *row = a.out + (size_t)my_row * (size_t)T;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tb = t_tile + 8 * q + 4 * g;
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int t = tb + i;
                    float val = (acc_h[4 * q + i] + acc_c[4 * q + i]) + anchor;
                    if (tile_bad && t >= my_warm && t < T) {
                        // plain FP32 dot, same taps as the FP32 kernel
                        const float *wrow = a.w_f32 + (size_t)(combo_base + m) * a.kmax;
                        float acc = 0.0f;
                        for (int k = 0; k < my_period; ++k)
                            acc = fmaf(wrow[k], a.prices[t - my_period + 1 + k], acc);
                        val = acc;
                    }
                    v[i] = (t >= my_warm) ? val : VTA_QNAN_F32;
                }
                if (((T & 3) == 0) && tb + 3 < T) {
                    *reinterpret_cast<float4 *>(row + tb) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (tb + i < T) row[tb + i] = v[i];
                }
            }
        }
    }
}

// grid.x = number of tile runs, grid.y = number of N-blocks; block = 64 lanes.
template <int R>
__device__ __forceinline__ void alma_mf_entry(const alma_mf_args &a) {
    __shared__ __attribute__((aligned(16))) uint32_t arr[MF_ARR_WORDS];
    __shared__ __attribute__((aligned(16))) float xs[MF_XS_MAX(R)];
    const int nb = blockIdx.y;
    const int tile_begin = blockIdx.x * R;
    const int ks = a.nblocks[nb].x;
    switch (ks) {
#define MF_CASE(N) case N: alma_mf_wave<N, R>(a, nb, tile_begin, arr, xs); break;
        MF_CASE(1) MF_CASE(2) MF_CASE(3) MF_CASE(4) MF_CASE(5) MF_CASE(6) MF_CASE(7) MF_CASE(8)
        MF_CASE(9) MF_CASE(10) MF_CASE(11) MF_CASE(12) MF_CASE(13) MF_CASE(14) MF_CASE(15)
        MF_CASE(16)
#undef MF_CASE
        default: break;
    }
}

extern "C" __global__ __launch_bounds__(64) void alma_batch_mfma_bf16x3(alma_mf_args a) {
    alma_mf_entry<16>(a);
}
extern "C" __global__ __launch_bounds__(64) void alma_batch_mfma_bf16x3_r8(alma_mf_args a) {
    alma_mf_entry<8>(a);
}
extern "C" __global__ __launch_bounds__(64) void alma_batch_mfma_bf16x3_r32(alma_mf_args a) {
    alma_mf_entry<32>(a);
}

// ---------------------------------------------------------------- MFMA, shared staging
//
// Same contraction, organised so the staging is shared: a workgroup of four
// waves takes up to four consecutive N-blocks (sorted by period, so their K
// differ little) over one run of R tiles.  The workgroup loads the run's prices
// once and stages every tile's hi/lo copies, sized for the group's largest K
// (K_g), into LDS with all 256 lanes; after one barrier each wave contracts its
// own N-block over all R tiles with no further synchronisation.  A wave whose
// K is below K_g starts its window K_g - K elements into the staged tile (a
// multiple of 16 elements, so the dword alignment of the two copies holds).
// The four arrays of a tile are `stride` dwords apart with stride = 16 mod 32,
// which puts the two copies a wave reads in one instruction on disjoint banks.

struct alma_mfg_args {
    const float *prices;
    const uint4 *frags;
    const int4 *nblocks;         // {ks, frag_off, combo_base, 0}
    const int4 *groups;          // {first N-block, count (<= 4), ks of the group, 0}
    const int32_t *sorted_period;
    const int32_t *sorted_row;
    const float *w_f32;
    float *out;
    int series_len;
    int first_valid;
    int n_groups;                // groups of this launch: group_base .. group_base + n_groups
    int n_tiles;
    int kmax;
    int group_base;
};

__device__ __forceinline__ int mfg_stride(int npair) { return ((npair + 15) >> 5) * 32 + 16; }

template <int KS, int R>
__device__ __forceinline__ void alma_mfg_wave(const alma_mfg_args &a, int nb, int tile_begin,
                                              int nt, int kg, int stride,
                                              const uint32_t *stage, const float *anch,
                                              const int *badf, float *tbuf) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int m = lane & 31;
    const int4 nbd = a.nblocks[nb];
    const int frag_off = nbd.y;
    const int combo_base = nbd.z;
    constexpr int K = KS * 16;
    uint4 bhi[KS], blo[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bhi[s] = a.frags[frag_off + (2 * s + 0) * 64 + lane];
        blo[s] = a.frags[frag_off + (2 * s + 1) * 64 + lane];
    }
    const int T = a.series_len;
    const int my_period = a.sorted_period[combo_base + m];
    const int my_row = a.sorted_row[combo_base + m];
    const int my_warm = a.first_valid + my_period - 1;
    int warm_max = my_warm;                       // wave max: tiles past it need no NaN select
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) warm_max = max(warm_max, __shfl_xor(warm_max, o));
    warm_max = __builtin_amdgcn_readfirstlane(warm_max);
    const int r = m & 1;
    const int d0 = (m >> 1) + 4 * g + ((kg - K) >> 1);
    const int hi_off = (r ? stride : 0) + d0;
    const int lo_off = 2 * stride + hi_off;
    float *row = a.out + (size_t)(my_row < 0 ? 0 : my_row) * (size_t)T;
    // Steady-state tiles leave through the wave's LDS tile `tbuf` ([32 combos]
    // [32 bars], rows 36 floats apart): each lane holds 16 bars of one combo
    // in four 16-byte pieces, and after the transpose each store instruction
    // writes 8 combos x 128 contiguous bytes (whole lines) instead of 32
    // combos x 32 bytes.  Lane L stores bars 4 (L & 7) .. +3 of combos
    // (L >> 3) + 8 s, s = 0..3.
    const int tcol = 4 * (lane & 7);
    int trow[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) trow[s] = __shfl(my_row, (lane >> 3) + 8 * s);

    // A fragments of K-step s of tile j; the next step's pair is loaded while
    // the current step's three MFMAs run (and tile j+1's first step during
    // tile j's epilogue).
    auto load_a = [&](int j, int s, uint4 &h, uint4 &l) {
        const uint32_t *hb = stage + j * 4 * stride + hi_off + 8 * s;
        const uint32_t *lb = stage + j * 4 * stride + lo_off + 8 * s;
        h.x = hb[0]; h.y = hb[1]; h.z = hb[2]; h.w = hb[3];
        l.x = lb[0]; l.y = lb[1]; l.z = lb[2]; l.w = lb[3];
    };
    uint4 ah_n, al_n;
    load_a(0, 0, ah_n, al_n);
    for (int j = 0; j < nt; ++j) {
        // all three products into one FP32 accumulator: the corrections are
        // rounded at the same 2^-24 |acc| step as the main terms either way
        vta_f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint4 ah = ah_n, al = al_n;
            if (s + 1 < KS) load_a(j, s + 1, ah_n, al_n);
            const vta_bf16x8 ahi = __builtin_bit_cast(vta_bf16x8, ah);
            const vta_bf16x8 alo = __builtin_bit_cast(vta_bf16x8, al);
            const vta_bf16x8 bh = __builtin_bit_cast(vta_bf16x8, bhi[s]);
            const vta_bf16x8 bl = __builtin_bit_cast(vta_bf16x8, blo[s]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bh, acc, 0, 0, 0);
        }
        if (j + 1 < nt) load_a(j + 1, 0, ah_n, al_n);
        const float anchor = anch[j];
        const int t_tile = (tile_begin + j) * 32;
        if (badf[j] == 0 && t_tile >= warm_max && t_tile + 32 <= T && (T & 3) == 0) {
            // steady state (wave-uniform): every bar of the tile is a finite
            // output; transpose through LDS, then whole-line stores
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *reinterpret_cast<float4 *>(tbuf + m * 36 + 8 * q + 4 * g) =
                    make_float4(acc[4 * q + 0] + anchor, acc[4 * q + 1] + anchor,
                                acc[4 * q + 2] + anchor, acc[4 * q + 3] + anchor);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const float4 v =
                    *reinterpret_cast<const float4 *>(tbuf + ((lane >> 3) + 8 * s) * 36 + tcol);
                if (trow[s] >= 0)
                    *reinterpret_cast<float4 *>(a.out + (size_t)trow[s] * (size_t)T + t_tile +
                                                tcol) = v;
            }
            continue;
        }
        if (my_row < 0) continue;
        const bool tile_bad = badf[j] != 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int tb = t_tile + 8 * q + 4 * g;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = tb + i;
                float val = acc[4 * q + i] + anchor;
                if (tile_bad && t >= my_warm && t < T) {
                    // plain FP32 dot, same taps as the FP32 kernel
                    const float *wrow = a.w_f32 + (size_t)(combo_base + m) * a.kmax;
                    float dot = 0.0f;
                    for (int k = 0; k < my_period; ++k)
                        dot = fmaf(wrow[k], a.prices[t - my_period + 1 + k], dot);
                    val = dot;
                }
                v[i] = (t >= my_warm) ? val : VTA_QNAN_F32;
            }
            if (((T & 3) == 0) && tb + 3 < T) {
                *reinterpret_cast<float4 *>(row + tb) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (tb + i < T) row[tb + i] = v[i];
            }
        }
    }
}

// Two tiles per pass: the two accumulators are independent MFMA chains, so the
// matrix core always has a second chain to issue while one waits, and the two
// epilogues put eight stores in flight.  An odd last tile recomputes tile
// nt-1 as its partner and skips the partner's stores.
template <int KS, int R>
__device__ __forceinline__ void alma_mfg_wave2(const alma_mfg_args &a, int nb, int tile_begin,
                                               int nt, int kg, int stride,
                                               const uint32_t *stage, const float *anch,
                                               const int *badf) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int m = lane & 31;
    const int4 nbd = a.nblocks[nb];
    const int frag_off = nbd.y;
    const int combo_base = nbd.z;
    constexpr int K = KS * 16;
    uint4 bhi[KS], blo[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bhi[s] = a.frags[frag_off + (2 * s + 0) * 64 + lane];
        blo[s] = a.frags[frag_off + (2 * s + 1) * 64 + lane];
    }
    const int T = a.series_len;
    const int my_period = a.sorted_period[combo_base + m];
    const int my_row = a.sorted_row[combo_base + m];
    const int my_warm = a.first_valid + my_period - 1;
    int warm_max = my_warm;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) warm_max = max(warm_max, __shfl_xor(warm_max, o));
    warm_max = __builtin_amdgcn_readfirstlane(warm_max);
    const int r = m & 1;
    const int d0 = (m >> 1) + 4 * g + ((kg - K) >> 1);
    const int hi_off = (r ? stride : 0) + d0;
    const int lo_off = 2 * stride + hi_off;
    float *row = a.out + (size_t)(my_row < 0 ? 0 : my_row) * (size_t)T;
    auto load_a = [&](int j, int s, uint4 &h, uint4 &l) {
        const uint32_t *hb = stage + j * 4 * stride + hi_off + 8 * s;
        const uint32_t *lb = stage + j * 4 * stride + lo_off + 8 * s;
        h.x = hb[0]; h.y = hb[1]; h.z = hb[2]; h.w = hb[3];
        l.x = lb[0]; l.y = lb[1]; l.z = lb[2]; l.w = lb[3];
    };
    auto epilogue = [&](const vta_f32x16 &acc, int j) {
        const float anchor = anch[j];
        const int t_tile = (tile_begin + j) * 32;
        if (badf[j] == 0 && t_tile >= warm_max && t_tile + 32 <= T && (T & 3) == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tb = t_tile + 8 * q + 4 * g;
                *reinterpret_cast<float4 *>(row + tb) =
                    make_float4(acc[4 * q + 0] + anchor, acc[4 * q + 1] + anchor,
                                acc[4 * q + 2] + anchor, acc[4 * q + 3] + anchor);
            }
            return;
        }
        const bool tile_bad = badf[j] != 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int tb = t_tile + 8 * q + 4 * g;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = tb + i;
                float val = acc[4 * q + i] + anchor;
                if (tile_bad && t >= my_warm && t < T) {
                    const float *wrow = a.w_f32 + (size_t)(combo_base + m) * a.kmax;
                    float dot = 0.0f;
                    for (int k = 0; k < my_period; ++k)
                        dot = fmaf(wrow[k], a.prices[t - my_period + 1 + k], dot);
                    val = dot;
                }
                v[i] = (t >= my_warm) ? val : VTA_QNAN_F32;
            }
            if (((T & 3) == 0) && tb + 3 < T) {
                *reinterpret_cast<float4 *>(row + tb) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (tb + i < T) row[tb + i] = v[i];
            }
        }
    };
    uint4 ah0_n, al0_n, ah1_n, al1_n;
    load_a(0, 0, ah0_n, al0_n);
    load_a(min(1, nt - 1), 0, ah1_n, al1_n);
    for (int j = 0; j < nt; j += 2) {
        const int j1 = min(j + 1, nt - 1);
        vta_f32x16 acc0 = {};
        vta_f32x16 acc1 = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint4 ah0 = ah0_n, al0 = al0_n, ah1 = ah1_n, al1 = al1_n;
            if (s + 1 < KS) {
                load_a(j, s + 1, ah0_n, al0_n);
                load_a(j1, s + 1, ah1_n, al1_n);
            }
            const vta_bf16x8 bh = __builtin_bit_cast(vta_bf16x8, bhi[s]);
            const vta_bf16x8 bl = __builtin_bit_cast(vta_bf16x8, blo[s]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, ah0), bh, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, ah1), bh, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, ah0), bl, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, ah1), bl, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, al0), bh, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, al1), bh, acc1, 0, 0, 0);
        }
        if (j + 2 < nt) {
            load_a(j + 2, 0, ah0_n, al0_n);
            load_a(min(j + 3, nt - 1), 0, ah1_n, al1_n);
        }
        if (my_row < 0) continue;
        epilogue(acc0, j);
        if (j1 != j) epilogue(acc1, j1);
    }
}

// grid.x = tile runs of R tiles, grid.y = N-block groups; block = 256 (4 waves).
// Dynamic LDS (dwords): XSN = 32R + kmax + 32 prices, R anchors, R flags, then
// R * 4 * mfg_stride(kmax/2 + 15) staged dwords.
template <int R, int TJ, int KSLO, int KSHI>
__device__ __forceinline__ void alma_mfg_entry(const alma_mfg_args &a) {
    extern __shared__ __attribute__((aligned(16))) uint32_t mfg_lds[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    // groups are sorted by K ascending; dispatch the heaviest first so the
    // last workgroups to start are the short ones
    const int4 gd = a.groups[a.group_base + a.n_groups - 1 - (int)blockIdx.y];
    const int kg = gd.z * 16;
    const int npair = kg / 2 + 15;
    const int stride = mfg_stride(npair);
    const int xsn = 32 * R + a.kmax + 32;
    float *xs = reinterpret_cast<float *>(mfg_lds);
    float *anch = xs + xsn;
    int *badf = reinterpret_cast<int *>(anch + R);
    uint32_t *stage = reinterpret_cast<uint32_t *>(badf + R);
    // per-wave output transpose tiles after the staged dwords (16-byte aligned)
    float *tbuf = reinterpret_cast<float *>(stage) +
                  ((R * 4 * mfg_stride(a.kmax / 2 + 15) + 3) & ~3) + wave * 32 * 36;
    const int T = a.series_len;
    const int fv = a.first_valid;
    const int tile_begin = blockIdx.x * R;
    const int nt = min(R, a.n_tiles - tile_begin);
    const int t_lo = tile_begin * 32 - kg + 1;
    const int nx = nt * 32 + kg;            // staged reads reach index 32 nt + kg - 2
    for (int i = tid; i < nx; i += 256) {
        const int idx = t_lo + i;
        const float x = vta_ld_clamped(a.prices, idx, fv, T);
        xs[i] = (idx >= fv && idx < T) ? x : 0.0f;
    }
    if (tid < R) badf[tid] = 0;
    __syncthreads();
    if (tid < nt) {
        // anchor: the tile's last in-range bar (xs index of bar t is t - t_lo)
        const int ai = min((tile_begin + tid) * 32 + 31, T - 1);
        const float an = xs[ai - t_lo];
        anch[tid] = vta_isfinite(an) ? an : 0.0f;
    }
    __syncthreads();
    // staging unit u: tile j = u / npair, pair i = u % npair (u < 16 * 143, so
    // the float quotient of u + 0.5 is never within rounding of an integer)
    const float inv_np = 1.0f / (float)npair;
    const int units = nt * npair;
    for (int u = tid; u < units; u += 256) {
        const int j = (int)(((float)u + 0.5f) * inv_np);
        const int i = u - j * npair;
        const float anchor = anch[j];
        const int xo = 32 * j + 2 * i;
        const float x0 = xs[xo], x1 = xs[xo + 1], x2 = xs[xo + 2];
        float v0 = x0 - anchor, v1 = x1 - anchor, v2 = x2 - anchor;
        bool bad = false;
        if (!vta_isfinite(x0)) { bad = true; v0 = 0.0f; }
        if (!vta_isfinite(x1)) { bad = true; v1 = 0.0f; }
        if (!vta_isfinite(x2)) { bad = true; v2 = 0.0f; }
        if (bad) badf[j] = 1;
        const uint16_t h0 = vta_bf16_rne(v0), h1 = vta_bf16_rne(v1), h2 = vta_bf16_rne(v2);
        const uint16_t l0 = vta_bf16_rne(v0 - vta_bf16_to_f32(h0));
        const uint16_t l1 = vta_bf16_rne(v1 - vta_bf16_to_f32(h1));
        const uint16_t l2 = vta_bf16_rne(v2 - vta_bf16_to_f32(h2));
        uint32_t *base = stage + j * 4 * stride + i;
        base[0] = mf_pack(h0, h1);
        base[stride] = mf_pack(h1, h2);
        base[2 * stride] = mf_pack(l0, l1);
        base[3 * stride] = mf_pack(l1, l2);
    }
    __syncthreads();
    if (wave >= gd.y) return;
    const int nb = gd.x + wave;
    switch (a.nblocks[nb].x) {
#define MFG_CASE(N)                                                                        \
    case N:                                                                                \
        if (N < KSLO || N > KSHI) break;                                                   \
        if (TJ == 2)                                                                       \
            alma_mfg_wave2<N, R>(a, nb, tile_begin, nt, kg, stride, stage, anch, badf);    \
        else                                                                               \
            alma_mfg_wave<N, R>(a, nb, tile_begin, nt, kg, stride, stage, anch, badf, tbuf); \
        break;
        MFG_CASE(1) MFG_CASE(2) MFG_CASE(3) MFG_CASE(4) MFG_CASE(5) MFG_CASE(6) MFG_CASE(7)
        MFG_CASE(8) MFG_CASE(9) MFG_CASE(10) MFG_CASE(11) MFG_CASE(12) MFG_CASE(13)
        MFG_CASE(14) MFG_CASE(15) MFG_CASE(16)
#undef MFG_CASE
        default: break;
    }
}

extern "C" __global__ __launch_bounds__(256) void alma_batch_mfma_bf16x3_wg8(alma_mfg_args a) {
    alma_mfg_entry<8, 1, 1, 16>(a);
}
extern "C" __global__ __launch_bounds__(256) void alma_batch_mfma_bf16x3_wg16(alma_mfg_args a) {
    alma_mfg_entry<16, 1, 1, 16>(a);
}
extern "C" __global__ __launch_bounds__(256) void alma_batch_mfma_bf16x3_wg16_t2(alma_mfg_args a) {
    alma_mfg_entry<16, 2, 1, 16>(a);
}
// Short-K kernels: groups whose largest K is <= 128 can run in a separate
// launch that is not held to the long-K register budget (the long-K launch
// uses the full-range kernels above, since its first group may hold short
// N-blocks too).
extern "C" __global__ __launch_bounds__(256) void alma_batch_mfma_bf16x3_wg16_klo(alma_mfg_args a) {
    alma_mfg_entry<16, 1, 1, 8>(a);
}
extern "C" __global__ __launch_bounds__(256) void alma_batch_mfma_bf16x3_wg16_t2_klo(alma_mfg_args a) {
    alma_mfg_entry<16, 2, 1, 8>(a);
}

// ---------------------------------------------------------------- direct fallbacks
//
// One output per lane, window read straight from global memory.  Used when a
// period is too long for the LDS tiles above (launch-time choice on the host).
extern "C" __global__ __launch_bounds__(256) void alma_batch_f32_direct(
    const float *__restrict__ prices, const float *__restrict__ weights,
    const int32_t *__restrict__ periods, const float *__restrict__ inv_norms, int max_period,
    int series_len, int n_combos, int first_valid, float *__restrict__ out, int time_major) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)series_len * (size_t)n_combos;
    if (i >= total) return;
    int c, t;
    if (time_major) { t = (int)(i / n_combos); c = (int)(i % n_combos); }
    else { c = (int)(i / series_len); t = (int)(i % series_len); }
    const int p = periods[c];
    const int warm = first_valid + p - 1;
    float r = VTA_QNAN_F32;
    if (t >= warm) {
        const float *wr = weights + (size_t)c * max_period;
        const float inv = inv_norms[c];
        float s = 0.0f;
        for (int k = 0; k < p; ++k) s = fmaf(wr[k] * inv, prices[t - p + 1 + k], s);
        r = s;
    }
    out[i] = r;
}

extern "C" __global__ __launch_bounds__(256) void alma_ms1p_f32_direct(
    const float *__restrict__ prices_tm, const float *__restrict__ weights, int period,
    float inv_norm, int num_series, int series_len, const int32_t *__restrict__ first_valids,
    float *__restrict__ out_tm) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)series_len * (size_t)num_series;
    if (i >= total) return;
    const int t = (int)(i / num_series), s = (int)(i % num_series);
    const int warm = first_valids[s] + period - 1;
    float r = VTA_QNAN_F32;
    if (t >= warm) {
        float acc = 0.0f;
        for (int k = 0; k < period; ++k)
            acc = fmaf(weights[k] * inv_norm, prices_tm[(size_t)(t - period + 1 + k) * num_series + s], acc);
        r = acc;
    }
    out_tm[i] = r;
}

// ---------------------------------------------------------------- MFMA sweep, diagonal form
//
// The same split contraction, organised so that every A fragment loaded from
// LDS feeds several tiles.  A workgroup of four waves takes up to four
// N-blocks of one group over a run of SW_R 32-bar tiles, cut into sub-runs of
// SW_RS tiles.  Each sub-run is staged once, with one anchor for its whole
// window (K_g + 32 SW_RS bars), instead of once per tile.  In the Hankel
// tile, the A fragment of tile j at K-step s is G(2j + s), the 16-tap slice n
// of the sub-run window (lane l: elements 16n + l%32 + 8(l/32) .. +7 from the
// window start).  A wave walks n = 0 .. KS + 2 SW_RS - 3, loads G(n) once
// (hi and lo) and applies it to every tile j with s = n - 2j in [0, KS),
// accumulating tile j in its own accumulator: SW_RS accumulators, B fragments
// (the N-block's weights) in registers for the whole run, and consecutive
// MFMAs on different accumulators.
//
// Element formats (template F16):
//   BF16 (false): x' = x - a, a = midpoint of the window's finite bars; hi =
//     bf16(x'), lo = bf16(x' - hi); weights unscaled.  The north-star form;
//     its split error is bounded by ~3 * 2^-16 * sum |w x'| relative terms.
//   F16 (true): x' = (x - a) * 2^-e with |x - a| < 2^e over the window, so
//     |x'| < 1, and each combo's weights scaled by 2^sw (largest in [0.5, 1));
//     hi/lo in half precision (11-bit significands).  Per output the split
//     error is at most (2^-22 sum w_s + eps_w + 2^-11 L_w) 2^-(sw + e) in
//     output units (eps_w = sum |w_s - hi - lo|, L_w = sum |lo| of the scaled
//     weights: the x residual |x' - hi - lo| <= 2^-22 for |x'| < 1, the weight
//     residual, and the dropped lo*lo product); the host folds it into one
//     factor per combo, cscale.y, so the bound is cscale.y * 2^e.  A tile whose
//     bound exceeds SW_GUARD * |y| for any of its outputs is recomputed with the
//     plain FP32 dot, so the split never adds more than SW_GUARD relative error.
// Tiles with a non-finite input at or after first_valid in their window are
// recomputed the same way.
#define SW_R 16
#define SW_GUARD 2.0e-6f

typedef __attribute__((ext_vector_type(8))) _Float16 vta_f16x8;

struct alma_sw_args {
    const float *prices;
    const uint4 *frags;          // [sum_b ks_b*2*64] 16-byte lane chunks (hi, lo per K-step)
    const int4 *nblocks;         // {ks, frag_off, combo_base, 0}
    const int4 *groups;          // {first N-block, count (<= 4), ks of the group, 0}
    const int32_t *sorted_period;
    const int32_t *sorted_row;
    const float *w_f32;          // [cpad][kmax] f32 weights (fallback path)
    const float2 *cscale;        // F16: per sorted combo {2^-sw, split-error factor}
    float *out;
    int series_len;
    int first_valid;
    int n_groups;
    int n_tiles;
    int kmax;
    int group_base;
};

// dwords per staged array of a sub-run of RS tiles: kg/2 + 16 RS pairs, padded
// to 16 mod 32 so the two copies a wave reads in one instruction fall on
// disjoint banks
__device__ __forceinline__ int sw_stride(int kg, int rs) { return mfg_stride(kg / 2 + 16 * rs); }

template <bool F16>
__device__ __forceinline__ uint32_t sw_pack2(float v0, float v1, uint32_t *lo_pair) {
    if constexpr (F16) {
        const _Float16 h0 = (_Float16)v0, h1 = (_Float16)v1;
        const _Float16 l0 = (_Float16)(v0 - (float)h0), l1 = (_Float16)(v1 - (float)h1);
        *lo_pair = (uint32_t)__builtin_bit_cast(uint16_t, l0) |
                   ((uint32_t)__builtin_bit_cast(uint16_t, l1) << 16);
        return (uint32_t)__builtin_bit_cast(uint16_t, h0) |
               ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
    } else {
        const uint16_t h0 = vta_bf16_rne(v0), h1 = vta_bf16_rne(v1);
        const uint16_t l0 = vta_bf16_rne(v0 - vta_bf16_to_f32(h0));
        const uint16_t l1 = vta_bf16_rne(v1 - vta_bf16_to_f32(h1));
        *lo_pair = mf_pack(l0, l1);
        return mf_pack(h0, h1);
    }
}

template <bool F16>
__device__ __forceinline__ vta_f32x16 sw_mfma(const uint4 &a, const uint4 &b, const vta_f32x16 &c) {
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(vta_f16x8, a),
                                                      __builtin_bit_cast(vta_f16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(vta_bf16x8, a),
                                                       __builtin_bit_cast(vta_bf16x8, b), c, 0, 0,
                                                       0);
}

// Split-F16 guard for one tile (wave-uniform result): true if, for any stored
// output of the tile, the split-error bound exceeds SW_GUARD |y|.  Lane (m, g)
// holds combo m, bars t_tile + 8q + 4g + i in y[4q + i].  In a tile past every
// combo's warm-up and inside the series all 16 values are stored (padding
// combos have zero weights and a zero bound), so one comparison against the
// smallest |y| decides; otherwise each output is checked with its mask.
__device__ __forceinline__ bool sw_guard_violated(const float (&y)[16], float bound, int t_tile,
                                                  int g, int my_row, int my_warm, int warm_max,
                                                  int T) {
    bool viol = false;
    if (t_tile >= warm_max && t_tile + 32 <= T) {
        float mn = fabsf(y[0]);
#pragma unroll
        for (int i = 1; i < 16; ++i) mn = fminf(mn, fabsf(y[i]));
        viol = !(bound <= SW_GUARD * mn);
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = t_tile + 8 * q + 4 * g + i;
                viol = viol || (my_row >= 0 && t >= my_warm && t < T &&
                                !(bound <= SW_GUARD * fabsf(y[4 * q + i])));
            }
    }
    return __any(viol);
}

// Stores (or recomputes) one 32-bar tile of this wave's N-block from its
// accumulator: lane (m, g) holds combo m, bars 8q + 4g + i.
template <bool F16, bool XPOSE>
__device__ __forceinline__ void sw_epilogue(const alma_sw_args &a, const vta_f32x16 &acc,
                                            int t_tile, float anchor, float unscale,
                                            float cmul, float cguard, bool tile_bad,
                                            int combo, int my_period, int my_row, int my_warm,
                                            int warm_max, float *tbuf, const int *trow) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int m = lane & 31;
    const int T = a.series_len;
    const float f = F16 ? cmul * unscale : 1.0f;
    float y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) y[i] = fmaf(acc[i], f, anchor);
    bool redo = tile_bad;
    if (F16 && !redo) redo = sw_guard_violated(y, cguard * unscale, t_tile, g, my_row, my_warm,
                                               warm_max, T);
    const bool steady = !redo && t_tile >= warm_max && t_tile + 32 <= T && (T & 3) == 0;
    if (steady && XPOSE) {
        // transpose through the wave's LDS tile: each store instruction then
        // writes 8 combos x 128 contiguous bytes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<float4 *>(tbuf + m * 36 + 8 * q + 4 * g) =
                make_float4(y[4 * q + 0], y[4 * q + 1], y[4 * q + 2], y[4 * q + 3]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int tcol = 4 * (lane & 7);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float4 v =
                *reinterpret_cast<const float4 *>(tbuf + ((lane >> 3) + 8 * s) * 36 + tcol);
            if (trow[s] >= 0)
                *reinterpret_cast<float4 *>(a.out + (size_t)trow[s] * (size_t)T + t_tile + tcol) =
                    v;
        }
        return;
    }
    if (my_row < 0) return;
    float *row = a.out + (size_t)my_row * (size_t)T;
    if (steady) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<float4 *>(row + t_tile + 8 * q + 4 * g) =
                make_float4(y[4 * q + 0], y[4 * q + 1], y[4 * q + 2], y[4 * q + 3]);
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int tb = t_tile + 8 * q + 4 * g;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int t = tb + i;
            float val = y[4 * q + i];
            if (redo && t >= my_warm && t < T) {
                // plain FP32 dot, same taps as the FP32 kernel
                const float *wrow = a.w_f32 + (size_t)combo * a.kmax;
                float dot = 0.0f;
                for (int k = 0; k < my_period; ++k)
                    dot = fmaf(wrow[k], a.prices[t - my_period + 1 + k], dot);
                val = dot;
            }
            v[i] = (t >= my_warm) ? val : VTA_QNAN_F32;
        }
        if (((T & 3) == 0) && tb + 3 < T) {
            *reinterpret_cast<float4 *>(row + tb) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (tb + i < T) row[tb + i] = v[i];
        }
    }
}

// Deferred-store form of one wave's run (DEFER kernels): the outputs of
// sub-run q are written, transposed, to the wave's RS LDS tiles at its end,
// and their global stores (4 RS float4 per lane: tile j, rows 8 s + lane / 8,
// bars 4 (lane % 8) .. + 3) are spread over the MFMA diagonals of sub-run
// q + 1, so the stores drain while the matrix cores run; the last sub-run's
// are issued after the loop.  Every tile goes through LDS: warm-up and
// recomputed tiles get their values in the epilogue, and a store is issued
// only for rows of real combos and bars below T (T % 4 == 0, so a float4 is
// wholly in or out).
template <bool F16, int RS, int KS>
__device__ __forceinline__ void sw_wave_deferred(const alma_sw_args &a, int nb, int tile_begin,
                                                 int nt, int kg, int stride,
                                                 const uint32_t *stage, const float *sub,
                                                 const int *badf, float *tbuf) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int m = lane & 31;
    const int T = a.series_len;
    const int4 nbd = a.nblocks[nb];
    const int frag_off = nbd.y;
    const int combo_base = nbd.z;
    constexpr int K = KS * 16;
    constexpr int NG = KS + 2 * RS - 2;
    constexpr int NOPS = 4 * RS;
    constexpr int TSLOT = 32 * 36;
    uint4 bh[KS], bl[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bh[s] = a.frags[frag_off + (2 * s + 0) * 64 + lane];
        bl[s] = a.frags[frag_off + (2 * s + 1) * 64 + lane];
    }
    const int combo = combo_base + m;
    const int my_period = a.sorted_period[combo];
    const int my_row = a.sorted_row[combo];
    const int my_warm = a.first_valid + my_period - 1;
    float cmul = 1.0f, cguard = 0.0f;
    if constexpr (F16) {
        const float2 cs = a.cscale[combo];
        cmul = cs.x;
        cguard = cs.y;
    }
    int warm_max = my_warm;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) warm_max = max(warm_max, __shfl_xor(warm_max, o));
    warm_max = __builtin_amdgcn_readfirstlane(warm_max);
    // row pointers of the transposed stores (null for padding combos)
    const int tcol = 4 * (lane & 7);
    float *rowp[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = __shfl(my_row, (lane >> 3) + 8 * s);
        rowp[s] = r >= 0 ? a.out + (size_t)r * (size_t)T + tcol : nullptr;
    }
    const int r = m & 1;
    const int d0 = (m >> 1) + 4 * g + ((kg - K) >> 1);
    const int hi_off = (r ? stride : 0) + d0;
    const int lo_off = 2 * stride + hi_off;
    const int nsub = (nt + RS - 1) / RS;
    int pend_t = 1 << 30;            // first bar of the pending sub-run (none yet)
    auto store_op = [&](int o) {
        const int j = o / 4, s = o % 4;
        const float4 v =
            *reinterpret_cast<const float4 *>(tbuf + j * TSLOT + ((lane >> 3) + 8 * s) * 36 + tcol);
        const int t = pend_t + 32 * j;
        if (rowp[s] != nullptr && t + tcol < T) *reinterpret_cast<float4 *>(rowp[s] + t) = v;
    };
    for (int q = 0; q < nsub; ++q) {
        const uint32_t *sb = stage + q * 4 * stride;
        auto load_g = [&](int n, uint4 &h, uint4 &l) {
            const uint32_t *hb = sb + hi_off + 8 * n;
            const uint32_t *lb = sb + lo_off + 8 * n;
            h.x = hb[0]; h.y = hb[1]; h.z = hb[2]; h.w = hb[3];
            l.x = lb[0]; l.y = lb[1]; l.z = lb[2]; l.w = lb[3];
        };
        vta_f32x16 acc[RS];
#pragma unroll
        for (int j = 0; j < RS; ++j) acc[j] = vta_f32x16{};
        uint4 gh, gl, nh, nl;
        load_g(0, gh, gl);
#pragma unroll
        for (int n = 0; n < NG; ++n) {
            if (n + 1 < NG) load_g(n + 1, nh, nl);
#pragma unroll
            for (int o = NOPS * n / NG; o < NOPS * (n + 1) / NG; ++o) store_op(o);
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gh, bh[s], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gh, bl[s], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gl, bh[s], acc[j]);
            }
            gh = nh;
            gl = nl;
        }
        const float anchor = sub[3 * q + 0];
        const float unscale = sub[3 * q + 2];
        const float f = F16 ? cmul * unscale : 1.0f;
        // the pending stores have read their LDS tiles: overwrite them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < RS; ++j) {
            const int tile = q * RS + j;
            if (tile >= nt) break;
            const int t_tile = (tile_begin + tile) * 32;
            float y[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) y[i] = fmaf(acc[j][i], f, anchor);
            bool redo = badf[tile] != 0;
            if (F16 && !redo)
                redo = sw_guard_violated(y, cguard * unscale, t_tile, g, my_row, my_warm, warm_max,
                                         T);
            if (redo || t_tile < warm_max) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int t = t_tile + 8 * (i >> 2) + 4 * g + (i & 3);
                    if (redo && my_row >= 0 && t >= my_warm && t < T) {
                        // plain FP32 dot, same taps as the FP32 kernel
                        const float *wrow = a.w_f32 + (size_t)combo * a.kmax;
                        float dot = 0.0f;
                        for (int k = 0; k < my_period; ++k)
                            dot = fmaf(wrow[k], a.prices[t - my_period + 1 + k], dot);
                        y[i] = dot;
                    }
                    if (t < my_warm) y[i] = VTA_QNAN_F32;
                }
            }
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
                *reinterpret_cast<float4 *>(tbuf + j * TSLOT + m * 36 + 8 * qq + 4 * g) =
                    make_float4(y[4 * qq + 0], y[4 * qq + 1], y[4 * qq + 2], y[4 * qq + 3]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        pend_t = (tile_begin + q * RS) * 32;
    }
#pragma unroll
    for (int o = 0; o < NOPS; ++o) store_op(o);
}

template <bool F16, bool XPOSE, int RS, int KS>
__device__ __forceinline__ void sw_wave(const alma_sw_args &a, int nb, int tile_begin, int nt,
                                        int kg, int stride, const uint32_t *stage,
                                        const float *sub, const int *badf, float *tbuf) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int m = lane & 31;
    const int4 nbd = a.nblocks[nb];
    const int frag_off = nbd.y;
    const int combo_base = nbd.z;
    constexpr int K = KS * 16;
    constexpr int NG = KS + 2 * RS - 2;     // G(n), n < NG, per sub-run
    uint4 bh[KS], bl[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        bh[s] = a.frags[frag_off + (2 * s + 0) * 64 + lane];
        bl[s] = a.frags[frag_off + (2 * s + 1) * 64 + lane];
    }
    const int combo = combo_base + m;
    const int my_period = a.sorted_period[combo];
    const int my_row = a.sorted_row[combo];
    const int my_warm = a.first_valid + my_period - 1;
    float cmul = 1.0f, cguard = 0.0f;
    if constexpr (F16) {
        const float2 cs = a.cscale[combo];
        cmul = cs.x;
        cguard = cs.y;
    }
    int warm_max = my_warm;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) warm_max = max(warm_max, __shfl_xor(warm_max, o));
    warm_max = __builtin_amdgcn_readfirstlane(warm_max);
    int trow[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) trow[s] = __shfl(my_row, (lane >> 3) + 8 * s);
    // lane's dword offset inside a staged copy at n = 0
    const int r = m & 1;
    const int d0 = (m >> 1) + 4 * g + ((kg - K) >> 1);
    const int hi_off = (r ? stride : 0) + d0;
    const int lo_off = 2 * stride + hi_off;
    const int nsub = (nt + RS - 1) / RS;
    for (int q = 0; q < nsub; ++q) {
        const uint32_t *sb = stage + q * 4 * stride;
        auto load_g = [&](int n, uint4 &h, uint4 &l) {
            const uint32_t *hb = sb + hi_off + 8 * n;
            const uint32_t *lb = sb + lo_off + 8 * n;
            h.x = hb[0]; h.y = hb[1]; h.z = hb[2]; h.w = hb[3];
            l.x = lb[0]; l.y = lb[1]; l.z = lb[2]; l.w = lb[3];
        };
        vta_f32x16 acc[RS];
#pragma unroll
        for (int j = 0; j < RS; ++j) acc[j] = vta_f32x16{};
        uint4 gh, gl, nh, nl;
        load_g(0, gh, gl);
#pragma unroll
        for (int n = 0; n < NG; ++n) {
            if (n + 1 < NG) load_g(n + 1, nh, nl);
            // products ordered hi*hi, hi*lo, lo*hi across the tiles of this
            // diagonal, so consecutive MFMAs write different accumulators
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gh, bh[s], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gh, bl[s], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < RS; ++j) {
                const int s = n - 2 * j;
                if (s >= 0 && s < KS) acc[j] = sw_mfma<F16>(gl, bh[s], acc[j]);
            }
            gh = nh;
            gl = nl;
        }
        const float anchor = sub[3 * q + 0];
        const float unscale = sub[3 * q + 2];
#pragma unroll
        for (int j = 0; j < RS; ++j) {
            const int tile = q * RS + j;
            if (tile >= nt) break;
            sw_epilogue<F16, XPOSE>(a, acc[j], (tile_begin + tile) * 32, anchor, unscale, cmul,
                                    cguard, badf[tile] != 0, combo, my_period, my_row, my_warm,
                                    warm_max, tbuf, trow);
        }
    }
}

// grid.x = runs of SW_R tiles, grid.y = N-block groups (heaviest first);
// block = 256 (4 waves).  Dynamic LDS, in floats: xs (32 SW_R + kmax + 33
// prices), 3 per sub-run (anchor, scale, unscale), SW_R tile flags, 3 x 32
// block statistics, then SW_NSUB x 4 x sw_stride(kmax) staged dwords (16-byte
// aligned) and, for XPOSE, four 32 x 36 transpose tiles.
template <bool F16, bool XPOSE, int RS, bool DEFER>
__device__ __forceinline__ void alma_sw_entry(const alma_sw_args &a, int run, int gidx) {
    constexpr int NSUB = SW_R / RS;
    extern __shared__ __attribute__((aligned(16))) uint32_t sw_lds[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int4 gd = a.groups[gidx];
    const int kg = gd.z * 16;
    const int stride = sw_stride(kg, RS);
    const int xsn = 32 * SW_R + a.kmax + 33;
    float *xs = reinterpret_cast<float *>(sw_lds);
    float *sub = xs + xsn;                          // 3 per sub-run
    int *badf = reinterpret_cast<int *>(sub + 3 * NSUB);
    float *bmin = reinterpret_cast<float *>(badf + SW_R);
    float *bmax = bmin + 32;
    int *bnf = reinterpret_cast<int *>(bmax + 32);
    const int head = xsn + 3 * NSUB + SW_R + 96;
    uint32_t *stage = sw_lds + ((head + 3) & ~3);
    float *tbuf = reinterpret_cast<float *>(stage) + NSUB * 4 * sw_stride(a.kmax, RS) +
                  wave * (DEFER ? RS : 1) * 32 * 36;
    const int T = a.series_len;
    const int fv = a.first_valid;
    const int tile_begin = run * SW_R;
    const int nt = min(SW_R, a.n_tiles - tile_begin);
    const int t_lo = tile_begin * 32 - kg + 1;      // xs index 0
    const int nx = nt * 32 + kg + 1;
    for (int i = tid; i < nx; i += 256) {
        const int idx = t_lo + i;
        const float x = vta_ld_clamped(a.prices, idx, fv, T);
        xs[i] = (idx >= fv && idx < T) ? x : 0.0f;
    }
    __syncthreads();
    // statistics of 32-bar blocks of xs: min / max of the finite in-range bars,
    // any non-finite in-range bar (lanes 0-31 and 32-63 take one block each)
    const int nblk = (nx + 31) >> 5;                // <= SW_R + kmax / 32 + 1 <= 25
    for (int b = 2 * wave + (lane >> 5); b < nblk; b += 8) {
        const int i = 32 * b + (lane & 31);
        const int idx = t_lo + i;
        const bool in = i < nx && idx >= fv && idx < T;
        const float x = in ? xs[i] : 0.0f;
        const bool fin = in && vta_isfinite(x);
        float mn = fin ? x : __builtin_inff(), mx = fin ? x : -__builtin_inff();
        int nf = (in && !fin) ? 1 : 0;
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, o));
            mx = fmaxf(mx, __shfl_xor(mx, o));
            nf |= __shfl_xor(nf, o);
        }
        if ((lane & 31) == 0) {
            bmin[b] = mn;
            bmax[b] = mx;
            bnf[b] = nf;
        }
    }
    __syncthreads();
    if (tid < nt) {
        // tile flag: a non-finite bar in the tile's window xs[32 j, 32 j + kg + 31)
        int bad = 0;
        for (int b = tid; b <= min(tid + (kg + 30) / 32, nblk - 1); ++b) bad |= bnf[b];
        badf[tid] = bad;
    } else if (tid >= 64 && tid < 64 + NSUB) {
        // sub-run q: window xs[32 RS q, 32 RS q + kg + 32 RS + 1): anchor at the
        // midpoint of its finite bars and (F16) the power of two bounding |x - a|
        const int q = tid - 64;
        float mn = __builtin_inff(), mx = -__builtin_inff();
        for (int b = RS * q; b <= min(RS * q + (kg + 32 * RS) / 32, nblk - 1); ++b) {
            mn = fminf(mn, bmin[b]);
            mx = fmaxf(mx, bmax[b]);
        }
        float anchor = 0.0f, scale = 1.0f, unscale = 1.0f;
        if (mn <= mx) {
            anchor = mn + 0.5f * (mx - mn);
            if (!vta_isfinite(anchor)) anchor = 0.0f;
            if constexpr (F16) {
                const float h = fmaxf(fabsf(mn - anchor), fabsf(mx - anchor));
                if (h > 0.0f && vta_isfinite(h)) {
                    int e;
                    (void)frexpf(h, &e);            // h = f 2^e, f in [0.5, 1)
                    scale = __builtin_ldexpf(1.0f, -e);
                    unscale = __builtin_ldexpf(1.0f, e);
                }
            }
        }
        sub[3 * q + 0] = anchor;
        sub[3 * q + 1] = scale;
        sub[3 * q + 2] = unscale;
    }
    __syncthreads();
    // staging: pair i of sub-run q, copy 0 = (x'[2i], x'[2i+1]), copy 1 =
    // (x'[2i+1], x'[2i+2]), x'[e] from xs[32 RS q + e]
    const int npair = kg / 2 + 16 * RS;
    const int nsub = (nt + RS - 1) / RS;
    const int units = nsub * npair;
    for (int u = tid; u < units; u += 256) {
        const int q = u / npair;
        const int i = u - q * npair;
        const float anchor = sub[3 * q + 0];
        const float scale = sub[3 * q + 1];
        const int x0i = 32 * RS * q + 2 * i;
        float v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int xi = x0i + k;
            const int idx = t_lo + xi;
            const float x = xi < nx ? xs[xi] : 0.0f;
            const bool use = xi < nx && idx >= fv && idx < T && vta_isfinite(x);
            v[k] = use ? (x - anchor) * scale : 0.0f;
        }
        uint32_t l01, l12;
        const uint32_t h01 = sw_pack2<F16>(v[0], v[1], &l01);
        const uint32_t h12 = sw_pack2<F16>(v[1], v[2], &l12);
        uint32_t *base = stage + q * 4 * stride + i;
        base[0] = h01;
        base[stride] = h12;
        base[2 * stride] = l01;
        base[3 * stride] = l12;
    }
    __syncthreads();
    if (wave >= gd.y) return;
    const int nb = gd.x + wave;
    switch (a.nblocks[nb].x) {
#define SW_CASE(N)                                                                          \
    case N:                                                                                 \
        if constexpr (DEFER)                                                                \
            sw_wave_deferred<F16, RS, N>(a, nb, tile_begin, nt, kg, stride, stage, sub, badf, \
                                         tbuf);                                             \
        else                                                                                \
            sw_wave<F16, XPOSE, RS, N>(a, nb, tile_begin, nt, kg, stride, stage, sub, badf, \
                                       tbuf);                                               \
        break;
        SW_CASE(1) SW_CASE(2) SW_CASE(3) SW_CASE(4) SW_CASE(5) SW_CASE(6) SW_CASE(7)
        SW_CASE(8) SW_CASE(9) SW_CASE(10) SW_CASE(11) SW_CASE(12) SW_CASE(13) SW_CASE(14)
        SW_CASE(15) SW_CASE(16)
#undef SW_CASE
        default: break;
    }
}

// RS = 2: two accumulators, fits two waves per SIMD; RS = 4: four, one wave.
// _p: deferred stores (sw_wave_deferred; T % 4 == 0 only).
#define SW_KERNEL(NAME, F16, XPOSE, RS, WPE, DEFER)                                           \
    extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) \
    void NAME(alma_sw_args a) {                                                               \
        alma_sw_entry<F16, XPOSE, RS, DEFER>(a, blockIdx.x,                                   \
                                             a.group_base + a.n_groups - 1 - (int)blockIdx.y); \
    }
SW_KERNEL(alma_sweep_bf16x3, false, true, 2, 2, false)
SW_KERNEL(alma_sweep_bf16x3_d, false, false, 2, 2, false)
SW_KERNEL(alma_sweep_bf16x3_r4, false, false, 4, 1, false)
SW_KERNEL(alma_sweep_bf16x3_p, false, true, 2, 2, true)
SW_KERNEL(alma_sweep_f16x3, true, true, 2, 2, false)
SW_KERNEL(alma_sweep_f16x3_d, true, false, 2, 2, false)
SW_KERNEL(alma_sweep_f16x3_r4, true, false, 4, 1, false)
SW_KERNEL(alma_sweep_f16x3_p, true, true, 2, 2, true)
#undef SW_KERNEL

// Packed sweeps: several one-series batches (uneven lengths and grids) in one
// launch of the direct-store diagonal form.  Block b runs 16-tile run
// items[b].y of group items[b].z of job items[b].x (the host orders the items
// heaviest group first), so every job's outputs are those of its own
// alma_sweep_*_d launch.
#define SW_PACK_KERNEL(NAME, F16)                                                             \
    extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))   \
    void NAME(const alma_sw_args *__restrict__ jobs, const int4 *__restrict__ items) {        \
        const int4 it = items[blockIdx.x];                                                    \
        const alma_sw_args a = jobs[it.x];                                                    \
        alma_sw_entry<F16, false, 2, false>(a, it.y, it.z);                                   \
    }
SW_PACK_KERNEL(alma_sweep_pack_bf16x3, false)
SW_PACK_KERNEL(alma_sweep_pack_f16x3, true)
#undef SW_PACK_KERNEL
