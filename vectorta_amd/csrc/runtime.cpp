// libvta_hip host runtime: context, code-object loading, memory helpers.
//
// Replaces the reference's per-wrapper CUDA plumbing (cust Context/Module/
// Stream in src/cuda/moving_averages/alma_wrapper.rs:133-169, the device probe
// in src/cuda/mod.rs:412-602, and build.rs's PTX staging) with HIP: one device,
// one non-blocking stream, gfx950 code objects loaded by hipModuleLoad.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>

#include "runtime_internal.h"

namespace vta {

static thread_local std::string g_err;

vta_status fail(vta_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

void clear_error() { g_err.clear(); }

static std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&vta_version), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash);
    }
    return ".";
}

static vta_status get_module(vta_ctx *ctx, const char *module, hipModule_t *out) {
    auto it = ctx->modules.find(module);
    if (it != ctx->modules.end()) {
        *out = it->second;
        return VTA_OK;
    }
    std::string path = ctx->kernel_dir + "/vta_" + module + ".hsaco";
    hipModule_t m = nullptr;
    hipError_t e = hipModuleLoad(&m, path.c_str());
    if (e != hipSuccess)
        return fail(VTA_ERR_MISSING_KERNEL_SYMBOL, "hipModuleLoad(%s) failed: %s", path.c_str(),
                    hipGetErrorString(e));
    ctx->modules[module] = m;
    *out = m;
    return VTA_OK;
}

vta_status get_function(vta_ctx *ctx, const char *module, const char *name, hipFunction_t *out) {
    std::string key = std::string(module) + ":" + name;
    auto it = ctx->funcs.find(key);
    if (it != ctx->funcs.end()) {
        *out = it->second;
        return VTA_OK;
    }
    hipModule_t m = nullptr;
    VTA_TRY(get_module(ctx, module, &m));
    hipFunction_t f = nullptr;
    hipError_t e = hipModuleGetFunction(&f, m, name);
    if (e != hipSuccess || !f)
        return fail(VTA_ERR_MISSING_KERNEL_SYMBOL, "missing kernel symbol: %s (module %s)", name,
                    module);
    ctx->funcs[key] = f;
    *out = f;
    return VTA_OK;
}

bool has_function(vta_ctx *ctx, const char *module, const char *name) {
    hipFunction_t f;
    std::string saved = g_err;
    bool ok = get_function(ctx, module, name, &f) == VTA_OK;
    g_err = saved;
    return ok;
}

vta_status validate_launch(vta_ctx *ctx, dim3 grid, dim3 block) {
    const hipDeviceProp_t &p = ctx->props;
    unsigned threads = block.x * block.y * block.z;
    if (threads == 0 || grid.x == 0 || grid.y == 0 || grid.z == 0 ||
        threads > (unsigned)p.maxThreadsPerBlock || block.x > (unsigned)p.maxThreadsDim[0] ||
        block.y > (unsigned)p.maxThreadsDim[1] || block.z > (unsigned)p.maxThreadsDim[2] ||
        grid.x > (unsigned)p.maxGridSize[0] || grid.y > (unsigned)p.maxGridSize[1] ||
        grid.z > (unsigned)p.maxGridSize[2])
        return fail(VTA_ERR_LAUNCH_CONFIG_TOO_LARGE,
                    "launch config too large: grid=(%u,%u,%u) block=(%u,%u,%u)", grid.x, grid.y,
                    grid.z, block.x, block.y, block.z);
    return VTA_OK;
}

vta_status launch_raw(vta_ctx *ctx, hipFunction_t f, dim3 grid, dim3 block, unsigned shmem,
                      void **args) {
    VTA_TRY(validate_launch(ctx, grid, block));
    VTA_HIP(hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, shmem,
                                  ctx->stream, args, nullptr));
    return VTA_OK;
}

vta_status check_fits(size_t required) {
    const char *env = getenv("VTA_MEM_CHECK");
    if (env && (strcmp(env, "0") == 0 || strcasecmp(env, "false") == 0)) return VTA_OK;
    size_t free_b = 0, total = 0;
    if (hipMemGetInfo(&free_b, &total) != hipSuccess) return VTA_OK;
    const size_t headroom = 64ull << 20;
    if (required > free_b || free_b - required < headroom)
        return fail(VTA_ERR_OUT_OF_MEMORY, "out of memory: required=%zu free=%zu headroom=%zu",
                    required, free_b, headroom);
    return VTA_OK;
}

bool mul_ok(size_t a, size_t b, size_t *out) { return !__builtin_mul_overflow(a, b, out); }
bool add_ok(size_t a, size_t b, size_t *out) { return !__builtin_add_overflow(a, b, out); }

vta_status DevBuf::alloc(size_t bytes) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(VTA_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes,
                    hipGetErrorString(e));
    }
    return VTA_OK;
}

vta_status upload(vta_ctx *ctx, DevBuf &buf, const void *src, size_t bytes) {
    VTA_TRY(buf.alloc(bytes));
    if (bytes) {
        // The source is pageable host memory that the caller may free right
        // after return, so the copy completes before we hand control back.
        VTA_HIP(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        VTA_HIP(hipStreamSynchronize(ctx->stream));
    }
    return VTA_OK;
}

bool expand_usize(vta_range_usize r, AxisKind kind, std::vector<size_t> &v) {
    v.clear();
    if (r.step == 0 || r.start == r.end) {
        v.push_back(r.start);
        return true;
    }
    size_t lo = r.start < r.end ? r.start : r.end;
    size_t hi = r.start < r.end ? r.end : r.start;
    if ((hi - lo) / r.step >= kMaxGridValues) return false;   // > 2^24 values
    if (r.start < r.end || kind == AXIS_SORTED || kind == AXIS_ATR) {
        for (size_t x = lo;; x += r.step) {
            v.push_back(x);
            if (x > hi - r.step || x + r.step > hi) break;
        }
        if (r.start > r.end && kind == AXIS_ATR) std::reverse(v.begin(), v.end());
        return !v.empty();
    }
    // AXIS_DOWN / AXIS_DOWN_END with start > end
    for (size_t x = r.start;; x -= r.step) {
        v.push_back(x);
        if (x < r.step || x - r.step < r.end) break;
    }
    if (kind == AXIS_DOWN_END && v.back() != r.end) v.push_back(r.end);
    return !v.empty();
}

bool expand_f64(vta_range_f64 r, std::vector<double> &v) {
    v.clear();
    if (std::fabs(r.step) < 1e-12 || std::fabs(r.start - r.end) < 1e-12) {
        v.push_back(r.start);
        return true;
    }
    double st = std::fabs(r.step);
    // an infinite bound never ends the loop; a NaN one (or a NaN step) ends it
    // at once, so only the span over a finite step is bounded here
    if (std::isinf(r.start) || std::isinf(r.end)) return false;
    if (std::fabs(r.end - r.start) / st >= (double)kMaxGridValues) return false;
    // The loops accumulate x += step as the reference does; a step below half
    // the rounding unit of x leaves x unchanged (1e17 + 1 == 1e17), so each
    // iteration checks that x moved and that the count stays in bounds.
    if (r.start < r.end) {
        for (double x = r.start; x <= r.end + 1e-12; x += st) {
            if (v.size() >= kMaxGridValues || x + st == x) return false;
            v.push_back(x);
        }
    } else {
        for (double x = r.start; x + 1e-12 >= r.end; x -= st) {
            if (v.size() >= kMaxGridValues || x - st == x) return false;
            v.push_back(x);
        }
    }
    return !v.empty();
}

vta_status host_first_valids(const float *tm, size_t cols, size_t rows, size_t need,
                             std::vector<int32_t> &fvs) {
    fvs.assign(cols, 0);
    for (size_t s = 0; s < cols; ++s) {
        size_t t = 0;
        while (t < rows && std::isnan(tm[t * cols + s])) ++t;
        if (t >= rows) return fail(VTA_ERR_INVALID_INPUT, "series %zu all NaN", s);
        if (rows - t < need)
            return fail(VTA_ERR_INVALID_INPUT,
                        "series %zu not enough valid data (needed >= %zu, valid = %zu)", s, need,
                        rows - t);
        fvs[s] = (int32_t)t;
    }
    return VTA_OK;
}

size_t first_valid_f32(const float *x, size_t n) {
    size_t i = 0;
    while (i < n && std::isnan(x[i])) ++i;
    return i;
}

uint16_t bf16_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

}  // namespace vta

using namespace vta;

extern "C" {

const char *vta_last_error(void) { return g_err.c_str(); }

const char *vta_version(void) { return "vta-hip 0.1.0 (gfx950)"; }

int vta_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int vta_device_available(void) {
    const char *skip = getenv("VTA_FORCE_SKIP");
    if (skip && (strcmp(skip, "1") == 0 || strcasecmp(skip, "true") == 0)) return 0;
    return vta_device_count() > 0 ? 1 : 0;
}

vta_status vta_ctx_create(int device_id, vta_ctx **out) {
    clear_error();
    if (!out) return fail(VTA_ERR_INVALID_INPUT, "out is null");
    *out = nullptr;
    if (!vta_device_available()) return fail(VTA_ERR_NO_DEVICE, "no HIP device available");
    int n = vta_device_count();
    if (device_id < 0 || device_id >= n)
        return fail(VTA_ERR_DEVICE_MISMATCH, "device mismatch: requested=%d count=%d", device_id,
                    n);
    VTA_HIP(hipSetDevice(device_id));
    vta_ctx *ctx = new vta_ctx();
    ctx->device = device_id;
    if (hipGetDeviceProperties(&ctx->props, device_id) != hipSuccess) {
        delete ctx;
        return fail(VTA_ERR_HIP, "hipGetDeviceProperties failed");
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(VTA_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    if (hipMalloc(&ctx->scratch, vta_ctx::kScratchBytes) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(VTA_ERR_OUT_OF_MEMORY, "hipMalloc of the context scratch failed");
    }
    // zeroed once: the argument-error word must start clear
    if (hipMemset(ctx->scratch, 0, vta_ctx::kScratchBytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(ctx->scratch);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(VTA_ERR_HIP, "clearing the context scratch failed");
    }
    if (hipMalloc(&ctx->carry, vta_ctx::kCarryBytes) != hipSuccess) {
        (void)hipFree(ctx->scratch);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(VTA_ERR_OUT_OF_MEMORY, "hipMalloc of the context carry buffer failed");
    }
    const char *kd = getenv("VTA_KERNEL_DIR");
    ctx->kernel_dir = (kd && *kd) ? std::string(kd) : library_dir();
    noalloc_setup(device_id);   // device address of the allocation-free scratch
    *out = ctx;
    return VTA_OK;
}

vta_status vta_ctx_destroy(vta_ctx *ctx) {
    if (!ctx) return VTA_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->modules) (void)hipModuleUnload(kv.second);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->carry) (void)hipFree(ctx->carry);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VTA_OK;
}

vta_status vta_ctx_synchronize(vta_ctx *ctx) {
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->arg_check_pending) {
        // device-side argument checks of enqueue-only calls since the last
        // synchronize (the reference's warm-index buffers)
        ctx->arg_check_pending = false;
        int32_t *err = static_cast<int32_t *>(ctx->scratch) + kArgErrorWord;
        int32_t h = 0;
        VTA_HIP(hipMemcpyAsync(&h, err, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        VTA_HIP(hipStreamSynchronize(ctx->stream));
        if (h != 0) {
            hipFunction_t fill;
            VTA_TRY(get_function(ctx, "rec", "vta_fill_i32", &fill));
            VTA_TRY(launch(ctx, fill, dim3(1), dim3(256), 0, err, 1, 0));
            VTA_HIP(hipStreamSynchronize(ctx->stream));
            return fail(VTA_ERR_INVALID_INPUT,
                        "invalid input in an enqueue-only call since the last synchronize: "
                        "a warm index differs from first_valid + period - 1 (or implies a "
                        "first_valid outside the series); that call's outputs are undefined");
        }
    }
    return VTA_OK;
}

void *vta_ctx_stream(vta_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int vta_ctx_device_id(const vta_ctx *ctx) { return ctx ? ctx->device : -1; }

vta_status vta_ctx_preload(vta_ctx *ctx) {
    static const char *const kSymbols[][2] = {
        {"alma", "alma_batch_f32"},
        {"alma", "alma_batch_f32_tm"},
        {"alma", "alma_ms1p_f32"},
        {"alma", "alma_batch_mfma_bf16x3"},
        {"alma", "alma_batch_mfma_bf16x3_r8"},
        {"alma", "alma_batch_mfma_bf16x3_r32"},
        {"alma", "alma_batch_mfma_bf16x3_wg8"},
        {"alma", "alma_batch_mfma_bf16x3_wg16"},
        {"alma", "alma_batch_mfma_bf16x3_wg16_t2"},
        {"alma", "alma_batch_mfma_bf16x3_wg16_klo"},
        {"alma", "alma_batch_mfma_bf16x3_wg16_t2_klo"},
        {"alma", "alma_batch_f32_direct"},
        {"alma", "alma_ms1p_f32_direct"},
        {"alma", "alma_sweep_bf16x3"},
        {"alma", "alma_sweep_bf16x3_d"},
        {"alma", "alma_sweep_bf16x3_r4"},
        {"alma", "alma_sweep_bf16x3_p"},
        {"alma", "alma_sweep_f16x3"},
        {"alma", "alma_sweep_f16x3_d"},
        {"alma", "alma_sweep_f16x3_r4"},
        {"alma", "alma_sweep_f16x3_p"},
        {"alma", "alma_batch_f32_pack"},
        {"alma", "alma_sweep_pack_bf16x3"},
        {"alma", "alma_sweep_pack_f16x3"},
        {"rec", "ema_batch_f32"},
        {"rec", "rsi_batch_f32"},
        {"rec", "atr_batch_f32"},
        {"rec", "adx_batch_f32"},
        {"rec", "ema_ms1p_f32"},
        {"rec", "rsi_ms1p_f32"},
        {"rec", "atr_ms1p_f32"},
        {"rec", "adx_ms1p_f32"},
        {"rec", "ema_scan_combine"},
        {"rec", "rsi_scan_combine"},
        {"rec", "atr_scan_combine"},
        {"rec", "smma_scan_combine"},
        {"rec", "wilders_scan_combine"},
        {"rec", "adx_scan_combine"},
        {"rec", "atr_tr_batch_f32"},
        {"rec", "atr_tr_scan_combine"},
        {"rec", "tr_from_hlc_f32"},
        {"rec", "smma_batch_f32"},
        {"rec", "wilders_batch_f32"},
        {"rec", "smma_ms1p_f32"},
        {"rec", "wilders_ms1p_f32"},
        {"fir", "fir_batch_f32"},
        {"fir", "fir_ms1p_sma_f32"},
        {"fir", "fir_ms1p_wma_f32"},
        {"fir", "fir_ms1p_swma_f32"},
        {"fir", "series_first_nonfinite"},
        {"fir", "series_first_nonfinite_dfv"},
        {"rec", "ms_ema_rsi_atr_f32"},
        {"rec", "ms_prepare_meta"},
        {"rec", "series_nonfinite_flag"},
        {"rec", "flag_to_clean"},
        {"rec", "vta_fill_i32"},
        {"rec", "vta_check_warms"},
        {"rec", "ms_clean_from_first_valids"},
        {"rec", "ms_first_nonfinite_from_first_valids"},
        {"rolling", "bb_batch_f32"},
        {"rolling", "bb_ms1p_f32"},
        {"rolling", "pnr_batch_f32"},
        {"rolling", "pnr_ms1p_f32"},
        {"rolling", "mad_batch_f32"},
        {"cross", "cross_rows_f64"},
        {"cross", "cross_pairs_f64"},
        {"cross", "cross_reduce"},
        {"rolling", "mad_ms1p_f32"},
        {"rolling", "mad_ms1p_reg16"},
    };
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    for (auto &s : kSymbols) {
        hipFunction_t f;
        VTA_TRY(get_function(ctx, s[0], s[1], &f));
    }
    // register-window PNR kernels (16..128 slots) and MAD kernels per period
    // (4..64)
    char name[32];
    for (int cap = 16; cap <= 128; cap += 16) {
        hipFunction_t f;
        snprintf(name, sizeof name, "pnr_ms1p_reg%d", cap);
        VTA_TRY(get_function(ctx, "rolling", name, &f));
    }
    for (int len = 4; len <= 64; ++len) {
        hipFunction_t f;
        snprintf(name, sizeof name, "mad_ms1p_len%d", len);
        VTA_TRY(get_function(ctx, "rolling", name, &f));
    }
    return VTA_OK;
}

vta_status vta_device_array_free(vta_ctx *ctx, vta_device_array_f32 *arr) {
    if (!arr) return VTA_OK;
    if (arr->ptr) {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        VTA_HIP(hipFree(arr->ptr));
    }
    arr->ptr = nullptr;
    arr->rows = arr->cols = 0;
    return VTA_OK;
}

vta_status vta_device_array_copy_to_host(vta_ctx *ctx, const vta_device_array_f32 *arr,
                                         float *host, size_t host_len) {
    if (!ctx || !arr || !host) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    size_t n;
    if (!mul_ok(arr->rows, arr->cols, &n)) return fail(VTA_ERR_ARITHMETIC_OVERFLOW, "rows*cols");
    if (host_len != n)
        return fail(VTA_ERR_INVALID_INPUT, "host buffer wrong length: got %zu, expected %zu",
                    host_len, n);
    VTA_HIP(hipMemcpyAsync(host, arr->ptr, n * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    return VTA_OK;
}

vta_status vta_device_alloc(vta_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    VTA_TRY(check_fits(bytes));
    VTA_HIP(hipMalloc(out, bytes ? bytes : 4));
    return VTA_OK;
}

vta_status vta_device_free(vta_ctx *ctx, void *ptr) {
    if (!ptr) return VTA_OK;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    VTA_HIP(hipFree(ptr));
    return VTA_OK;
}

vta_status vta_memcpy_htod(vta_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    VTA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    return VTA_OK;
}

vta_status vta_memcpy_dtoh(vta_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    VTA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    return VTA_OK;
}

}  // extern "C"

namespace vta {

vta_status check_warms_device(vta_ctx *ctx, const int32_t *d_periods, const int32_t *d_warms,
                              int n, int T, int first_valid, int max_period,
                              int32_t *d_fv_out) {
    hipFunction_t f;
    VTA_TRY(get_function(ctx, "rec", "vta_check_warms", &f));
    int32_t *err = static_cast<int32_t *>(ctx->scratch) + kArgErrorWord;
    ctx->arg_check_pending = true;
    return launch(ctx, f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d_periods, d_warms,
                  n, T, first_valid, max_period, d_fv_out, err);
}

}  // namespace vta

namespace {

vta_status set_policy(vta_ctx *ctx, int slot, const vta_kernel_policy *p) {
    clear_error();
    if (!ctx || !p) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    const int32_t kinds[2] = {p->batch, p->many_series};
    const uint32_t bx[2] = {p->batch_block_x, p->many_series_block_x};
    for (int i = 0; i < 2; ++i) {
        if (kinds[i] != VTA_KERNEL_AUTO && kinds[i] != VTA_KERNEL_PLAIN)
            return fail(VTA_ERR_INVALID_POLICY, "invalid policy kind %d", (int)kinds[i]);
        if (kinds[i] == VTA_KERNEL_PLAIN && bx[i] != 0 && bx[i] != 64)
            return fail(VTA_ERR_INVALID_POLICY,
                        "invalid policy: block_x %u (these kernels run 64-lane waves; use 0 or "
                        "64)", bx[i]);
    }
    vta_kernel_policy q = *p;
    if (q.batch == VTA_KERNEL_PLAIN) q.batch_block_x = 64;
    if (q.many_series == VTA_KERNEL_PLAIN) q.many_series_block_x = 64;
    if (q.batch == VTA_KERNEL_AUTO) q.batch_block_x = 0;
    if (q.many_series == VTA_KERNEL_AUTO) q.many_series_block_x = 0;
    ctx->policy[slot] = q;
    return VTA_OK;
}

vta_status get_policy(const vta_ctx *ctx, int slot, vta_kernel_policy *out) {
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = ctx->policy[slot];
    return VTA_OK;
}

vta_status get_selected(const vta_ctx *ctx, const vta_kernel_selected *v,
                        vta_kernel_selected *out) {
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = *v;
    return VTA_OK;
}

}  // namespace

extern "C" {

#define VTA_POLICY_ENTRY_POINTS(NAME, SLOT)                                                 \
    vta_status vta_##NAME##_set_policy(vta_ctx *ctx, const vta_kernel_policy *policy) {     \
        return set_policy(ctx, SLOT, policy);                                               \
    }                                                                                       \
    vta_status vta_##NAME##_policy(const vta_ctx *ctx, vta_kernel_policy *out) {            \
        return get_policy(ctx, SLOT, out);                                                  \
    }                                                                                       \
    vta_status vta_##NAME##_selected_batch_kernel(const vta_ctx *ctx,                       \
                                                  vta_kernel_selected *out) {               \
        return get_selected(ctx, ctx ? &ctx->sel_batch[SLOT] : nullptr, out);               \
    }                                                                                       \
    vta_status vta_##NAME##_selected_many_series_kernel(const vta_ctx *ctx,                 \
                                                        vta_kernel_selected *out) {         \
        return get_selected(ctx, ctx ? &ctx->sel_ms[SLOT] : nullptr, out);                  \
    }

VTA_POLICY_ENTRY_POINTS(ema, VTA_POLICY_SLOT_EMA)
VTA_POLICY_ENTRY_POINTS(rsi, VTA_POLICY_SLOT_RSI)
VTA_POLICY_ENTRY_POINTS(adx, VTA_POLICY_SLOT_ADX)
VTA_POLICY_ENTRY_POINTS(pnr, VTA_POLICY_SLOT_PNR)
#undef VTA_POLICY_ENTRY_POINTS

}  // extern "C"
