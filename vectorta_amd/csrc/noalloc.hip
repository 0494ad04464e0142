// Allocation-free ALMA sweep from device prices: the variant of
// vta_alma_batch_from_device_prices for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The caller passes the prices and the output buffer; the
// per-combo tables the sweep kernels read (weights, MFMA operand fragments,
// N-block and group descriptors), which the existing entry point allocates
// per call, go into `scratch`, a static device array of kScratchBytes that
// each GPU holds until the process ends.
//
// The array's name is a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for it, looked up once per
// device with that device current (at context creation, else at the entry
// point's first call on the device) and kept per device.
//
// Every call on a GPU shares that GPU's array and leaves its last tables in
// it; calls that overlap on one GPU (two host threads, or two contexts of the
// same device) race on it and must be serialised by the caller.  A call
// returns after its sweep has completed, so consecutive calls from one thread
// never overlap.
#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current).
hipError_t scratch_address(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

namespace vta {

void noalloc_setup(int device) {
    void *p;
    (void)scratch_address(device, &p);   // a failure is reported by the entry point
}

}  // namespace vta

using namespace vta;

extern "C" {

vta_status vta_alma_batch_from_device_prices_noalloc(vta_ctx *ctx, const float *d_prices,
                                                     size_t len, size_t first_valid,
                                                     const vta_alma_batch_range *range,
                                                     float *d_out, size_t out_len) {
    clear_error();
    if (!ctx || !d_prices || !d_out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (len == 0) return fail(VTA_ERR_INVALID_INPUT, "series_len is zero");
    AlmaTables t;
    VTA_TRY(alma_tables_for_range(ctx, range, len, first_valid, t));
    if (t.blob.size() > kScratchBytes)
        return fail(VTA_ERR_INVALID_INPUT,
                    "sweep tables need %zu bytes, above the %zu-byte allocation-free scratch "
                    "(%zu combos, longest period %zu); use alma_batch_from_device_prices",
                    t.blob.size(), kScratchBytes, t.n_combos, t.max_period);
    size_t expected;
    if (!mul_ok(t.n_combos, len, &expected))
        return fail(VTA_ERR_INVALID_INPUT, "combos*series_len overflow");
    if (out_len != expected)
        return fail(VTA_ERR_INVALID_INPUT, "out buffer wrong length: got %zu, expected %zu",
                    out_len, expected);
    VTA_HIP(hipSetDevice(ctx->device));
    void *base = nullptr;
    hipError_t e = scratch_address(ctx->device, &base);
    if (e != hipSuccess)
        return fail(VTA_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                    hipGetErrorString(e));
    VTA_HIP(hipMemcpyAsync(base, t.blob.data(), t.blob.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    VTA_TRY(alma_tables_launch(ctx, t, base, d_prices, len, first_valid, d_out));
    VTA_HIP(hipStreamSynchronize(ctx->stream));
    return VTA_OK;
}

}  // extern "C"
