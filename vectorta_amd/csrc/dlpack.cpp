// DLPack export of device-resident results: the HIP counterpart of
// export_f32_cuda_dlpack_2d (src/utilities/dlpack_cuda.rs:13-276).
//
// The structs below are the public DLPack ABI (dlpack.h, v1.x): DLDevice,
// DLDataType, DLTensor, DLManagedTensor (legacy capsule "dltensor") and
// DLManagedTensorVersioned (capsule "dltensor_versioned").  The layouts must
// match the standard byte for byte, so they are written out here field by
// field.  Device type kDLROCM = 10 (the reference exports kDLCUDA = 2).
#include <cstdint>
#include <new>

#include "runtime_internal.h"

using namespace vta;

namespace {

struct DLDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DLDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DLTensor {
    void *data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(DLManagedTensor *);
};
struct DLPackVersion {
    uint32_t major;
    uint32_t minor;
};
struct DLManagedTensorVersioned {
    DLPackVersion version;
    void *manager_ctx;
    void (*deleter)(DLManagedTensorVersioned *);
    uint64_t flags;
    DLTensor dl_tensor;
};
static_assert(sizeof(DLTensor) == 48, "DLTensor layout");
static_assert(sizeof(DLManagedTensor) == 64, "DLManagedTensor layout");
static_assert(sizeof(DLManagedTensorVersioned) == 80, "DLManagedTensorVersioned layout");

constexpr int32_t kDLROCM = 10;
constexpr uint8_t kDLFloat = 2;

// Owner of the exported buffer and of the shape/strides arrays.
struct Manager {
    void *buf;
    int device;
    int64_t shape[2];
    int64_t strides[2];
};

void release(Manager *m) {
    if (!m) return;
    if (m->buf) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(m->device);
        (void)hipFree(m->buf);
        (void)hipSetDevice(prev);
    }
    delete m;
}

void delete_legacy(DLManagedTensor *t) {
    if (!t) return;
    release(static_cast<Manager *>(t->manager_ctx));
    delete t;
}

void delete_versioned(DLManagedTensorVersioned *t) {
    if (!t) return;
    release(static_cast<Manager *>(t->manager_ctx));
    delete t;
}

void fill_tensor(DLTensor &d, Manager *m, const vta_device_array_f32 &arr) {
    d.data = (arr.rows && arr.cols) ? arr.ptr : nullptr;
    d.device = {kDLROCM, m->device};
    d.ndim = 2;
    d.dtype = {kDLFloat, 32, 1};
    d.shape = m->shape;
    d.strides = m->strides;
    d.byte_offset = 0;
}

}  // namespace

extern "C" {

vta_status vta_device_array_to_dlpack(vta_ctx *ctx, vta_device_array_f32 *arr, int versioned,
                                      void **out_managed) {
    clear_error();
    if (!arr || !out_managed) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (arr->ptr && !ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    if (arr->rows > (size_t)INT64_MAX || arr->cols > (size_t)INT64_MAX)
        return fail(VTA_ERR_INVALID_INPUT, "shape exceeds int64");
    if (ctx && arr->ptr) VTA_HIP(hipStreamSynchronize(ctx->stream));
    Manager *m = new (std::nothrow) Manager{arr->ptr, ctx ? ctx->device : 0,
                                            {(int64_t)arr->rows, (int64_t)arr->cols},
                                            {(int64_t)arr->cols, 1}};
    if (!m) return fail(VTA_ERR_OUT_OF_MEMORY, "host allocation failed");
    if (versioned) {
        auto *t = new (std::nothrow) DLManagedTensorVersioned{};
        if (!t) {
            delete m;
            return fail(VTA_ERR_OUT_OF_MEMORY, "host allocation failed");
        }
        t->version = {1, 0};
        t->manager_ctx = m;
        t->deleter = delete_versioned;
        t->flags = 0;
        fill_tensor(t->dl_tensor, m, *arr);
        *out_managed = t;
    } else {
        auto *t = new (std::nothrow) DLManagedTensor{};
        if (!t) {
            delete m;
            return fail(VTA_ERR_OUT_OF_MEMORY, "host allocation failed");
        }
        t->manager_ctx = m;
        t->deleter = delete_legacy;
        fill_tensor(t->dl_tensor, m, *arr);
        *out_managed = t;
    }
    // ownership moved into the managed tensor
    arr->ptr = nullptr;
    arr->rows = arr->cols = 0;
    return VTA_OK;
}

void vta_dlpack_delete(void *managed, int versioned) {
    if (!managed) return;
    if (versioned) {
        auto *t = static_cast<DLManagedTensorVersioned *>(managed);
        if (t->deleter) t->deleter(t);
    } else {
        auto *t = static_cast<DLManagedTensor *>(managed);
        if (t->deleter) t->deleter(t);
    }
}

}  // extern "C"
