// Captured batches: a sequence of enqueue-only calls on one context recorded
// with stream capture and replayed as one hipGraph launch.
//
// No counterpart in the reference, whose wrappers launch every kernel eagerly
// (src/cuda/moving_averages/alma_wrapper.rs:1005-1181).  A batch of several
// indicators over resident series is 5-20 kernel launches (the chunked scans
// alone are 3-5); replaying them as one graph removes the per-launch host
// work, which dominates for small requests.
#include "runtime_internal.h"

struct vta_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    size_t n_nodes = 0;
};

using namespace vta;

extern "C" {

vta_status vta_ctx_begin_capture(vta_ctx *ctx) {
    clear_error();
    if (!ctx) return fail(VTA_ERR_INVALID_INPUT, "ctx is null");
    VTA_HIP(hipSetDevice(ctx->device));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    VTA_HIP(hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(VTA_ERR_INVALID_INPUT, "the context's stream is already capturing");
    // every code object loaded before the capture starts
    VTA_TRY(vta_ctx_preload(ctx));
    VTA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    return VTA_OK;
}

vta_status vta_ctx_end_capture(vta_ctx *ctx, vta_graph **out) {
    clear_error();
    if (!ctx || !out) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    *out = nullptr;
    VTA_HIP(hipSetDevice(ctx->device));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    VTA_HIP(hipStreamIsCapturing(ctx->stream, &cs));
    if (cs == hipStreamCaptureStatusNone)
        return fail(VTA_ERR_INVALID_INPUT, "the context's stream is not capturing");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return fail(VTA_ERR_HIP,
                    "hipStreamEndCapture: %s (an entry point that allocates, copies from the "
                    "host or synchronises was called while capturing)",
                    hipGetErrorString(e));
    }
    size_t n = 0;
    e = hipGraphGetNodes(g, nullptr, &n);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(VTA_ERR_HIP, "hipGraphGetNodes: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(VTA_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    vta_graph *r = new vta_graph;
    r->graph = g;
    r->exec = x;
    r->device = ctx->device;
    r->n_nodes = n;
    *out = r;
    return VTA_OK;
}

vta_status vta_graph_launch(vta_ctx *ctx, const vta_graph *graph) {
    clear_error();
    if (!ctx || !graph) return fail(VTA_ERR_INVALID_INPUT, "null argument");
    if (graph->device != ctx->device)
        return fail(VTA_ERR_DEVICE_MISMATCH, "graph captured on device %d, context on %d",
                    graph->device, ctx->device);
    VTA_HIP(hipSetDevice(ctx->device));
    VTA_HIP(hipGraphLaunch(graph->exec, ctx->stream));
    return VTA_OK;
}

size_t vta_graph_num_nodes(const vta_graph *graph) { return graph ? graph->n_nodes : 0; }

vta_status vta_graph_destroy(vta_graph *graph) {
    if (!graph) return VTA_OK;
    (void)hipSetDevice(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return VTA_OK;
}

}  // extern "C"
