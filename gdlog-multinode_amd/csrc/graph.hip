// graph.hip -- capture of a context's stream into a HIP graph and its replay
// (the hipGraph capture / launch entry points of include/gdl_hip.h).
//
// The capture runs in thread-local mode: an entry point that synchronizes the
// stream, reads a result back or allocates while capturing fails (HIP refuses
// the call) and the capture ends in an error, instead of recording a graph
// that would replay something else.  Kernel timing is refused while
// capturing (its events would be recorded into the graph).  Nothing here adds
// memset nodes.
#include "gdl_internal.hpp"

struct gdl_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using namespace gdl;

extern "C" {

gdl_status gdl_graph_begin(gdl_ctx *ctx) {
    if (!ctx) return GDL_ERR_INVALID_ARG;
    if (ctx->capturing) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_graph_begin: already capturing");
    if (ctx->stats_on)
        return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_graph_begin: kernel statistics are on");
    GDL_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return GDL_OK;
}

gdl_status gdl_graph_end(gdl_ctx *ctx, gdl_graph **out) {
    if (!ctx || !out) return GDL_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ctx->capturing) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_graph_end: not capturing");
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return hip_fail(ctx, e, "hipStreamEndCapture (a captured call synchronized or allocated)");
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(ctx, e, "hipGraphInstantiate");
    }
    *out = new gdl_graph{g, x};
    return GDL_OK;
}

gdl_status gdl_graph_launch(gdl_ctx *ctx, const gdl_graph *graph) {
    if (!ctx || !graph || !graph->exec) return GDL_ERR_INVALID_ARG;
    if (ctx->capturing) return fail(ctx, GDL_ERR_INVALID_ARG, "gdl_graph_launch: capturing");
    GDL_HIP(ctx, hipGraphLaunch(graph->exec, ctx->stream));
    return GDL_OK;
}

gdl_status gdl_graph_destroy(gdl_graph *graph) {
    if (!graph) return GDL_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return GDL_OK;
}

}  // extern "C"
