// HIP graph capture and replay of a stream's work, for the GMRES chunks:
// the iterations of a restart cycle (preconditioner, operator, Arnoldi and
// Givens kernels) are enqueued once under capture and replayed as one graph
// per chunk.  No counterpart in the reference (PETSc's KSPSolve launches
// nothing on a device).
#include "rtx_internal.h"

using namespace rtx;

/* Begin capturing the work enqueued on `stream` (not the legacy default
   stream) by this thread. */
extern "C" int rtx_graph_capture_begin(void *stream) {
    RTX_CHECK(stream != nullptr, "rtx_graph_capture_begin: capture needs a created stream");
    RTX_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                  hipStreamCaptureModeThreadLocal));
    return RTX_OK;
}

/* End the capture and instantiate it: *graph_exec receives the executable
   graph (release with rtx_graph_destroy). */
extern "C" int rtx_graph_capture_end(void *stream, void **graph_exec) {
    RTX_CHECK(stream != nullptr && graph_exec != nullptr, "rtx_graph_capture_end: null argument");
    hipGraph_t g = nullptr;
    RTX_HIP(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        return RTX_ERR_HIP;
    }
    *graph_exec = ex;
    return RTX_OK;
}

/* Replay an instantiated graph on `stream`. */
extern "C" int rtx_graph_launch(void *graph_exec, void *stream) {
    RTX_CHECK(graph_exec != nullptr, "rtx_graph_launch: null graph");
    RTX_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                           static_cast<hipStream_t>(stream)));
    return RTX_OK;
}

extern "C" int rtx_graph_destroy(void *graph_exec) {
    if (graph_exec) RTX_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
    return RTX_OK;
}
