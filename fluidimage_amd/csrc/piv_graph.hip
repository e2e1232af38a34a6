// Capture and replay of a sequence of libpiv_hip calls as one HIP graph.
//
// Every compute entry point of the library only launches kernels on the
// caller's stream (no allocation, no memset, no synchronisation), so the
// calls issued on a stream between piv_graph_begin and piv_graph_end are
// recorded as graph nodes; piv_graph_launch replays them in one submission.
// The reference has no counterpart: its per-window loop
// (src/fluidimage/works/piv/singlepass.py:262-361) pays the host's cost per
// window; here a repeated batch pays the launch cost of its kernels once.

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/piv_hip.h"
#include "last_error.hpp"

struct piv_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int graph_fail(int code, const std::string& msg) {
  fia::last_error() = msg;
  return code;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return PIV_OK;
  return graph_fail(PIV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int piv_graph_begin(void* stream) {
  if (stream == nullptr)
    return graph_fail(PIV_EINVAL, "piv_graph_begin: the null stream cannot be captured");
  // thread-local mode: the capturing thread may not call allocating or
  // synchronising runtime functions until piv_graph_end; other threads may
  return hip_status(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                          hipStreamCaptureModeThreadLocal),
                    "piv_graph_begin");
}

int piv_graph_end(void* stream, piv_graph** graph) {
  if (stream == nullptr || graph == nullptr)
    return graph_fail(PIV_EINVAL, "piv_graph_end: NULL argument");
  *graph = nullptr;
  hipGraph_t g = nullptr;
  int rc = hip_status(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g),
                      "piv_graph_end");
  if (rc != PIV_OK) {
    if (g != nullptr) (void)hipGraphDestroy(g);
    return rc;
  }
  hipGraphExec_t exec = nullptr;
  rc = hip_status(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "piv_graph_end");
  if (rc != PIV_OK) {
    (void)hipGraphDestroy(g);
    return rc;
  }
  auto* out = new piv_graph;
  out->graph = g;
  out->exec = exec;
  *graph = out;
  return PIV_OK;
}

int piv_graph_launch(piv_graph* graph, void* stream) {
  if (graph == nullptr || graph->exec == nullptr)
    return graph_fail(PIV_EINVAL, "piv_graph_launch: NULL graph");
  return hip_status(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)),
                    "piv_graph_launch");
}

int piv_graph_destroy(piv_graph* graph) {
  if (graph == nullptr) return PIV_OK;
  int rc = PIV_OK;
  if (graph->exec != nullptr) rc = hip_status(hipGraphExecDestroy(graph->exec), "piv_graph_destroy");
  if (graph->graph != nullptr) {
    const int rc2 = hip_status(hipGraphDestroy(graph->graph), "piv_graph_destroy");
    if (rc == PIV_OK) rc = rc2;
  }
  delete graph;
  return rc;
}

}  // extern "C"
