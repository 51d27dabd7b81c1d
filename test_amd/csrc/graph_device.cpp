// Device lifecycle wrappers of include/kcds_graph.h (the reference's
// memoryAllocationGraph / freeGraph, src/gpuMemoryAllocation.cu:5-27 and
// :159-168) over one kcds_ctx.
#include "../../include/kcds_graph.h"

int memoryAllocationGraph(deviceGraphPointers &G, const Graph &graph, int device, uint64_t hbm_budget_bytes) {
  if (G.ctx) freeGraph(G);
  int rc = kcds_create(&G.ctx, device, hbm_budget_bytes);
  if (rc != KCDS_OK) {
    G.ctx = nullptr;
    return rc;
  }
  rc = kcds_upload_csr(G.ctx, graph.n, graph.offset.data(), graph.neighbors.data());
  if (rc != KCDS_OK) return rc;
  G.n = graph.n;
  G.m = graph.m;
  return KCDS_OK;
}

void freeGraph(deviceGraphPointers &G) {
  if (G.ctx) kcds_destroy(G.ctx);
  G.ctx = nullptr;
  G.n = G.m = 0;
}
