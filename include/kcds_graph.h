// kcds_graph.h -- the reference's host-side C++ surface, kept for drop-in use.
//
//   class Graph            inc/graph.h:12-30, src/graph.cpp:1-113
//   memoryAllocation* /    inc/gpuMemoryAllocation.cuh:7-31,
//   free*                  src/gpuMemoryAllocation.cu:1-227
//
// Graph loads the reference's adjacency-list format (header "n m", then one
// line "v nbr nbr ..." per vertex in order) and also 'u v' edge lists; both
// are symmetrised, deduplicated and self-loop free, every neighbour list
// sorted ascending.  getListingOrder / coreDecompose compute the same
// Batagelj-Zaversnik peel as src/graph.cpp:53-113 (same peel sequence).
//
// The device side is one kcds_ctx (include/kcds.h) per graph: every device
// buffer the reference allocates with its own alloc/free pair is owned by the
// context, allocated (grow-only, against the context's HBM budget) by the
// stage that needs it and released by freeGraph (kcds_destroy).  Each pair of
// inc/gpuMemoryAllocation.cuh:7-31 maps to:
//   memoryAllocationGraph / freeGraph          memoryAllocationGraph /
//                                              freeGraph below: kcds_create +
//                                              kcds_upload_csr / kcds_destroy
//   memoryAllocationDAG / freeDAG              kcds_orient (ranks, DAG offsets
//                                              and arcs) / the context
//   memoryAllocationTrie / freeTrie            kcds_list (clique list, vertex
//                                              incidence, T * (8k + 4) bytes);
//                                              kcds_count needs only counters
//   memoryAllocationlevelData / freeLevelData  none: partial cliques live in
//                                              LDS frames, not level buffers
//   memoryAllocationDensestCore /              kcds_peel (core numbers, clique
//     freeDensestCore                          statuses, peel records)
//   memoryAllocationPrunnedNeighbors /         kcds_components (arcs of the
//     freePruneneighbors                       core-clique graph)
//   memoryAllocationComponent /                kcds_components (labels and
//     freeComponents                           component-grouped vertices)
//   memoryAllocationFlowNetwork /              kcds_maxflow / kcds_solve (one
//     freeFlownetwork                          network per component, built on
//                                              the device)
// and every one of them is released by freeGraph.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "kcds.h"

typedef unsigned int ui;

class Graph {
 public:
  ui n = 0;
  ui m = 0;  // undirected edges
  ui kmax = 0;

  std::vector<ui> offset;
  std::vector<ui> neighbors;
  std::vector<ui> degree;
  std::vector<ui> core;
  std::vector<ui> corePeelSequence;
  std::string filePath;

 public:
  Graph();
  // Throws std::runtime_error when the file cannot be opened or parsed.
  explicit Graph(std::string path);
  // Builds from an undirected CSR (lists need not be sorted or deduplicated).
  Graph(ui n, const std::vector<ui> &offset, const std::vector<ui> &neighbors);
  // arr[v] = 1 + position of v in corePeelSequence (src/graph.cpp:41-51).
  void getListingOrder(std::vector<ui> &arr);
  // Fills core[] and kmax; arr receives the peel sequence (src/graph.cpp:53-113).
  // Unlike the reference, degree[] is left unchanged.
  void coreDecompose(std::vector<ui> &arr);
};

// Device-side graph: the context that owns every device buffer of a solve.
struct deviceGraphPointers {
  kcds_ctx *ctx = nullptr;
  ui n = 0;
  ui m = 0;
};

// Creates the context on `device` and copies the CSR (memoryAllocationGraph,
// src/gpuMemoryAllocation.cu:5-27).  Returns a kcds_status.
int memoryAllocationGraph(deviceGraphPointers &G, const Graph &graph, int device = 0,
                          uint64_t hbm_budget_bytes = 0);
// Releases the context and every buffer it owns (freeGraph and the other free*
// functions of src/gpuMemoryAllocation.cu:159-227).
void freeGraph(deviceGraphPointers &G);
