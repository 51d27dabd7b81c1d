// Drop-in CLI for the reference driver (main.cu:1272-1522).
//
//   kcds <graph> <k> [pSize cpSize glBufferSize partitionSize]
//
// The four sizing arguments of the reference (main.cu:1280-1293) are
// accepted and ignored: buffers here are sized from the graph.  The graph may
// be the reference adjacency list or a 'u v' edge list.  Output keeps the
// reference's two final lines (main.cu:1269, 1519):
//   Max Density <d> Size <s>
//   Time taken: <ms> ms
// preceded by the exact rational density and stage details.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "../../include/kcds_graph.h"

int main(int argc, const char *argv[]) {
  if (argc != 3 && argc != 7) {
    std::cout << "usage: " << argv[0] << " <graph> <k> [pSize cpSize glBufferSize partitionSize]" << std::endl;
    return 1;
  }
  const std::string path = argv[1];
  const long k = std::atol(argv[2]);
  if (k < 2 || k > 8) {
    std::cout << "k must be in 2..8" << std::endl;
    return 1;
  }
  Graph graph;
  try {
    graph = Graph(path);
  } catch (const std::exception &e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  std::cout << "n =" << graph.n << ", m=" << graph.m << std::endl;
  deviceGraphPointers G;
  // the timed window starts after loading, as main.cu:1349
  auto start = std::chrono::high_resolution_clock::now();
  int rc = memoryAllocationGraph(G, graph);
  if (rc != KCDS_OK) {
    std::cout << "device error " << rc << ": " << (G.ctx ? kcds_last_error(G.ctx) : "no GPU context") << std::endl;
    freeGraph(G);
    return 1;
  }
  kcds_result res;
  rc = kcds_solve(G.ctx, (uint32_t)k, &res, nullptr);
  auto end = std::chrono::high_resolution_clock::now();
  if (rc != KCDS_OK) {
    std::cout << "solve error " << rc << ": " << kcds_last_error(G.ctx) << std::endl;
    freeGraph(G);
    return 1;
  }
  std::cout << "Total cliques " << res.T << std::endl;
  if (res.T) {
    std::cout << "k prime " << res.k_prime << " core vertices " << res.core_vertices << " components "
              << res.ncomp << " flows " << res.flow_calls << std::endl;
    std::cout << "Exact Density " << res.rho_num << "/" << res.rho_den << std::endl;
  }
  const double density = res.rho_den ? (double)res.rho_num / (double)res.rho_den : 0.0;
  std::cout << "Max Density " << density << " Size " << res.rho_den << std::endl;
  std::chrono::duration<double, std::milli> ms = end - start;
  std::cout << "Time taken: " << ms.count() << " ms" << std::endl;
  freeGraph(G);
  return 0;
}
