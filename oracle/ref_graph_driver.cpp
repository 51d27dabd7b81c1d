// Driver for the reference's own host graph code (src/graph.cpp and
// src/common.cpp of wajidmanzoor/test, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/ref_graph).  TEST INFRASTRUCTURE ONLY.
//
// Usage: ref_graph <adjacency-list file>
// Prints, one per line, space-separated:
//   n m kmax
//   offset[0..n]
//   neighbors[0..2m)        (as the reference loader stores them)
//   degree[0..n) after coreDecompose (graph.cpp overwrites it)
//   core[0..n)
//   corePeelSequence[0..n)
//   listingOrder[0..n)      (main.cu:1299-1301 via Graph::getListingOrder)
#include <cstdio>
#include <vector>

#include "inc/graph.h"

static void dump(const std::vector<ui> &v, size_t len) {
  for (size_t i = 0; i < len; ++i) printf(i ? " %u" : "%u", v[i]);
  printf("\n");
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s graph\n", argv[0]);
    return 2;
  }
  Graph g(argv[1]);
  std::vector<ui> degree0 = g.degree;
  std::vector<ui> lo(g.n);
  g.getListingOrder(lo);
  printf("%u %u %u\n", g.n, g.m, g.kmax);
  dump(g.offset, g.n + 1);
  dump(g.neighbors, g.offset[g.n]);
  dump(g.degree, g.n);
  dump(g.core, g.n);
  dump(g.corePeelSequence, g.n);
  dump(lo, g.n);
  (void)degree0;
  return 0;
}
