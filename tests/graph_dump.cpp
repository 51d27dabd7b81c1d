// Test driver: loads a graph file with the product's Graph (test_amd/csrc/graph.cpp),
// runs getListingOrder (and so coreDecompose), and prints every member as JSON.
#include <cstdio>
#include <exception>
#include <vector>

#include "../include/kcds_graph.h"

static void arr(const char *name, const std::vector<ui> &v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ", %u" : "%u", v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  try {
    Graph g(argv[1]);
    std::vector<ui> lo;
    g.getListingOrder(lo);
    std::printf("{\"n\": %u, \"m\": %u, \"kmax\": %u,\n", g.n, g.m, g.kmax);
    arr("offset", g.offset);
    arr("neighbors", g.neighbors);
    arr("degree", g.degree);
    arr("core", g.core);
    arr("corePeelSequence", g.corePeelSequence);
    arr("listingOrder", lo, true);
    std::printf("}\n");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
