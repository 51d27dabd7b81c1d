// Host graph loader, CSR build and Batagelj-Zaversnik ordering (the
// reference's inc/graph.h / src/graph.cpp surface).  Host code only: the
// device lifecycle wrappers of include/kcds_graph.h are in graph_device.cpp.
#include "../../include/kcds_graph.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace {

// Builds a sorted, deduplicated, symmetric CSR from undirected pairs.
void csr_from_pairs(ui n, std::vector<uint64_t> &e, std::vector<ui> &off, std::vector<ui> &nbr) {
  for (auto &x : e) {
    ui u = (ui)(x >> 32), v = (ui)x;
    if (u > v) x = ((uint64_t)v << 32) | u;
  }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  off.assign((size_t)n + 1, 0);
  for (uint64_t x : e) {
    off[(x >> 32) + 1]++;
    off[(ui)x + 1]++;
  }
  for (ui v = 0; v < n; ++v) off[v + 1] += off[v];
  nbr.resize(off[n]);
  std::vector<ui> pos(off.begin(), off.end() - 1);
  for (uint64_t x : e) nbr[pos[(ui)x]++] = (ui)(x >> 32);
  for (uint64_t x : e) nbr[pos[x >> 32]++] = (ui)x;
}

bool parse_u64(const char *&p, uint64_t &out) {
  while (*p == ' ' || *p == '\t' || *p == '\r' || *p == ',') ++p;
  if (*p < '0' || *p > '9') return false;
  uint64_t x = 0;
  while (*p >= '0' && *p <= '9') x = x * 10 + (uint64_t)(*p++ - '0');
  out = x;
  return true;
}

}  // namespace

Graph::Graph() {}

Graph::Graph(ui n_, const std::vector<ui> &off, const std::vector<ui> &nb) : n(n_) {
  std::vector<uint64_t> e;
  e.reserve(nb.size());
  for (ui v = 0; v < n; ++v)
    for (ui i = off[v]; i < off[v + 1]; ++i)
      if (nb[i] != v) e.push_back(((uint64_t)v << 32) | nb[i]);
  csr_from_pairs(n, e, offset, neighbors);
  m = (ui)(neighbors.size() / 2);
  degree.resize(n);
  for (ui v = 0; v < n; ++v) degree[v] = offset[v + 1] - offset[v];
}

// Format detection: the reference adjacency list (src/graph.cpp:7-40) has a
// header "n m" followed by exactly one line per vertex whose first token is
// the vertex id, in order 0..n-1, and its lists are symmetric with at most
// 2m entries besides self-loops (the reference sizes neighbors 2m).  A file
// that looks like that but fails the symmetry or size check is an edge list
// whose first lines happen to fit the pattern.  Anything else is read as
// 'u v' pairs ('#' and '%' lines are comments).  A file that passes every
// check reads as an adjacency list.
Graph::Graph(std::string path) : filePath(path) {
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("Graph file Open Failed: " + path);
  std::vector<std::string> lines;
  {
    std::string line;
    while (std::getline(in, line)) {
      if (!line.empty() && line.back() == '\r') line.pop_back();
      size_t b = line.find_first_not_of(" \t");
      if (b == std::string::npos) continue;
      if (line[b] == '#' || line[b] == '%') continue;
      lines.push_back(line);
    }
  }
  if (lines.empty()) throw std::runtime_error("empty graph file: " + path);
  std::vector<uint64_t> e;
  bool adjacency = false;
  uint64_t hn = 0, hm = 0;
  {
    const char *p = lines[0].c_str();
    uint64_t extra;
    if (parse_u64(p, hn) && parse_u64(p, hm) && !parse_u64(p, extra) && lines.size() == hn + 1 && hn > 0) {
      adjacency = true;
      for (uint64_t v = 0; v < hn && adjacency; ++v) {
        const char *q = lines[v + 1].c_str();
        uint64_t id;
        adjacency = parse_u64(q, id) && id == v;
      }
    }
  }
  uint64_t nn = 0;
  if (adjacency) {
    for (uint64_t v = 0; v < hn && adjacency; ++v) {
      const char *q = lines[v + 1].c_str();
      uint64_t id, w;
      parse_u64(q, id);
      while (parse_u64(q, w)) {
        if (w >= hn) {
          adjacency = false;
          break;
        }
        if (w != v) e.push_back((v << 32) | w);  // self-loops skipped (graph.cpp:28-29)
      }
    }
    if (adjacency && e.size() > 2 * hm) adjacency = false;
    if (adjacency) {  // every arc v -> w has its w -> v
      std::vector<uint64_t> a(e), t(e.size());
      for (size_t i = 0; i < e.size(); ++i) t[i] = (e[i] << 32) | (e[i] >> 32);
      std::sort(a.begin(), a.end());
      std::sort(t.begin(), t.end());
      adjacency = a == t;
    }
    if (adjacency) nn = hn;
    else e.clear();
  }
  if (!adjacency) {
    for (const std::string &ln : lines) {
      const char *q = ln.c_str();
      uint64_t u, v;
      if (!parse_u64(q, u) || !parse_u64(q, v)) throw std::runtime_error("malformed edge line in " + path);
      if (u > 0xFFFFFFFEull || v > 0xFFFFFFFEull) throw std::runtime_error("vertex id exceeds 32 bits");
      nn = std::max(nn, std::max(u, v) + 1);
      if (u != v) e.push_back((u << 32) | v);
    }
  }
  n = (ui)nn;
  csr_from_pairs(n, e, offset, neighbors);
  m = (ui)(neighbors.size() / 2);
  degree.resize(n);
  for (ui v = 0; v < n; ++v) degree[v] = offset[v + 1] - offset[v];
}

void Graph::getListingOrder(std::vector<ui> &arr) {
  corePeelSequence.resize(n);
  coreDecompose(corePeelSequence);
  arr.resize(n);
  for (size_t i = 0; i < n; ++i) arr[corePeelSequence[i]] = (ui)i + 1;
}

// Bin-sorted peel: vertices sit in degree bins; removing the lowest vertex
// moves each higher-degree neighbour to the front of its bin and shrinks that
// bin (the same swap order as src/graph.cpp:80-106: for neighbour lists in
// the same order it yields the same peel sequence).
void Graph::coreDecompose(std::vector<ui> &arr) {
  core.assign(n, 0);
  arr.resize(n);
  kmax = 0;
  if (n == 0) return;
  std::vector<ui> deg(degree);
  ui maxd = *std::max_element(deg.begin(), deg.end());
  std::vector<ui> bins((size_t)maxd + 1, 0);
  for (ui d : deg) bins[d]++;
  std::vector<int64_t> start((size_t)maxd + 1, 0);
  std::partial_sum(bins.begin(), bins.end(), start.begin());
  std::vector<ui> pos(n), vert(n);
  for (ui v = 0; v < n; ++v) {
    pos[v] = (ui)--start[deg[v]];
    vert[pos[v]] = v;
  }
  for (ui i = 0; i < n; ++i) {
    const ui v = vert[i];
    core[v] = deg[v];
    arr[n - i - 1] = v;
    for (ui j = offset[v]; j < offset[v + 1]; ++j) {
      const ui u = neighbors[j];
      if (deg[u] <= deg[v]) continue;
      const ui du = deg[u], pu = pos[u], pw = (ui)start[du], w = vert[pw];
      if (u != w) {
        pos[u] = pw;
        vert[pu] = w;
        pos[w] = pu;
        vert[pw] = u;
      }
      start[du]++;
      deg[u]--;
    }
  }
  kmax = *std::max_element(core.begin(), core.end());
}
