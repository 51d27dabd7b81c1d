// Seeded synthetic graph generators for the k-CDS workloads.
//
// These produce the inputs of the benchmark and test workloads: Erdos-Renyi
// G(n,p) (Batagelj-Brandes geometric skipping), Graph500-style R-MAT without
// noise, Chung-Lu power-law graphs (Miller-Hagberg skipping), each with planted
// cliques or near-cliques on uniformly chosen vertices.  Every graph is
// undirected, has no self-loops and no duplicate edges, and is returned as a
// CSR whose neighbour lists are sorted ascending (int32 ids 0..n-1).
//
// The RNG is SplitMix64 seeding xoshiro256** so a (seed, stream) pair always
// gives the same graph on any host.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0,1)
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, bound)
  uint64_t below(uint64_t bound) {
    // Lemire's multiply-shift with rejection: unbiased
    __uint128_t mlt = (__uint128_t)next() * bound;
    uint64_t lo = (uint64_t)mlt;
    if (lo < bound) {
      uint64_t t = (0 - bound) % bound;
      while (lo < t) {
        mlt = (__uint128_t)next() * bound;
        lo = (uint64_t)mlt;
      }
    }
    return (uint64_t)(mlt >> 64);
  }
};

struct EdgeSet {
  std::vector<uint64_t> e;  // packed (min<<32)|max
  void add(uint32_t u, uint32_t v) {
    if (u == v) return;
    if (u > v) std::swap(u, v);
    e.push_back(((uint64_t)u << 32) | v);
  }
};

// Fisher-Yates over 0..n-1; returns the first `take` entries.
std::vector<uint32_t> shuffled_prefix(Rng &rng, uint32_t n, uint32_t take) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  if (take > n) take = n;
  for (uint32_t i = 0; i < take; ++i) {
    uint32_t j = i + (uint32_t)rng.below(n - i);
    std::swap(p[i], p[j]);
  }
  p.resize(take);
  return p;
}

void er_edges(Rng &rng, uint32_t n, double p, EdgeSet &es) {
  if (p <= 0.0 || n < 2) return;
  if (p >= 1.0) {
    for (uint32_t v = 1; v < n; ++v)
      for (uint32_t w = 0; w < v; ++w) es.add(v, w);
    return;
  }
  const double lp = std::log(1.0 - p);
  int64_t v = 1, w = -1;
  while (v < (int64_t)n) {
    double r = rng.uni();
    w = w + 1 + (int64_t)std::floor(std::log(1.0 - r) / lp);
    while (w >= v && v < (int64_t)n) {
      w -= v;
      ++v;
    }
    if (v < (int64_t)n) es.add((uint32_t)v, (uint32_t)w);
  }
}

void plant_clique(const uint32_t *vs, uint32_t s, double density, Rng &rng, EdgeSet &es) {
  for (uint32_t i = 0; i < s; ++i)
    for (uint32_t j = i + 1; j < s; ++j)
      if (density >= 1.0 || rng.uni() < density) es.add(vs[i], vs[j]);
}

}  // namespace

extern "C" {

// A CSR graph owned by this library; release with kcdsgen_free.
typedef struct kcdsgen_graph {
  uint32_t n;
  uint64_t m;          // undirected edges
  uint32_t *offset;    // n+1
  uint32_t *nbr;       // 2m, each list sorted ascending
  uint32_t nplanted;   // vertices listed in `planted`
  uint32_t *planted;   // planted (near-)clique members, clique after clique
  double weight_scale; // Chung-Lu weight scale factor actually used (1 otherwise)
} kcdsgen_graph;

static int finish(uint32_t n, EdgeSet &es, kcdsgen_graph *out) {
  std::sort(es.e.begin(), es.e.end());
  es.e.erase(std::unique(es.e.begin(), es.e.end()), es.e.end());
  uint64_t m = es.e.size();
  if (2 * m >= 0xFFFFFFFFull) return -1;
  out->n = n;
  out->m = m;
  out->offset = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
  out->nbr = (uint32_t *)malloc(std::max<size_t>(1, 2 * m) * sizeof(uint32_t));
  if (!out->offset || !out->nbr) return -2;
  std::vector<uint32_t> deg(n, 0);
  for (uint64_t x : es.e) {
    deg[x >> 32]++;
    deg[(uint32_t)x]++;
  }
  for (uint32_t v = 0; v < n; ++v) out->offset[v + 1] = out->offset[v] + deg[v];
  std::vector<uint32_t> pos(out->offset, out->offset + n);
  // edges sorted by (u,v) with u<v: appending v to u's list and u to v's list
  // keeps every list ascending (u's partners > u come in order; v's partners
  // < v arrive in increasing u order before any partner > v is appended).
  for (uint64_t x : es.e) {
    uint32_t u = (uint32_t)(x >> 32), v = (uint32_t)x;
    out->nbr[pos[v]++] = u;
  }
  for (uint64_t x : es.e) {
    uint32_t u = (uint32_t)(x >> 32), v = (uint32_t)x;
    out->nbr[pos[u]++] = v;
  }
  // lists of v: first all partners u<v (ascending), then partners > v ascending
  std::vector<uint64_t>().swap(es.e);
  return 0;
}

static void set_planted(kcdsgen_graph *out, const std::vector<uint32_t> &pl) {
  out->nplanted = (uint32_t)pl.size();
  out->planted = (uint32_t *)malloc(std::max<size_t>(1, pl.size()) * sizeof(uint32_t));
  if (!pl.empty()) memcpy(out->planted, pl.data(), pl.size() * sizeof(uint32_t));
}

void kcdsgen_free(kcdsgen_graph *g) {
  if (!g) return;
  free(g->offset);
  free(g->nbr);
  free(g->planted);
  memset(g, 0, sizeof(*g));
}

// Erdos-Renyi G(n, p) plus one planted clique of size `clique` (0 = none).
int kcdsgen_er(uint32_t n, double p, uint32_t clique, uint64_t seed, kcdsgen_graph *out) {
  memset(out, 0, sizeof(*out));
  out->weight_scale = 1.0;
  Rng rng(seed);
  EdgeSet es;
  er_edges(rng, n, p, es);
  std::vector<uint32_t> pl = shuffled_prefix(rng, n, clique);
  plant_clique(pl.data(), (uint32_t)pl.size(), 1.0, rng, es);
  set_planted(out, pl);
  return finish(n, es, out);
}

// R-MAT (Graph500 recursive quadrants, no noise) on 2^scale ids with
// edge_factor * 2^scale samples; ids >= n fold back by modulo.  Then plants
// `ncliques` disjoint cliques with sizes `sizes[i]` on shuffled vertices.
int kcdsgen_rmat(uint32_t n, uint32_t scale, double edge_factor, double a, double b, double c,
                 const uint32_t *sizes, uint32_t ncliques, uint64_t seed, kcdsgen_graph *out) {
  memset(out, 0, sizeof(*out));
  out->weight_scale = 1.0;
  if (scale > 31 || n == 0) return -1;
  Rng rng(seed);
  EdgeSet es;
  uint64_t samples = (uint64_t)std::llround(edge_factor * (double)(1ull << scale));
  es.e.reserve(samples);
  const double ab = a + b, abc = a + b + c;
  for (uint64_t s = 0; s < samples; ++s) {
    uint32_t u = 0, v = 0;
    for (uint32_t bit = 0; bit < scale; ++bit) {
      double r = rng.uni();
      uint32_t bu, bv;
      if (r < a) { bu = 0; bv = 0; }
      else if (r < ab) { bu = 0; bv = 1; }
      else if (r < abc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u >= n) u %= n;
    if (v >= n) v %= n;
    es.add(u, v);
  }
  uint32_t total = 0;
  for (uint32_t i = 0; i < ncliques; ++i) total += sizes[i];
  std::vector<uint32_t> pl = shuffled_prefix(rng, n, total);
  uint32_t at = 0;
  for (uint32_t i = 0; i < ncliques && at < pl.size(); ++i) {
    uint32_t s = std::min<uint32_t>(sizes[i], (uint32_t)pl.size() - at);
    plant_clique(pl.data() + at, s, 1.0, rng, es);
    at += s;
  }
  set_planted(out, pl);
  return finish(n, es, out);
}

// Chung-Lu: weights from a Pareto density w^-exponent truncated to [wmin,wmax]
// (inverse CDF), then all weights scaled by c so that sum(w)/2 equals
// target_edges (c = 1 when target_edges == 0).  Edge {i,j} appears with
// probability min(1, w_i w_j / sum w), sampled by Miller-Hagberg skipping over
// vertices sorted by weight.  Plants one near-clique of size `nc` in which each
// pair is an edge with probability `nc_density`.
int kcdsgen_chunglu(uint32_t n, double exponent, double wmin, double wmax, double target_edges,
                    uint32_t nc, double nc_density, uint64_t seed, kcdsgen_graph *out) {
  memset(out, 0, sizeof(*out));
  Rng rng(seed);
  const double al = exponent - 1.0;  // tail index of the Pareto CDF
  const double lo = std::pow(wmin, -al), hi = std::pow(wmax, -al);
  std::vector<std::pair<double, uint32_t>> w(n);
  double sum = 0;
  for (uint32_t i = 0; i < n; ++i) {
    double u = rng.uni();
    double x = std::pow(lo - u * (lo - hi), -1.0 / al);
    w[i] = {x, i};
    sum += x;
  }
  double cs = 1.0;
  if (target_edges > 0) cs = 2.0 * target_edges / sum;
  for (auto &x : w) x.first *= cs;
  sum *= cs;
  out->weight_scale = cs;
  std::sort(w.begin(), w.end(), [](const auto &x, const auto &y) {
    return x.first > y.first || (x.first == y.first && x.second < y.second);
  });
  EdgeSet es;
  es.e.reserve((size_t)(sum / 2 * 1.05) + 16);
  for (uint32_t u = 0; u + 1 < n; ++u) {
    uint64_t v = u + 1;
    double p = std::min(w[u].first * w[v].first / sum, 1.0);
    while (v < n && p > 0) {
      if (p != 1.0) {
        double r = rng.uni();
        v += (uint64_t)std::floor(std::log(1.0 - r) / std::log(1.0 - p));
      }
      if (v < n) {
        double q = std::min(w[u].first * w[v].first / sum, 1.0);
        double r = rng.uni();
        if (r < q / p) es.add(w[u].second, w[v].second);
        p = q;
        v += 1;
      }
    }
  }
  std::vector<uint32_t> pl = shuffled_prefix(rng, n, nc);
  plant_clique(pl.data(), (uint32_t)pl.size(), nc_density, rng, es);
  set_planted(out, pl);
  return finish(n, es, out);
}

// Writes the graph as one 'u v' line per undirected edge (u < v).
int kcdsgen_write_edgelist(const char *path, uint32_t n, const uint32_t *offset, const uint32_t *nbr) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t i = offset[u]; i < offset[u + 1]; ++i)
      if (nbr[i] > u) fprintf(f, "%u %u\n", u, nbr[i]);
  return fclose(f) == 0 ? 0 : -1;
}

// Writes the reference's adjacency-list format: a header "n m", then one line
// "v nbr nbr ..." per vertex in order 0..n-1 with both directions listed.
int kcdsgen_write_adjlist(const char *path, uint32_t n, const uint32_t *offset, const uint32_t *nbr) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "%u %u\n", n, offset[n] / 2);
  for (uint32_t u = 0; u < n; ++u) {
    fprintf(f, "%u", u);
    for (uint32_t i = offset[u]; i < offset[u + 1]; ++i) fprintf(f, " %u", nbr[i]);
    fputc('\n', f);
  }
  return fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"
