// CPU oracle for the k-CDS pipeline -- TEST INFRASTRUCTURE ONLY (checker for
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never part
// of the product library).  See oracle.h for the interface.  Each function
// names the reference file:line whose semantics it restates.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   loader, core_decompose : the reference's own src/graph.cpp compiled into
//                            oracle/_ref (byte-equal CSR, core, peel sequence,
//                            listing order) and hand-derived values for
//                            data/test.txt.
//   count / list           : brute force over all k-subsets (n <= 24),
//                            networkx clique enumeration, closed forms
//                            C(s,k) / C(s-1,k-1) for planted cliques,
//                            sum(per_vertex) == k*T.
//   peel                   : brute-force definition of the clique-core
//                            (largest L such that v lies in an induced
//                            subgraph where every vertex is in >= L cliques),
//                            closed form C(s-1,k-1) on disjoint s-cliques.
//   clique_maxflow         : networkx maximum_flow on the same network, and
//                            max-flow == min-cut value.
//   solve                  : brute force over all vertex subsets (n <= 16),
//                            closed form 247 = C(40,3)/40 for config C1.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <queue>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef unsigned __int128 u128;

template <typename T> T *dup_array(const std::vector<T> &v) {
  T *p = (T *)malloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// a/b >= c/d for non-negative rationals with positive denominators
bool ratio_ge(uint64_t a, uint64_t b, uint64_t c, uint64_t d) { return (u128)a * d >= (u128)c * b; }
bool ratio_gt(uint64_t a, uint64_t b, uint64_t c, uint64_t d) { return (u128)a * d > (u128)c * b; }

// Degeneracy DAG used by the oracle's own listing: arc u->v iff u is removed
// before v by the BZ peel, out-lists sorted by id.  Any acyclic orientation
// lists every clique exactly once; this one bounds out-degree by the core.
struct Dag {
  uint32_t n = 0;
  std::vector<uint32_t> off, nbr;
};

Dag make_dag(uint32_t n, const uint32_t *offset, const uint32_t *nbr) {
  std::vector<uint32_t> core(n), seq(n), lo(n);
  uint32_t kmax = 0;
  oracle_core_decompose(n, offset, nbr, core.data(), seq.data(), lo.data(), &kmax);
  // seq[n-1-i] is the i-th removed vertex, so lo[v] = n - removal index.
  Dag d;
  d.n = n;
  d.off.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t c = 0;
    for (uint32_t i = offset[v]; i < offset[v + 1]; ++i)
      if (lo[nbr[i]] < lo[v]) ++c;
    d.off[v + 1] = d.off[v] + c;
  }
  d.nbr.resize(d.off[n]);
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t at = d.off[v];
    for (uint32_t i = offset[v]; i < offset[v + 1]; ++i)
      if (lo[nbr[i]] < lo[v]) d.nbr[at++] = nbr[i];
    std::sort(d.nbr.begin() + d.off[v], d.nbr.begin() + d.off[v + 1]);
  }
  return d;
}

void intersect(const std::vector<uint32_t> &a, const uint32_t *b, uint32_t nb,
               std::vector<uint32_t> &out) {
  out.clear();
  size_t i = 0, j = 0;
  while (i < a.size() && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      out.push_back(a[i]);
      ++i;
      ++j;
    }
  }
}

struct Lister {
  const Dag &d;
  uint32_t k;
  int64_t *pv;                       // per-vertex counts (may be null)
  std::vector<uint32_t> *sink;       // materialised cliques (may be null)
  uint64_t T = 0;
  std::vector<uint32_t> prefix;
  std::vector<std::vector<uint32_t>> cand;
  Lister(const Dag &d_, uint32_t k_, int64_t *pv_, std::vector<uint32_t> *sink_)
      : d(d_), k(k_), pv(pv_), sink(sink_), cand(k_ + 1) {}

  // prefix holds l vertices; cand[l] the common out-neighbours of the prefix
  void rec(uint32_t l) {
    const std::vector<uint32_t> &C = cand[l];
    if (l == k - 1) {
      T += C.size();
      if (pv) {
        for (uint32_t p : prefix) pv[p] += (int64_t)C.size();
        for (uint32_t c : C) pv[c] += 1;
      }
      if (sink) {
        for (uint32_t c : C) {
          size_t base = sink->size();
          sink->insert(sink->end(), prefix.begin(), prefix.end());
          sink->push_back(c);
          std::sort(sink->begin() + base, sink->end());
        }
      }
      return;
    }
    for (uint32_t c : C) {
      intersect(C, d.nbr.data() + d.off[c], d.off[c + 1] - d.off[c], cand[l + 1]);
      if (cand[l + 1].size() + l + 1 < k) continue;
      prefix.push_back(c);
      rec(l + 1);
      prefix.pop_back();
    }
  }
  void root(uint32_t r) {
    cand[1].assign(d.nbr.begin() + d.off[r], d.nbr.begin() + d.off[r + 1]);
    if (cand[1].size() + 1 < k) return;
    prefix.assign(1, r);
    rec(1);
  }
};

uint64_t count_impl(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                    int64_t *per_vertex, uint32_t rb, uint32_t re, int threads) {
  if (per_vertex) memset(per_vertex, 0, sizeof(int64_t) * n);
  if (k < 2 || n == 0) return 0;
  Dag d = make_dag(n, offset, nbr);
  re = std::min(re, n);
  uint64_t T = 0;
#ifdef _OPENMP
  int nt = threads > 0 ? threads : omp_get_max_threads();
#else
  int nt = 1;
  (void)threads;
#endif
  std::vector<std::vector<int64_t>> local(per_vertex ? nt : 0);
#pragma omp parallel num_threads(nt) reduction(+ : T)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    int64_t *pv = nullptr;
    if (per_vertex) {
      local[tid].assign(n, 0);
      pv = local[tid].data();
    }
    Lister L(d, k, pv, nullptr);
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = rb; r < (int64_t)re; ++r) L.root((uint32_t)r);
    T += L.T;
  }
  if (per_vertex)
    for (auto &l : local)
      for (uint32_t v = 0; v < n; ++v) per_vertex[v] += l[v];
  return T;
}

// ----------------------------------------------------------------------------
// Dinic max-flow on int64 capacities.
struct Flow {
  struct Arc {
    uint32_t to;
    int64_t cap;  // residual
  };
  uint32_t N;
  std::vector<std::vector<uint32_t>> adj;  // arc indices
  std::vector<Arc> arcs;
  std::vector<int32_t> level;
  std::vector<uint32_t> it;
  explicit Flow(uint32_t n) : N(n), adj(n), level(n), it(n) {}
  void add(uint32_t u, uint32_t v, int64_t c) {
    adj[u].push_back((uint32_t)arcs.size());
    arcs.push_back({v, c});
    adj[v].push_back((uint32_t)arcs.size());
    arcs.push_back({u, 0});
  }
  bool bfs(uint32_t s, uint32_t t) {
    std::fill(level.begin(), level.end(), -1);
    std::vector<uint32_t> q;
    q.reserve(N);
    q.push_back(s);
    level[s] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      uint32_t u = q[h];
      for (uint32_t a : adj[u])
        if (arcs[a].cap > 0 && level[arcs[a].to] < 0) {
          level[arcs[a].to] = level[u] + 1;
          q.push_back(arcs[a].to);
        }
    }
    return level[t] >= 0;
  }
  int64_t dfs(uint32_t s, uint32_t t) {
    // iterative augmenting along the level graph
    int64_t total = 0;
    std::vector<uint32_t> path;  // arc indices
    uint32_t u = s;
    while (true) {
      if (u == t) {
        int64_t f = INT64_MAX;
        for (uint32_t a : path) f = std::min(f, arcs[a].cap);
        for (uint32_t a : path) {
          arcs[a].cap -= f;
          arcs[a ^ 1].cap += f;
        }
        total += f;
        path.clear();
        u = s;
        continue;
      }
      bool adv = false;
      for (; it[u] < adj[u].size(); ++it[u]) {
        uint32_t a = adj[u][it[u]];
        uint32_t v = arcs[a].to;
        if (arcs[a].cap > 0 && level[v] == level[u] + 1) {
          path.push_back(a);
          u = v;
          adv = true;
          break;
        }
      }
      if (adv) continue;
      if (u == s) break;
      level[u] = -1;  // dead end
      uint32_t a = path.back();
      path.pop_back();
      u = arcs[a ^ 1].to;
      ++it[u];
    }
    return total;
  }
  int64_t maxflow(uint32_t s, uint32_t t) {
    int64_t f = 0;
    while (bfs(s, t)) {
      std::fill(it.begin(), it.end(), 0);
      f += dfs(s, t);
    }
    return f;
  }
  // nodes that can reach t through residual arcs
  std::vector<uint8_t> reach_t(uint32_t t) const {
    std::vector<uint8_t> r(N, 0);
    std::vector<uint32_t> q{t};
    r[t] = 1;
    for (size_t h = 0; h < q.size(); ++h) {
      uint32_t v = q[h];
      for (uint32_t a : adj[v]) {
        // arc a: v -> w; its partner a^1: w -> v with residual arcs[a^1].cap
        uint32_t w = arcs[a].to;
        if (!r[w] && arcs[a ^ 1].cap > 0) {
          r[w] = 1;
          q.push_back(w);
        }
      }
    }
    return r;
  }
  std::vector<uint8_t> reach_s(uint32_t s) const {
    std::vector<uint8_t> r(N, 0);
    std::vector<uint32_t> q{s};
    r[s] = 1;
    for (size_t h = 0; h < q.size(); ++h) {
      uint32_t v = q[h];
      for (uint32_t a : adj[v])
        if (!r[arcs[a].to] && arcs[a].cap > 0) {
          r[arcs[a].to] = 1;
          q.push_back(arcs[a].to);
        }
    }
    return r;
  }
};

// Builds the clique network of helpers.cu:1148-1264 with capacities scaled by
// q for alpha = p/q: s->v q*deg(v), v->t k*p, v->c q, c->v q*(k-1).
Flow build_network(uint32_t nc, uint32_t k, uint64_t Tc, const uint32_t *cl, int64_t p, int64_t q) {
  uint32_t s = nc + (uint32_t)Tc, t = s + 1;
  Flow F(t + 1);
  std::vector<int64_t> deg(nc, 0);
  for (uint64_t c = 0; c < Tc; ++c)
    for (uint32_t j = 0; j < k; ++j) deg[cl[c * k + j]]++;
  for (uint32_t v = 0; v < nc; ++v) {
    F.add(s, v, q * deg[v]);
    F.add(v, t, (int64_t)k * p);
  }
  for (uint64_t c = 0; c < Tc; ++c)
    for (uint32_t j = 0; j < k; ++j) {
      uint32_t v = cl[c * k + j];
      F.add(v, nc + (uint32_t)c, q);
      F.add(nc + (uint32_t)c, v, q * (int64_t)(k - 1));
    }
  return F;
}

struct Dsu {
  std::vector<uint32_t> p;
  explicit Dsu(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
  }
  void unite(uint32_t a, uint32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b) p[b] = a;  // root = smallest id
    else p[a] = b;
  }
};

}  // namespace

extern "C" {

void oracle_free(void *p) { free(p); }

// src/graph.cpp:7-40
int oracle_load_adjlist(const char *path, uint32_t *n_out, uint32_t *m_out, uint32_t **offset_out,
                        uint32_t **nbr_out) {
  std::ifstream in(path);
  if (!in.is_open()) return -1;
  std::string line;
  if (!std::getline(in, line)) return -2;
  std::istringstream hs(line);
  uint64_t n = 0, m = 0;
  if (!(hs >> n >> m)) return -2;
  std::vector<uint32_t> off(n + 1, 0), nbr(2 * m);
  uint64_t expect = 0;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    int64_t v, w;
    if (!(is >> v)) continue;  // blank line
    if (v != (int64_t)expect || v >= (int64_t)n) return -3;  // lines must be in order 0..n-1
    uint64_t d = 0;
    while (is >> w) {
      if (w == v) continue;  // graph.cpp:28-29 skips self-loops
      if (w < 0 || w >= (int64_t)n) return -4;
      if (off[v] + d >= nbr.size()) return -5;  // more than 2m entries
      nbr[off[v] + d] = (uint32_t)w;
      ++d;
    }
    off[v + 1] = off[v] + (uint32_t)d;
    ++expect;
  }
  for (uint64_t v = expect; v < n; ++v) off[v + 1] = off[v];
  nbr.resize(off[n]);
  *n_out = (uint32_t)n;
  *m_out = (uint32_t)m;
  *offset_out = dup_array(off);
  *nbr_out = dup_array(nbr);
  return 0;
}

int oracle_load_edgelist(const char *path, uint32_t n_hint, uint32_t *n_out, uint32_t *m_out,
                         uint32_t **offset_out, uint32_t **nbr_out) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  std::vector<uint64_t> e;
  uint64_t n = n_hint;
  char buf[256];
  while (fgets(buf, sizeof buf, f)) {
    char *p = buf;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
    unsigned long long u, v;
    if (sscanf(p, "%llu %llu", &u, &v) != 2) {
      fclose(f);
      return -2;
    }
    if (u > 0xFFFFFFFEull || v > 0xFFFFFFFEull) {
      fclose(f);
      return -4;
    }
    n = std::max<uint64_t>(n, std::max(u, v) + 1);
    if (u == v) continue;
    if (u > v) std::swap(u, v);
    e.push_back((u << 32) | v);
  }
  fclose(f);
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  std::vector<uint32_t> off(n + 1, 0), nbr(2 * e.size());
  for (uint64_t x : e) {
    off[(x >> 32) + 1]++;
    off[(uint32_t)x + 1]++;
  }
  for (uint64_t v = 0; v < n; ++v) off[v + 1] += off[v];
  std::vector<uint32_t> pos(off.begin(), off.end() - 1);
  for (uint64_t x : e) nbr[pos[(uint32_t)x]++] = (uint32_t)(x >> 32);
  for (uint64_t x : e) nbr[pos[x >> 32]++] = (uint32_t)x;
  *n_out = (uint32_t)n;
  *m_out = (uint32_t)e.size();
  *offset_out = dup_array(off);
  *nbr_out = dup_array(nbr);
  return 0;
}

// src/graph.cpp:53-113 (Batagelj-Zaversnik bin sort; identical swap order)
void oracle_core_decompose(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t *core,
                           uint32_t *peel_seq, uint32_t *listing_order, uint32_t *kmax) {
  if (n == 0) {
    if (kmax) *kmax = 0;
    return;
  }
  std::vector<uint32_t> degree(n);
  for (uint32_t v = 0; v < n; ++v) degree[v] = offset[v + 1] - offset[v];
  uint32_t maxd = *std::max_element(degree.begin(), degree.end());
  std::vector<uint32_t> bins(maxd + 1, 0);
  for (uint32_t d : degree) bins[d]++;
  std::vector<int64_t> binpos(maxd + 1, 0);
  std::partial_sum(bins.begin(), bins.end(), binpos.begin());
  std::vector<uint32_t> position(n), sorted(n);
  for (uint32_t v = 0; v < n; ++v) {
    position[v] = (uint32_t)--binpos[degree[v]];
    sorted[position[v]] = v;
  }
  std::vector<uint32_t> seq(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t v = sorted[i];
    core[v] = degree[v];
    seq[n - i - 1] = v;
    for (uint32_t j = offset[v]; j < offset[v + 1]; ++j) {
      uint32_t u = nbr[j];
      if (degree[u] > degree[v]) {
        uint32_t du = degree[u], pu = position[u];
        uint32_t pw = (uint32_t)binpos[du];
        uint32_t w = sorted[pw];
        if (u != w) {
          position[u] = pw;
          sorted[pu] = w;
          position[w] = pu;
          sorted[pw] = u;
        }
        binpos[du]++;
        degree[u]--;
      }
    }
  }
  uint32_t km = core[0];
  for (uint32_t v = 0; v < n; ++v) km = std::max(km, core[v]);
  if (kmax) *kmax = km;
  if (peel_seq) memcpy(peel_seq, seq.data(), sizeof(uint32_t) * n);
  if (listing_order)
    for (uint32_t i = 0; i < n; ++i) listing_order[seq[i]] = i + 1;
}

uint64_t oracle_count(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                      int64_t *per_vertex, int threads) {
  return count_impl(n, offset, nbr, k, per_vertex, 0, n, threads);
}

uint64_t oracle_count_roots(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                            uint32_t root_begin, uint32_t root_end, int threads) {
  return count_impl(n, offset, nbr, k, nullptr, root_begin, root_end, threads);
}

uint64_t oracle_list(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                     uint32_t *out, uint64_t cap) {
  if (k < 2 || n == 0) return 0;
  Dag d = make_dag(n, offset, nbr);
  std::vector<uint32_t> all;
  Lister L(d, k, nullptr, &all);
  for (uint32_t r = 0; r < n; ++r) L.root(r);
  uint64_t T = L.T;
  std::vector<uint64_t> idx(T);
  std::iota(idx.begin(), idx.end(), 0ull);
  std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
    return std::lexicographical_compare(all.begin() + a * k, all.begin() + a * k + k,
                                        all.begin() + b * k, all.begin() + b * k + k);
  });
  for (uint64_t i = 0; i < std::min(T, cap); ++i)
    memcpy(out + i * k, all.data() + idx[i] * k, sizeof(uint32_t) * k);
  return T;
}

// main.cu:329-420 with helpers.cu:587-726.  Level-synchronous semantics: at
// level L every vertex whose live-clique count is <= L is removed (cascading
// within the level), its live cliques are destroyed with status L.  After a
// level, density = live cliques / remaining vertices; the maximum (ties keep
// the later core, main.cu:407) is maxDensity.  Implemented as an exact
// min-count peel: core(v) = max(L, count(v)) at removal.
int oracle_peel(uint32_t n, uint32_t k, uint64_t T, const uint32_t *cl, int32_t *core_out,
                int32_t *status_out, oracle_peel_stats *st) {
  std::vector<uint64_t> ioff(n + 1, 0);
  for (uint64_t c = 0; c < T; ++c)
    for (uint32_t j = 0; j < k; ++j) ioff[cl[c * k + j] + 1]++;
  for (uint32_t v = 0; v < n; ++v) ioff[v + 1] += ioff[v];
  std::vector<uint64_t> inc(ioff[n]);
  {
    std::vector<uint64_t> pos(ioff.begin(), ioff.end() - 1);
    for (uint64_t c = 0; c < T; ++c)
      for (uint32_t j = 0; j < k; ++j) inc[pos[cl[c * k + j]]++] = c;
  }
  std::vector<int64_t> cnt(n);
  for (uint32_t v = 0; v < n; ++v) cnt[v] = (int64_t)(ioff[v + 1] - ioff[v]);
  std::vector<int32_t> core(n, -1);
  std::vector<int32_t> status(T, -1);
  typedef std::pair<int64_t, uint32_t> E;
  std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
  for (uint32_t v = 0; v < n; ++v) pq.push({cnt[v], v});
  int64_t L = 0;
  uint64_t live = T, remaining = n;
  uint64_t mnum = T, mden = n ? n : 1;  // main.cu:361-366: initial T/n
  uint32_t levels = 0;
  int64_t last_level = -1;
  while (!pq.empty()) {
    E e = pq.top();
    pq.pop();
    uint32_t v = e.second;
    if (core[v] >= 0 || e.first != std::max(cnt[v], L)) continue;  // stale
    if (e.first > L) {
      // level L complete: density of the remaining (L+1)-core (main.cu:398-413);
      // empty levels up to e.first-1 repeat the same density.
      if (remaining > 0 && ratio_ge(live, remaining, mnum, mden)) {
        mnum = live;
        mden = remaining;
      }
      L = e.first;
    }
    core[v] = (int32_t)L;
    if (L != last_level) {
      ++levels;
      last_level = L;
    }
    --remaining;
    for (uint64_t i = ioff[v]; i < ioff[v + 1]; ++i) {
      uint64_t c = inc[i];
      if (status[c] != -1) continue;
      status[c] = (int32_t)L;
      --live;
      for (uint32_t j = 0; j < k; ++j) {
        uint32_t u = cl[c * k + j];
        if (u == v || core[u] >= 0) continue;
        --cnt[u];
        pq.push({std::max(cnt[u], L), u});
      }
    }
  }
  if (st) {
    uint64_t g = gcd64(mnum, mden);
    if (g == 0) g = 1;
    st->max_num = mnum / g;
    st->max_den = mden / g;
    st->k_prime = (uint32_t)((mnum + mden - 1) / mden);  // ceil (main.cu:1476)
    int32_t mc = 0;
    for (uint32_t v = 0; v < n; ++v) mc = std::max(mc, core[v]);
    st->max_core = mc;
    uint32_t cs = 0;
    for (uint32_t v = 0; v < n; ++v)
      if (core[v] >= (int32_t)st->k_prime) ++cs;
    st->core_size = cs;
    uint64_t cc = 0;
    for (uint64_t c = 0; c < T; ++c)
      if (status[c] >= (int32_t)st->k_prime) ++cc;
    st->core_cliques = cc;
    st->levels = levels;
  }
  if (core_out) memcpy(core_out, core.data(), sizeof(int32_t) * n);
  if (status_out) memcpy(status_out, status.data(), sizeof(int32_t) * T);
  return 0;
}

int64_t oracle_clique_maxflow(uint32_t nc, uint32_t k, uint64_t Tc, const uint32_t *cl, int64_t p,
                              int64_t q, uint8_t *side) {
  Flow F = build_network(nc, k, Tc, cl, p, q);
  uint32_t s = nc + (uint32_t)Tc, t = s + 1;
  int64_t f = F.maxflow(s, t);
  if (side) {
    std::vector<uint8_t> r = F.reach_t(t);
    for (uint32_t v = 0; v < nc; ++v) side[v] = r[v] ? 0 : 1;
  }
  return f;
}

// main.cu:1272-1522 (intended semantics; Appendix-A defects not reproduced).
int oracle_solve(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                 oracle_result *res, int64_t *per_vertex, int32_t *core_out, uint32_t *comp_out,
                 uint8_t *in_set) {
  if (k < 2) return -1;  // k = 2: edge density (main.cu:1365-1423)
  oracle_result R;
  memset(&R, 0, sizeof R);
  // listing (main.cu:93-327)
  Dag d = make_dag(n, offset, nbr);
  std::vector<uint32_t> cl;
  {
    Lister L(d, k, nullptr, &cl);
    for (uint32_t r = 0; r < n; ++r) L.root(r);
    R.T = L.T;
  }
  uint64_t T = R.T;
  if (per_vertex) {
    memset(per_vertex, 0, sizeof(int64_t) * n);
    for (uint64_t i = 0; i < T * k; ++i) per_vertex[cl[i]]++;
  }
  if (comp_out)
    for (uint32_t v = 0; v < n; ++v) comp_out[v] = UINT32_MAX;
  if (in_set) memset(in_set, 0, n);
  if (T == 0) {  // main.cu:1431-1438 prints "Max Density 0 Size 0"
    if (core_out) memset(core_out, 0, sizeof(int32_t) * n);
    if (res) *res = R;
    return 0;
  }
  // peeling (main.cu:329-420)
  std::vector<int32_t> core(n), status(T);
  oracle_peel(n, k, T, cl.data(), core.data(), status.data(), &R.peel);
  if (core_out) memcpy(core_out, core.data(), sizeof(int32_t) * n);
  const int32_t kp = (int32_t)R.peel.k_prime;
  R.k_prime = (uint32_t)kp;
  R.core_vertices = R.peel.core_size;
  // densest core + prune + components (main.cu:422-728): vertices with
  // core >= k'; an edge survives iff it lies in a clique whose status >= k'
  // (all members in the core); components of the surviving edges.
  Dsu dsu(n);
  for (uint64_t c = 0; c < T; ++c)
    if (status[c] >= kp)
      for (uint32_t j = 1; j < k; ++j) dsu.unite(cl[c * k], cl[c * k + j]);
  std::vector<uint32_t> comp(n, UINT32_MAX);
  std::vector<uint32_t> comps;  // canonical ids in increasing order
  for (uint32_t v = 0; v < n; ++v)
    if (core[v] >= kp) {
      comp[v] = dsu.find(v);
      if (comp[v] == v) comps.push_back(v);
    }
  R.ncomp = (uint32_t)comps.size();
  if (comp_out) memcpy(comp_out, comp.data(), sizeof(uint32_t) * n);
  // per-component vertex lists and clique lists in local ids
  std::vector<uint32_t> cidx(n, UINT32_MAX), local(n, UINT32_MAX);
  for (uint32_t i = 0; i < comps.size(); ++i) cidx[comps[i]] = i;
  std::vector<std::vector<uint32_t>> cverts(comps.size()), ccl(comps.size());
  for (uint32_t v = 0; v < n; ++v)
    if (comp[v] != UINT32_MAX) {
      uint32_t c = cidx[comp[v]];
      local[v] = (uint32_t)cverts[c].size();
      cverts[c].push_back(v);
    }
  for (uint64_t c = 0; c < T; ++c)
    if (status[c] >= kp) {
      uint32_t ci = cidx[comp[cl[c * k]]];
      for (uint32_t j = 0; j < k; ++j) ccl[ci].push_back(local[cl[c * k + j]]);
    }
  // bounds (helpers.cu:1115-1146, main.cu:862-876): lb = max_c Tc/nc; a
  // component can hold a set of density >= alpha only if
  // min(maxCore, Tc^((k-1)/k) / (k!)^(1/k)) >= alpha.
  uint64_t an = 0, ad = 1;
  for (size_t c = 0; c < comps.size(); ++c) {
    uint64_t tc = ccl[c].size() / k, nc = cverts[c].size();
    if (ratio_gt(tc, nc, an, ad)) {
      an = tc;
      ad = nc;
    }
  }
  double kfact = 1;
  for (uint32_t i = 2; i <= k; ++i) kfact *= i;
  auto may_reach = [&](size_t c, uint64_t num, uint64_t den) {
    double tc = (double)(ccl[c].size() / k);
    double ub = std::min((double)R.peel.max_core, std::pow(tc, (k - 1.0) / k) / std::pow(kfact, 1.0 / k));
    return ub * (1 + 1e-9) + 1e-9 >= (double)num / (double)den;
  };
  // exact search: from alpha = an/ad, each min cut whose source side A is
  // non-empty yields a strictly denser set; alpha := density(A) until no
  // component admits a denser set.
  for (size_t c = 0; c < comps.size(); ++c) {
    uint64_t tc = ccl[c].size() / k;
    uint32_t nc = (uint32_t)cverts[c].size();
    if (tc == 0) continue;
    while (may_reach(c, an, ad)) {
      Flow F = build_network(nc, k, tc, ccl[c].data(), (int64_t)an, (int64_t)ad);
      uint32_t s = nc + (uint32_t)tc, t = s + 1;
      int64_t f = F.maxflow(s, t);
      R.flow_calls++;
      if (f >= (int64_t)(ad * k * tc)) break;  // no set denser than alpha
      std::vector<uint8_t> rs = F.reach_s(s);
      uint64_t sz = 0, tin = 0;
      for (uint32_t v = 0; v < nc; ++v) sz += rs[v];
      for (uint64_t q = 0; q < tc; ++q) {
        bool all = true;
        for (uint32_t j = 0; j < k; ++j) all = all && rs[ccl[c][q * k + j]];
        tin += all;
      }
      if (sz == 0 || !ratio_gt(tin, sz, an, ad)) return -2;  // cannot happen
      uint64_t g = gcd64(tin, sz);
      an = tin / g;
      ad = sz / g;
    }
  }
  // the maximal densest set: per component, the source side of the maximal
  // min cut at alpha = rho* (vertices that cannot reach t).
  uint64_t snum = 0, sden = 0;
  for (size_t c = 0; c < comps.size(); ++c) {
    uint64_t tc = ccl[c].size() / k;
    uint32_t nc = (uint32_t)cverts[c].size();
    if (tc == 0 || !may_reach(c, an, ad)) continue;
    std::vector<uint8_t> side(nc);
    oracle_clique_maxflow(nc, k, tc, ccl[c].data(), (int64_t)an, (int64_t)ad, side.data());
    R.flow_calls++;
    for (uint32_t v = 0; v < nc; ++v)
      if (side[v]) {
        sden++;
        if (in_set) in_set[cverts[c][v]] = 1;
      }
    for (uint64_t q = 0; q < tc; ++q) {
      bool all = true;
      for (uint32_t j = 0; j < k; ++j) all = all && side[ccl[c][q * k + j]];
      snum += all;
    }
  }
  if (sden == 0 || (u128)snum * ad != (u128)an * sden) return -3;  // cannot happen
  R.rho_num = snum;
  R.rho_den = sden;
  if (res) *res = R;
  return 0;
}

}  // extern "C"
