// CPU oracle for the k-clique densest subgraph (k-CDS) pipeline.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the reference
// algorithm (wajidmanzoor/test: src/graph.cpp, main.cu, src/helpers.cu) used
// as the checker by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  The product path (test_amd/libkcds.so) never links,
// loads or calls it.
//
// Every stage uses exact integer arithmetic: counts are int64, densities are
// rationals compared by cross multiplication, max-flow capacities are integers
// scaled by the denominator of alpha = p/q.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Reference loader semantics (src/graph.cpp:7-40): header "n m", then one
// line "v nbr nbr ..." per vertex, self-loops skipped, no deduplication.
// Returns 0, or <0 on open/parse failure.  Free the arrays with oracle_free.
int oracle_load_adjlist(const char *path, uint32_t *n, uint32_t *m, uint32_t **offset,
                        uint32_t **nbr);
// 'u v' edge list: symmetrised, deduplicated, self-loops dropped, lists sorted.
// n = 1 + max id seen (or n_hint if larger).
int oracle_load_edgelist(const char *path, uint32_t n_hint, uint32_t *n, uint32_t *m,
                         uint32_t **offset, uint32_t **nbr);
void oracle_free(void *p);

// Batagelj-Zaversnik core decomposition exactly as src/graph.cpp:53-113
// (same bin layout and swap order, so the same peel sequence).  `degree`
// is read-only here (the reference overwrites it).  listing_order[v] =
// 1 + index of v in corePeelSequence (graph.cpp:48-50).
void oracle_core_decompose(uint32_t n, const uint32_t *offset, const uint32_t *nbr,
                           uint32_t *core, uint32_t *peel_seq, uint32_t *listing_order,
                           uint32_t *kmax);

// Number of k-cliques T and, if per_vertex != NULL, the number of k-cliques
// containing each vertex (int64).  threads <= 0 uses all cores.  k >= 2.
// Restates the listing semantics of main.cu:93-327 (T of countCliques,
// cliqueDegree of writeFinalCliques, helpers.cu:489-550).
uint64_t oracle_count(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                      int64_t *per_vertex, int threads);
// Same count restricted to roots r in [root_begin, root_end) of the oracle's
// own degeneracy DAG (for bounded CPU timing samples).
uint64_t oracle_count_roots(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                            uint32_t root_begin, uint32_t root_end, int threads);

// Materialises all k-cliques, each as k ascending ids, in lexicographic order.
// Returns T; writes min(T, cap) cliques into out (k*cap words).
uint64_t oracle_list(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                     uint32_t *out, uint64_t cap);

typedef struct oracle_peel_stats {
  uint64_t max_num, max_den;    // maxDensity of the densest clique-core (exact rational)
  uint32_t k_prime;             // ceil(maxDensity) (main.cu:1476)
  uint32_t core_size;           // coreSize[k_prime] = |{v : core(v) >= k_prime}|
  uint64_t core_cliques;        // live cliques of the maxDensity core
  int32_t max_core;             // largest clique-core number
  uint32_t levels;              // non-empty peeling levels
} oracle_peel_stats;

// k-clique-core peeling (main.cu:329-420, helpers.cu:587-726) over the clique
// list `cliques` (T rows of k ids).  core[v] = level at which v is removed;
// status[c] = level at which clique c is destroyed (= min member core).
// Either output may be NULL.
int oracle_peel(uint32_t n, uint32_t k, uint64_t T, const uint32_t *cliques, int32_t *core,
                int32_t *status, oracle_peel_stats *st);

// Exact max-flow on the clique flow network of one component
// (helpers.cu:1148-1264): nodes = nc vertices, Tc cliques, s, t;
// s->v cap q*deg(v), v->t cap k*p, v->c cap q, c->v cap q*(k-1).
// `cliques` holds Tc rows of k local ids in [0,nc).  Returns the flow value
// (<= q*k*Tc).  If side != NULL, side[v] = 1 for vertices on the source side
// of the MAXIMAL min cut (vertices that cannot reach t in the residual).
int64_t oracle_clique_maxflow(uint32_t nc, uint32_t k, uint64_t Tc, const uint32_t *cliques,
                              int64_t p, int64_t q, uint8_t *side);

typedef struct oracle_result {
  uint64_t T;             // k-cliques in the graph
  uint64_t rho_num;       // cliques inside the answer set S
  uint64_t rho_den;       // |S|; density = rho_num / rho_den
  uint32_t k_prime;
  uint32_t core_vertices; // |k'-core|
  uint32_t ncomp;         // components of the pruned k'-core
  uint32_t flow_calls;    // max-flow solves performed
  oracle_peel_stats peel;
} oracle_result;

// Whole pipeline (main.cu:1272-1522 intended semantics): count/list, peel,
// densest core, prune, components, exact density search.  S is the unique
// maximal densest set.  Optional outputs (any may be NULL):
//   per_vertex[n] counts, core[n], comp[n] (canonical component id = smallest
//   original vertex id of the component, UINT32_MAX outside the k'-core),
//   in_set[n] (1 for members of S).
int oracle_solve(uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                 oracle_result *res, int64_t *per_vertex, int32_t *core, uint32_t *comp,
                 uint8_t *in_set);

#ifdef __cplusplus
}
#endif
