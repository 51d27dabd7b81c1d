/* kcds.h -- C ABI of the MI355X k-clique densest subgraph (k-CDS) solver.
 *
 * One kcds_ctx owns one HIP stream on one device and every device buffer of a
 * solve (graph CSR, DAG, clique list, incidence index, peeling state, flow
 * networks).  All entry points return KCDS_OK (0) or a negative kcds_status;
 * none calls exit().  kcds_last_error() describes the last failure.  Contexts
 * are independent; use one per host thread.
 *
 * Mapping to the reference (wajidmanzoor/test) path each entry point replaces:
 *   kcds_upload_csr        memoryAllocationGraph   src/gpuMemoryAllocation.cu:5-27
 *   kcds_orient            generateDAG             main.cu:22-91 (+ src/graph.cpp:41-51
 *                                                  listing order when computed on the GPU)
 *   kcds_count             listAllCliques + countCliques (T, cliqueDegree)
 *                                                  main.cu:93-244, src/helpers.cu:1074-1113
 *   kcds_list              listAllCliques + writeFinalCliques (trie, status)
 *                                                  main.cu:246-327, src/helpers.cu:489-550
 *   kcds_peel              cliqueCoreDecompose     main.cu:329-420
 *   kcds_components        generateDensestCore + prune + componentDecompose
 *                                                  main.cu:422-728
 *   kcds_maxflow           one max-flow of dynamicExactAlgo  main.cu:1037-1237
 *   kcds_solve             main's stage sequence   main.cu:1348-1519
 *   kcds_create/destroy    the buffer lifecycle of src/gpuMemoryAllocation.cu (alloc/free pairs)
 */
#ifndef KCDS_H
#define KCDS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kcds_ctx kcds_ctx;

typedef enum kcds_status {
  KCDS_OK = 0,
  KCDS_E_INVALID = -1, /* bad argument: NULL pointer, k outside 2..8, not a permutation, ... */
  KCDS_E_HIP = -2,     /* HIP runtime error (message in kcds_last_error) */
  KCDS_E_NOMEM = -3,   /* HBM budget of the context exceeded or allocation failed */
  KCDS_E_STATE = -4,   /* call out of order, e.g. kcds_count before kcds_orient */
  KCDS_E_LIMIT = -5,   /* size beyond a supported limit (message names it) */
  KCDS_E_PENDING = -6  /* kcds_poll: the request has not finished yet */
} kcds_status;

typedef struct kcds_peel_stats {
  uint64_t max_num, max_den; /* maxDensity of the densest clique-core, exact (main.cu:398-413) */
  uint64_t k_prime;          /* ceil(maxDensity) (main.cu:1476) */
  uint64_t core_cliques;     /* cliques with every member in the k'-core */
  int64_t max_core;          /* largest clique-core number */
  uint32_t core_size;        /* |{v : core(v) >= k_prime}| (main.cu:1482) */
  uint32_t levels;           /* non-empty peeling levels processed */
} kcds_peel_stats;

typedef struct kcds_result {
  uint64_t T;            /* k-cliques in the graph */
  uint64_t rho_num;      /* cliques inside the answer set S */
  uint64_t rho_den;      /* |S|: density = rho_num / rho_den exactly */
  uint64_t k_prime;
  uint32_t core_vertices;
  uint32_t ncomp;        /* components of the pruned k'-core */
  uint32_t flow_calls;   /* max-flow solves */
  kcds_peel_stats peel;
  float ms_orient, ms_list, ms_peel, ms_components, ms_flow, ms_total; /* stage times (hipEvents) */
  uint32_t dense_passes;   /* greedy passes of the dense exact search (0: clique flow networks were used) */
  uint32_t dense_critical; /* critical vertices its certificate ended with (members of S given first claim
                              on the shares of their cliques) */
} kcds_result;

/* Creates a context on `device` with an HBM budget (0 = 90% of free memory). */
int kcds_create(kcds_ctx **ctx, int device, uint64_t hbm_budget_bytes);
void kcds_destroy(kcds_ctx *ctx);
const char *kcds_last_error(const kcds_ctx *ctx);

/* Copies an undirected CSR (host arrays; offset[n+1], nbr[offset[n]], no
 * self-loops, each edge listed in both directions) to the device.  Resets all
 * later stages. */
int kcds_upload_csr(kcds_ctx *ctx, uint32_t n, const uint32_t *offset, const uint32_t *nbr);

/* Builds the device CSR from an edge list on the GPU: uv[2m] holds m pairs
 * (u, v) with ids < n, in either direction, repeated or self-loops allowed.
 * Each pair is taken in both directions, self-loops are dropped and every
 * neighbour list is sorted ascending without duplicates -- the form
 * kcds_upload_csr takes, and what the CPU loaders make of a `u v` file.
 * *m_out (may be NULL) receives the number of distinct undirected edges.
 * KCDS_E_INVALID for an id >= n; KCDS_E_LIMIT above 2^31 - 2^15 pairs.
 * Resets all later stages.  Replaces the host-side parse-and-build of
 * Graph::Graph (src/graph.cpp:7-40) for callers that hold the pairs in
 * memory. */
int kcds_upload_edges(kcds_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *uv, uint64_t *m_out);

/* Copies the device CSR to the host: offset[n+1] (if not NULL) and the first
 * min(cap, nnz) neighbour ids (if nbr is not NULL); *nnz receives the total. */
int kcds_download_csr(kcds_ctx *ctx, uint32_t *offset, uint32_t *nbr, uint64_t cap, uint64_t *nnz);

/* Builds the DAG used for listing.  listing_order == NULL: an order computed
 * on the GPU -- by default the degree order (vertices sorted by degree, ties
 * by id), or the degeneracy order (KCDS_TUNE_DEGENERACY_ORDER, with slack
 * unless KCDS_TUNE_EXACT_ORDER is set; see kcds_set_tuning); the degree order
 * falls back to the degeneracy order when it leaves an out-list longer than
 * 512.  Otherwise listing_order[v] in 1..n is the reference's
 * listing order (src/graph.cpp:48-50: 1 + position in corePeelSequence, the
 * reverse removal order); arcs point from the smaller listing order to the
 * larger, i.e. from the vertex peeled later (higher core) to the one peeled
 * earlier, as the reference's generateDAG does.  Every
 * acyclic orientation yields the same cliques, counts and solve results: an
 * order that leaves an out-list longer than 512 (the LDS frame rows; the
 * reference's order does on hub-heavy graphs, where hubs point to their
 * low-core neighbours) is replaced by the GPU order, which
 * kcds_get_list_info reports.  When no GPU order fits 512 (degeneracy above
 * 512) the exact degeneracy order is used and its wider roots list with their
 * bitmask rows in HBM; KCDS_E_LIMIT when an out-list exceeds 4,096
 * (degeneracy above 4,096). */
int kcds_orient(kcds_ctx *ctx, const uint32_t *listing_order);

/* Counts k-cliques: *T and, if per_vertex != NULL, the number of k-cliques
 * containing each vertex (int64, original ids).  3 <= k <= 8. */
int kcds_count(kcds_ctx *ctx, uint32_t k, uint64_t *T, int64_t *per_vertex);

/* Materialises every k-clique on the device (clique list + vertex->clique
 * incidence index); *T receives the count. */
int kcds_list(kcds_ctx *ctx, uint32_t k, uint64_t *T);
/* Copies up to cap cliques (k ascending original ids each, rows sorted
 * lexicographically) to the host; *T receives the total. */
int kcds_download_cliques(kcds_ctx *ctx, uint32_t *out, uint64_t cap, uint64_t *T);

/* k-clique-core peeling.  After kcds_list it runs over the materialised
 * cliques; after kcds_count alone (or with KCDS_TUNE_ENUM_PEEL) it enumerates
 * each removed vertex's live cliques from its neighbourhood instead, which
 * needs O(n + m) memory whatever T is.  Both give the same core numbers and
 * statistics.  Core numbers are 64-bit: a vertex's count can exceed 2^31 (the
 * reference's cliqueCore is a 32-bit int, main.cu:329-420).  clique_core[n]
 * (original ids) may be NULL. */
int kcds_peel(kcds_ctx *ctx, int64_t *clique_core, kcds_peel_stats *stats);

/* Densest core, edge pruning and connected components.  comp[n] receives the
 * canonical component id of each vertex (smallest original vertex id in its
 * component; UINT32_MAX outside the k'-core).  comp may be NULL. */
int kcds_components(kcds_ctx *ctx, uint32_t *comp, uint32_t *ncomp);

/* Max-flow value of the clique flow network of component `comp_id` (its
 * canonical id) at alpha = p/q, capacities scaled by q:
 * s->v q*deg(v), v->t k*p, v->c q, c->v q*(k-1).  side[n] (original ids, may
 * be NULL) receives 1 for vertices on the source side of the maximal min cut. */
int kcds_maxflow(kcds_ctx *ctx, uint32_t comp_id, uint64_t p, uint64_t q, int64_t *flow_value,
                 uint8_t *side);

/* Whole pipeline on the uploaded graph: orient (GPU degeneracy order), list,
 * peel, components, exact density search.  in_set[n] (may be NULL) receives
 * 1 for members of the maximal densest set S. */
int kcds_solve(kcds_ctx *ctx, uint32_t k, kcds_result *res, uint8_t *in_set);

/* Per-graph result of a batched solve. */
typedef struct kcds_graph_result {
  uint64_t T;          /* k-cliques of the graph */
  uint64_t rho_num;    /* cliques inside the graph's maximal densest set */
  uint64_t rho_den;    /* its size (0 when the graph has no k-clique) */
  uint64_t k_prime;
  uint32_t core_vertices;
  uint32_t ncomp;
  uint32_t flow_calls;
  uint32_t status;     /* KCDS_GRAPH_SOLVED, or KCDS_GRAPH_OVER_BUDGET (nothing solved; T is then a lower
                          bound above the cap: counting stops once the graph is over budget) */
  uint32_t pad;
} kcds_graph_result;

#define KCDS_GRAPH_SOLVED 0u
#define KCDS_GRAPH_OVER_BUDGET 1u

/* Per-graph HBM admission: a graph whose materialised cliques, incidence
 * and status (T * (8k + 4) bytes) exceed bytes is not solved -- kcds_solve
 * fails with KCDS_E_NOMEM, a batch marks the graph KCDS_GRAPH_OVER_BUDGET
 * and solves the others.  It is decided on the count pass, which stops
 * counting a graph as soon as its count passes bytes / (8k + 4).  0 = no
 * per-graph limit (the context budget still applies).  No counterpart in
 * the reference. */
int kcds_set_graph_budget(kcds_ctx *ctx, uint64_t bytes);

/* Batched solve: the uploaded CSR is the disjoint union of ngraphs graphs,
 * graph g owning vertex ids [graph_voff[g], graph_voff[g+1]) (graph_voff[0] =
 * 0, graph_voff[ngraphs] = n, no edge between graphs).  Every stage runs once
 * over the union; peeling keeps per-graph density bookkeeping and every
 * component's exact search runs in one persistent launch.  results[ngraphs]
 * get each graph's answer (identical to solving it alone); in_set[n] (may be
 * NULL) marks the members of each graph's maximal densest set.  No
 * counterpart in the reference, which solves one graph per process. */
int kcds_solve_batch(kcds_ctx *ctx, uint32_t k, uint32_t ngraphs, const uint32_t *graph_voff,
                     kcds_graph_result *results, uint8_t *in_set);

/* ---- Serving: continuous batching behind submit / poll ----------------------
 * No counterpart in the reference (one graph per process).  A server owns
 * `slots` batch contexts and `large_slots` large-lane contexts, each with
 * its own host thread and HIP stream.  Requests wait in per-k queues; a free
 * batch slot takes the k whose oldest request has waited longest and solves
 * that k's waiting requests, oldest first, up to max_batch graphs and
 * batch_edges adjacency entries, in one kcds_solve_batch.  Requests submitted
 * meanwhile join the next batch, and batches of different slots overlap on
 * the device.  A graph whose clique list exceeds graph_budget bytes (T *
 * (8k + 4)) moves to the large lane, which solves one graph at a time
 * without a clique list; one with more than max_cliques cliques is answered
 * KCDS_GRAPH_OVER_BUDGET.  kcds_submit copies the CSR.  Results equal
 * kcds_solve on the graph alone. */
typedef struct kcds_server kcds_server;

typedef struct kcds_server_config {
  int32_t device;
  uint32_t slots;        /* batch slots (>= 1) */
  uint32_t large_slots;  /* large-lane slots; 0: graphs over graph_budget are answered over budget */
  uint32_t max_batch;    /* graphs per batch (0 = 512) */
  uint64_t batch_edges;  /* adjacency entries per batch (0 = 4,000,000) */
  uint64_t graph_budget; /* clique-list bytes per graph in a batch (0 = no per-graph cap) */
  uint64_t max_cliques;  /* graphs with more cliques are answered over budget (0 = no cap) */
  uint64_t hbm_budget;   /* bytes shared by all contexts (0 = each context 90% of free memory) */
  uint32_t want_sets;    /* keep each answer's vertex set for kcds_poll */
  uint32_t pad;
} kcds_server_config;

typedef struct kcds_served {
  kcds_graph_result result;
  double t_submit, t_done; /* kcds_server_clock() seconds */
  uint32_t batch;          /* graphs in the batch that answered it (1 in the large lane) */
  uint32_t lane;           /* 0 batch slot, 1 large lane */
  int32_t code;            /* KCDS_OK, or the status that left it unsolved (KCDS_E_NOMEM: over budget) */
  uint32_t pad;
} kcds_served;

int kcds_server_create(kcds_server **srv, const kcds_server_config *cfg);
/* Queues a request (undirected CSR as for kcds_upload_csr) and returns its ticket. */
int kcds_submit(kcds_server *srv, uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                uint64_t *ticket);
/* Waits up to wait_ms (0: no wait, < 0: until done) for the request; KCDS_OK
 * fills out (and in_set[n], if both it and want_sets are set) and forgets the
 * ticket, KCDS_E_PENDING if it has not finished. */
int kcds_poll(kcds_server *srv, uint64_t ticket, int32_t wait_ms, kcds_served *out, uint8_t *in_set);
/* Seconds since the server was created (the clock of t_submit / t_done). */
double kcds_server_clock(const kcds_server *srv);
/* Answers every request that no slot has started yet with code
 * KCDS_E_STATE (status KCDS_GRAPH_OVER_BUDGET, nothing solved); requests in
 * flight finish normally.  Returns how many were cancelled. */
int kcds_server_cancel_pending(kcds_server *srv);
/* Serves every queued request, then stops the threads and frees the contexts. */
void kcds_server_destroy(kcds_server *srv);

/* Blocks until all work queued on the context's stream has finished. */
int kcds_sync(kcds_ctx *ctx);

/* Allocation-free k-clique count.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, code that
 * must not call hipMalloc, harnesses that forbid it).  Counts what kcds_count
 * counts, over a DAG the caller already holds in device memory in listing
 * form: doff[n+1] and dnbr[doff[n]] with every arc u -> w having u < w < n,
 * each out-list strictly ascending and at most 512 long (the form of the DAG
 * kcds_orient builds, vertices numbered by listing rank).  per_vertex[n]
 * (device memory) receives the number of k-cliques containing each vertex
 * and *T (host memory) their total; 2 <= k <= 8, n <= KCDS_NOALLOC_MAX_N.
 * stream: a hipStream_t of `device`, or NULL for its null stream; the call
 * returns when the count is done.  It allocates nothing: its temporaries live
 * in one static device array of KCDS_NOALLOC_SCRATCH_BYTES, which every device
 * the library runs on holds until the process ends and which every call on
 * that device shares, so calls that overlap on one device race.  Returns
 * KCDS_E_LIMIT above n or the out-degree limit, KCDS_E_INVALID for bad
 * arguments or a DAG not in listing form (messages on stderr with KCDS_LOG). */
#define KCDS_NOALLOC_MAX_N (1u << 20)
#define KCDS_NOALLOC_SCRATCH_BYTES (25u << 20)
int kcds_count_dag_noalloc(int device, uint32_t n, const uint32_t *doff, const uint32_t *dnbr, uint32_t k,
                           uint64_t *per_vertex, uint64_t *T, void *stream);

/* Listing-kernel launch shape: waves per CU for the persistent listing
 * kernels (0 = library default), and flags:
 *   KCDS_TUNE_GRID_PEEL  peel a batch with the grid-wide step kernels even
 *                        when every graph fits the workgroup-per-graph peel
 *                        (both give the same results; used to test one
 *                        against the other).  A single graph always takes
 *                        the grid peel. */
#define KCDS_TUNE_GRID_PEEL 1u
/*   KCDS_TUNE_GRID_FLOW  solve every component with the grid-wide max-flow
 *                        (normally only components of >= 49,152 cliques,
 *                        or >= 4,096 when a solve has fewer components than
 *                        a quarter of the GPU's compute units). */
#define KCDS_TUNE_GRID_FLOW 2u
/*   KCDS_TUNE_EXACT_ORDER  orient by the exact degeneracy order.  By default
 *                        the GPU order peels with each threshold the minimum
 *                        live degree rounded up by a half (x -> x + x/2):
 *                        out-degrees stay within 3/2 of the degeneracy and
 *                        the peel takes far fewer steps; when that leaves an out-list wider than
 *                        512 the exact order is used instead.  Every acyclic
 *                        orientation lists each clique exactly once, so all
 *                        results are the same either way. */
#define KCDS_TUNE_EXACT_ORDER 4u
/*   KCDS_TUNE_NO_GRAPHS  launch the grid-wide max-flow's relabel and sweep
 *                        sequences one kernel at a time instead of replaying
 *                        them as captured HIP graphs (same results).  The
 *                        environment variable KCDS_NO_GRAPHS does the same for
 *                        every context, and so does running under rocprofv3
 *                        (graph replay under its kernel tracing has crashed
 *                        the host process inside hipGraphLaunch). */
#define KCDS_TUNE_NO_GRAPHS 8u
/*   KCDS_TUNE_BLOCK_FLOW  solve every component with the workgroup max-flow
 *                        (same results; used to test one solver against the
 *                        other).  KCDS_TUNE_GRID_FLOW takes precedence. */
#define KCDS_TUNE_BLOCK_FLOW 16u
/*   KCDS_TUNE_ENUM_PEEL  peel by enumerating each removed vertex's live
 *                        cliques from its neighbourhood (no clique list), even
 *                        when the clique list would fit; the solve then lists
 *                        only the k'-core's cliques.  Without the flag a solve
 *                        takes this path by itself when the clique list would
 *                        not fit in HBM (T * (8k + 4) bytes, or T >= 2^32),
 *                        and for a single graph from 2^25 cliques up, where
 *                        enumerating is the faster peel. */
#define KCDS_TUNE_ENUM_PEEL 32u
/*   KCDS_TUNE_DENSE      take the dense exact search for a single graph whose
 *                        k'-core has at most 1,024 vertices, whatever its
 *                        clique count (by default only from 2^24 core cliques
 *                        up, where a clique flow network stops fitting): the
 *                        core is held as a bitmask matrix, greedy clique
 *                        splits (one listing pass each) are combined by
 *                        Wolfe's minimum-norm-point method, and the answer
 *                        is certified by an exact integer max-flow over
 *                        vertex pairs. */
#define KCDS_TUNE_DENSE 64u
/*   KCDS_TUNE_DENSE_PEEL  let a single graph's enumerating peel go over to a
 *                        dense mode once at most 1,024 vertices are live: the
 *                        live subgraph as a bitmask matrix, each sub-round's
 *                        destroyed cliques counted in one pass rooted at the
 *                        frontier, levels kept on the host.  Same cores either
 *                        way; off by default, where the step kernels are as
 *                        fast or faster (measured on MI355X: C2 peel 0.52 s
 *                        either way, C3 0.99 s against 1.09 s). */
#define KCDS_TUNE_DENSE_PEEL 128u
/*   KCDS_TUNE_FLOW_ALL   run the max-flow search on every component.  By
 *                        default a component whose clique count is C(n_c, k)
 *                        (a complete graph on its n_c vertices) is answered
 *                        without one: its k-clique density C(t, k)/t grows
 *                        with the size t of a subset, so the whole component
 *                        is its unique densest subset.  Same results either
 *                        way. */
#define KCDS_TUNE_FLOW_ALL 256u
/*   KCDS_TUNE_DEGENERACY_ORDER  orient by the degeneracy order (with slack
 *                        unless KCDS_TUNE_EXACT_ORDER is set) instead of the
 *                        default degree order.  By default kcds_orient(ctx,
 *                        NULL) sorts the vertices by (degree, id) and points
 *                        every arc to the later one; if that leaves an
 *                        out-list wider than 512 it falls back to the slack
 *                        degeneracy order, then to the exact one. */
#define KCDS_TUNE_DEGENERACY_ORDER 512u
/*   KCDS_TUNE_BUILD_FRAMES  let the enumerating peel build every removed
 *                        vertex's frame from its neighbours' out-lists.  By
 *                        default the frames of vertices of degree <= 512
 *                        that lie in a clique are built once, in parallel,
 *                        before the first level (about 8 bytes per
 *                        neighbour, up to 8 words per neighbour for degree
 *                        257..512; skipped when that does not fit the HBM
 *                        budget), and each step restricts and renumbers the
 *                        stored rows instead.  Same cores either way. */
#define KCDS_TUNE_BUILD_FRAMES 1024u
/*   KCDS_TUNE_BLOCK_PEEL  peel a single graph that fits the workgroup-per-graph
 *                        peel (n <= 32,768, T <= 2^18) with it, as a batch's
 *                        small graphs are; by default a single graph takes the
 *                        grid peel.  Same cores either way. */
#define KCDS_TUNE_BLOCK_PEEL 2048u
int kcds_set_tuning(kcds_ctx *ctx, uint32_t list_waves_per_cu, uint32_t flags);

/* Major cycles the dense exact search may run before a solve whose k'-core
 * clique list fits in HBM lists it and takes the clique max-flow instead
 * (same result; 0 restores the default, 512).  A core too large to list, or
 * KCDS_TUNE_DENSE, keeps the search going up to 12,000 cycles.  No
 * counterpart in the reference (it always builds the clique network). */
int kcds_set_dense_budget(kcds_ctx *ctx, uint32_t major_cycles);

/* Listing launch shape: waves per root (workgroup size / 64, 1..8) of each
 * width class W = 1, 2, 4, 8 words (roots of out-degree <= 64, 128, 256,
 * 512); NULL restores the defaults {1, 2, 4, 8}.  Replaces the reference's
 * hand-set sizing arguments pSize, cpSize, glBufferSize, partitionSize
 * (main.cu:1273-1293), which size its per-warp partitions. */
int kcds_set_list_shape(kcds_ctx *ctx, const uint32_t *waves_per_root /* 4, or NULL */);

/* Autotuner: on the oriented graph, times every waves-per-root choice of
 * every width class alone at clique size k, then refines the shape by
 * coordinate descent on the time of the whole pass with the classes running
 * at once; keeps the result and returns it in waves_per_root[4] and each
 * class's time alone at its chosen shape (ms) in class_ms[4] (either may be
 * NULL).  No counterpart in the reference (its launch shape is fixed). */
int kcds_autotune_list(kcds_ctx *ctx, uint32_t k, uint32_t *waves_per_root, float *class_ms);

/* Per-stage facts of the last kcds_count/kcds_list call, for the roofline:
 * n, m (DAG arcs), W2 = sum over the roots the pass stages (out-degree >=
 * k-1) of the out-degrees of their out-neighbours, max out-degree, the
 * duration of the listing pass (ms, hipEvents around its launches), the
 * number of launches, the roots of each width class and their waves per
 * root. */
typedef struct kcds_list_info {
  uint32_t n;
  uint64_t arcs;
  uint64_t w2;
  uint32_t max_outdeg;
  float kernel_ms;
  uint32_t kernel_launches;
  uint32_t class_roots[4];
  uint32_t waves_per_root[4];
  uint32_t order_replaced;  /* 1: kcds_orient's listing_order left an out-list
                               wider than 512, so the GPU order was used
                               instead (same cliques and results) */
  uint32_t order_kind;      /* the DAG's orientation: 0 degree order, 1 slack
                               degeneracy order, 2 exact degeneracy order,
                               3 the caller's listing_order */
} kcds_list_info;
int kcds_get_list_info(kcds_ctx *ctx, kcds_list_info *info);

#ifdef __cplusplus
}
#endif
#endif /* KCDS_H */
