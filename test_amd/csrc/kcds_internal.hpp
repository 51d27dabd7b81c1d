// Host-side internals of libkcds: the context, device buffers and the launch
// wrappers each kernel file exports.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/kcds.h"

struct Scatter;

#ifndef KCDS_ULL_DEFINED
#define KCDS_ULL_DEFINED
typedef unsigned long long ull;
#endif

namespace kcds {

// Thrown inside the library, turned into a kcds_status at the C boundary.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const char *fmt, ...);

#define HIP_CHECK(expr)                                                          \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      ::kcds::fail(KCDS_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                          \
  } while (0)

#define KCDS_LAUNCH_CHECK(name) HIP_CHECK(hipGetLastError())

struct Arena;

// A growable device buffer charged against the context's HBM budget.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <typename T> T *as() const { return (T *)p; }
};

struct Ctx;
// A launch sequence on the context stream captured once as a HIP graph and
// replayed; freed with the object.  capture() ends the capture (and
// rethrows) if a launch in it fails.
struct StreamGraph {
  Ctx &c;
  hipGraphExec_t exec = nullptr;
  explicit StreamGraph(Ctx &cc) : c(cc) {}
  ~StreamGraph();
  template <typename F> void capture(F &&launches);
  void launch();
};
void buf_reserve(Ctx &c, DevBuf &b, size_t bytes);  // grows, never shrinks; contents lost on growth
void buf_free(Ctx &c, DevBuf &b);

// Per-stage timing with hipEvents on the context stream.
struct StageTimer {
  hipEvent_t a = nullptr, b = nullptr;
};

// Per-graph peeling bookkeeping (a batch peels many disjoint graphs at once).
struct GraphPeel {
  ull live;            // live cliques during the peel; T of the graph before it
  ull max_num, max_den;  // maxDensity of the graph's clique-cores
  ull T;
  ull core_cliques;    // cliques with every member in the k'-core
  ull k_prime;         // ceil(maxDensity); NO_KPRIME for a graph without cliques
  ull max_core;        // largest clique-core number
  uint32_t rem;        // remaining vertices during the peel
  uint32_t touch;      // touched in the current level
  uint32_t n;
  uint32_t core_size;
  uint32_t levels;     // non-empty peeling levels
  uint32_t pad;
};
constexpr ull NO_KPRIME = 0x7FFFFFFFFFFFFFFFull;  // k' of a graph without cliques (no vertex reaches it)

// Core vertices and core cliques grouped by component (dense index d in
// order of the canonical label = smallest original id of the component).
struct CompData {
  DevBuf par, flag, dense, cid, local, vcount, voff, verts, ccount, coff, cfill, ccl;
  uint32_t ncomp = 0;
  std::vector<uint32_t> h_voff;  // C+1 vertex offsets
  std::vector<ull> h_coff;       // C+1 clique offsets
  std::vector<uint32_t> h_label; // canonical label of each component
};

// Buffers of one clique flow network (reused across max-flow calls).
struct FlowBufs {
  DevBuf scal;
  DevBuf cside;  // uint8 per core vertex (component-grouped): min-cut side of each solved component
  DevBuf b_off, b_excess, b_height, b_fr0, b_fr1, b_head, b_rev, b_res, b_tasks;  // block solver slices
  DevBuf g_off, g_head, g_rev, g_res, g_excess, g_height, g_fr, g_ctl;           // grid solver (one component)
  DevBuf g_nch, g_choff, g_chv, g_vmin, g_csum, g_esnap, g_cmin, g_hsnap;        // grid solver: vertex arc chunks
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t lstream[3] = {nullptr, nullptr, nullptr};  // listing classes run beside the context stream
  hipEvent_t lfork = nullptr, ljoin[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ltime[2] = {nullptr, nullptr};  // around the listing pass
  hipEvent_t ffork = nullptr;                // workgroup max-flow launch beside the grid solver
  uint64_t budget = 0, used = 0;
  int num_cus = 0;
  std::string err;
  uint32_t list_waves_per_cu = 0;
  uint32_t tune_flags = 0;
  uint32_t list_waves_per_root[4] = {0, 0, 0, 0};  // 0 = default of the class
  float tune_ms[4] = {0, 0, 0, 0};
  uint64_t order_steps = 0;  // work/control steps of the last GPU ordering
  bool order_exact = true;   // the last GPU ordering was the exact degeneracy order
  uint32_t order_kind = 0;   // orientation of the DAG: 0 degree, 1 slack degeneracy, 2 exact degeneracy, 3 caller's
  bool order_replaced = false;  // kcds_orient's listing order was too wide; the degeneracy order is used
  uint64_t peel_steps = 0;   // steps of the last grid peel
  uint64_t graph_budget = 0; // per-graph cap on T * (8k + 4) bytes (0 = none)
  uint32_t dense_budget = 0; // dense search cycles before a listable core takes the flow (0: DENSE_LISTABLE_ITERS)

  // pinned host staging for small readbacks
  void *pinned = nullptr;
  size_t pinned_bytes = 0;

  // input graph (original ids)
  uint32_t n = 0;
  uint64_t nnz = 0;
  uint32_t maxdeg = 0;       // largest degree of the uploaded graph
  DevBuf off, nbr;
  bool have_graph = false;

  // orientation: rank[v] = position of v in the listing order (new id),
  // inv[r] = original id; DAG over new ids, lists sorted ascending
  DevBuf rank, inv, doff, dnbr;
  uint32_t dmax = 0;
  uint64_t darcs = 0;
  uint64_t w2 = 0;
  bool oriented = false;

  // root buckets by bitmask width class
  DevBuf roots, qctr;
  uint32_t class_count[8] = {0};
  uint32_t class_start[8] = {0};

  // listing outputs
  uint32_t k = 0;
  DevBuf cnt;     // ull[n] per-vertex counts in new-id space (copy 0 of the listing's copies)
  DevBuf lsum;    // ull sharded listing totals
  DevBuf gsum;    // ull per graph: admission lower bound of the count pass
  ull count_cap = 0;  // cliques per graph above which the last count pass stopped a graph (0 = none)
  DevBuf scal;    // ull scratch scalars (T, counters)
  DevBuf rootT;   // ull[n+1] per-root clique counts -> offsets
  DevBuf cliques; // uint32[T*k] new ids
  DevBuf incoff;  // ull[n+1]
  DevBuf inc;     // uint32[T*k] clique ids per vertex
  uint64_t T = 0;
  bool counted = false, listed = false;

  // peeling
  DevBuf core;    // int32[n] new-id space
  DevBuf status;  // int32[T]
  // enumerating peel: full adjacency over new ids (in-list, then out-list)
  // and the step at which each vertex leaves (0xFFFFFFFF while live)
  DevBuf foff, fnbr, erem;
  DevBuf fofs, fstore;      // the enumerating peel's stored frames (word offsets per vertex, rows)
  DevBuf kdoff, kdnbr;      // the k'-core's DAG (listing of the core cliques)
  bool enum_peel = false;   // the last peel enumerated cliques (no clique list)
  bool core_counted = false; // ... and the k'-cores' cliques have been counted since
  bool core_listed = false;  // ... and materialised (they fit in HBM)
  kcds_peel_stats peel{};
  bool peeled = false;

  // components
  DevBuf comp;    // uint32[n] new-id space: canonical component label (min original id) or ~0
  uint32_t ncomp = 0;
  bool have_comp = false;

  // batch partition: graph g owns original ids [h_gvoff[g], h_gvoff[g+1])
  uint32_t ngraphs = 1;
  std::vector<uint32_t> h_gvoff;
  DevBuf gvoff, gid, gpeel, tlist;  // gid[r]: graph of new id r
  std::vector<GraphPeel> h_gpeel;

  CompData cd;
  FlowBufs fb;
  DevBuf inset;   // uint8[n] new-id space: members of the answer set

  // generic temporaries
  DevBuf tmp0, tmp1, tmp2, tmp3;
  DevBuf wide;              // HBM bitmask rows of the wide roots (one slice per workgroup)
  DevBuf scan_ws;  // block sums of the multi-level scans

  // last listing kernel timing
  float list_kernel_ms = 0;
  uint32_t list_kernel_launches = 0;

  void *host_scratch(size_t bytes);
  void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

// ---- kernel launchers (k_util.hip) ----
// out[i] = exclusive prefix sum of in[0..i), out[len] = total; returns total.
void scan_u32_to_u32(Ctx &c, const uint32_t *in, uint32_t *out, uint64_t len);
void scan_u64_to_u64(Ctx &c, const ull *in, ull *out, uint64_t len);
void fill_u32(Ctx &c, uint32_t *p, uint32_t v, uint64_t len);
// Zeroes up to four ull ranges with one launch (null or empty ranges skipped).
void zero_u64(Ctx &c, ull *p0, uint64_t n0, ull *p1 = nullptr, uint64_t n1 = 0, ull *p2 = nullptr, uint64_t n2 = 0,
              ull *p3 = nullptr, uint64_t n3 = 0);
void fill_u64(Ctx &c, ull *p, ull v, uint64_t len);
void max_u32(Ctx &c, const uint32_t *in, uint64_t len, uint32_t *out_dev);
ull sum_u64_host(Ctx &c, const ull *in, uint64_t len);
ull max_u64_host(Ctx &c, const ull *in, uint64_t len);

// ---- orientation (k_orient.hip) ----
// rank/inv set -> doff, dnbr, dmax, darcs; false (nothing failed) when approx
// and an out-list is wider than LIST_MAX_OUTDEG
bool orient_from_rank(Ctx &c, bool approx = false);
// fills rank/inv from a GPU k-core peel; thresholds x -> x + (x >> slack_shift)
// (NO_SLACK: the exact degeneracy order)
void degeneracy_rank_gpu(Ctx &c, uint32_t slack_shift);
// pairs uv[2m] (host) -> c.off, c.nbr, c.nnz, c.maxdeg (k_edges.hip)
void build_csr_from_edges(Ctx &c, uint32_t n, uint64_t m, const uint32_t *uv_host);
void orient_gpu(Ctx &c);  // degeneracy_rank_gpu + orient_from_rank, slack order first

// ---- listing (k_list.hip) ----
void list_count(Ctx &c, uint32_t k, ull cap = 0);  // cnt, T (cap: per-graph admission cap, 0 = none)
void list_autotune(Ctx &c, uint32_t k);  // sets list_waves_per_root, tune_ms
// Per-graph clique counts of the last count pass (needs graph ids for G > 1).
void graph_counts(Ctx &c, std::vector<ull> &gT);
// Drops the cliques of the flagged graphs from the count pass (roots and
// per-vertex counts zeroed); returns the remaining T.
ull drop_graphs(Ctx &c, const std::vector<uint8_t> &drop, const std::vector<ull> &gT);
uint32_t list_waves_per_root(const Ctx &c, int cls, uint32_t k);  // effective waves per root of a class
void list_materialize(Ctx &c, uint32_t k);  // cliques, incoff, inc (requires list_count)

// ---- peeling (k_peel.hip, k_epeel.hip) ----
void peel_run(Ctx &c);
void peel_graphs_begin(Ctx &c, GraphPeel *gp);
void peel_graphs_end(Ctx &c, GraphPeel *gp, uint32_t levels);
// Clique-core peel from the per-vertex counts of the count pass alone: each
// removed vertex's live cliques are enumerated from its neighbourhood.
void epeel_run(Ctx &c);
// After epeel_run: counts the k'-cores' cliques (status >= k') and, when they
// fit, materialises them, so the components and flow stages run as after a
// materialised peel.
void core_materialize(Ctx &c, bool materialize = true);
uint64_t build_core_dag(Ctx &c);  // c.kdoff/c.kdnbr: the DAG restricted to the k'-cores
// Exact search on graph 0's k'-core when it has at most 1,024 vertices (false
// otherwise, or when the certificate is not reached within max_iters greedy
// passes): the minimum-norm-point / integer-certificate search of k_dense.hip.
// num/den = cliques in / size of the maximal densest set, marked in c.inset;
// critical = the critical set its certificate ended with.
bool dense_core_solve(Ctx &c, uint32_t k, ull &num, ull &den, uint32_t &ncomp, uint32_t &passes, uint32_t &critical,
                      uint32_t max_iters);
// Dense mode of the enumerating peel (k_dense.hip) for a single graph with at
// most 1,024 live vertices; -1 when not applicable.
long long dense_peel_finish(Ctx &c, const ull *cnt, Scatter sc, GraphPeel *gp, long long last_level);
constexpr ull DENSE_MIN_T = 1ull << 24;
// A single graph with at least this many cliques takes the enumerating peel.
// Measured on MI355X with stored frames, over the R-MAT graphs of a served
// stream (k = 3..6), per graph: from 3.5e7 cliques up enumerating took 0.3 to
// 0.8 of the listed-and-peeled time, below 2.5e7 about 1.0 to 1.2.
constexpr ull ENUM_MIN_T = 1ull << 25;
constexpr uint32_t DENSE_MAX_ITERS = 12000;  // major cycles of the minimum-norm-point search
// ... when the k'-core's clique list would fit in HBM, so the clique flow
// network can take over: a certificate usually holds within 100 cycles, and
// one that needs thousands (an outsider whose best load sits within 1e-6 of
// rho*) costs more than the flow search on the listed cliques
constexpr uint32_t DENSE_LISTABLE_ITERS = 512;

// ---- components (k_comp.hip) ----
void components_run(Ctx &c);

// ---- flow (k_flow.hip) ----
// Marks component d's slice of c.fb.cside (its last min-cut side) in c.inset.
void mark_component_side(Ctx &c, uint32_t d);
// The same for every component d with sel[d] (one launch).
void mark_component_sides(Ctx &c, const std::vector<uint8_t> &sel);
// Sets the min-cut side flag of every vertex of the components flagged in all
// (components answered without a max-flow: the whole component).
void mark_component_sides_all(Ctx &c, const std::vector<uint8_t> &all);

// Exact search (or one max-flow when max_iters == 1) of many components at
// once, one workgroup each.
struct BlockTask {
  uint32_t d;     // component dense index
  ull p0, q0;     // starting alpha
};
struct BlockResult {
  ull p = 0, q = 1;            // final alpha (reduced)
  ull side_n = 0, side_T = 0;  // maximal min-cut side at the final alpha
  long long flow = 0;          // value of the last max-flow
  uint32_t flows = 0;
};
void dinkelbach_blocks(Ctx &c, const std::vector<BlockTask> &in, std::vector<BlockResult> &out, uint32_t max_iters);
// The same in two halves: the launch on stream st, then the wait and results.
struct BlockRun {
  std::vector<uint32_t> order;  // task j solves in[order[j]]
  void *dout = nullptr;
  hipStream_t st = nullptr;
};
void block_launch(Ctx &c, const std::vector<BlockTask> &in, uint32_t max_iters, hipStream_t st, BlockRun &run);
void block_finish(Ctx &c, const std::vector<BlockTask> &in, BlockRun &run, std::vector<BlockResult> &out);
void dinkelbach_grid(Ctx &c, const BlockTask &task, BlockResult &out, uint32_t max_iters);
// Reserves the grid solver's buffers for components of up to nc vertices and
// Tc cliques, so that solving them frees and allocates nothing.
void grid_reserve(Ctx &c, uint32_t nc, uint64_t Tc);
// Components with at least this many cliques are solved by the grid solver;
// with fewer components than a quarter of the CUs, from GRID_FLOW_ALONE_T up.
// (49,152: with the grid solves beside the workgroup launch, the C4 batch
// took 145-151 ms against 166-188 ms at 65,536 and 170-204 ms at 32-40K.)
constexpr uint64_t GRID_FLOW_MIN_T = 3u << 14, GRID_FLOW_ALONE_T = 1u << 12;
constexpr uint32_t NO_SLACK = 32, ORDER_SLACK_SHIFT = 1;  // slack order: thresholds x + x/2
constexpr uint32_t LIST_MAX_OUTDEG = 512;                  // widest out-list the listing stages in LDS
constexpr uint32_t LIST_WIDE_MAX = 4096;                   // widest out-list at all (wider roots: rows in HBM)
// All tasks: large components one after another on the whole GPU, the rest in
// one workgroup-per-component launch.
void dinkelbach_components(Ctx &c, const std::vector<BlockTask> &in, std::vector<BlockResult> &out,
                           uint32_t max_iters);
void graph_ids(Ctx &c);
void solve_graphs(Ctx &c, uint32_t k, kcds_result *agg, kcds_graph_result *per_graph, uint8_t *in_set);
void graph_ids(Ctx &c);

template <typename F> void StreamGraph::capture(F &&launches) {
  HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    launches();
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(c.stream, &g));
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_CHECK(e);
}

}  // namespace kcds

struct kcds_ctx {
  kcds::Ctx c;
};
