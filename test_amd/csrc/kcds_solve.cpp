// Stage sequence of a solve (one graph or a batch of disjoint graphs) and the
// exact density search.
//
// Reference: main (main.cu:1348-1519) runs generateDAG, listAllCliques,
// cliqueCoreDecompose, generateDensestCore, prune, componentDecompose and
// dynamicExactAlgo (main.cu:729-1270).  dynamicExactAlgo bounds every
// component (getLbUbandSize, src/helpers.cu:1115-1146: lb = Tc/nc,
// ub = min(max clique-core, Tc^((k-1)/k) / (k!)^(1/k))), skips components with
// ub <= max lb (main.cu:915) and bisects alpha with a double-precision
// push-relabel, one component after another on the host.
//
// Here the search is exact and rational, and components run concurrently.
// For every graph, alpha0 = its best whole-component density (max lb); every
// component whose upper bound reaches alpha0 runs a Dinkelbach iteration
// from alpha0: max-flow at alpha = p/q either certifies that no set is denser
// (flow == q*k*Tc) or yields a strictly denser set A (the source side of the
// maximal minimum cut) and alpha := density(A).  A graph's rho* is the largest
// final alpha of its components, and its answer set is the union of the
// maximal minimum-cut sides of the components that end at rho* -- the unique
// maximal densest set.  Every component runs its whole iteration inside one
// workgroup (one launch for all components of all graphs).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kcds_internal.hpp"

namespace kcds {

StreamGraph::~StreamGraph() {
  // a replay may still run when an error unwinds past its owner: the
  // executable graph is destroyed only once the stream has drained
  if (exec) {
    (void)hipStreamSynchronize(c.stream);
    (void)hipGraphExecDestroy(exec);
  }
}

void StreamGraph::launch() { HIP_CHECK(hipGraphLaunch(exec, c.stream)); }


void check_partition(Ctx &c);

namespace {

typedef unsigned __int128 u128;

constexpr uint32_t ALL_ITERATIONS = 0xFFFFFFFFu;

ull gcd(ull a, ull b) {
  while (b) {
    ull t = a % b;
    a = b;
    b = t;
  }
  return a;
}

bool ratio_gt(ull a, ull b, ull c, ull d) { return (u128)a * d > (u128)c * b; }

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, s));
  }
  float stop() {
    HIP_CHECK(hipEventRecord(b, s));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

// Upper bound of any set density inside a component with Tc cliques
// (helpers.cu:1126-1135), widened by a relative 1e-9 so that rounding never
// skips a component that could hold a set of density >= num/den.
bool may_reach(uint32_t k, ull max_core, ull Tc, ull num, ull den) {
  double kf = 1;
  for (uint32_t i = 2; i <= k; ++i) kf *= i;
  const double ub = std::min((double)max_core, std::pow((double)Tc, (k - 1.0) / k) / std::pow(kf, 1.0 / k));
  return ub * (1 + 1e-9) + 1e-9 >= (double)num / (double)den;
}

// True when Tc == C(nc, k): every k-subset of the component's nc vertices is
// a clique, so the component is a complete graph.  Its k-clique density
// C(t, k)/t = C(t-1, k-1)/k grows with the size t of a subset, so the whole
// component is its unique densest subset and no max-flow is needed.
bool complete_component(ull Tc, ull nc, uint32_t k) {
  if (nc < k) return false;
  u128 b = 1;  // C(nc, i) for i = 1..k, stopped once it passes Tc
  for (uint32_t i = 1; i <= k; ++i) {
    b = b * (u128)(nc - k + i) / i;
    if (b > (u128)Tc) return false;
  }
  return b == (u128)Tc;
}

}  // namespace

void dinkelbach_components(Ctx &c, const std::vector<BlockTask> &in, std::vector<BlockResult> &out,
                           uint32_t max_iters) {
  CompData &g = c.cd;
  FlowBufs &f = c.fb;
  out.assign(in.size(), BlockResult{});
  if (f.cside.bytes < (size_t)g.h_voff.back() + 8) {
    buf_reserve(c, f.cside, (size_t)g.h_voff.back() + 8);
    HIP_CHECK(hipMemsetAsync(f.cside.p, 0, (size_t)g.h_voff.back() + 8, c.stream));
  }
  const bool all_grid = (c.tune_flags & KCDS_TUNE_GRID_FLOW) != 0;
  // A batch of many components fills the GPU with one workgroup each; a few
  // components would leave most of it idle, so there the grid solver takes
  // them from GRID_FLOW_ALONE_T cliques up (a lone 36,000-clique component:
  // 39 ms in one workgroup, 7 ms on the grid; the first 100 C5 requests'
  // flow stage 522 -> 441 ms).
  const ull grid_min = (c.tune_flags & KCDS_TUNE_BLOCK_FLOW) ? ~0ull
                       : in.size() * 4 < (size_t)c.num_cus ? GRID_FLOW_ALONE_T
                                                          : GRID_FLOW_MIN_T;
  std::vector<BlockTask> small;
  std::vector<size_t> small_at, big_at;
  uint32_t big_nc = 0;
  ull big_T = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    const uint32_t d = in[i].d;
    const ull Tc = g.h_coff[d + 1] - g.h_coff[d];
    if (all_grid || Tc >= grid_min) {
      big_at.push_back(i);
      big_nc = std::max(big_nc, g.h_voff[d + 1] - g.h_voff[d]);
      big_T = std::max(big_T, Tc);
    } else {
      small.push_back(in[i]);
      small_at.push_back(i);
    }
  }
  // Both kinds: the workgroup launch runs on a stream of its own while the
  // grid solver takes the big components one after another; its buffers are
  // reserved at their largest first, so nothing is freed (which would wait for
  // the whole device) meanwhile.  The workgroup launch's tail is its largest
  // component, which leaves most CUs to the grid solver.
  const bool overlap = !small.empty() && !big_at.empty();
  hipStream_t bst = c.stream;
  if (overlap) {
    grid_reserve(c, big_nc, big_T);
    if (!c.lstream[0]) HIP_CHECK(hipStreamCreateWithFlags(&c.lstream[0], hipStreamNonBlocking));
    bst = c.lstream[0];
  }
  BlockRun run;
  block_launch(c, small, max_iters, bst, run);
  // the workgroup launch on its own stream still reads the component data and
  // writes the cut sides: when the grid solver throws, wait for it before the
  // error leaves (the caller may reuse the context at once)
  struct JoinOnThrow {
    hipStream_t st;
    bool armed;
    ~JoinOnThrow() {
      if (armed) (void)hipStreamSynchronize(st);
    }
  } join{bst, overlap};
  for (size_t i : big_at) dinkelbach_grid(c, in[i], out[i], max_iters);
  join.armed = false;
  std::vector<BlockResult> r;
  block_finish(c, small, run, r);
  for (size_t j = 0; j < small.size(); ++j) out[small_at[j]] = r[j];
}

void maxflow_component(Ctx &c, uint32_t comp_id, uint64_t p, uint64_t q, int64_t *flow, uint8_t *side) {
  CompData &g = c.cd;
  auto it = std::lower_bound(g.h_label.begin(), g.h_label.end(), comp_id);
  if (it == g.h_label.end() || *it != comp_id) fail(KCDS_E_INVALID, "no component with id %u", comp_id);
  const uint32_t d = (uint32_t)(it - g.h_label.begin());
  std::vector<BlockResult> r;
  dinkelbach_components(c, {BlockTask{d, p, q}}, r, 1);
  if (flow) *flow = r[0].flow;
  if (side && c.n) {
    buf_reserve(c, c.inset, (size_t)c.n + 8);
    HIP_CHECK(hipMemsetAsync(c.inset.p, 0, c.n, c.stream));
    mark_component_side(c, d);
    std::vector<uint8_t> tmp(c.n);
    std::vector<uint32_t> rank(c.n);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), c.inset.p, c.n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(rank.data(), c.rank.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (uint32_t v = 0; v < c.n; ++v) side[v] = tmp[rank[v]];
  }
}

// Solves every graph of the current partition (c.ngraphs, c.h_gvoff).
void solve_graphs(Ctx &c, uint32_t k, kcds_result *agg, kcds_graph_result *per_graph, uint8_t *in_set) {
  const uint32_t G = c.ngraphs;
  kcds_result R;
  memset(&R, 0, sizeof R);
  std::vector<kcds_graph_result> out(G);
  memset(out.data(), 0, sizeof(kcds_graph_result) * G);
  Timer total(c.stream);
  {
    Timer t(c.stream);
    orient_gpu(c);
    c.oriented = true;
    R.ms_orient = t.stop();
  }
  if (G > 1) {
    check_partition(c);
    graph_ids(c);
  }
  std::vector<ull> gT;
  std::vector<uint8_t> drop(G, 0);
  bool enumerate = false;
  {
    Timer t(c.stream);
    // per-graph HBM admission on the count pass: cliques, incidence and
    // status take T * (8k + 4) bytes, so a graph is admitted iff T <= cap;
    // the count pass stops counting a graph once it has passed cap
    const ull cap = c.graph_budget ? c.graph_budget / (8ull * k + 4ull) : 0;
    list_count(c, k, cap);
    R.T = c.T;
    graph_counts(c, gT);
    uint32_t ndrop = 0;
    if (cap)
      for (uint32_t g = 0; g < G; ++g)
        if (gT[g] > cap) {
          drop[g] = 1;
          ++ndrop;
        }
    if (ndrop && G == 1) {
      c.counted = false;  // the counts stopped at the cap
      fail(KCDS_E_NOMEM, "graph has more than %llu %u-cliques (counted to %llu): over %.0f bytes to materialise, "
           "over the graph budget %llu", (unsigned long long)cap, k, (unsigned long long)gT[0],
           (double)gT[0] * (8.0 * k + 4.0), (unsigned long long)c.graph_budget);
    }
    if (ndrop) drop_graphs(c, drop, gT);
    // the clique list, its incidence and status take T * (8k + 4) bytes and
    // 32-bit clique ids; when that does not fit, the peel enumerates instead
    // and only the k'-cores' cliques are listed afterwards.  A single graph
    // from ENUM_MIN_T cliques up enumerates as well: listing and peeling the
    // list costs more than enumerating each removed vertex's cliques there.
    const size_t avail = c.budget > c.used ? c.budget - c.used : 0;
    enumerate = (c.tune_flags & KCDS_TUNE_ENUM_PEEL) || (G == 1 && c.T >= ENUM_MIN_T) ||
                (c.T && (c.T >= 0xFFFFFFFFull || (double)c.T * (8.0 * k + 4.0) > 0.85 * (double)avail));
    // the materialising peel keeps int32 clique statuses (destroy levels, at
    // most a vertex's count): larger counts take the enumerating peel
    if (!enumerate && c.T > 0x7FFFFFFFull && max_u64_host(c, c.cnt.as<ull>(), c.n) > 0x7FFFFFFFull) enumerate = true;
    // the enumerating peel stages a vertex's live neighbours in 512-row
    // frames: a DAG with wider out-lists peels over the clique list
    if (enumerate && c.dmax > LIST_MAX_OUTDEG) {
      if (c.T >= 0xFFFFFFFFull || (double)c.T * (8.0 * k + 4.0) > 0.85 * (double)avail)
        fail(KCDS_E_LIMIT, "out-lists up to %u (above %u) and %llu %u-cliques: too many to list for the peel",
             c.dmax, LIST_MAX_OUTDEG, (unsigned long long)c.T, k);
      enumerate = false;
    }
    if (c.T && !enumerate) list_materialize(c, k);
    R.ms_list = t.stop();
  }
  if (in_set && c.n) memset(in_set, 0, c.n);
  auto finish_status = [&]() {
    for (uint32_t gi = 0; gi < G; ++gi)
      if (drop[gi]) {
        memset(&out[gi], 0, sizeof(kcds_graph_result));
        out[gi].T = gT[gi];
        out[gi].status = KCDS_GRAPH_OVER_BUDGET;
      }
  };
  if (c.T == 0) {  // main.cu:1431-1438: "Max Density 0 Size 0" for every graph
    for (uint32_t gi = 0; gi < G; ++gi) out[gi].T = gT[gi];
    finish_status();
    R.ms_total = total.stop();
    if (agg) *agg = R;
    if (per_graph) memcpy(per_graph, out.data(), sizeof(kcds_graph_result) * G);
    return;
  }
  {
    Timer t(c.stream);
    if (enumerate) {
      epeel_run(c);
    } else {
      c.enum_peel = false;
      peel_run(c);
    }
    R.peel = c.peel;
    R.k_prime = c.peel.k_prime;
    R.core_vertices = c.peel.core_size;
    R.ms_peel = t.stop();
  }
  // A single graph whose k'-core is small in vertices but huge in cliques
  // (from DENSE_MIN_T up, or always with KCDS_TUNE_DENSE) takes the dense
  // exact search: no clique list, no clique flow network (k_dense.hip).
  const bool dense_try = G == 1 && k >= 3 && c.peel.core_size > 0 && c.peel.core_size <= 1024 &&
                         ((c.tune_flags & KCDS_TUNE_DENSE) || c.peel.core_cliques >= DENSE_MIN_T || enumerate);
  if (dense_try) {
    Timer t(c.stream);
    if (enumerate) {
      core_materialize(c, false);  // counts the core's cliques, lists nothing
      R.peel.core_cliques = c.peel.core_cliques;
    } else {
      build_core_dag(c);
    }
    if ((c.tune_flags & KCDS_TUNE_DENSE) || c.peel.core_cliques >= DENSE_MIN_T) {
      ull num = 0, den = 0;
      uint32_t ncomp = 0, passes = 0, critical = 0;
      const size_t avail = c.budget > c.used ? c.budget - c.used : 0;
      const bool listable = c.peel.core_cliques < 0xFFFFFFFFull / (2ull * k) &&
                            (double)c.peel.core_cliques * (8.0 * k + 4.0) <= 0.85 * (double)avail;
      const uint32_t iters = (c.tune_flags & KCDS_TUNE_DENSE) || !listable ? DENSE_MAX_ITERS
                             : c.dense_budget ? c.dense_budget : DENSE_LISTABLE_ITERS;
      if (dense_core_solve(c, k, num, den, ncomp, passes, critical, iters)) {
        R.ms_components = t.stop();
        Timer tf(c.stream);
        R.ncomp = ncomp;
        R.dense_passes = passes;
        R.dense_critical = critical;
        R.rho_num = num;
        R.rho_den = den;
        out[0].T = c.h_gpeel[0].T;
        out[0].rho_num = num;
        out[0].rho_den = den;
        out[0].k_prime = c.h_gpeel[0].k_prime;
        out[0].core_vertices = c.h_gpeel[0].core_size;
        out[0].ncomp = ncomp;
        if (in_set && c.n) {
          std::vector<uint8_t> tmp(c.n);
          std::vector<uint32_t> rank(c.n);
          HIP_CHECK(hipMemcpyAsync(tmp.data(), c.inset.p, c.n, hipMemcpyDeviceToHost, c.stream));
          HIP_CHECK(hipMemcpyAsync(rank.data(), c.rank.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost,
                                   c.stream));
          c.sync();
          for (uint32_t v = 0; v < c.n; ++v) in_set[v] = tmp[rank[v]];
        }
        R.ms_flow = tf.stop();
        R.ms_total = total.stop();
        if (agg) *agg = R;
        if (per_graph) memcpy(per_graph, out.data(), sizeof(kcds_graph_result) * G);
        return;
      }
      if (c.peel.core_cliques >= 0xFFFFFFFFull / (2ull * k))
        fail(KCDS_E_LIMIT, "dense exact search on the %u-vertex k'-core (%llu cliques) not certified within %u "
             "greedy passes", c.peel.core_size, (unsigned long long)c.peel.core_cliques, passes);
    }
    (void)t.stop();
  }
  {
    Timer t(c.stream);
    if (enumerate) {
      core_materialize(c);
      R.peel.core_cliques = c.peel.core_cliques;
      if (!c.core_listed)
        fail(KCDS_E_LIMIT, "the k'-core holds %llu cliques, too many to list in HBM",
             (unsigned long long)c.peel.core_cliques);
    }
    components_run(c);
    R.ncomp = c.ncomp;
    R.ms_components = t.stop();
  }
  Timer tf(c.stream);
  CompData &g = c.cd;
  const uint32_t C = g.ncomp;
  // graph of each component: labels are original ids, graphs own id ranges
  std::vector<uint32_t> cg(C);
  for (uint32_t d = 0; d < C; ++d)
    cg[d] = (uint32_t)(std::upper_bound(c.h_gvoff.begin(), c.h_gvoff.end(), g.h_label[d]) - c.h_gvoff.begin()) - 1;
  // alpha0 per graph: the best whole-component density
  std::vector<ull> an(G, 0), ad(G, 1);
  for (uint32_t d = 0; d < C; ++d) {
    const ull tc = g.h_coff[d + 1] - g.h_coff[d], nc = g.h_voff[d + 1] - g.h_voff[d];
    const uint32_t gi = cg[d];
    out[gi].ncomp++;
    if (nc && ratio_gt(tc, nc, an[gi], ad[gi])) {
      an[gi] = tc;
      ad[gi] = nc;
    }
  }
  // the peel's density record is that of a vertex set too (the vertices
  // left at some level), often rho* itself: starting there, one max-flow
  // certifies a component whose densest set the peel already found
  for (uint32_t gi = 0; gi < G; ++gi) {
    const GraphPeel &P = c.h_gpeel[gi];
    if (P.max_den && ratio_gt(P.max_num, P.max_den, an[gi], ad[gi])) {
      an[gi] = P.max_num;
      ad[gi] = P.max_den;
    }
  }
  for (uint32_t gi = 0; gi < G; ++gi) {
    ull gg = gcd(an[gi], ad[gi]);
    if (gg) {
      an[gi] /= gg;
      ad[gi] /= gg;
    }
  }
  uint64_t core_vertices = g.h_voff[C];
  buf_reserve(c, c.fb.cside, core_vertices + 8);
  HIP_CHECK(hipMemsetAsync(c.fb.cside.p, 0, core_vertices + 8, c.stream));
  std::vector<BlockTask> tasks;
  // final alpha and side size of every solved component
  std::vector<BlockResult> res(C);
  std::vector<uint8_t> solved(C, 0), complete(C, 0);
  bool any_complete = false;
  for (uint32_t d = 0; d < C; ++d) {
    const ull tc = g.h_coff[d + 1] - g.h_coff[d], nc = g.h_voff[d + 1] - g.h_voff[d];
    const uint32_t gi = cg[d];
    if (tc == 0 || !may_reach(k, c.h_gpeel[gi].max_core, tc, an[gi], ad[gi])) continue;
    if (!(c.tune_flags & KCDS_TUNE_FLOW_ALL) && complete_component(tc, nc, k)) {  // its answer is itself: alpha = Tc/nc, side = all of it
      const ull gg = gcd(tc, nc);
      res[d].p = tc / gg;
      res[d].q = nc / gg;
      res[d].side_n = nc;
      res[d].side_T = tc;
      solved[d] = complete[d] = 1;
      any_complete = true;
      continue;
    }
    tasks.push_back(BlockTask{d, an[gi], ad[gi]});
  }
  if (any_complete) mark_component_sides_all(c, complete);
  {
    std::vector<BlockResult> r;
    dinkelbach_components(c, tasks, r, ALL_ITERATIONS);
    for (size_t i = 0; i < tasks.size(); ++i) {
      res[tasks[i].d] = r[i];
      solved[tasks[i].d] = 1;
      out[cg[tasks[i].d]].flow_calls += r[i].flows;
    }
  }
  // per graph: rho* = largest final alpha among components with a non-empty
  // side; the answer is the union of those components' sides at rho*
  std::vector<ull> bn(G, 0), bd(G, 1);
  std::vector<uint8_t> have(G, 0);
  for (uint32_t d = 0; d < C; ++d) {
    if (!solved[d] || res[d].side_n == 0) continue;
    const uint32_t gi = cg[d];
    if (!have[gi] || ratio_gt(res[d].p, res[d].q, bn[gi], bd[gi])) {
      bn[gi] = res[d].p;
      bd[gi] = res[d].q;
      have[gi] = 1;
    }
  }
  buf_reserve(c, c.inset, (size_t)c.n + 8);
  HIP_CHECK(hipMemsetAsync(c.inset.p, 0, c.n, c.stream));
  std::vector<uint8_t> sel(C, 0);
  bool any = false;
  for (uint32_t d = 0; d < C; ++d) {
    if (!solved[d] || res[d].side_n == 0) continue;
    const uint32_t gi = cg[d];
    if ((u128)res[d].p * bd[gi] != (u128)bn[gi] * res[d].q) continue;
    sel[d] = 1;
    any = true;
    out[gi].rho_num += res[d].side_T;
    out[gi].rho_den += res[d].side_n;
  }
  if (any) mark_component_sides(c, sel);
  for (uint32_t gi = 0; gi < G; ++gi) {
    const GraphPeel &P = c.h_gpeel[gi];
    out[gi].T = P.T;
    out[gi].k_prime = P.T ? P.k_prime : 0;
    out[gi].core_vertices = P.T ? P.core_size : 0;
    if (P.T && (out[gi].rho_den == 0 || (u128)out[gi].rho_num * bd[gi] != (u128)bn[gi] * out[gi].rho_den))
      fail(KCDS_E_HIP, "densest set check failed for graph %u", gi);
  }
  R.rho_num = out[0].rho_num;
  R.rho_den = out[0].rho_den;
  for (uint32_t gi = 0; gi < G; ++gi) R.flow_calls += out[gi].flow_calls;
  R.ms_flow = tf.stop();
  finish_status();
  if (in_set && c.n) {
    std::vector<uint8_t> tmp(c.n);
    std::vector<uint32_t> rank(c.n);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), c.inset.p, c.n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(rank.data(), c.rank.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (uint32_t v = 0; v < c.n; ++v) in_set[v] = tmp[rank[v]];
  }
  R.ms_total = total.stop();
  if (agg) *agg = R;
  if (per_graph) memcpy(per_graph, out.data(), sizeof(kcds_graph_result) * G);
}

}  // namespace kcds
