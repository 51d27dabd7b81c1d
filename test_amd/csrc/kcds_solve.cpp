// Stage sequence of a whole solve and the exact density search.
//
// Reference: main (main.cu:1348-1519) runs generateDAG, listAllCliques,
// cliqueCoreDecompose, generateDensestCore, prune, componentDecompose and
// dynamicExactAlgo (main.cu:729-1270).  dynamicExactAlgo bounds every
// component (getLbUbandSize, src/helpers.cu:1115-1146: lb = Tc/nc,
// ub = min(max clique-core, Tc^((k-1)/k) / (k!)^(1/k))), skips components with
// ub <= max lb (main.cu:915) and bisects alpha with a double-precision
// push-relabel per component.
//
// Here the search is exact and rational.  alpha starts at the best whole
// component density (max lb); for each component (in canonical order) whose
// upper bound reaches alpha, max-flow at alpha = p/q either certifies that
// no set is denser (flow == q*k*Tc) or yields a strictly denser set A (the
// source side of the maximal minimum cut), and alpha := density(A).  When no
// component admits a denser set, alpha = rho* exactly; the answer set is the
// union over components of the maximal minimum-cut source sides at rho*,
// i.e. the unique maximal densest set.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kcds_internal.hpp"

namespace kcds {

void gather_by_rank_u32(Ctx &c, const uint32_t *src_new, uint32_t *dst_orig);
void rank_from_listing_order(Ctx &c, const uint32_t *lo_dev);

namespace {

typedef unsigned __int128 u128;

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

// upper bound of any set density inside component d (helpers.cu:1126-1135);
// conservative by a relative 1e-9 so rounding never skips a component that
// could hold a set of density >= num/den
bool may_reach(const Ctx &c, uint32_t d, ull num, ull den) {
  const CompData &g = c.cd;
  const double tc = (double)(g.h_coff[d + 1] - g.h_coff[d]);
  double kf = 1;
  for (uint32_t i = 2; i <= c.k; ++i) kf *= i;
  const double ub = std::min((double)c.peel.max_core, std::pow(tc, (c.k - 1.0) / c.k) / std::pow(kf, 1.0 / c.k));
  return ub * (1 + 1e-9) + 1e-9 >= (double)num / (double)den;
}

}  // namespace

void maxflow_component(Ctx &c, uint32_t comp_id, uint64_t p, uint64_t q, int64_t *flow, uint8_t *side) {
  CompData &g = c.cd;
  auto it = std::lower_bound(g.h_label.begin(), g.h_label.end(), comp_id);
  if (it == g.h_label.end() || *it != comp_id) fail(KCDS_E_INVALID, "no component with id %u", comp_id);
  const uint32_t d = (uint32_t)(it - g.h_label.begin());
  FlowOut o = flow_component(c, d, p, q);
  if (flow) *flow = o.value;
  if (side && c.n) {
    buf_reserve(c, c.inset, (size_t)c.n + 8);
    HIP_CHECK(hipMemsetAsync(c.inset.p, 0, c.n, c.stream));
    flow_mark_side(c, d);
    std::vector<uint8_t> tmp(c.n);
    std::vector<uint32_t> rank(c.n);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), c.inset.p, c.n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(rank.data(), c.rank.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (uint32_t v = 0; v < c.n; ++v) side[v] = tmp[rank[v]];
  }
}

void solve_pipeline(Ctx &c, uint32_t k, kcds_result *res, uint8_t *in_set) {
  kcds_result R;
  memset(&R, 0, sizeof R);
  Timer total(c.stream);
  {
    Timer t(c.stream);
    degeneracy_rank_gpu(c);
    orient_from_rank(c);
    c.oriented = true;
    R.ms_orient = t.stop();
  }
  {
    Timer t(c.stream);
    list_count(c, k);
    R.T = c.T;
    if (c.T) list_materialize(c, k);
    R.ms_list = t.stop();
  }
  if (in_set && c.n) memset(in_set, 0, c.n);
  if (c.T == 0) {  // main.cu:1431-1438: "Max Density 0 Size 0"
    R.ms_total = total.stop();
    if (res) *res = R;
    return;
  }
  {
    Timer t(c.stream);
    peel_run(c);
    R.peel = c.peel;
    R.k_prime = c.peel.k_prime;
    R.core_vertices = c.peel.core_size;
    R.ms_peel = t.stop();
  }
  {
    Timer t(c.stream);
    components_run(c);
    R.ncomp = c.ncomp;
    R.ms_components = t.stop();
  }
  Timer tf(c.stream);
  CompData &g = c.cd;
  const uint32_t C = g.ncomp;
  ull an = 0, ad = 1;
  for (uint32_t d = 0; d < C; ++d) {
    const ull tc = g.h_coff[d + 1] - g.h_coff[d], nc = g.h_voff[d + 1] - g.h_voff[d];
    if (nc && ratio_gt(tc, nc, an, ad)) {
      an = tc;
      ad = nc;
    }
  }
  {
    ull gg = gcd(an, ad);
    if (gg) {
      an /= gg;
      ad /= gg;
    }
  }
  // alpha at which each component's last flow ran (its side is then reusable)
  std::vector<std::pair<ull, ull>> last(C, {0, 0});
  for (uint32_t d = 0; d < C; ++d) {
    const ull tc = g.h_coff[d + 1] - g.h_coff[d];
    if (tc == 0) continue;
    while (may_reach(c, d, an, ad)) {
      FlowOut o = flow_component(c, d, an, ad);
      R.flow_calls++;
      if ((u128)o.value >= (u128)ad * k * tc) break;  // no set denser than alpha
      if (o.side_n == 0 || !ratio_gt(o.side_T, o.side_n, an, ad))
        fail(KCDS_E_HIP, "min cut below q*k*Tc without a denser source side (component %u)", d);
      ull gg = gcd(o.side_T, o.side_n);
      an = o.side_T / gg;
      ad = o.side_n / gg;
    }
  }
  // maximal densest set at alpha = rho*
  buf_reserve(c, c.inset, (size_t)c.n + 8);
  HIP_CHECK(hipMemsetAsync(c.inset.p, 0, c.n, c.stream));
  ull snum = 0, sden = 0;
  for (uint32_t d = 0; d < C; ++d) {
    const ull tc = g.h_coff[d + 1] - g.h_coff[d];
    if (tc == 0 || !may_reach(c, d, an, ad)) continue;
    FlowOut o = flow_component(c, d, an, ad);
    R.flow_calls++;
    if (o.side_n == 0) continue;
    flow_mark_side(c, d);
    snum += o.side_T;
    sden += o.side_n;
  }
  if (sden == 0 || (u128)snum * ad != (u128)an * sden)
    fail(KCDS_E_HIP, "densest set check failed: %llu/%llu vs %llu/%llu", (unsigned long long)snum,
         (unsigned long long)sden, (unsigned long long)an, (unsigned long long)ad);
  R.rho_num = snum;
  R.rho_den = sden;
  R.ms_flow = tf.stop();
  if (in_set && c.n) {
    std::vector<uint8_t> tmp(c.n);
    std::vector<uint32_t> rank(c.n);
    HIP_CHECK(hipMemcpyAsync(tmp.data(), c.inset.p, c.n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(rank.data(), c.rank.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (uint32_t v = 0; v < c.n; ++v) in_set[v] = tmp[rank[v]];
  }
  R.ms_total = total.stop();
  if (res) *res = R;
}

}  // namespace kcds
