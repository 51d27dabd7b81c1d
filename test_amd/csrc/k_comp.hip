// Densest core, edge pruning and connected components.
//
// Reference: generateDensestCore (main.cu:422-527, src/helpers.cu:797-872)
// keeps {v : cliqueCore[v] >= k'} with k' = ceil(maxDensity); prune
// (main.cu:528-613, src/helpers.cu:874-1017) keeps an edge iff it lies in a
// clique whose status >= k' (all members in the core); componentDecompose
// (main.cu:615-728, src/helpers.cu:913-954) propagates minimum labels to a
// fixpoint and renumbers components in label order.
//
// Here the surviving edges are exactly the member pairs of core cliques, so
// components are formed directly over core cliques by lock-free union-find:
// hooking always attaches the root with the larger key (original vertex id)
// under the one with the smaller key, and finds halve paths (pointer
// jumping).  The root of each component is therefore its smallest original
// id, which is the canonical component id.  Components are then numbered
// densely in order of that id, and the core vertices and core cliques are
// grouped per component (local vertex ids) for the flow stage.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr int CC_E = 8;        // cliques per thread in a component-count round
constexpr int CC_TB = 4096;    // its LDS key table (2 * 256 * CC_E slots)

__device__ __forceinline__ uint32_t ld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t *par, uint32_t x) {
  while (true) {
    const uint32_t p = ld(&par[x]);
    if (p == x) return x;
    const uint32_t gp = ld(&par[p]);
    if (gp != p) atomicCAS(&par[x], p, gp);  // path halving: gp is an ancestor of x
    x = gp;
  }
}

// Root of x followed through plain, cacheable loads.  Every value par[y] ever
// held is y or a vertex of smaller key (hooks point down in key, path halving
// to a grandparent), so the walk ends, at a vertex that was a root of x's set
// when it was read; sets never split, so equal results for two vertices mean
// one set.  (Unequal results may be stale: the caller then unites exactly.)
__device__ uint32_t uf_find_ro(const uint32_t *par, uint32_t x) {
  while (true) {
    const uint32_t p = par[x];
    if (p == x) return x;
    x = p;
  }
}

__device__ void uf_unite(uint32_t *par, const uint32_t *key, uint32_t a, uint32_t b) {
  while (true) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (key[a] < key[b]) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // a has the larger key: hook it under b (keys strictly decrease upward)
    if (atomicCAS(&par[a], a, b) == a) return;
  }
}

__global__ void k_uf_init(uint32_t n, uint32_t *par) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) par[v] = v;
}

// Points every vertex straight at its root (no unions run meanwhile).
__global__ void k_uf_compress(uint32_t n, uint32_t *par) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t r = uf_find(par, v);
    if (r != v) __hip_atomic_store(&par[v], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// a clique is in its graph's k'-core iff its status (min member core) >= k'
__device__ __forceinline__ bool core_clique(uint64_t c, uint32_t k, const uint32_t *cl, const int32_t *status,
                                            const GraphPeel *gp, const uint32_t *gid) {
  return (long long)status[c] >= (long long)gp[gid[cl[c * k]]].k_prime;
}

// Unites the members of every stride-th clique.
__global__ void k_uf_hook(uint64_t T, uint64_t stride, uint32_t k, const uint32_t *cl, const int32_t *status,
                          const GraphPeel *gp, const uint32_t *gid, uint32_t *par, const uint32_t *key) {
  const uint64_t ns = (T + stride - 1) / stride;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = s * stride;
    if (!core_clique(c, k, cl, status, gp, gid)) continue;
    const uint32_t a = cl[c * k];
    // members with one common parent are in one set already: most cliques,
    // once the big components have formed, and this test reads no root (a
    // root's entry is read by every find and would be one hot L2 line).  The
    // test uses plain, cacheable loads: every value par[x] ever held is an
    // ancestor of x at that time and sets never split, so equal parents seen
    // in any (possibly stale) copies still mean one set; a stale copy can
    // only send the clique to the exact unite below.
    const uint32_t pa = par[a];
    bool same = true;
    for (uint32_t t = 1; t < k; ++t) same = same && par[cl[c * k + t]] == pa;
    if (same) continue;
    // then common roots through plain loads (a hub core makes the parents of
    // members differ inside one set long before they all point at the root)
    const uint32_t ra = uf_find_ro(par, a);
    same = true;
    for (uint32_t t = 1; t < k; ++t) same = same && uf_find_ro(par, cl[c * k + t]) == ra;
    if (same) continue;
    for (uint32_t t = 1; t < k; ++t) uf_unite(par, key, a, cl[c * k + t]);
  }
}

// comp[r] = key of the root for core vertices, ~0 elsewhere; roots flagged in
// original-id order for dense numbering
__global__ void k_uf_label(uint32_t n, const long long *core, const GraphPeel *gp, const uint32_t *gid, uint32_t *par,
                           const uint32_t *key, uint32_t *comp, uint32_t *isroot_orig) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const bool in = core[r] >= (long long)gp[gid[r]].k_prime;
    const uint32_t root = uf_find(par, r);
    comp[r] = in ? key[root] : 0xFFFFFFFFu;
    isroot_orig[key[r]] = (in && root == r) ? 1u : 0u;
  }
}

// dense component index, vertex counts and local vertex ids
__global__ void k_group_vertices(uint32_t n, const uint32_t *comp, const uint32_t *dense_of_orig,
                                 const uint32_t *rank, uint32_t *cid, uint32_t *vcount, uint32_t *local) {
  // one atomic per group of lanes in the same component (one giant
  // component is the common case)
  for (uint32_t r0 = blockIdx.x * blockDim.x; r0 < n; r0 += gridDim.x * blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    const uint32_t lab = r < n ? comp[r] : 0xFFFFFFFFu;
    const bool in = lab != 0xFFFFFFFFu;
    const uint32_t d = in ? dense_of_orig[lab] : 0u;
    if (r < n) cid[r] = in ? d : 0xFFFFFFFFu;
    const uint32_t slot = wave_slot_by_key<uint32_t>(in, d, vcount);
    if (in) local[r] = slot;
  }
}

__global__ void k_scatter_vertices(uint32_t n, const uint32_t *cid, const uint32_t *local, const uint32_t *voff,
                                   uint32_t *verts) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t d = cid[r];
    if (d != 0xFFFFFFFFu) verts[voff[d] + local[r]] = r;
  }
}

// Core cliques per component.  A batch has few components and a graph
// usually one, so the cliques of a workgroup round (CC_E per thread) take
// their component's count with one atomic per distinct component.
__global__ __launch_bounds__(256) void k_count_comp_cliques(uint64_t T, uint32_t k, const uint32_t *cl,
                                                            const int32_t *status, const GraphPeel *gp,
                                                            const uint32_t *gid, const uint32_t *cid, ull *ccount) {
  __shared__ KeyTable<CC_TB> tb;
  const uint64_t round = 256ull * CC_E, stride = (uint64_t)gridDim.x * round;
  for (uint64_t c0 = (uint64_t)blockIdx.x * round; c0 < T; c0 += stride) {
    bool want[CC_E];
    uint32_t key[CC_E];
    ull slot[CC_E];
#pragma unroll
    for (int e = 0; e < CC_E; ++e) {
      const uint64_t c = c0 + (uint64_t)e * 256 + threadIdx.x;
      want[e] = c < T && core_clique(c, k, cl, status, gp, gid);
      key[e] = want[e] ? cid[cl[c * k]] : 0u;
    }
    block_slot_by_key<CC_E, CC_TB, ull>(tb, want, key, slot, ccount);
  }
}

// core cliques grouped by component, members as local vertex ids
__global__ __launch_bounds__(256) void k_scatter_cliques(uint64_t T, uint32_t k, const uint32_t *cl,
                                                         const int32_t *status, const GraphPeel *gp,
                                                         const uint32_t *gid, const uint32_t *cid,
                                                         const uint32_t *local, const ull *coff, ull *cfill,
                                                         uint32_t *ccl) {
  __shared__ KeyTable<CC_TB> tb;
  const uint64_t round = 256ull * CC_E, stride = (uint64_t)gridDim.x * round;
  for (uint64_t c0 = (uint64_t)blockIdx.x * round; c0 < T; c0 += stride) {
    bool want[CC_E];
    uint32_t key[CC_E];
    ull slot[CC_E];
#pragma unroll
    for (int e = 0; e < CC_E; ++e) {
      const uint64_t c = c0 + (uint64_t)e * 256 + threadIdx.x;
      want[e] = c < T && core_clique(c, k, cl, status, gp, gid);
      key[e] = want[e] ? cid[cl[c * k]] : 0u;
    }
    block_slot_by_key<CC_E, CC_TB, ull>(tb, want, key, slot, cfill);
#pragma unroll
    for (int e = 0; e < CC_E; ++e) {
      if (!want[e]) continue;
      const uint64_t c = c0 + (uint64_t)e * 256 + threadIdx.x;
      const uint32_t d = key[e];
      const ull at = coff[d] + slot[e];
      if (at >= coff[d + 1]) continue;  // never taken: counts and scatter agree
      for (uint32_t t = 0; t < k; ++t) ccl[at * k + t] = local[cl[c * k + t]];
    }
  }
}

// counts edges whose endpoints lie in different graphs of the batch
__global__ void k_check_partition(uint32_t n, const uint32_t *off, const uint32_t *nbr, const uint32_t *gvoff,
                                  uint32_t G, uint32_t *bad) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t g = upper_bound_u32(gvoff, G + 1, v) - 1;
    const uint32_t lo = gvoff[g], hi = gvoff[g + 1];
    for (uint32_t i = off[v]; i < off[v + 1]; ++i)
      if (nbr[i] < lo || nbr[i] >= hi) {
        atomicAdd(bad, 1u);
        break;
      }
  }
}

uint32_t grid1d(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 16));
}

constexpr uint64_t UF_STRIDE0 = 4096, UF_STRIDE1 = 64;  // union-find sample passes

}  // namespace

void check_partition(Ctx &c) {
  const uint32_t n = c.n, G = c.ngraphs;
  buf_reserve(c, c.gvoff, sizeof(uint32_t) * ((size_t)G + 1));
  buf_reserve(c, c.scal, 64 * sizeof(ull));
  uint32_t *bad = (uint32_t *)(c.scal.as<ull>() + 40);
  HIP_CHECK(hipMemcpyAsync(c.gvoff.p, c.h_gvoff.data(), sizeof(uint32_t) * ((size_t)G + 1), hipMemcpyHostToDevice,
                           c.stream));
  HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(uint32_t), c.stream));
  if (n) {
    k_check_partition<<<grid1d(n), 256, 0, c.stream>>>(n, c.off.as<uint32_t>(), c.nbr.as<uint32_t>(),
                                                       c.gvoff.as<uint32_t>(), G, bad);
    KCDS_LAUNCH_CHECK();
  }
  uint32_t *h = (uint32_t *)c.host_scratch(8);
  HIP_CHECK(hipMemcpyAsync(h, bad, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (*h) fail(KCDS_E_INVALID, "%u vertices have edges into another graph of the batch", *h);
}

void components_run(Ctx &c) {
  if (!c.peeled) fail(KCDS_E_STATE, "components before peel");
  const uint32_t n = c.n, k = c.k;
  const uint64_t T = c.T;
  const GraphPeel *gp = c.gpeel.as<GraphPeel>();
  const uint32_t *gidp = c.gid.as<uint32_t>();
  uint64_t core_vertices = 0, core_cliques = 0;
  for (const GraphPeel &P : c.h_gpeel) {
    core_vertices += P.core_size;
    core_cliques += P.core_cliques;
  }
  CompData &g = c.cd;
  buf_reserve(c, c.comp, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, g.par, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, g.flag, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, g.dense, sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t *par = g.par.as<uint32_t>();
  const uint32_t *key = c.inv.as<uint32_t>();
  k_uf_init<<<grid1d(n), 256, 0, c.stream>>>(n, par);
  KCDS_LAUNCH_CHECK();
  // Every 4096th clique, then every 64th, then all, with every path
  // compressed after each sample pass.  The samples, spread over the whole
  // clique range, form the big components with few competing unions; after
  // them a member's parent is its root, and almost every clique of the full
  // pass leaves at the common-parent test without a find.  (One launch over
  // all 6M cliques of a 697-vertex core had hundreds of thousands of threads
  // in finds and root CASes at once: 1.8-2.1 ms.)
  for (uint64_t stride : {UF_STRIDE0, UF_STRIDE1, (uint64_t)1}) {
    if (stride > 1 && T < stride * 64) continue;  // small T: the full pass alone
    k_uf_hook<<<grid1d((T + stride - 1) / stride), 256, 0, c.stream>>>(T, stride, k, c.cliques.as<uint32_t>(),
                                                                        c.status.as<int32_t>(), gp, gidp, par, key);
    KCDS_LAUNCH_CHECK();
    if (stride > 1) {
      k_uf_compress<<<grid1d(n), 256, 0, c.stream>>>(n, par);
      KCDS_LAUNCH_CHECK();
    }
  }
  k_uf_label<<<grid1d(n), 256, 0, c.stream>>>(n, c.core.as<long long>(), gp, gidp, par, key, c.comp.as<uint32_t>(),
                                              g.flag.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
  // dense[orig] = number of component roots with a smaller original id
  scan_u32_to_u32(c, g.flag.as<uint32_t>(), g.dense.as<uint32_t>(), n);
  uint32_t *h = (uint32_t *)c.host_scratch(16);
  HIP_CHECK(hipMemcpyAsync(h, g.dense.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  const uint32_t C = h[0];
  c.ncomp = C;
  g.ncomp = C;
  // vertices per component
  buf_reserve(c, g.cid, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, g.local, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, g.vcount, sizeof(uint32_t) * ((size_t)C + 2));
  buf_reserve(c, g.voff, sizeof(uint32_t) * ((size_t)C + 2));
  buf_reserve(c, g.verts, sizeof(uint32_t) * ((size_t)core_vertices + 1));
  HIP_CHECK(hipMemsetAsync(g.vcount.p, 0, sizeof(uint32_t) * ((size_t)C + 2), c.stream));
  k_group_vertices<<<grid1d(n), 256, 0, c.stream>>>(n, c.comp.as<uint32_t>(), g.dense.as<uint32_t>(),
                                                    c.rank.as<uint32_t>(), g.cid.as<uint32_t>(),
                                                    g.vcount.as<uint32_t>(), g.local.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
  scan_u32_to_u32(c, g.vcount.as<uint32_t>(), g.voff.as<uint32_t>(), C);
  k_scatter_vertices<<<grid1d(n), 256, 0, c.stream>>>(n, g.cid.as<uint32_t>(), g.local.as<uint32_t>(),
                                                      g.voff.as<uint32_t>(), g.verts.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
  // cliques per component
  buf_reserve(c, g.ccount, sizeof(ull) * ((size_t)C + 2));
  buf_reserve(c, g.coff, sizeof(ull) * ((size_t)C + 2));
  buf_reserve(c, g.cfill, sizeof(ull) * ((size_t)C + 2));
  HIP_CHECK(hipMemsetAsync(g.ccount.p, 0, sizeof(ull) * ((size_t)C + 2), c.stream));
  HIP_CHECK(hipMemsetAsync(g.cfill.p, 0, sizeof(ull) * ((size_t)C + 2), c.stream));
  if (T) {
    k_count_comp_cliques<<<grid1d((T + CC_E - 1) / CC_E), 256, 0, c.stream>>>(T, k, c.cliques.as<uint32_t>(), c.status.as<int32_t>(),
                                                          gp, gidp, g.cid.as<uint32_t>(), g.ccount.as<ull>());
    KCDS_LAUNCH_CHECK();
  }
  scan_u64_to_u64(c, g.ccount.as<ull>(), g.coff.as<ull>(), C);
  buf_reserve(c, g.ccl, sizeof(uint32_t) * (core_cliques * k + 1));
  if (T) {
    k_scatter_cliques<<<grid1d((T + CC_E - 1) / CC_E), 256, 0, c.stream>>>(T, k, c.cliques.as<uint32_t>(), c.status.as<int32_t>(), gp,
                                                       gidp, g.cid.as<uint32_t>(), g.local.as<uint32_t>(),
                                                       g.coff.as<ull>(), g.cfill.as<ull>(), g.ccl.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  // host copies of the per-component sizes drive the flow stage
  g.h_voff.resize((size_t)C + 1);
  g.h_coff.resize((size_t)C + 1);
  g.h_label.resize(C);
  HIP_CHECK(hipMemcpyAsync(g.h_voff.data(), g.voff.p, sizeof(uint32_t) * ((size_t)C + 1), hipMemcpyDeviceToHost,
                           c.stream));
  HIP_CHECK(hipMemcpyAsync(g.h_coff.data(), g.coff.p, sizeof(ull) * ((size_t)C + 1), hipMemcpyDeviceToHost,
                           c.stream));
  c.sync();
  // canonical labels: the original id of each component's first-listed root,
  // recovered from the dense numbering (k_uf_label flags roots by original id)
  std::vector<uint32_t> flags(n);
  HIP_CHECK(hipMemcpyAsync(flags.data(), g.flag.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  for (uint32_t v = 0, d = 0; v < n; ++v)
    if (flags[v]) g.h_label[d++] = v;
  c.have_comp = true;
}

}  // namespace kcds
