// k-clique-core peeling.
//
// Reference: cliqueCoreDecompose (main.cu:329-420) drives selectNodes and
// processNodesByWarp (src/helpers.cu:587-726) from the host, one level per
// iteration: level L removes every vertex whose live-clique count is L
// (cascading), destroys its live cliques (status = L), decrements the other
// members (clamped to L) and after the level records density = live cliques /
// remaining vertices; the largest density (ties keep the later core,
// main.cu:407) is maxDensity.
//
// Here the whole peel is one persistent cooperative kernel over the
// materialised cliques and the vertex->clique incidence index:
//   * a new level starts only when the frontier is empty; its value is the
//     minimum live count (empty levels change neither density nor cores, so
//     they are skipped by a min-reduction);
//   * sub-rounds remove the frontier, claim each live clique of a removed
//     vertex exactly once with atomicCAS(status, -1, L) and decrement the
//     surviving members; a member whose count falls from L+1 to L joins the
//     next sub-round (each vertex crosses that value once);
//   * one thread keeps the exact rational maximum of live/remaining.
// Frontier vertices with long incidence lists are processed by the whole
// grid, the rest wave-per-vertex.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t BIG_LIST = 1u << 14;

struct CorePeelState {
  ull level;
  ull minval;
  ull live;        // live cliques
  ull max_num, max_den;
  uint32_t removed;
  uint32_t fsize[2];
  uint32_t nbig;
  uint32_t levels;
  uint32_t barrier;
  uint32_t ntouch;
};

// a/b >= c/d for 64-bit non-negative integers, b, d > 0 (exact, 128-bit products)
__device__ __forceinline__ bool ratio_ge(ull a, ull b, ull c, ull d) {
  ull hi1 = __umul64hi(a, d), lo1 = a * d;
  ull hi2 = __umul64hi(c, b), lo2 = c * b;
  return hi1 > hi2 || (hi1 == hi2 && lo1 >= lo2);
}

__device__ __forceinline__ void claim_clique(uint32_t c, uint32_t v, uint32_t k, ull L, const uint32_t *cl,
                                             int32_t *status, const int32_t *core, ull *cnt, uint32_t *nx,
                                             uint32_t *nx_size, ull &destroyed, GraphPeel *gp,
                                             const uint32_t *gid) {
  if (__hip_atomic_load(&status[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != -1) return;
  if (atomicCAS(&status[c], -1, (int32_t)L) != -1) return;
  destroyed += 1;
  if (gp) atomicSub(&gp[gid[v]].live, 1ull);
  for (uint32_t t = 0; t < k; ++t) {
    const uint32_t u = cl[(ull)c * k + t];
    if (u == v) continue;
    if (__hip_atomic_load(&core[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != -1) continue;  // removed
    const ull old = atomicSub(&cnt[u], 1ull);
    if (old == L + 1) nx[atomicAdd(nx_size, 1u)] = u;
  }
}

// gp == nullptr: one graph, bookkeeping on the global live count.  Otherwise
// the vertices form several graphs (gid[r]); each graph keeps its own live
// cliques / remaining vertices and, at the end of every level that touched
// it, its own maximum density -- exactly what that graph's own peel records,
// because a graph's vertices leave at their core numbers whatever else is
// peeled alongside.
__global__ __launch_bounds__(1024) void k_clique_peel(uint32_t n, uint32_t k, const uint32_t *cl,
                                                      const ull *incoff, const uint32_t *inc, ull *cnt,
                                                      int32_t *core, int32_t *status, uint32_t *fr0,
                                                      uint32_t *fr1, uint32_t *big, CorePeelState *st,
                                                      GraphPeel *gp, const uint32_t *gid, uint32_t *tlist) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / KCDS_WAVE, nwaves = nthreads / KCDS_WAVE;
  uint32_t epoch = 0;
  int cur = 0;
  while (true) {
    const uint32_t removed = __hip_atomic_load(&st->removed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (removed >= n) break;
    uint32_t fs = __hip_atomic_load(&st->fsize[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fs == 0) {
      // previous level complete: record its density, then open the next one
      if (!gp) {
        if (tid == 0) {
          const ull live = st->live, rem = (ull)(n - removed);
          if (rem > 0 && ratio_ge(live, rem, st->max_num, st->max_den)) {
            st->max_num = live;
            st->max_den = rem;
          }
        }
      } else {
        const uint32_t nt = __hip_atomic_load(&st->ntouch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (uint32_t i = tid; i < nt; i += nthreads) {
          GraphPeel &G = gp[tlist[i]];
          if (G.rem > 0 && ratio_ge(G.live, G.rem, G.max_num, G.max_den)) {
            G.max_num = G.live;
            G.max_den = G.rem;
          }
          G.touch = 0;
        }
      }
      ull m = ~0ull;
      for (uint32_t v = tid; v < n; v += nthreads)
        if (core[v] == -1) m = min(m, cnt[v]);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = min(m, (ull)__shfl_xor(m, o, 64));
      if (lane == 0 && m != ~0ull) atomicMin(&st->minval, m);
      grid_barrier(&st->barrier, gridDim.x, epoch);
      const ull L = max(__hip_atomic_load(&st->minval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(&st->level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      uint32_t *fr = cur ? fr1 : fr0;
      for (uint32_t v = tid; v < n; v += nthreads)
        if (core[v] == -1 && cnt[v] <= L) fr[atomicAdd(&st->fsize[cur], 1u)] = v;
      grid_barrier(&st->barrier, gridDim.x, epoch);
      if (tid == 0) {
        st->level = L;
        st->minval = ~0ull;
        st->levels += 1;
        st->ntouch = 0;  // every thread read it before the first barrier of this branch
      }
      fs = __hip_atomic_load(&st->fsize[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      grid_barrier(&st->barrier, gridDim.x, epoch);
    }
    const ull L = __hip_atomic_load(&st->level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t *fr = cur ? fr1 : fr0;
    uint32_t *nx = cur ? fr0 : fr1;
    uint32_t *nx_size = &st->fsize[cur ^ 1];
    // remove the frontier; long incidence lists go to the grid-wide list
    for (uint32_t i = tid; i < fs; i += nthreads) {
      const uint32_t v = fr[i];
      core[v] = (int32_t)L;
      if (incoff[v + 1] - incoff[v] >= BIG_LIST) big[atomicAdd(&st->nbig, 1u)] = v;
      if (gp) {
        const uint32_t g = gid[v];
        atomicSub(&gp[g].rem, 1u);
        if (atomicExch(&gp[g].touch, 1u) == 0u) tlist[atomicAdd(&st->ntouch, 1u)] = g;
      }
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    ull destroyed = 0;
    for (uint32_t i = wid; i < fs; i += nwaves) {
      const uint32_t v = fr[i];
      const ull b = incoff[v], e = incoff[v + 1];
      if (e - b >= BIG_LIST) continue;
      for (ull j = b + lane; j < e; j += KCDS_WAVE)
        claim_clique(inc[j], v, k, L, cl, status, core, cnt, nx, nx_size, destroyed, gp, gid);
    }
    const uint32_t nbig = __hip_atomic_load(&st->nbig, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t bi = 0; bi < nbig; ++bi) {
      const uint32_t v = big[bi];
      const ull b = incoff[v], e = incoff[v + 1];
      for (ull j = b + tid; j < e; j += nthreads)
        claim_clique(inc[j], v, k, L, cl, status, core, cnt, nx, nx_size, destroyed, gp, gid);
    }
    destroyed = wave_sum_u64(destroyed);
    if (lane == 0 && destroyed) atomicSub(&st->live, destroyed);
    grid_barrier(&st->barrier, gridDim.x, epoch);
    if (tid == 0) {
      st->removed = removed + fs;
      st->fsize[cur] = 0;
      st->nbig = 0;
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    cur ^= 1;
  }
}

__global__ void k_peel_init(uint32_t n, uint64_t T, const ull *cnt0, ull *cnt, int32_t *core, int32_t *status) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    cnt[i] = cnt0[i];
    core[i] = -1;
  }
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += stride) status[i] = -1;
}

__global__ void k_gid(uint32_t n, const uint32_t *inv, const uint32_t *gvoff, uint32_t G, uint32_t *gid) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    gid[r] = upper_bound_u32(gvoff, G + 1, inv[r]) - 1;
}

__global__ void k_graph_accum(uint32_t n, const uint32_t *gid, const ull *cnt, GraphPeel *gp) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    GraphPeel &G = gp[gid[r]];
    if (cnt[r]) atomicAdd(&G.live, cnt[r]);
    atomicAdd(&G.n, 1u);
  }
}

__global__ void k_graph_init(uint32_t G, uint32_t k, GraphPeel *gp) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    GraphPeel &P = gp[g];
    P.live /= k;  // sum of per-vertex counts = k * T
    P.T = P.live;
    P.rem = P.n;
    P.max_num = P.T;  // main.cu:361-366: the whole graph's density first
    P.max_den = P.n ? P.n : 1;
    P.touch = 0;
  }
}

__global__ void k_graph_single(const CorePeelState *st, GraphPeel *gp) {
  gp[0].max_num = st->max_num;
  gp[0].max_den = st->max_den;
}

// k' = ceil(maxDensity) per graph (main.cu:1476)
__global__ void k_graph_kprime(uint32_t G, GraphPeel *gp) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    GraphPeel &P = gp[g];
    P.k_prime = P.T == 0 ? 0x7FFFFFFFu : (uint32_t)((P.max_num + P.max_den - 1) / P.max_den);
    P.core_size = 0;
    P.core_cliques = 0;
    P.max_core = 0;
  }
}

// k'-core size, cliques inside it and the largest core, per graph
__global__ void k_graph_summary(uint32_t n, uint64_t T, uint32_t k, const uint32_t *gid, const int32_t *core,
                                const int32_t *status, const uint32_t *cl, GraphPeel *gp) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    GraphPeel &P = gp[gid[i]];
    const int32_t cv = core[i];
    if (cv >= (int32_t)P.k_prime) atomicAdd(&P.core_size, 1u);
    if (cv > 0) atomicMax(&P.max_core, (uint32_t)cv);
  }
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T; c += stride) {
    GraphPeel &P = gp[gid[cl[c * k]]];
    if (status[c] >= (int32_t)P.k_prime) atomicAdd(&P.core_cliques, 1ull);
  }
}

}  // namespace

void graph_ids(Ctx &c) {
  const uint32_t n = c.n, G = c.ngraphs;
  buf_reserve(c, c.gid, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.gvoff, sizeof(uint32_t) * ((size_t)G + 1));
  HIP_CHECK(hipMemcpyAsync(c.gvoff.p, c.h_gvoff.data(), sizeof(uint32_t) * ((size_t)G + 1), hipMemcpyHostToDevice,
                           c.stream));
  if (n) {
    k_gid<<<(n + 255) / 256, 256, 0, c.stream>>>(n, c.inv.as<uint32_t>(), c.gvoff.as<uint32_t>(), G,
                                                 c.gid.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.sync();  // h_gvoff may change after return
}

void peel_run(Ctx &c) {
  const uint32_t n = c.n, k = c.k, G = c.ngraphs;
  const uint64_t T = c.T;
  const bool multi = G > 1;
  // core numbers are int32 (as the reference's cliqueCore); they never exceed
  // the largest per-vertex count
  if (T > 0x7FFFFFFFull && max_u64_host(c, c.cnt.as<ull>(), n) > 0x7FFFFFFFull)
    fail(KCDS_E_LIMIT, "a per-vertex clique count exceeds the int32 clique-core range");
  graph_ids(c);
  buf_reserve(c, c.core, sizeof(int32_t) * ((size_t)n + 1));
  buf_reserve(c, c.status, sizeof(int32_t) * (T + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.tlist, sizeof(uint32_t) * ((size_t)G + 1));
  GraphPeel *gp = c.gpeel.as<GraphPeel>();
  HIP_CHECK(hipMemsetAsync(gp, 0, sizeof(GraphPeel) * G, c.stream));
  // working counts, two frontiers, the big-vertex list and the state
  size_t words = (size_t)n * 2 /*cnt ull*/ + (size_t)n * 3 + 64;
  buf_reserve(c, c.tmp0, words * sizeof(uint32_t) + sizeof(CorePeelState) + 256);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fr0 = (uint32_t *)(cnt + n), *fr1 = fr0 + n, *big = fr1 + n;
  CorePeelState *st = (CorePeelState *)(((uintptr_t)(big + n) + 255) & ~(uintptr_t)255);
  CorePeelState h0{};
  h0.level = 0;
  h0.minval = ~0ull;
  h0.live = T;
  h0.max_num = T;
  h0.max_den = n ? n : 1;  // main.cu:361-366: the whole graph's density T/n
  HIP_CHECK(hipMemcpyAsync(st, &h0, sizeof h0, hipMemcpyHostToDevice, c.stream));
  const uint64_t mx = std::max<uint64_t>(n, T);
  uint32_t grid = (uint32_t)std::min<uint64_t>((mx + 255) / 256, 65535ull * 16);
  grid = std::max<uint32_t>(grid, 1);
  k_peel_init<<<grid, 256, 0, c.stream>>>(n, T, c.cnt.as<ull>(), cnt, c.core.as<int32_t>(),
                                          c.status.as<int32_t>());
  KCDS_LAUNCH_CHECK();
  const uint32_t *gidp = c.gid.as<uint32_t>();
  if (n) {
    k_graph_accum<<<(n + 255) / 256, 256, 0, c.stream>>>(n, gidp, c.cnt.as<ull>(), gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_init<<<(G + 255) / 256, 256, 0, c.stream>>>(G, k, gp);
  KCDS_LAUNCH_CHECK();
  if (n) {
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_clique_peel, 1024, 0));
    if (per_cu < 1) fail(KCDS_E_HIP, "k_clique_peel cannot be resident");
    uint32_t pgrid = (uint32_t)c.num_cus * (uint32_t)std::min(per_cu, 2);
    const uint32_t *cl = c.cliques.as<uint32_t>(), *inc = c.inc.as<uint32_t>();
    const ull *incoff = c.incoff.as<ull>();
    int32_t *core = c.core.as<int32_t>(), *status = c.status.as<int32_t>();
    uint32_t nn = n, kk = k;
    GraphPeel *gpa = multi ? gp : nullptr;
    uint32_t *tl = c.tlist.as<uint32_t>();
    void *args[] = {(void *)&nn, (void *)&kk, (void *)&cl, (void *)&incoff, (void *)&inc, (void *)&cnt,
                    (void *)&core, (void *)&status, (void *)&fr0, (void *)&fr1, (void *)&big, (void *)&st,
                    (void *)&gpa, (void *)&gidp, (void *)&tl};
    HIP_CHECK(hipLaunchCooperativeKernel((const void *)k_clique_peel, dim3(pgrid), dim3(1024), args, 0, c.stream));
  }
  if (!multi) {
    k_graph_single<<<1, 1, 0, c.stream>>>(st, gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_kprime<<<(G + 255) / 256, 256, 0, c.stream>>>(G, gp);
  KCDS_LAUNCH_CHECK();
  k_graph_summary<<<grid, 256, 0, c.stream>>>(n, T, k, gidp, c.core.as<int32_t>(), c.status.as<int32_t>(),
                                              c.cliques.as<uint32_t>(), gp);
  KCDS_LAUNCH_CHECK();
  CorePeelState *hs = (CorePeelState *)c.host_scratch(sizeof(CorePeelState) + 64);
  HIP_CHECK(hipMemcpyAsync(hs, st, sizeof(CorePeelState), hipMemcpyDeviceToHost, c.stream));
  c.h_gpeel.resize(G);
  HIP_CHECK(hipMemcpyAsync(c.h_gpeel.data(), gp, sizeof(GraphPeel) * G, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  const GraphPeel &P = c.h_gpeel[0];
  kcds_peel_stats ps{};
  ull g = P.max_num, b = P.max_den;
  while (b) {
    ull t = g % b;
    g = b;
    b = t;
  }
  if (!g) g = 1;
  ps.max_num = P.max_num / g;
  ps.max_den = P.max_den / g;
  ps.k_prime = P.k_prime;
  ps.levels = hs->levels;
  ps.core_size = P.core_size;
  ps.core_cliques = P.core_cliques;
  ps.max_core = (int32_t)P.max_core;
  c.peel = ps;
  c.peeled = true;
  c.have_comp = false;
}

}  // namespace kcds
