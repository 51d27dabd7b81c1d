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
// Here one peeling sub-round is a fixed sequence of step kernels whose work
// is decided by device-side control state, so the host enqueues many
// sub-rounds back to back and only reads the "done" flag now and then:
//   touch / min / open / collect   (no-ops unless the frontier is empty)
//     a new level starts only when the frontier is empty; its value is the
//     minimum live count (empty levels change neither density nor cores, so
//     they are skipped by a min-reduction); the density of the level just
//     completed is recorded first, as an exact rational;
//   mark      frontier vertices take the level as their core number;
//   process   each live clique of a removed vertex is claimed exactly once
//             (atomicCAS(status, -1, L)) and its surviving members are
//             decremented; a member whose count falls from L+1 to L joins the
//             next sub-round (each vertex crosses that value once);
//   finish    advance to the next frontier.
// Kernel boundaries order every step, so no step needs a grid-wide barrier.
// Frontier vertices with long incidence lists are processed grid-wide, the
// rest wave-per-vertex.  With several graphs in one batch, every graph keeps
// its own live cliques / remaining vertices and records its own densities at
// the end of each level that touched it -- exactly what its own peel records,
// because a graph's vertices leave at their core numbers whatever else is
// peeled alongside.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t BIG_LIST = 1u << 14;

struct PeelCtl {
  ull level;      // current level
  ull minval;     // min live count over remaining vertices (reduction target)
  ull live;       // live cliques (one-graph bookkeeping)
  ull max_num, max_den;
  ull destroyed;  // cliques destroyed in the current sub-round
  uint32_t removed;
  uint32_t fsize[2];
  uint32_t cur;
  uint32_t nbig;
  uint32_t levels;
  uint32_t ntouch;
  uint32_t done;
  uint32_t opening;  // a level was opened in this sub-round: collect its frontier
};

// a/b >= c/d for 64-bit non-negative integers, b, d > 0 (exact, 128-bit products)
__device__ __forceinline__ bool ratio_ge(ull a, ull b, ull c, ull d) {
  ull hi1 = __umul64hi(a, d), lo1 = a * d;
  ull hi2 = __umul64hi(c, b), lo2 = c * b;
  return hi1 > hi2 || (hi1 == hi2 && lo1 >= lo2);
}

__device__ __forceinline__ bool level_end(const PeelCtl *ctl) {
  return !ctl->done && ctl->fsize[ctl->cur] == 0;
}

// multi-graph: record the density of every graph touched by the level that
// just completed
__global__ void k_pl_touch(PeelCtl *ctl, GraphPeel *gp, uint32_t *tlist) {
  if (!level_end(ctl)) return;
  const uint32_t nt = ctl->ntouch;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += gridDim.x * blockDim.x) {
    GraphPeel &G = gp[tlist[i]];
    const ull rem = G.rem;
    if (rem > 0 && ratio_ge(G.live, rem, G.max_num, G.max_den)) {
      G.max_num = G.live;
      G.max_den = rem;
    }
    G.touch = 0;
  }
}

__global__ void k_pl_min(PeelCtl *ctl, uint32_t n, const ull *cnt, const int32_t *core) {
  if (!level_end(ctl)) return;
  ull m = ~0ull;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (core[v] == -1) m = min(m, cnt[v]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = min(m, (ull)__shfl_xor(m, o, 64));
  if (lane_id() == 0 && m != ~0ull) atomicMin(&ctl->minval, m);
}

// one thread: close the completed level, open the next one
__global__ void k_pl_open(PeelCtl *ctl, uint32_t n, uint32_t multi) {
  if (!level_end(ctl)) return;
  if (!multi) {
    const ull rem = (ull)(n - ctl->removed);
    if (rem > 0 && ratio_ge(ctl->live, rem, ctl->max_num, ctl->max_den)) {
      ctl->max_num = ctl->live;
      ctl->max_den = rem;
    }
  }
  ctl->ntouch = 0;
  if (ctl->minval == ~0ull) {  // nothing left (cannot happen while removed < n)
    ctl->done = 1;
    return;
  }
  ctl->level = max(ctl->level, ctl->minval);
  ctl->minval = ~0ull;
  ctl->levels += 1;
  ctl->opening = 1;
}

__global__ void k_pl_collect(PeelCtl *ctl, uint32_t n, const ull *cnt, const int32_t *core, uint32_t *fr0,
                             uint32_t *fr1) {
  if (ctl->done || !ctl->opening) return;
  const ull L = ctl->level;
  const uint32_t cur = ctl->cur;
  uint32_t *fr = cur ? fr1 : fr0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (core[v] == -1 && cnt[v] <= L) fr[atomicAdd(&ctl->fsize[cur], 1u)] = v;
}

__global__ void k_pl_mark(PeelCtl *ctl, const uint32_t *fr0, const uint32_t *fr1, const ull *incoff, int32_t *core,
                          uint32_t *big, GraphPeel *gp, const uint32_t *gid, uint32_t *tlist) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur, fs = ctl->fsize[cur];
  const uint32_t *fr = cur ? fr1 : fr0;
  const int32_t L = (int32_t)ctl->level;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < fs; i += gridDim.x * blockDim.x) {
    const uint32_t v = fr[i];
    core[v] = L;
    if (incoff[v + 1] - incoff[v] >= BIG_LIST) big[atomicAdd(&ctl->nbig, 1u)] = v;
    if (gp) {
      const uint32_t g = gid[v];
      atomicSub(&gp[g].rem, 1u);
      if (atomicExch(&gp[g].touch, 1u) == 0u) tlist[atomicAdd(&ctl->ntouch, 1u)] = g;
    }
  }
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
    if (core[u] != -1) continue;  // removed (core numbers were set by the mark step)
    const ull old = atomicSub(&cnt[u], 1ull);
    if (old == L + 1) nx[atomicAdd(nx_size, 1u)] = u;
  }
}

// wave per frontier vertex with a short incidence list
__global__ void k_pl_proc(PeelCtl *ctl, uint32_t k, const uint32_t *cl, const ull *incoff, const uint32_t *inc,
                          ull *cnt, const int32_t *core, int32_t *status, uint32_t *fr0, uint32_t *fr1,
                          GraphPeel *gp, const uint32_t *gid) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur, fs = ctl->fsize[cur];
  const uint32_t *fr = cur ? fr1 : fr0;
  uint32_t *nx = cur ? fr0 : fr1;
  const ull L = ctl->level;
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / KCDS_WAVE;
  const uint32_t nw = gridDim.x * blockDim.x / KCDS_WAVE;
  ull destroyed = 0;
  for (uint32_t i = wid; i < fs; i += nw) {
    const uint32_t v = fr[i];
    const ull b = incoff[v], e = incoff[v + 1];
    if (e - b >= BIG_LIST) continue;
    for (ull j = b + lane; j < e; j += KCDS_WAVE)
      claim_clique(inc[j], v, k, L, cl, status, core, cnt, nx, &ctl->fsize[cur ^ 1], destroyed, gp, gid);
  }
  destroyed = wave_sum_u64(destroyed);
  if (lane == 0 && destroyed) atomicAdd(&ctl->destroyed, destroyed);
}

// the whole grid over each frontier vertex with a long incidence list
__global__ void k_pl_proc_big(PeelCtl *ctl, uint32_t k, const uint32_t *cl, const ull *incoff, const uint32_t *inc,
                              ull *cnt, const int32_t *core, int32_t *status, uint32_t *fr0, uint32_t *fr1,
                              const uint32_t *big, GraphPeel *gp, const uint32_t *gid) {
  if (ctl->done) return;
  const uint32_t nbig = ctl->nbig;
  if (nbig == 0) return;
  const uint32_t cur = ctl->cur;
  uint32_t *nx = cur ? fr0 : fr1;
  const ull L = ctl->level;
  const ull tid = (ull)blockIdx.x * blockDim.x + threadIdx.x, nt = (ull)gridDim.x * blockDim.x;
  ull destroyed = 0;
  for (uint32_t bi = 0; bi < nbig; ++bi) {
    const uint32_t v = big[bi];
    for (ull j = incoff[v] + tid; j < incoff[v + 1]; j += nt)
      claim_clique(inc[j], v, k, L, cl, status, core, cnt, nx, &ctl->fsize[cur ^ 1], destroyed, gp, gid);
  }
  destroyed = wave_sum_u64(destroyed);
  if (lane_id() == 0 && destroyed) atomicAdd(&ctl->destroyed, destroyed);
}

// one thread: advance to the next frontier
__global__ void k_pl_finish(PeelCtl *ctl, uint32_t n) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur;
  ctl->removed += ctl->fsize[cur];
  ctl->live -= ctl->destroyed;
  ctl->destroyed = 0;
  ctl->fsize[cur] = 0;
  ctl->nbig = 0;
  ctl->opening = 0;
  ctl->cur = cur ^ 1;
  if (ctl->removed >= n) ctl->done = 1;
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

__global__ void k_graph_single(const PeelCtl *ctl, GraphPeel *gp) {
  gp[0].max_num = ctl->max_num;
  gp[0].max_den = ctl->max_den;
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

// ---------------------------------------------------------------------------
// Workgroup-per-graph peel.  For graphs small enough for one workgroup, a
// whole peel -- every level and sub-round, the density records and the k'
// summary -- runs inside one 1024-thread workgroup with workgroup barriers
// only; the workgroups of a launch take graphs from a sharded queue.  The
// semantics are those of the step kernels above (levels skip to the minimum
// live count, a level cascades in sub-rounds, each live clique of a removed
// vertex is claimed once and its surviving members are decremented, the
// density live/remaining is recorded after each level with ties kept).
// Graph g owns original ids [gvoff[g], gvoff[g+1]); its vertex v has new id
// rank[v]; its frontiers live in fr0/fr1 at the same original-id range.
constexpr uint32_t BP_THREADS = 1024;

__device__ __forceinline__ ull block_min_u64(ull v, ull *slot) {
  if (threadIdx.x == 0) *slot = ~0ull;
  block_sync();
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (ull)__shfl_xor(v, o, 64));
  if (lane_id() == 0) atomicMin(slot, v);
  block_sync();
  const ull r = *slot;
  block_sync();
  return r;
}

__device__ __forceinline__ ull block_sum_u64(ull v, ull *slot) {
  if (threadIdx.x == 0) *slot = 0;
  block_sync();
  v = wave_sum_u64(v);
  if (lane_id() == 0 && v) atomicAdd(slot, v);
  block_sync();
  const ull r = *slot;
  block_sync();
  return r;
}

__global__ __launch_bounds__(BP_THREADS) void k_peel_block(uint32_t G, const uint32_t *__restrict__ gvoff,
                                                          const uint32_t *__restrict__ rank, uint32_t k,
                                                          const ull *__restrict__ cnt0, ull *cnt, int32_t *core,
                                                          int32_t *status, const uint32_t *__restrict__ cl,
                                                          const ull *__restrict__ incoff,
                                                          const uint32_t *__restrict__ inc, uint32_t *fr0,
                                                          uint32_t *fr1, uint32_t *queue, GraphPeel *gp) {
  __shared__ ull wsum[16];
  __shared__ ull s_red;
  __shared__ ull s_pre[BP_THREADS + 1];
  __shared__ uint32_t s_r[BP_THREADS];
  __shared__ ull s_b[BP_THREADS];
  __shared__ uint32_t s_g, s_fsize[2];
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  QueueCursor qc = queue_start(blockIdx.x);
  while (true) {
    if (tid == 0) s_g = queue_next(queue, G, qc);
    block_sync();
    const uint32_t g = s_g;
    block_sync();
    if (g >= G) break;
    const uint32_t vb = gvoff[g], ng = gvoff[g + 1] - vb;
    uint32_t *F0 = fr0 + vb, *F1 = fr1 + vb;
    ull mine = 0;
    for (uint32_t i = tid; i < ng; i += nt) {
      const uint32_t r = rank[vb + i];
      const ull c0 = cnt0[r];
      stc_u64(&cnt[r], c0);
      stc_i32(&core[r], -1);
      mine += c0;
    }
    const ull T = block_sum_u64(mine, &s_red) / k;  // sum of per-vertex counts = k T
    ull live = T, max_num = T, max_den = ng ? ng : 1, L = 0;
    uint32_t rem = ng, levels = 0;
    while (rem > 0) {
      ull m = ~0ull;
      for (uint32_t i = tid; i < ng; i += nt) {
        const uint32_t r = rank[vb + i];
        if (ldc_i32(&core[r]) == -1) m = min(m, ldc_u64(&cnt[r]));
      }
      m = block_min_u64(m, &s_red);
      if (m == ~0ull) break;
      L = max(L, m);
      ++levels;
      if (tid == 0) s_fsize[0] = 0;
      block_sync();
      for (uint32_t i0 = 0; i0 < ng; i0 += nt) {  // the level's frontier
        const uint32_t i = i0 + tid;
        bool want = false;
        uint32_t r = 0;
        if (i < ng) {
          r = rank[vb + i];
          want = ldc_i32(&core[r]) == -1 && ldc_u64(&cnt[r]) <= L;
        }
        const ull bal = __ballot(want);
        uint32_t at = 0;
        if (bal && lane_id() == (uint32_t)__builtin_ctzll(bal)) at = atomicAdd(&s_fsize[0], (uint32_t)__popcll(bal));
        at = __shfl(at, bal ? __builtin_ctzll(bal) : 0, 64);
        if (want) stc_u32(&F0[at + lanes_below(bal)], r);
      }
      block_sync();
      uint32_t cur = 0;
      while (true) {  // sub-rounds of level L
        const uint32_t fs = s_fsize[cur];
        if (fs == 0) break;
        uint32_t *Fc = cur ? F1 : F0, *Fn = cur ? F0 : F1;
        for (uint32_t i = tid; i < fs; i += nt) stc_i32(&core[ldc_u32(&Fc[i])], (int32_t)L);
        rem -= fs;
        if (tid == 0) s_fsize[cur ^ 1] = 0;
        block_sync();
        ull destroyed = 0;
        for (uint32_t f0 = 0; f0 < fs; f0 += nt) {
          ull len = 0;
          if (f0 + tid < fs) {
            const uint32_t v = ldc_u32(&Fc[f0 + tid]);
            s_r[tid] = v;
            s_b[tid] = incoff[v];
            len = incoff[v + 1] - incoff[v];
          }
          ull total = 0;
          const ull ex = block_excl_scan_u64(len, &total, wsum);
          const uint32_t nv = min(nt, fs - f0);
          s_pre[tid] = ex;
          if (tid == 0) s_pre[nv] = total;
          block_sync();
          for (ull j = tid; j < total; j += nt) {
            uint32_t lo = 0, hi = nv;  // owner t: s_pre[t] <= j < s_pre[t+1]
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (s_pre[mid] <= j) lo = mid;
              else hi = mid;
            }
            const uint32_t v = s_r[lo];
            const uint32_t c = inc[s_b[lo] + (j - s_pre[lo])];
            if (ldc_i32(&status[c]) != -1) continue;
            if (atomicCAS(&status[c], -1, (int32_t)L) != -1) continue;
            ++destroyed;
            for (uint32_t t = 0; t < k; ++t) {
              const uint32_t u = cl[(ull)c * k + t];
              if (u == v || ldc_i32(&core[u]) != -1) continue;
              if (atomicSub(&cnt[u], 1ull) == L + 1) stc_u32(&Fn[atomicAdd(&s_fsize[cur ^ 1], 1u)], u);
            }
          }
          block_sync();  // s_pre, s_r, s_b are reused by the next chunk
        }
        live -= block_sum_u64(destroyed, &s_red);
        cur ^= 1;
      }
      // density of the level just completed (ties keep the later core)
      if (rem > 0 && ratio_ge(live, rem, max_num, max_den)) {
        max_num = live;
        max_den = rem;
      }
    }
    // k' = ceil(maxDensity) and the k'-core summary
    const uint32_t kp = T == 0 ? 0x7FFFFFFFu : (uint32_t)((max_num + max_den - 1) / max_den);
    ull csize = 0, ccl = 0, mc = 0;
    for (uint32_t i = tid; i < ng; i += nt) {
      const uint32_t r = rank[vb + i];
      const int32_t cv = ldc_i32(&core[r]);
      if (cv >= (int32_t)kp) csize += 1;
      if (cv > 0) mc = max(mc, (ull)cv);
      for (ull j = incoff[r]; j < incoff[r + 1]; ++j) ccl += ldc_i32(&status[inc[j]]) >= (int32_t)kp;
    }
    csize = block_sum_u64(csize, &s_red);
    ccl = block_sum_u64(ccl, &s_red) / k;  // every clique is seen from its k members
    mc = ~block_min_u64(~mc, &s_red);
    if (tid == 0) {
      GraphPeel P;
      P.live = live;
      P.max_num = max_num;
      P.max_den = max_den;
      P.T = T;
      P.core_cliques = ccl;
      P.rem = rem;
      P.touch = 0;
      P.n = ng;
      P.k_prime = kp;
      P.core_size = (uint32_t)csize;
      P.max_core = (uint32_t)mc;
      P.levels = levels;
      P.pad = 0;
      gp[g] = P;
    }
  }
}

uint32_t grid1d(uint64_t len, uint32_t cap = 65535u * 16) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, cap));
}

}  // namespace

void graph_ids(Ctx &c) {
  const uint32_t n = c.n, G = c.ngraphs;
  buf_reserve(c, c.gid, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.gvoff, sizeof(uint32_t) * ((size_t)G + 1));
  HIP_CHECK(hipMemcpyAsync(c.gvoff.p, c.h_gvoff.data(), sizeof(uint32_t) * ((size_t)G + 1), hipMemcpyHostToDevice,
                           c.stream));
  if (n) {
    k_gid<<<grid1d(n), 256, 0, c.stream>>>(n, c.inv.as<uint32_t>(), c.gvoff.as<uint32_t>(), G, c.gid.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.sync();  // h_gvoff may change after return
}

// Graphs up to this size are peeled by one workgroup each.
constexpr uint32_t BLOCK_PEEL_MAX_N = 1u << 15;

void peel_blocks(Ctx &c) {
  const uint32_t n = c.n, k = c.k, G = c.ngraphs;
  const uint64_t T = c.T;
  buf_reserve(c, c.core, sizeof(int32_t) * ((size_t)n + 1));
  buf_reserve(c, c.status, sizeof(int32_t) * (T + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.gvoff, sizeof(uint32_t) * ((size_t)G + 1));
  HIP_CHECK(hipMemcpyAsync(c.gvoff.p, c.h_gvoff.data(), sizeof(uint32_t) * ((size_t)G + 1), hipMemcpyHostToDevice,
                           c.stream));
  const size_t words = (size_t)n * 2 /*cnt*/ + (size_t)n * 2 /*frontiers*/ + QUEUE_WORDS + 64;
  buf_reserve(c, c.tmp0, words * sizeof(uint32_t) + 256);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fr0 = (uint32_t *)(cnt + n), *fr1 = fr0 + n, *queue = fr1 + n;
  HIP_CHECK(hipMemsetAsync(queue, 0, QUEUE_WORDS * sizeof(uint32_t), c.stream));
  if (T) {
    k_peel_init<<<grid1d(T), 256, 0, c.stream>>>(0, T, nullptr, nullptr, nullptr, c.status.as<int32_t>());
    KCDS_LAUNCH_CHECK();
  }
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(G, (uint32_t)c.num_cus * 2));
  k_peel_block<<<grid, BP_THREADS, 0, c.stream>>>(G, c.gvoff.as<uint32_t>(), c.rank.as<uint32_t>(), k,
                                                  c.cnt.as<ull>(), cnt, c.core.as<int32_t>(), c.status.as<int32_t>(),
                                                  c.cliques.as<uint32_t>(), c.incoff.as<ull>(), c.inc.as<uint32_t>(),
                                                  fr0, fr1, queue, c.gpeel.as<GraphPeel>());
  KCDS_LAUNCH_CHECK();
  c.h_gpeel.resize(G);
  HIP_CHECK(hipMemcpyAsync(c.h_gpeel.data(), c.gpeel.p, sizeof(GraphPeel) * G, hipMemcpyDeviceToHost, c.stream));
  c.sync();
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
  uint32_t largest = 0;
  for (uint32_t g = 0; g < G; ++g) largest = std::max(largest, c.h_gvoff[g + 1] - c.h_gvoff[g]);
  if (largest <= BLOCK_PEEL_MAX_N && !(c.tune_flags & KCDS_TUNE_GRID_PEEL)) {
    peel_blocks(c);
    const GraphPeel &P = c.h_gpeel[0];
    kcds_peel_stats ps{};
    ull g0 = P.max_num, b0 = P.max_den;
    while (b0) {
      ull t = g0 % b0;
      g0 = b0;
      b0 = t;
    }
    if (!g0) g0 = 1;
    ps.max_num = P.max_num / g0;
    ps.max_den = P.max_den / g0;
    ps.k_prime = P.k_prime;
    ps.levels = P.levels;
    ps.core_size = P.core_size;
    ps.core_cliques = P.core_cliques;
    ps.max_core = (int32_t)P.max_core;
    c.peel = ps;
    c.peeled = true;
    c.have_comp = false;
    return;
  }
  buf_reserve(c, c.core, sizeof(int32_t) * ((size_t)n + 1));
  buf_reserve(c, c.status, sizeof(int32_t) * (T + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.tlist, sizeof(uint32_t) * ((size_t)G + 1));
  GraphPeel *gp = c.gpeel.as<GraphPeel>();
  HIP_CHECK(hipMemsetAsync(gp, 0, sizeof(GraphPeel) * G, c.stream));
  // working counts, two frontiers, the big-vertex list and the control block
  const size_t words = (size_t)n * 2 /*cnt ull*/ + (size_t)n * 3 + 64;
  buf_reserve(c, c.tmp0, words * sizeof(uint32_t) + sizeof(PeelCtl) + 256);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fr0 = (uint32_t *)(cnt + n), *fr1 = fr0 + n, *big = fr1 + n;
  PeelCtl *ctl = (PeelCtl *)(((uintptr_t)(big + n) + 255) & ~(uintptr_t)255);
  PeelCtl h0{};
  h0.minval = ~0ull;
  h0.live = T;
  h0.max_num = T;
  h0.max_den = n ? n : 1;  // main.cu:361-366: the whole graph's density T/n
  h0.done = n == 0;
  HIP_CHECK(hipMemcpyAsync(ctl, &h0, sizeof h0, hipMemcpyHostToDevice, c.stream));
  const uint32_t grid = grid1d(std::max<uint64_t>(n, T));
  k_peel_init<<<grid, 256, 0, c.stream>>>(n, T, c.cnt.as<ull>(), cnt, c.core.as<int32_t>(), c.status.as<int32_t>());
  KCDS_LAUNCH_CHECK();
  const uint32_t *gidp = c.gid.as<uint32_t>();
  if (n) {
    k_graph_accum<<<grid1d(n), 256, 0, c.stream>>>(n, gidp, c.cnt.as<ull>(), gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_init<<<grid1d(G), 256, 0, c.stream>>>(G, k, gp);
  KCDS_LAUNCH_CHECK();
  const uint32_t *cl = c.cliques.as<uint32_t>(), *inc = c.inc.as<uint32_t>();
  const ull *incoff = c.incoff.as<ull>();
  int32_t *core = c.core.as<int32_t>(), *status = c.status.as<int32_t>();
  GraphPeel *gpa = multi ? gp : nullptr;
  uint32_t *tl = c.tlist.as<uint32_t>();
  const uint32_t vgrid = grid1d(n, (uint32_t)c.num_cus * 8);
  const uint32_t pgrid = (uint32_t)c.num_cus * 8;  // 256-thread blocks, 4 waves each
  uint32_t *hdone = (uint32_t *)c.host_scratch(64);
  // enqueue sub-rounds in growing batches; every step is a no-op once done
  uint32_t batch = 4;
  uint64_t rounds = 0;
  while (n) {
    for (uint32_t r = 0; r < batch; ++r) {
      if (multi) k_pl_touch<<<grid1d(G, 1024), 256, 0, c.stream>>>(ctl, gp, tl);
      k_pl_min<<<vgrid, 256, 0, c.stream>>>(ctl, n, cnt, core);
      k_pl_open<<<1, 1, 0, c.stream>>>(ctl, n, multi ? 1u : 0u);
      k_pl_collect<<<vgrid, 256, 0, c.stream>>>(ctl, n, cnt, core, fr0, fr1);
      k_pl_mark<<<vgrid, 256, 0, c.stream>>>(ctl, fr0, fr1, incoff, core, big, gpa, gidp, tl);
      k_pl_proc<<<pgrid, 256, 0, c.stream>>>(ctl, k, cl, incoff, inc, cnt, core, status, fr0, fr1, gpa, gidp);
      k_pl_proc_big<<<pgrid, 256, 0, c.stream>>>(ctl, k, cl, incoff, inc, cnt, core, status, fr0, fr1, big, gpa,
                                                 gidp);
      k_pl_finish<<<1, 1, 0, c.stream>>>(ctl, n);
      KCDS_LAUNCH_CHECK();
    }
    rounds += batch;
    HIP_CHECK(hipMemcpyAsync(hdone, &ctl->done, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (*hdone) break;
    // every sub-round removes at least one vertex, so n + 1 sub-rounds suffice
    if (rounds > (uint64_t)n + 1)
      fail(KCDS_E_HIP, "peeling did not finish within %llu sub-rounds", (unsigned long long)rounds);
    batch = std::min<uint32_t>(batch * 2, 256);
  }
  if (!multi) {
    k_graph_single<<<1, 1, 0, c.stream>>>(ctl, gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_kprime<<<grid1d(G), 256, 0, c.stream>>>(G, gp);
  KCDS_LAUNCH_CHECK();
  k_graph_summary<<<grid, 256, 0, c.stream>>>(n, T, k, gidp, core, status, cl, gp);
  KCDS_LAUNCH_CHECK();
  PeelCtl *hs = (PeelCtl *)c.host_scratch(sizeof(PeelCtl) + 64);
  HIP_CHECK(hipMemcpyAsync(hs, ctl, sizeof(PeelCtl), hipMemcpyDeviceToHost, c.stream));
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
