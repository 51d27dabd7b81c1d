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
// Here the peel of large graphs is one kernel per step (k_pl_step: OPEN a
// level or PROCESS a sub-round) driven by device-side control state; the
// last workgroup to finish a step runs the control step, and the host
// enqueues steps in batches, reading the "done" flag now and then.  A
// frontier carries each removed vertex's incidence list cut into pieces, one
// wave per piece, so a vertex in millions of cliques is spread over the GPU;
// decrements of one vertex by several lanes of a wave are one atomic.
// Workgroups share only atomic counters inside a step; kernel boundaries
// order the steps.  Graphs with few vertices and cliques are instead peeled
// one workgroup each (k_peel_block), all in one launch.  With several graphs
// in one batch, every graph keeps its own live cliques / remaining vertices
// and records its own densities at the end of each level that touched it --
// exactly what its own peel records, because a graph's vertices leave at
// their core numbers whatever else is peeled alongside.
#include <cstring>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t PSEG = 4;        // frontier segments / counter shards
// Incidence entries per work item: small pieces keep a level's critical path
// short (a piece is one wave's chain of dependent memory steps), larger ones
// bound the piece frontiers when T is large (peel_piece()).
constexpr uint32_t PPIECE_MIN = 64;
constexpr uint32_t KMAX = 8;        // largest k

// Counters 64 words apart (same-address device atomics are serialised).
struct PeelCtl {
  ull level, target;
  uint32_t cur, done, phase, steps, levels, pad0[64 - 9];
  uint32_t fsize[2][PSEG][64];  // vertex frontier segments
  uint32_t psize[2][PSEG][64];  // incidence-piece frontier segments
  ull gmin[PSEG][8];            // minimum live count above target (OPEN)
  uint32_t gblocks[PSEG][64];
  uint32_t blocks[64];
  uint32_t ntouch[64];          // graphs touched by the current level
  ull dmin[PSEG][8];            // minimum count above the level left by this level's decrements (PROCESS)
  ull nextmin[8];               // minimum live count above the level when the level opened
};

struct PItem {
  ull st, en;  // incidence entries [st, en) of vertex v
  uint32_t v, pad;
};

// a/b >= c/d for 64-bit non-negative integers, b, d > 0 (exact, 128-bit products)
__device__ __forceinline__ bool ratio_ge(ull a, ull b, ull c, ull d) {
  ull hi1 = __umul64hi(a, d), lo1 = a * d;
  ull hi2 = __umul64hi(c, b), lo2 = c * b;
  return hi1 > hi2 || (hi1 == hi2 && lo1 >= lo2);
}

// Appends v (where want) to a vertex frontier segment and its incidence list,
// cut in piece-entry pieces, to a piece segment; one counter atomic of each
// kind per wave.  Every lane calls it.
__device__ __forceinline__ void pl_append(bool want, uint32_t v, const ull *__restrict__ incoff, uint32_t *fv,
                                          uint32_t *fvsize, PItem *fp, uint32_t *fpsize, uint32_t piece) {
  const ull bal = __ballot(want);
  if (!bal) return;
  const uint32_t lane = lane_id(), lead = (uint32_t)__builtin_ctzll(bal);
  ull b = 0, e = 0;
  uint32_t np = 0;
  if (want) {
    b = incoff[v];
    e = incoff[v + 1];
    np = (uint32_t)((e - b + piece - 1) / piece);
  }
  const uint32_t incl = wave_incl_scan_u32(np);
  const uint32_t tot = __shfl(incl, 63, 64);
  uint32_t vb = 0, pb = 0;
  if (lane == lead) {
    vb = atomicAdd(fvsize, (uint32_t)__popcll(bal));
    if (tot) pb = atomicAdd(fpsize, tot);
  }
  vb = __shfl(vb, lead, 64);
  pb = __shfl(pb, lead, 64);
  if (want) {
    fv[vb + lanes_below(bal)] = v;
    uint32_t p = pb + incl - np;
    for (ull st = b; st < e; st += piece) {
      PItem it;
      it.st = st;
      it.en = min(st + piece, e);
      it.v = v;
      it.pad = 0;
      fp[p++] = it;
    }
  }
}

// Lanes with want grouped by key (the first eight groups; later lanes each
// alone): returns the group size on the group's leader lane and 0 elsewhere;
// leader = the leader's lane, rank = the lane's position in its group.
__device__ __forceinline__ uint32_t wave_group(bool want, uint32_t key, uint32_t &rank, uint32_t &leader) {
  const uint32_t lane = lane_id();
  uint32_t size = 0;
  rank = 0;
  leader = lane;
  ull pending = __ballot(want);
#pragma unroll 1
  for (int round = 0; round < 8 && pending; ++round) {
    const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
    const uint32_t kl = __shfl(key, lead, 64);
    const ull same = __ballot(want && key == kl) & pending;
    if ((same >> lane) & 1ull) {
      rank = lanes_below(same);
      leader = lead;
    }
    if (lane == lead) size = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  if ((pending >> lane) & 1ull) size = 1;
  return size;
}

// The last workgroup of a launch (two-level ticket) returns true.
// What the control step reads of a workgroup (the frontier-size counters,
// whose atomics returned their values before the appends, and thread 0's
// minimum atomic, drained below) is complete at the ticket; the rest of
// the step's writes are read only by later launches.
__device__ bool pl_last_block(PeelCtl *ctl, uint32_t *s_last) {
  lds_sync();
  if (threadIdx.x == 0) {
    wave_drain();
    const uint32_t seg = blockIdx.x % PSEG;
    const uint32_t in_seg = (gridDim.x - seg + PSEG - 1) / PSEG;
    uint32_t last = 0;
    if (atomicAdd(&ctl->gblocks[seg][0], 1u) == in_seg - 1)
      last = atomicAdd(&ctl->blocks[0], 1u) == min(gridDim.x, PSEG) - 1;
    *s_last = last;
  }
  lds_sync();
  return *s_last != 0;
}

// One peeling step.
//   OPEN    (frontier empty) records the density of every graph the level
//           just completed touched, then collects every live vertex whose
//           count is <= target and the minimum count of the others; the
//           control step starts level = target, or retries at that minimum
//           (empty levels cost one pass), or ends the peel.
//   PROCESS frontier vertices take the level as their core number; each live
//           clique of a removed vertex is claimed exactly once
//           (atomicCAS(status, -1, L)) and its surviving members are
//           decremented; a member whose count falls from L+1 to L joins the
//           next sub-round (each vertex crosses that value once).  Members
//           removed in this same step may still be decremented: their counts
//           no longer matter.
// Per graph: remaining vertices and live cliques (batches peel many graphs
// at once; each records its own densities, exactly what its own peel
// records, because a graph's vertices leave at their core numbers whatever
// else is peeled alongside).
__global__ __launch_bounds__(256) void k_pl_step(PeelCtl *ctl, uint32_t n, uint32_t k, const uint32_t *__restrict__ cl,
                                                 const ull *__restrict__ incoff, const uint32_t *__restrict__ inc,
                                                 ull *cnt, long long *core, int32_t *status, uint32_t *fv0, uint32_t *fv1,
                                                 PItem *fp0, PItem *fp1, uint32_t seg_v, uint32_t seg_p,
                                                 GraphPeel *gp, const uint32_t *__restrict__ gid, uint32_t *tlist,
                                                 Scatter sc, uint32_t piece) {
  __shared__ ull s_hdr[2];
  __shared__ uint32_t s_hdr2[3], s_f[PSEG + 1], s_p[PSEG + 1], s_last, s_sz[2][2][PSEG];
  __shared__ ull s_min;
  const uint32_t tid = threadIdx.x, lane = lane_id(), seg = blockIdx.x % PSEG;
  // the header and both frontiers' segment sizes in one round trip
  if (tid < 2) s_hdr[tid] = (&ctl->level)[tid];
  if (tid < 3) s_hdr2[tid] = (&ctl->cur)[tid];
  if (tid >= 64 && tid < 64 + 4 * PSEG) {
    const uint32_t i = tid - 64, f = i / (2 * PSEG), kind = (i / PSEG) & 1, g = i % PSEG;
    s_sz[f][kind][g] = kind ? ctl->psize[f][g][0] : ctl->fsize[f][g][0];
  }
  lds_sync();
  const ull L = s_hdr[0], target = s_hdr[1];
  const uint32_t cur = s_hdr2[0], done = s_hdr2[1], phase = s_hdr2[2];
  if (done) return;
  uint32_t *fv = cur ? fv1 : fv0, *nv = cur ? fv0 : fv1;
  PItem *fp = cur ? fp1 : fp0, *np = cur ? fp0 : fp1;
  const uint32_t gtid = blockIdx.x * blockDim.x + tid, gstride = gridDim.x * blockDim.x;
  if (phase == 0) {
    // densities of the graphs the completed level touched
    const uint32_t nt = ctl->ntouch[0];
    for (uint32_t i = gtid; i < nt; i += gstride) {
      GraphPeel &P = gp[tlist[i]];
      const ull rem = P.rem;
      if (rem > 0 && ratio_ge(P.live, rem, P.max_num, P.max_den)) {
        P.max_num = P.live;
        P.max_den = rem;
      }
      P.touch = 0;
    }
    if (tid == 0) s_min = ~0ull;
    lds_sync();
    ull m = ~0ull;
    const uint32_t K = sc.mask + 1;
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < K; p0 += gstride) {  // counter slots in order
      const uint32_t p = p0 + tid, v = vertex_of(sc, p);
      const bool live = p < K && v < n && core[v] == -1;
      const ull cv = live ? cnt[p] : ~0ull;
      pl_append(live && cv <= target, v, incoff, fv + (size_t)seg * seg_v, &ctl->fsize[cur][seg][0],
                fp + (size_t)seg * seg_p, &ctl->psize[cur][seg][0], piece);
      if (live && cv > target) m = min(m, cv);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = min(m, (ull)__shfl_xor(m, o, 64));
    if (lane == 0 && m != ~0ull) atomicMin(&s_min, m);
    lds_sync();
    if (tid == 0 && s_min != ~0ull) atomicMin(&ctl->gmin[seg][0], s_min);
  } else {
    if (tid == 0) {
      s_f[0] = 0;
      s_p[0] = 0;
      for (uint32_t g = 0; g < PSEG; ++g) {
        s_f[g + 1] = s_f[g] + s_sz[cur][0][g];
        s_p[g + 1] = s_p[g] + s_sz[cur][1][g];
      }
    }
    lds_sync();
    const uint32_t fs = s_f[PSEG], ps = s_p[PSEG];
    // mark: core numbers, remaining vertices and the touched list per graph
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < fs; i0 += gstride) {
      const uint32_t i = i0 + tid;
      uint32_t v = 0, g = 0;
      const bool valid = i < fs;
      if (valid) {
        uint32_t sg = 0;
        while (i >= s_f[sg + 1]) ++sg;
        v = fv[(size_t)sg * seg_v + (i - s_f[sg])];
        __hip_atomic_store(&core[v], (long long)L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        g = gid[v];
      }
      uint32_t rank, leader;
      const uint32_t gsize = wave_group(valid, g, rank, leader);
      if (gsize) {
        atomicSub(&gp[g].rem, gsize);
        if (atomicExch(&gp[g].touch, 1u) == 0u) tlist[atomicAdd(&ctl->ntouch[0], 1u)] = g;
      }
    }
    // incidence pieces, wave per piece.  dm: the lowest count above L that a
    // decrement leaves (the next level starts at min(dm, the minimum when
    // this level opened), so it needs no failed OPEN step to find it)
    ull dm = ~0ull;
    const uint32_t wid = gtid / KCDS_WAVE, nw = gstride / KCDS_WAVE;
    for (uint32_t it = wid; it < ps; it += nw) {
      uint32_t sg = 0;
      while (it >= s_p[sg + 1]) ++sg;
      const PItem item = fp[(size_t)sg * seg_p + (it - s_p[sg])];
      const uint32_t v = item.v;
      uint32_t destroyed = 0;
      for (ull j0 = item.st; j0 < item.en; j0 += KCDS_WAVE) {
        const ull j = j0 + lane;
        bool mine = false;
        uint32_t c = 0;
        if (j < item.en) {
          c = inc[j];
          mine = atomicCAS(&status[c], -1, (int32_t)L) == -1;
        }
        destroyed += mine;
        // the k members' ids are loaded together, then every decrement is
        // issued before any result is used.  A clique still live when claimed
        // has no member removed in an earlier sub-round (removing a vertex
        // destroys all its live cliques), so a member's state needs no read:
        // one in the current frontier has count <= L, and its decrement can
        // neither join nor lower the next level's start
        uint32_t u[KMAX];
        bool dec[KMAX];
#pragma unroll
        for (uint32_t t = 0; t < KMAX; ++t) u[t] = mine && t < k ? cl[(ull)c * k + t] : v;
#pragma unroll
        for (uint32_t t = 0; t < KMAX; ++t) dec[t] = u[t] != v;
        ull old[KMAX];
        uint32_t rank[KMAX], leader[KMAX];
#pragma unroll
        for (uint32_t t = 0; t < KMAX; ++t) {
          // one atomic per group of lanes decrementing the same (hub) vertex;
          // each lane's own old value is the leader's minus its rank
          const uint32_t gsize = wave_group(dec[t], u[t], rank[t], leader[t]);
          old[t] = 0;
          if (gsize) old[t] = atomicSub(&cnt[slot_of(sc, u[t])], (ull)gsize);
        }
#pragma unroll
        for (uint32_t t = 0; t < KMAX; ++t) {
          const ull lold = __shfl(old[t], leader[t], 64);
          const bool join = dec[t] && lold - rank[t] == L + 1;
          const ull now = lold - rank[t] - 1;  // this lane's count after its decrement
          if (dec[t] && now > L) dm = min(dm, now);
          pl_append(join, u[t], incoff, nv + (size_t)seg * seg_v, &ctl->fsize[cur ^ 1][seg][0],
                    np + (size_t)seg * seg_p, &ctl->psize[cur ^ 1][seg][0], piece);
        }
      }
      destroyed = wave_sum_u32(destroyed);
      if (lane == 0 && destroyed) atomicSub(&gp[gid[v]].live, (ull)destroyed);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dm = min(dm, (ull)__shfl_xor(dm, o, 64));
    if (tid == 0) s_min = ~0ull;
    lds_sync();
    if (lane == 0 && dm != ~0ull) atomicMin(&s_min, dm);
    lds_sync();
    if (tid == 0 && s_min != ~0ull) atomicMin(&ctl->dmin[seg][0], s_min);
  }
  // control step: the last workgroup to finish
  // wave 0 of the last workgroup: lanes g < PSEG read and reset segment g's
  // counters together; L, target, cur and phase are this step's header
  // (only the control step writes them)
  if (pl_last_block(ctl, &s_last) && tid < KCDS_WAVE) {
    uint32_t fc = 0, fn = 0;
    ull mn = ~0ull, dn = ~0ull;
    if (lane < PSEG) {
      fc = ldc_u32(&ctl->fsize[cur][lane][0]);
      fn = ldc_u32(&ctl->fsize[cur ^ 1][lane][0]);
      // the other workgroups' atomics on these completed before their
      // tickets, so dmin is read here with the rest and reset below only
      // when the level ends (no second round trip after fnext is known)
      if (phase == 1) dn = ldc_u64(&ctl->dmin[lane][0]);
      mn = atomicExch(&ctl->gmin[lane][0], ~0ull);
      ctl->gblocks[lane][0] = 0;
      if (phase == 1) {
        ctl->fsize[cur][lane][0] = 0;
        ctl->psize[cur][lane][0] = 0;
      }
    }
    const uint32_t fcur = wave_sum_u32(fc), fnext = wave_sum_u32(fn);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, (ull)__shfl_xor(mn, o, 64));
    // end of a level: the lowest count its decrements left above it
    if (!(phase == 1 && fnext == 0)) dn = ~0ull;
    else if (lane < PSEG) stc_u64(&ctl->dmin[lane][0], ~0ull);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dn = min(dn, (ull)__shfl_xor(dn, o, 64));
    if (lane == 0) {
      ctl->blocks[0] = 0;
      ++ctl->steps;
      if (phase == 1) {
        ctl->cur = cur ^ 1;
        if (fnext == 0) {
          // every live count is >= the minimum when the level opened (untouched
          // vertices) or >= the lowest a decrement left (touched ones): start
          // the next level there.  A stale minimum (its vertex left in the
          // cascade) only costs one OPEN that finds nothing and the true one.
          const ull nm = min(ctl->nextmin[0], dn);
          ctl->phase = 0;
          ctl->target = nm == ~0ull ? L + 1 : max(nm, L + 1);
        }
      } else {
        ctl->ntouch[0] = 0;  // the densities were recorded by this step
        if (fcur > 0) {
          ctl->level = target;
          ctl->phase = 1;
          ctl->nextmin[0] = mn;
          ++ctl->levels;
        } else if (mn == ~0ull) {
          ctl->done = 1;
        } else {
          ctl->target = mn;
        }
      }
    }
  }
}

// cnt: the working counts, in the scattered layout of sc when scattered
// (grid peel), else by vertex
__global__ void k_peel_init(uint32_t n, uint64_t T, const ull *cnt0, ull *cnt, long long *core, int32_t *status,
                            Scatter sc, uint32_t scattered) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    cnt[scattered ? slot_of(sc, (uint32_t)i) : i] = cnt0[i];
    core[i] = -1;
  }
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T; i += stride) status[i] = -1;
}

__global__ void k_gid(uint32_t n, const uint32_t *inv, const uint32_t *gvoff, uint32_t G, uint32_t *gid) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    gid[r] = upper_bound_u32(gvoff, G + 1, inv[r]) - 1;
}

__global__ void k_graph_accum(uint32_t n, const uint32_t *gid, const ull *cnt, GraphPeel *gp) {
  for (uint32_t r0 = blockIdx.x * blockDim.x; r0 < n; r0 += gridDim.x * blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    const bool valid = r < n;
    const uint32_t g = valid ? gid[r] : 0u;
    const ull cv = valid ? cnt[r] : 0ull;
    const uint32_t g0 = __shfl(g, 0, 64);
    if (__ballot(valid && g != g0) == 0ull) {  // one graph for the whole wave
      const ull sum = wave_sum_u64(cv);
      const uint32_t nv = (uint32_t)__popcll(__ballot(valid));
      if (lane_id() == 0) {
        if (sum) atomicAdd(&gp[g0].live, sum);
        if (nv) atomicAdd(&gp[g0].n, nv);
      }
    } else if (valid) {
      if (cv) atomicAdd(&gp[g].live, cv);
      atomicAdd(&gp[g].n, 1u);
    }
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

// k' = ceil(maxDensity) per graph (main.cu:1476)
__global__ void k_graph_kprime(uint32_t G, GraphPeel *gp) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    GraphPeel &P = gp[g];
    P.k_prime = P.T == 0 ? NO_KPRIME : (P.max_num + P.max_den - 1) / P.max_den;
    P.core_size = 0;
    P.core_cliques = 0;
    P.max_core = 0;
  }
}

// k'-core size, cliques inside it and the largest core, per graph.  Lanes of
// one graph are combined first (one atomic per wave when they agree).
__global__ void k_graph_summary(uint32_t n, uint64_t T, uint32_t k, const uint32_t *gid, const long long *core,
                                const int32_t *status, const uint32_t *cl, GraphPeel *gp) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t lane = lane_id();
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
    const uint64_t i = i0 + threadIdx.x;
    const bool valid = i < n;
    const uint32_t g = valid ? gid[i] : 0u;
    const long long cv = valid ? core[i] : -1;
    const bool in = valid && cv >= (long long)gp[g].k_prime;
    const uint32_t g0 = __shfl(g, 0, 64);
    if (__ballot(valid && g != g0) == 0ull) {  // the whole wave is in graph g0
      const uint32_t cs = (uint32_t)__popcll(__ballot(in));
      ull mc = cv > 0 ? (ull)cv : 0ull;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mc = max(mc, (ull)__shfl_xor(mc, o, 64));
      if (lane == 0) {
        if (cs) atomicAdd(&gp[g0].core_size, cs);
        if (mc) atomicMax(&gp[g0].max_core, mc);
      }
    } else if (valid) {
      if (in) atomicAdd(&gp[g].core_size, 1u);
      if (cv > 0) atomicMax(&gp[g].max_core, (ull)cv);
    }
  }
  // core cliques: a wave whose cliques all lie in one graph adds to a
  // running count and flushes it when the graph changes (a lone graph's
  // count then takes one atomic per wave, not one per 64 cliques).  Without
  // a clique list (enumerating peel) they are counted later, on the core.
  if (!cl) return;
  uint32_t acc_g = 0xFFFFFFFFu;
  ull acc = 0;
  for (uint64_t c0 = (uint64_t)blockIdx.x * blockDim.x; c0 < T; c0 += stride) {
    const uint64_t c = c0 + threadIdx.x;
    const bool valid = c < T;
    const uint32_t g = valid ? gid[cl[c * k]] : 0u;
    const bool in = valid && (long long)status[c] >= (long long)gp[g].k_prime;
    const uint32_t g0 = __shfl(g, 0, 64);
    if (__ballot(valid && g != g0) == 0ull) {
      const uint32_t cc = (uint32_t)__popcll(__ballot(in));
      if (g0 != acc_g) {
        if (lane == 0 && acc) atomicAdd(&gp[acc_g].core_cliques, acc);
        acc_g = g0;
        acc = 0;
      }
      acc += cc;
    } else if (in) {
      atomicAdd(&gp[g].core_cliques, 1ull);
    }
  }
  if (lane == 0 && acc) atomicAdd(&gp[acc_g].core_cliques, acc);
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

__global__ __launch_bounds__(BP_THREADS) void k_peel_block(uint32_t G, const uint32_t *__restrict__ glist,
                                                          const uint32_t *__restrict__ gvoff,
                                                          const uint32_t *__restrict__ rank, uint32_t k,
                                                          const ull *__restrict__ cnt0, ull *cnt, long long *core,
                                                          int32_t *status, const uint32_t *__restrict__ cl,
                                                          const ull *__restrict__ incoff,
                                                          const uint32_t *__restrict__ inc, uint32_t *fr0,
                                                          uint32_t *fr1, uint32_t *queue, GraphPeel *gp,
                                                          uint32_t init_status) {
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
    const uint32_t gi = s_g;
    block_sync();
    if (gi >= G) break;
    const uint32_t g = glist[gi];
    const uint32_t vb = gvoff[g], ng = gvoff[g + 1] - vb;
    uint32_t *F0 = fr0 + vb, *F1 = fr1 + vb;
    ull mine = 0;
    for (uint32_t i = tid; i < ng; i += nt) {
      const uint32_t r = rank[vb + i];
      const ull c0 = cnt0[r];
      stc_u64(&cnt[r], c0);
      stc_i64(&core[r], -1);
      mine += c0;
      if (init_status)  // the graph's cliques (when not reset grid-wide before the launch)
        for (ull j = incoff[r]; j < incoff[r + 1]; ++j) stc_i32(&status[inc[j]], -1);
    }
    const ull T = block_sum_u64(mine, &s_red) / k;  // sum of per-vertex counts = k T
    ull live = T, max_num = T, max_den = ng ? ng : 1, L = 0;
    uint32_t rem = ng, levels = 0;
    while (rem > 0) {
      ull m = ~0ull;
      for (uint32_t i = tid; i < ng; i += nt) {
        const uint32_t r = rank[vb + i];
        if (ldc_i64(&core[r]) == -1) m = min(m, ldc_u64(&cnt[r]));
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
          want = ldc_i64(&core[r]) == -1 && ldc_u64(&cnt[r]) <= L;
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
        for (uint32_t i = tid; i < fs; i += nt) stc_i64(&core[ldc_u32(&Fc[i])], (long long)L);
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
              if (u == v || ldc_i64(&core[u]) != -1) continue;
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
    const ull kp = T == 0 ? NO_KPRIME : (max_num + max_den - 1) / max_den;
    ull csize = 0, ccl = 0, mc = 0;
    for (uint32_t i = tid; i < ng; i += nt) {
      const uint32_t r = rank[vb + i];
      const long long cv = ldc_i64(&core[r]);
      if (cv >= (long long)kp) csize += 1;
      if (cv > 0) mc = max(mc, (ull)cv);
      for (ull j = incoff[r]; j < incoff[r + 1]; ++j) ccl += (long long)ldc_i32(&status[inc[j]]) >= (long long)kp;
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
      P.max_core = mc;
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
constexpr ull BLOCK_PEEL_MAX_T = 1u << 18;

// Graph 0's peel summary into c.peel (the single-graph API reports it).
void set_peel_stats(Ctx &c, uint32_t levels) {
  const GraphPeel &P = c.h_gpeel[0];
  kcds_peel_stats ps{};
  ull g0 = P.max_num, b0 = P.max_den;
  while (b0) {
    const ull t = g0 % b0;
    g0 = b0;
    b0 = t;
  }
  if (!g0) g0 = 1;
  ps.max_num = P.max_num / g0;
  ps.max_den = P.max_den / g0;
  ps.k_prime = P.k_prime;
  ps.levels = levels;
  ps.core_size = P.core_size;
  ps.core_cliques = P.core_cliques;
  ps.max_core = (int64_t)P.max_core;
  c.peel = ps;
  c.peeled = true;
  c.have_comp = false;
}

// Workgroup-per-graph peel of the graphs in ids (all graphs when ids is
// empty); writes their GraphPeel entries on the device and in c.h_gpeel.
void peel_blocks(Ctx &c, const std::vector<uint32_t> &ids) {
  const uint32_t n = c.n, k = c.k, G = c.ngraphs;
  const uint64_t T = c.T;
  std::vector<uint32_t> all;
  if (ids.empty()) {
    all.resize(G);
    for (uint32_t g = 0; g < G; ++g) all[g] = g;
  }
  const std::vector<uint32_t> &list = ids.empty() ? all : ids;
  const uint32_t L = (uint32_t)list.size();
  buf_reserve(c, c.core, sizeof(long long) * ((size_t)n + 1));
  buf_reserve(c, c.status, sizeof(int32_t) * (T + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.gvoff, sizeof(uint32_t) * ((size_t)G + 1));
  HIP_CHECK(hipMemcpyAsync(c.gvoff.p, c.h_gvoff.data(), sizeof(uint32_t) * ((size_t)G + 1), hipMemcpyHostToDevice,
                           c.stream));
  const size_t words = (size_t)n * 2 /*cnt*/ + (size_t)n * 2 /*frontiers*/ + QUEUE_WORDS + (size_t)L + 64;
  buf_reserve(c, c.tmp0, words * sizeof(uint32_t) + 256);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fr0 = (uint32_t *)(cnt + n), *fr1 = fr0 + n, *queue = fr1 + n, *glist = queue + QUEUE_WORDS;
  HIP_CHECK(hipMemsetAsync(queue, 0, QUEUE_WORDS * sizeof(uint32_t), c.stream));
  HIP_CHECK(hipMemcpyAsync(glist, list.data(), sizeof(uint32_t) * L, hipMemcpyHostToDevice, c.stream));
  if (ids.empty() && T) {  // every graph: reset all clique statuses at once
    k_peel_init<<<grid1d(T), 256, 0, c.stream>>>(0, T, nullptr, nullptr, nullptr, c.status.as<int32_t>(),
                                                 make_scatter(1), 0u);
    KCDS_LAUNCH_CHECK();
  }
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(L, (uint32_t)c.num_cus * 2));
  k_peel_block<<<grid, BP_THREADS, 0, c.stream>>>(L, glist, c.gvoff.as<uint32_t>(), c.rank.as<uint32_t>(), k,
                                                  c.cnt.as<ull>(), cnt, c.core.as<long long>(), c.status.as<int32_t>(),
                                                  c.cliques.as<uint32_t>(), c.incoff.as<ull>(), c.inc.as<uint32_t>(),
                                                  fr0, fr1, queue, c.gpeel.as<GraphPeel>(), ids.empty() ? 0u : 1u);
  KCDS_LAUNCH_CHECK();
  c.h_gpeel.resize(G);
  HIP_CHECK(hipMemcpyAsync(c.h_gpeel.data(), c.gpeel.p, sizeof(GraphPeel) * G, hipMemcpyDeviceToHost, c.stream));
  c.sync();  // list lives on the host
}

// vertices of graphs peeled elsewhere take core -2: the grid peel treats
// them as neither live nor removed
__global__ void k_exclude(uint32_t n, const uint32_t *gid, const uint8_t *excl, long long *core) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    if (excl[gid[r]]) core[r] = -2;
}

void peel_run(Ctx &c) {
  const uint32_t n = c.n, k = c.k, G = c.ngraphs;
  const uint64_t T = c.T;
  const bool multi = G > 1;
  // clique statuses (destroy levels) are int32; they never exceed the
  // largest per-vertex count (larger counts take the enumerating peel)
  if (T > 0x7FFFFFFFull && max_u64_host(c, c.cnt.as<ull>(), n) > 0x7FFFFFFFull)
    fail(KCDS_E_LIMIT, "a per-vertex clique count exceeds the int32 clique-core range");
  graph_ids(c);
  // in a batch, graphs small enough (vertices and cliques) for one workgroup
  // each; a single graph takes the grid peel, which used the whole GPU to
  // beat one workgroup on every small single graph measured (C1 peel 0.49 ->
  // 0.22 ms, 24 small served graphs 344 -> 200 ms of solve time)
  std::vector<ull> gT;
  graph_counts(c, gT);
  std::vector<uint32_t> small_ids;
  std::vector<uint8_t> small(G, 0);
  if (!(c.tune_flags & KCDS_TUNE_GRID_PEEL) && (G > 1 || (c.tune_flags & KCDS_TUNE_BLOCK_PEEL)))
    for (uint32_t g = 0; g < G; ++g)
      if (c.h_gvoff[g + 1] - c.h_gvoff[g] <= BLOCK_PEEL_MAX_N && gT[g] <= BLOCK_PEEL_MAX_T) {
        small[g] = 1;
        small_ids.push_back(g);
      }
  if (small_ids.size() == G) {
    peel_blocks(c, {});
    set_peel_stats(c, c.h_gpeel[0].levels);
    return;
  }
  buf_reserve(c, c.core, sizeof(long long) * ((size_t)n + 1));
  buf_reserve(c, c.status, sizeof(int32_t) * (T + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.tlist, sizeof(uint32_t) * ((size_t)G + 1));
  GraphPeel *gp = c.gpeel.as<GraphPeel>();
  HIP_CHECK(hipMemsetAsync(gp, 0, sizeof(GraphPeel) * G, c.stream));
  // working counts, two frontiers (vertices, incidence pieces) in PSEG
  // segments that can each hold any frontier, and the control block
  const uint32_t seg_v = ((n + 63) / 64) * 64;
  // pieces of 64 incidences while the piece frontiers stay under ~2^24
  // entries per segment, larger for very large T
  uint32_t piece = PPIECE_MIN;
  while ((T * k) / piece > (1ull << 24)) piece *= 2;
  const uint64_t seg_p64 = (((uint64_t)n + (T * k) / piece + 64 + 63) / 64) * 64;
  if (seg_p64 * PSEG >= (1ull << 32)) fail(KCDS_E_LIMIT, "too many cliques for the peeling frontiers");
  const uint32_t seg_p = (uint32_t)seg_p64;
  const Scatter sc = make_scatter(n);  // working counts: hub counters on lines of their own
  const size_t kslots = ((size_t)sc.mask + 1 + 63) / 64 * 64;
  const size_t bytes = sizeof(ull) * kslots + sizeof(uint32_t) * 2 * PSEG * (size_t)seg_v +
                       sizeof(PItem) * 2 * PSEG * (size_t)seg_p + sizeof(PeelCtl) + 1024;
  buf_reserve(c, c.tmp0, bytes);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fv0 = (uint32_t *)(cnt + kslots), *fv1 = fv0 + (size_t)PSEG * seg_v;
  PItem *fp0 = (PItem *)(((uintptr_t)(fv1 + (size_t)PSEG * seg_v) + 255) & ~(uintptr_t)255);
  PItem *fp1 = fp0 + (size_t)PSEG * seg_p;
  PeelCtl *ctl = (PeelCtl *)(((uintptr_t)(fp1 + (size_t)PSEG * seg_p) + 255) & ~(uintptr_t)255);
  PeelCtl *h0 = (PeelCtl *)c.host_scratch(sizeof(PeelCtl));
  memset(h0, 0, sizeof(PeelCtl));
  for (uint32_t g = 0; g < PSEG; ++g) {
    h0->gmin[g][0] = ~0ull;
    h0->dmin[g][0] = ~0ull;
  }
  h0->nextmin[0] = ~0ull;
  h0->done = n == 0;
  HIP_CHECK(hipMemcpyAsync(ctl, h0, sizeof(PeelCtl), hipMemcpyHostToDevice, c.stream));
  c.sync();  // h0 is reused below
  const uint32_t grid = grid1d(std::max<uint64_t>(n, T));
  k_peel_init<<<grid, 256, 0, c.stream>>>(n, T, c.cnt.as<ull>(), cnt, c.core.as<long long>(), c.status.as<int32_t>(),
                                          sc, 1u);
  KCDS_LAUNCH_CHECK();
  if (!small_ids.empty() && n) {
    buf_reserve(c, c.tmp3, (size_t)G + 8);
    HIP_CHECK(hipMemcpyAsync(c.tmp3.p, small.data(), G, hipMemcpyHostToDevice, c.stream));
    k_exclude<<<grid1d(n), 256, 0, c.stream>>>(n, c.gid.as<uint32_t>(), c.tmp3.as<uint8_t>(), c.core.as<long long>());
    KCDS_LAUNCH_CHECK();
  }
  const uint32_t *gidp = c.gid.as<uint32_t>();
  if (n) {
    k_graph_accum<<<grid1d(n), 256, 0, c.stream>>>(n, gidp, c.cnt.as<ull>(), gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_init<<<grid1d(G), 256, 0, c.stream>>>(G, k, gp);
  KCDS_LAUNCH_CHECK();
  const uint32_t *cl = c.cliques.as<uint32_t>(), *inc = c.inc.as<uint32_t>();
  const ull *incoff = c.incoff.as<ull>();
  long long *core = c.core.as<long long>();
  int32_t *status = c.status.as<int32_t>();
  uint32_t *tl = c.tlist.as<uint32_t>();
  // at least a workgroup per CU: the work of a step is the frontier's
  // incidence pieces (up to k*T), not its vertices
  const uint32_t sgrid = (uint32_t)std::max<uint64_t>(c.num_cus, std::min<uint64_t>((n + 255) / 256, c.num_cus * 2ull));
  // the number of steps is a property of the graph: start from the last count
  uint32_t batch = c.peel_steps ? (uint32_t)std::min<uint64_t>(c.peel_steps, 1u << 20) : 32;
  uint64_t rounds = 0;
  while (n) {
    for (uint32_t r = 0; r < batch; ++r) {
      k_pl_step<<<sgrid, 256, 0, c.stream>>>(ctl, n, k, cl, incoff, inc, cnt, core, status, fv0, fv1, fp0, fp1, seg_v,
                                             seg_p, gp, gidp, tl, sc, piece);
      KCDS_LAUNCH_CHECK();
    }
    rounds += batch;
    HIP_CHECK(hipMemcpyAsync(h0, ctl, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (h0->done) break;
    // every PROCESS step removes at least one vertex; an OPEN step is
    // followed by a PROCESS step or by one OPEN step that succeeds
    if (rounds > 3ull * n + 8)
      fail(KCDS_E_HIP, "peeling did not finish within %llu steps", (unsigned long long)rounds);
    batch = std::min<uint32_t>(std::max<uint32_t>(batch, 32) * 2, 128);
  }
  c.peel_steps = h0->steps;
  k_graph_kprime<<<grid1d(G), 256, 0, c.stream>>>(G, gp);
  KCDS_LAUNCH_CHECK();
  k_graph_summary<<<std::min<uint32_t>(grid, c.num_cus * 8), 256, 0, c.stream>>>(n, T, k, gidp, core, status, cl,
                                                                                gp);
  KCDS_LAUNCH_CHECK();
  PeelCtl *hs = (PeelCtl *)c.host_scratch(sizeof(PeelCtl) + 64);
  HIP_CHECK(hipMemcpyAsync(hs, ctl, sizeof(PeelCtl), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  const uint32_t levels = hs->levels;
  if (!small_ids.empty()) {
    peel_blocks(c, small_ids);  // also brings every graph's entry to the host
  } else {
    c.h_gpeel.resize(G);
    HIP_CHECK(hipMemcpyAsync(c.h_gpeel.data(), gp, sizeof(GraphPeel) * G, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  }
  set_peel_stats(c, levels);
}

// Shared by the enumerating peel (k_epeel.hip): per-graph live cliques and
// vertices from the per-vertex counts before a peel, and k', the k'-core
// size and the largest core after it (cl = nullptr: no clique list, so the
// core cliques are left at 0).
void peel_graphs_begin(Ctx &c, GraphPeel *gp) {
  const uint32_t n = c.n, G = c.ngraphs;
  HIP_CHECK(hipMemsetAsync(gp, 0, sizeof(GraphPeel) * G, c.stream));
  if (n) {
    k_graph_accum<<<grid1d(n), 256, 0, c.stream>>>(n, c.gid.as<uint32_t>(), c.cnt.as<ull>(), gp);
    KCDS_LAUNCH_CHECK();
  }
  k_graph_init<<<grid1d(G), 256, 0, c.stream>>>(G, c.k, gp);
  KCDS_LAUNCH_CHECK();
}

void peel_graphs_end(Ctx &c, GraphPeel *gp, uint32_t levels) {
  const uint32_t n = c.n, G = c.ngraphs;
  k_graph_kprime<<<grid1d(G), 256, 0, c.stream>>>(G, gp);
  KCDS_LAUNCH_CHECK();
  if (n) {
    k_graph_summary<<<std::min<uint32_t>(grid1d(n), c.num_cus * 8), 256, 0, c.stream>>>(
        n, 0, c.k, c.gid.as<uint32_t>(), c.core.as<long long>(), nullptr, nullptr, gp);
    KCDS_LAUNCH_CHECK();
  }
  c.h_gpeel.resize(G);
  HIP_CHECK(hipMemcpyAsync(c.h_gpeel.data(), gp, sizeof(GraphPeel) * G, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  set_peel_stats(c, levels);
}

}  // namespace kcds
