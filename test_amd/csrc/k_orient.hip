// Degeneracy ordering and DAG orientation.
//
// Reference: src/graph.cpp:71-143 computes a Batagelj-Zaversnik peel on the
// host and main.cu:22-91 (generateDegreeDAG / generateNeighborDAG,
// src/helpers.cu:112-173) keeps arc u->v iff listingOrder[u] <
// listingOrder[v], i.e. from the vertex peeled LATER to the one peeled
// earlier.  Here every arc points from the vertex peeled EARLIER to the one
// peeled later, so out-degree is bounded by the core number; every acyclic
// orientation lists each clique exactly once, so counts, cores and the
// densest set do not depend on this choice.
//
// The DAG is stored over new ids r = rank[v] (position in the peel order):
// N+(r) = { rank[u] : u in N(v), rank[u] > r }, each list sorted ascending.
// Relabelling makes the out-neighbourhood of every vertex an increasing run
// of ids, which the listing kernel uses to derive directed rows from
// symmetric bitmask rows.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t SORT_WAVE = 1024;  // longest out-list a wave sorts in LDS

__global__ void k_rank_from_lo(const uint32_t *lo, uint32_t n, uint32_t *rank, uint32_t *inv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    uint32_t r = n - lo[v];  // lo in 1..n, checked on the host to be a permutation
    rank[v] = r;
    inv[r] = v;
  }
}

// Wave per vertex v (4 waves per block, 4 KB of LDS each): gathers the new
// ids of its later neighbours, sorts them by rank counting (position = number
// of smaller entries; entries are distinct) and writes them to
// tmp[off[v] ..), the start of v's undirected list; dout[rank[v]] = count.
// A list longer than SORT_WAVE is counted but not written (the caller fails).
__global__ __launch_bounds__(256) void k_dag_gather(const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ rank, uint32_t n,
                                                    uint32_t *__restrict__ dout, uint32_t *__restrict__ tmp) {
  __shared__ uint32_t sbuf[4][SORT_WAVE];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  uint32_t *buf = sbuf[wv];
  const uint32_t gw = blockIdx.x * 4 + wv, nwv = gridDim.x * 4;
  for (uint32_t v = gw; v < n; v += nwv) {
    const uint32_t b = off[v], e = off[v + 1], r = rank[v];
    uint32_t cnt = 0;
    for (uint32_t i0 = b; i0 < e; i0 += KCDS_WAVE) {
      const uint32_t i = i0 + lane;
      const uint32_t x = (i < e) ? rank[nbr[i]] : 0u;
      const bool keep = (i < e) && x > r;
      const ull bal = __ballot(keep);
      if (keep && cnt + lanes_below(bal) < SORT_WAVE) buf[cnt + lanes_below(bal)] = x;
      cnt += (uint32_t)__popcll(bal);
    }
    if (lane == 0) dout[r] = cnt;
    wave_sync();
    if (cnt <= SORT_WAVE) {
      for (uint32_t i = lane; i < cnt; i += KCDS_WAVE) {
        const uint32_t x = buf[i];
        uint32_t pos = 0;
        for (uint32_t j = 0; j < cnt; ++j) pos += buf[j] < x;
        tmp[b + pos] = x;
      }
    }
    wave_sync();
  }
}

// Wave per new id r: copies the sorted out-list of v = inv[r] into the DAG.
__global__ __launch_bounds__(256) void k_dag_copy(const uint32_t *__restrict__ off, const uint32_t *__restrict__ inv,
                                                  uint32_t n, const uint32_t *__restrict__ doff,
                                                  const uint32_t *__restrict__ tmp, uint32_t *__restrict__ dnbr) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwv = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t r = gw; r < n; r += nwv) {
    const uint32_t src = off[inv[r]], dst = doff[r], d = doff[r + 1] - dst;
    for (uint32_t i = lane; i < d; i += KCDS_WAVE) dnbr[dst + i] = tmp[src + i];
  }
}

// ---------------------------------------------------------------------------
// GPU degeneracy order: level-synchronous k-core peel as a fixed pair of
// kernels per sub-round (work / control) driven by device-side control
// state and enqueued by the host in batches.
//   OPEN    collect every live vertex of degree <= target into the frontier
//           and take the minimum degree of the others; the control step
//           either starts level = target or, when nothing was collected,
//           retries with target = that minimum (empty levels cost one pass).
//   PROCESS the frontier takes the next ranks (in frontier order) and leaves
//           the graph, decrementing its live neighbours; a neighbour whose
//           degree falls from L+1 to L joins the next sub-round of level L.
// A vertex leaves by having its degree replaced with DEAD; decrements that
// race with that replacement keep the value >= DEAD_MIN, which every reader
// treats as dead.  Kernel boundaries order the steps; no grid-wide barrier
// is needed.

struct OrderCtl {
  uint32_t level;
  uint32_t target;
  uint32_t minval;
  uint32_t next_rank;
  uint32_t fsize[2];
  uint32_t cur;
  uint32_t done;
  uint32_t phase;  // 0 = OPEN, 1 = PROCESS
};

constexpr uint32_t DEAD = 0xFFFFFFFFu;
constexpr uint32_t DEAD_MIN = 0x80000000u;

// Appends v (where want) to q with one counter atomic per wave.
__device__ __forceinline__ void wave_append(bool want, uint32_t v, uint32_t *q, uint32_t *qsize) {
  const ull bal = __ballot(want);
  if (!bal) return;
  uint32_t base = 0;
  if (lane_id() == (uint32_t)__builtin_ctzll(bal)) base = atomicAdd(qsize, (uint32_t)__popcll(bal));
  base = __shfl(base, __builtin_ctzll(bal), 64);
  if (want) q[base + lanes_below(bal)] = v;
}

__global__ __launch_bounds__(256) void k_ko_work(OrderCtl *ctl, uint32_t n, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ nbr, uint32_t *deg, uint32_t *fr0,
                                                 uint32_t *fr1, uint32_t *rank, uint32_t *inv) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur;
  uint32_t *fr = cur ? fr1 : fr0;
  uint32_t *nx = cur ? fr0 : fr1;
  const uint32_t lane = lane_id();
  if (ctl->phase == 0) {
    const uint32_t target = ctl->target;
    uint32_t m = DEAD;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v0 = blockIdx.x * blockDim.x; v0 < n; v0 += stride) {
      const uint32_t v = v0 + threadIdx.x;
      const uint32_t d = v < n ? deg[v] : DEAD;
      const bool live = d < DEAD_MIN;
      wave_append(live && d <= target, v, fr, &ctl->fsize[cur]);
      if (live && d > target) m = min(m, d);
    }
    m = wave_min_u32(m);
    if (lane == 0 && m != DEAD) atomicMin(&ctl->minval, m);
    return;
  }
  const uint32_t fs = ctl->fsize[cur], L = ctl->level, base = ctl->next_rank;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / KCDS_WAVE;
  const uint32_t nw = gridDim.x * blockDim.x / KCDS_WAVE;
  for (uint32_t i = wid; i < fs; i += nw) {
    const uint32_t v = fr[i];
    if (lane == 0) {
      rank[v] = base + i;
      inv[base + i] = v;
      atomicExch(&deg[v], DEAD);
    }
    const uint32_t b = off[v], e = off[v + 1];
    for (uint32_t j0 = b; j0 < e; j0 += KCDS_WAVE) {
      const uint32_t j = j0 + lane;
      bool join = false;
      uint32_t u = 0;
      if (j < e) {
        u = nbr[j];
        if (__hip_atomic_load(&deg[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < DEAD_MIN)
          join = atomicSub(&deg[u], 1u) == L + 1;
      }
      wave_append(join, u, nx, &ctl->fsize[cur ^ 1]);
    }
  }
}

__global__ void k_ko_ctl(OrderCtl *ctl, uint32_t n) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur;
  if (ctl->phase == 1) {
    ctl->next_rank += ctl->fsize[cur];
    ctl->fsize[cur] = 0;
    ctl->cur = cur ^ 1;
    if (ctl->next_rank >= n) {
      ctl->done = 1;
    } else if (ctl->fsize[cur ^ 1] == 0) {
      ctl->phase = 0;
      ctl->target = ctl->level + 1;
      ctl->minval = DEAD;
    }
    return;
  }
  if (ctl->fsize[cur] > 0) {
    ctl->level = ctl->target;
    ctl->phase = 1;
  } else if (ctl->minval == DEAD) {
    ctl->done = 1;
  } else {
    ctl->target = ctl->minval;
    ctl->minval = DEAD;
  }
}

__global__ void k_init_deg(const uint32_t *off, uint32_t n, uint32_t *deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    deg[v] = off[v + 1] - off[v];
}

}  // namespace

void degeneracy_rank_gpu(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.rank, sizeof(uint32_t) * (size_t)n + 4);
  buf_reserve(c, c.inv, sizeof(uint32_t) * (size_t)n + 4);
  if (n == 0) return;
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * (size_t)n * 3 + sizeof(OrderCtl) + 256);
  uint32_t *deg = c.tmp0.as<uint32_t>();
  uint32_t *fr0 = deg + n, *fr1 = fr0 + n;
  OrderCtl *ctl = (OrderCtl *)(((uintptr_t)(fr1 + n) + 255) & ~(uintptr_t)255);
  const uint32_t vgrid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, c.num_cus * 8ull));
  k_init_deg<<<(n + 255) / 256, 256, 0, c.stream>>>(c.off.as<uint32_t>(), n, deg);
  KCDS_LAUNCH_CHECK();
  OrderCtl h0{};
  h0.minval = DEAD;
  HIP_CHECK(hipMemcpyAsync(ctl, &h0, sizeof h0, hipMemcpyHostToDevice, c.stream));
  const uint32_t *off = c.off.as<uint32_t>(), *nbr = c.nbr.as<uint32_t>();
  uint32_t *rank = c.rank.as<uint32_t>(), *inv = c.inv.as<uint32_t>();
  uint32_t *hdone = (uint32_t *)c.host_scratch(64);
  // the number of steps is a property of the graph: start from the last count
  uint32_t batch = c.order_steps ? (uint32_t)std::min<uint64_t>(c.order_steps, 1u << 20) : 32;
  uint64_t rounds = 0;
  while (true) {
    for (uint32_t r = 0; r < batch; ++r) {
      k_ko_work<<<vgrid, 256, 0, c.stream>>>(ctl, n, off, nbr, deg, fr0, fr1, rank, inv);
      k_ko_ctl<<<1, 1, 0, c.stream>>>(ctl, n);
      KCDS_LAUNCH_CHECK();
    }
    rounds += batch;
    HIP_CHECK(hipMemcpyAsync(hdone, &ctl->done, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (*hdone) break;
    // every PROCESS step ranks at least one vertex; an OPEN step is followed
    // by a PROCESS step or by one OPEN step that succeeds
    if (rounds > 3ull * n + 8) fail(KCDS_E_HIP, "ordering did not finish within %llu sub-rounds",
                                    (unsigned long long)rounds);
    batch = std::min<uint32_t>(std::max<uint32_t>(batch, 32) * 2, 128);
  }
  c.order_steps = rounds;
}

void orient_from_rank(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.doff, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * ((size_t)c.nnz + 1));
  uint32_t *dout = c.tmp1.as<uint32_t>(), *tmp = c.tmp0.as<uint32_t>();
  if (n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 3) / 4, (uint64_t)c.num_cus * 40);
    k_dag_gather<<<grid, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(), c.rank.as<uint32_t>(), n,
                                             dout, tmp);
    KCDS_LAUNCH_CHECK();
  }
  scan_u32_to_u32(c, dout, c.doff.as<uint32_t>(), n);
  // read back arcs and max out-degree
  buf_reserve(c, c.scal, 64 * sizeof(ull));
  uint32_t *dmax_d = (uint32_t *)(c.scal.as<ull>() + 60);
  max_u32(c, dout, n, dmax_d);
  uint32_t *h = (uint32_t *)c.host_scratch(16);
  HIP_CHECK(hipMemcpyAsync(h, c.doff.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(h + 1, dmax_d, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  c.darcs = h[0];
  c.dmax = h[1];
  if (c.dmax > SORT_WAVE)
    fail(KCDS_E_LIMIT, "DAG out-degree %u exceeds %u (orientation not degeneracy-bounded)", c.dmax, SORT_WAVE);
  buf_reserve(c, c.dnbr, sizeof(uint32_t) * (c.darcs + 1));
  if (n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 3) / 4, (uint64_t)c.num_cus * 40);
    k_dag_copy<<<grid, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.inv.as<uint32_t>(), n, c.doff.as<uint32_t>(),
                                           tmp, c.dnbr.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.w2 = 0;  // the listing pass reports the out-wedge sum it stages
}

void rank_from_listing_order(Ctx &c, const uint32_t *lo_dev) {
  const uint32_t n = c.n;
  if (!n) return;
  k_rank_from_lo<<<(n + 255) / 256, 256, 0, c.stream>>>(lo_dev, n, c.rank.as<uint32_t>(), c.inv.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
}

}  // namespace kcds
