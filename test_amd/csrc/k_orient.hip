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

constexpr uint32_t SORT_LDS_MAX = 8192;  // out-lists sorted in LDS (32 KB)

__global__ void k_rank_from_lo(const uint32_t *lo, uint32_t n, uint32_t *rank, uint32_t *inv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    uint32_t r = n - lo[v];  // lo in 1..n, checked on the host to be a permutation
    rank[v] = r;
    inv[r] = v;
  }
}

// wave per vertex: out-degree in the DAG, indexed by new id
__global__ __launch_bounds__(256) void k_outdeg(const uint32_t *off, const uint32_t *nbr,
                                                const uint32_t *rank, uint32_t n, uint32_t *dout) {
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / KCDS_WAVE;
  const uint32_t nw = gridDim.x * blockDim.x / KCDS_WAVE;
  for (uint32_t v = wid; v < n; v += nw) {
    uint32_t b = off[v], e = off[v + 1], r = rank[v];
    uint32_t c = 0;
    for (uint32_t i = b + lane; i < e; i += KCDS_WAVE) c += rank[nbr[i]] > r;
    c = wave_sum_u32(c);
    if (lane == 0) dout[r] = c;
  }
}

// wave per vertex: gather the out-list into LDS, sort it by rank counting
// (position = number of smaller entries; entries are distinct), write it.
__global__ __launch_bounds__(64) void k_dag_fill(const uint32_t *off, const uint32_t *nbr,
                                                 const uint32_t *rank, uint32_t n, const uint32_t *doff,
                                                 uint32_t *dnbr) {
  __shared__ uint32_t buf[SORT_LDS_MAX];
  const uint32_t lane = lane_id();
  for (uint32_t v = blockIdx.x; v < n; v += gridDim.x) {
    uint32_t b = off[v], e = off[v + 1], r = rank[v];
    uint32_t ob = doff[r], d = doff[r + 1] - ob;
    if (d == 0) continue;
    uint32_t cnt = 0;
    for (uint32_t i0 = b; i0 < e; i0 += KCDS_WAVE) {
      uint32_t i = i0 + lane;
      uint32_t x = (i < e) ? rank[nbr[i]] : 0u;
      bool keep = (i < e) && x > r;
      ull bal = __ballot(keep);
      if (keep) buf[cnt + lanes_below(bal)] = x;
      cnt += (uint32_t)__popcll(bal);
    }
    wave_sync();
    for (uint32_t i = lane; i < d; i += KCDS_WAVE) {
      uint32_t x = buf[i], pos = 0;
      for (uint32_t j = 0; j < d; ++j) {
        uint32_t y = buf[j];
        pos += (y < x) || (y == x && j < i);  // stable: equal entries keep their order
      }
      dnbr[ob + pos] = x;
    }
    wave_sync();
  }
}

// out-wedge sum W2 = sum over arcs (r,u) of outdeg(u), and a per-root work
// estimate (used by the listing kernels to order roots)
__global__ void k_w2(const uint32_t *doff, const uint32_t *dnbr, uint64_t arcs, ull *w2) {
  ull s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < arcs;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = dnbr[i];
    s += doff[u + 1] - doff[u];
  }
  s = wave_sum_u64(s);
  if (lane_id() == 0) atomicAdd(w2, s);
}

// ---------------------------------------------------------------------------
// GPU degeneracy order: level-synchronous k-core peel in one persistent
// cooperative kernel.  Level L removes every live vertex of degree <= L in
// sub-rounds; each sub-round's vertices take the next ranks (in frontier
// order), then decrement their live neighbours; a neighbour whose degree
// falls from L+1 to L joins the next sub-round.  Empty levels are skipped by
// a min-reduction over live degrees.

struct PeelState {
  uint32_t level;
  uint32_t next_rank;
  uint32_t fsize[2];
  uint32_t minval;
  uint32_t barrier;
};

__global__ __launch_bounds__(1024) void k_kcore_order(const uint32_t *off, const uint32_t *nbr, uint32_t n,
                                                      uint32_t *deg, uint32_t *rank, uint32_t *inv,
                                                      uint32_t *fr0, uint32_t *fr1, PeelState *st) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / KCDS_WAVE, nwaves = nthreads / KCDS_WAVE;
  const uint32_t DEAD = 0xFFFFFFFFu;
  uint32_t epoch = 0;
  int cur = 0;
  while (true) {
    // all ranks assigned?
    uint32_t done = __hip_atomic_load(&st->next_rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (done >= n) break;
    uint32_t fs = __hip_atomic_load(&st->fsize[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fs == 0) {
      // new level: L = min live degree (never below the current level)
      for (uint32_t v = tid; v < n; v += nthreads) {
        uint32_t d = deg[v];
        if (d != DEAD) atomicMin(&st->minval, d);
      }
      grid_barrier(&st->barrier, gridDim.x, epoch);
      uint32_t L = max(__hip_atomic_load(&st->minval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(&st->level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      uint32_t *fr = cur ? fr1 : fr0;
      for (uint32_t v = tid; v < n; v += nthreads) {
        uint32_t d = deg[v];
        if (d != DEAD && d <= L) fr[atomicAdd(&st->fsize[cur], 1u)] = v;
      }
      grid_barrier(&st->barrier, gridDim.x, epoch);
      if (tid == 0) {
        st->level = L;
        st->minval = DEAD;
      }
      fs = __hip_atomic_load(&st->fsize[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      grid_barrier(&st->barrier, gridDim.x, epoch);
    }
    const uint32_t L = __hip_atomic_load(&st->level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t base = __hip_atomic_load(&st->next_rank, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t *fr = cur ? fr1 : fr0;
    uint32_t *nx = cur ? fr0 : fr1;
    // take ranks and leave the graph
    for (uint32_t i = tid; i < fs; i += nthreads) {
      uint32_t v = fr[i];
      rank[v] = base + i;
      inv[base + i] = v;
      deg[v] = DEAD;
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    // decrement live neighbours (wave per removed vertex)
    for (uint32_t i = wid; i < fs; i += nwaves) {
      uint32_t v = fr[i];
      for (uint32_t j = off[v] + lane; j < off[v + 1]; j += KCDS_WAVE) {
        uint32_t u = nbr[j];
        if (__hip_atomic_load(&deg[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == DEAD) continue;
        uint32_t old = atomicSub(&deg[u], 1u);
        if (old == L + 1) nx[atomicAdd(&st->fsize[cur ^ 1], 1u)] = u;
      }
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    if (tid == 0) {
      st->next_rank = base + fs;
      st->fsize[cur] = 0;
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    cur ^= 1;
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
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * (size_t)n * 3 + sizeof(PeelState) + 64);
  uint32_t *deg = c.tmp0.as<uint32_t>();
  uint32_t *fr0 = deg + n, *fr1 = fr0 + n;
  PeelState *st = (PeelState *)(((uintptr_t)(fr1 + n) + 63) & ~(uintptr_t)63);
  k_init_deg<<<(n + 255) / 256, 256, 0, c.stream>>>(c.off.as<uint32_t>(), n, deg);
  KCDS_LAUNCH_CHECK();
  HIP_CHECK(hipMemsetAsync(st, 0, sizeof(PeelState), c.stream));
  fill_u32(c, &st->minval, 0xFFFFFFFFu, 1);
  int per_cu = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_kcore_order, 1024, 0));
  if (per_cu < 1) fail(KCDS_E_HIP, "k_kcore_order cannot be resident");
  uint32_t grid = (uint32_t)c.num_cus * (uint32_t)std::min(per_cu, 2);
  const uint32_t *off = c.off.as<uint32_t>(), *nbr = c.nbr.as<uint32_t>();
  uint32_t *rank = c.rank.as<uint32_t>(), *inv = c.inv.as<uint32_t>();
  void *args[] = {(void *)&off, (void *)&nbr, (void *)&n, (void *)&deg, (void *)&rank,
                  (void *)&inv, (void *)&fr0, (void *)&fr1, (void *)&st};
  HIP_CHECK(hipLaunchCooperativeKernel((const void *)k_kcore_order, dim3(grid), dim3(1024), args, 0, c.stream));
}

void orient_from_rank(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.doff, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t *dout = c.tmp1.as<uint32_t>();
  if (n) {
    uint32_t grid = std::min<uint32_t>((n + 3) / 4, 65535u * 8);
    k_outdeg<<<grid, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(), c.rank.as<uint32_t>(), n,
                                         dout);
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
  if (c.dmax > SORT_LDS_MAX)
    fail(KCDS_E_LIMIT, "DAG out-degree %u exceeds %u (orientation not degeneracy-bounded)", c.dmax,
         SORT_LDS_MAX);
  buf_reserve(c, c.dnbr, sizeof(uint32_t) * (c.darcs + 1));
  if (n) {
    uint32_t grid = std::min<uint32_t>(n, 65535u * 4);
    k_dag_fill<<<grid, 64, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(), c.rank.as<uint32_t>(), n,
                                          c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  ull *w2d = c.scal.as<ull>() + 59;
  fill_u64(c, w2d, 0, 1);
  if (c.darcs) {
    uint32_t grid = (uint32_t)std::min<uint64_t>((c.darcs + 255) / 256, 65535u * 4);
    k_w2<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(), c.darcs, w2d);
    KCDS_LAUNCH_CHECK();
  }
  ull *hw = (ull *)c.host_scratch(8);
  HIP_CHECK(hipMemcpyAsync(hw, w2d, 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  c.w2 = *hw;
}

void rank_from_listing_order(Ctx &c, const uint32_t *lo_dev) {
  const uint32_t n = c.n;
  if (!n) return;
  k_rank_from_lo<<<(n + 255) / 256, 256, 0, c.stream>>>(lo_dev, n, c.rank.as<uint32_t>(), c.inv.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
}

}  // namespace kcds
