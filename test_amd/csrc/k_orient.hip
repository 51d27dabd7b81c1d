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
// GPU degeneracy order: level-synchronous k-core peel as a fixed sequence of
// step kernels (min / open / collect / rank / decrement / finish) driven by
// device-side control state, enqueued by the host in batches.  Level L
// removes every live vertex of degree <= L in sub-rounds; each sub-round's
// vertices take the next ranks (in frontier order), then decrement their live
// neighbours; a neighbour whose degree falls from L+1 to L joins the next
// sub-round.  Empty levels are skipped by a min-reduction over live degrees.
// Kernel boundaries order the steps; no grid-wide barrier is needed.

struct OrderCtl {
  uint32_t level;
  uint32_t minval;
  uint32_t next_rank;
  uint32_t fsize[2];
  uint32_t cur;
  uint32_t done;
  uint32_t opening;
};

constexpr uint32_t DEAD = 0xFFFFFFFFu;

__global__ void k_ko_min(OrderCtl *ctl, uint32_t n, const uint32_t *deg) {
  if (ctl->done || ctl->fsize[ctl->cur] != 0) return;
  uint32_t m = DEAD;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) m = min(m, deg[v]);
  m = wave_min_u32(m);
  if (lane_id() == 0 && m != DEAD) atomicMin(&ctl->minval, m);
}

__global__ void k_ko_open(OrderCtl *ctl) {
  if (ctl->done || ctl->fsize[ctl->cur] != 0) return;
  if (ctl->minval == DEAD) {
    ctl->done = 1;
    return;
  }
  ctl->level = max(ctl->level, ctl->minval);
  ctl->minval = DEAD;
  ctl->opening = 1;
}

__global__ void k_ko_collect(OrderCtl *ctl, uint32_t n, const uint32_t *deg, uint32_t *fr0, uint32_t *fr1) {
  if (ctl->done || !ctl->opening) return;
  const uint32_t L = ctl->level, cur = ctl->cur;
  uint32_t *fr = cur ? fr1 : fr0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t d = deg[v];
    if (d != DEAD && d <= L) fr[atomicAdd(&ctl->fsize[cur], 1u)] = v;
  }
}

// the frontier takes the next ranks and leaves the graph
__global__ void k_ko_rank(OrderCtl *ctl, const uint32_t *fr0, const uint32_t *fr1, uint32_t *deg, uint32_t *rank,
                          uint32_t *inv) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur, fs = ctl->fsize[cur], base = ctl->next_rank;
  const uint32_t *fr = cur ? fr1 : fr0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < fs; i += gridDim.x * blockDim.x) {
    const uint32_t v = fr[i];
    rank[v] = base + i;
    inv[base + i] = v;
    deg[v] = DEAD;
  }
}

// wave per removed vertex: decrement live neighbours
__global__ void k_ko_decrement(OrderCtl *ctl, const uint32_t *off, const uint32_t *nbr, uint32_t *deg,
                               uint32_t *fr0, uint32_t *fr1) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur, fs = ctl->fsize[cur], L = ctl->level;
  const uint32_t *fr = cur ? fr1 : fr0;
  uint32_t *nx = cur ? fr0 : fr1;
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / KCDS_WAVE;
  const uint32_t nw = gridDim.x * blockDim.x / KCDS_WAVE;
  for (uint32_t i = wid; i < fs; i += nw) {
    const uint32_t v = fr[i];
    for (uint32_t j = off[v] + lane; j < off[v + 1]; j += KCDS_WAVE) {
      const uint32_t u = nbr[j];
      if (__hip_atomic_load(&deg[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == DEAD) continue;
      const uint32_t old = atomicSub(&deg[u], 1u);
      if (old == L + 1) nx[atomicAdd(&ctl->fsize[cur ^ 1], 1u)] = u;
    }
  }
}

__global__ void k_ko_finish(OrderCtl *ctl, uint32_t n) {
  if (ctl->done) return;
  const uint32_t cur = ctl->cur;
  ctl->next_rank += ctl->fsize[cur];
  ctl->fsize[cur] = 0;
  ctl->opening = 0;
  ctl->cur = cur ^ 1;
  if (ctl->next_rank >= n) ctl->done = 1;
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
  uint32_t batch = 8;
  uint64_t rounds = 0;
  while (true) {
    for (uint32_t r = 0; r < batch; ++r) {
      k_ko_min<<<vgrid, 256, 0, c.stream>>>(ctl, n, deg);
      k_ko_open<<<1, 1, 0, c.stream>>>(ctl);
      k_ko_collect<<<vgrid, 256, 0, c.stream>>>(ctl, n, deg, fr0, fr1);
      k_ko_rank<<<vgrid, 256, 0, c.stream>>>(ctl, fr0, fr1, deg, rank, inv);
      k_ko_decrement<<<vgrid, 256, 0, c.stream>>>(ctl, off, nbr, deg, fr0, fr1);
      k_ko_finish<<<1, 1, 0, c.stream>>>(ctl, n);
      KCDS_LAUNCH_CHECK();
    }
    rounds += batch;
    HIP_CHECK(hipMemcpyAsync(hdone, &ctl->done, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (*hdone) break;
    // every sub-round ranks at least one vertex
    if (rounds > (uint64_t)n + 1) fail(KCDS_E_HIP, "ordering did not finish within %llu sub-rounds",
                                       (unsigned long long)rounds);
    batch = std::min<uint32_t>(batch * 2, 512);
  }
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
