// Clique flow network and exact integer push-relabel max-flow.
//
// Reference: createFlowNetworkOffset / createFlowNetwork
// (src/helpers.cu:1148-1264) build, per component, the network with nodes
// [vertices | cliques | s | t] and arcs s->v (cap clique-degree), v->t
// (cap k*alpha), v->c (cap 1), c->v (cap k-1), each paired with its reverse
// (flowIndex); preFlow / pushRelabel / globalRelabel (src/helpers.cu:
// 1266-1422) run a tile-per-node push-relabel in double precision.
//
// Here capacities are int64 and scaled by q for alpha = p/q:
//   s->v q*deg(v), v->t k*p, v->c q, c->v q*(k-1),
// so min cut < q*k*Tc iff some vertex set is denser than alpha.  The arc
// layout needs only one scan (of the vertex arc counts):
//   vertex v : [v->t, v->s, v->c ...]   (2 + deg(v) arcs)
//   clique c : [c->member_0 .. c->member_{k-1}]
//   s        : [s->v for every v],  t : [t->v for every v]
// One persistent cooperative kernel runs the max-preflow phase: a global
// relabel (level-synchronous BFS from t over residual arcs gives exact
// distance labels, unreachable nodes get height N), then synchronous
// push/relabel sweeps (wave per vertex node, thread per clique node), until
// a global relabel finds no node with excess that can still reach t.  The
// excess at t is then the max-flow value, and the vertices that cannot reach
// t form the source side of the maximal minimum cut.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

typedef long long i64;

struct FlowState {
  uint32_t barrier;
  uint32_t fsize[2];
  uint32_t active;
  uint32_t changed[2];  // double-buffered: slot of sweep i is cleared during sweep i-1
  uint32_t phases;
  uint32_t status;  // 0 running, 1 done, 2 phase cap reached
};

constexpr uint32_t SWEEPS = 24;
constexpr uint32_t MAX_PHASES = 1u << 16;

uint32_t grid1d(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 16));
}

__global__ void k_flow_deg(uint64_t Tc, uint32_t k, const uint32_t *ccl, uint32_t *deg) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Tc * k;
       i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&deg[ccl[i]], 1u);
}

__global__ void k_flow_varcs(uint32_t nc, const uint32_t *deg, uint32_t *cntv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x)
    cntv[v] = 2 + deg[v];
}

// node offsets for cliques, s, t; vertex arcs to t and s; s and t arcs
__global__ void k_flow_fixed(uint32_t nc, uint64_t Tc, uint32_t k, i64 p, i64 q, uint32_t *off, const uint32_t *deg,
                             uint32_t *head, uint32_t *rev, i64 *res, i64 *excess) {
  const uint32_t CA = off[nc];
  const uint32_t SA = CA + (uint32_t)(Tc * k), TA = SA + nc;
  const uint32_t s = nc + (uint32_t)Tc, t = s + 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= Tc; c += stride)
    off[nc + c] = CA + (uint32_t)(c * k);
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += stride) {
    const uint32_t a = off[v];
    const i64 cd = (i64)deg[v];
    // v->t, cap k*p; reverse t->v
    head[a] = t;
    rev[a] = TA + (uint32_t)v;
    res[a] = (i64)k * p;
    head[TA + v] = (uint32_t)v;
    rev[TA + v] = a;
    res[TA + v] = 0;
    // s->v saturated by the preflow: residual 0 forward, q*deg backward
    head[a + 1] = s;
    rev[a + 1] = SA + (uint32_t)v;
    res[a + 1] = q * cd;
    head[SA + v] = (uint32_t)v;
    rev[SA + v] = a + 1;
    res[SA + v] = 0;
    excess[v] = q * cd;
  }
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < Tc + 2; c += stride)
    excess[nc + c] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    off[s] = SA;
    off[t] = TA;
    off[t + 1] = TA + nc;
  }
}

// v->c (cap q) and c->v (cap q(k-1)) pairs
__global__ void k_flow_cliques(uint32_t nc, uint64_t Tc, uint32_t k, i64 q, const uint32_t *ccl, const uint32_t *off,
                               uint32_t *fill, uint32_t *head, uint32_t *rev, i64 *res) {
  const uint32_t CA = off[nc];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Tc * k;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = i / k;
    const uint32_t v = ccl[i];
    const uint32_t slot = atomicAdd(&fill[v], 1u);
    const uint32_t a = off[v] + 2 + slot;
    const uint32_t b = CA + (uint32_t)i;
    head[a] = nc + (uint32_t)c;
    rev[a] = b;
    res[a] = q;
    head[b] = v;
    rev[b] = a;
    res[b] = q * (i64)(k - 1);
  }
}

__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ i64 ld64(const i64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add64(i64 *p, i64 v) { atomicAdd((ull *)p, (ull)v); }

// Pushes from u along the admissible arc with the lowest head, or relabels.
// best = (height << 32) | local arc index of the lowest residual neighbour.
__device__ __forceinline__ bool discharge_step(uint32_t u, ull best, const uint32_t *off, const uint32_t *head,
                                               const uint32_t *rev, i64 *res, i64 *excess, uint32_t *height) {
  const i64 e = ld64(&excess[u]);
  if (best == ~0ull) return false;
  const uint32_t hmin = (uint32_t)(best >> 32);
  const uint32_t a = off[u] + (uint32_t)(best & 0xFFFFFFFFu);
  const uint32_t hu = ld32(&height[u]);
  if (hu > hmin) {
    const i64 r = ld64(&res[a]);
    const i64 d = e < r ? e : r;
    if (d <= 0) return false;
    add64(&res[a], -d);
    add64(&res[rev[a]], d);
    add64(&excess[u], -d);
    add64(&excess[head[a]], d);
  } else {
    __hip_atomic_store(&height[u], hmin + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

__global__ __launch_bounds__(1024) void k_push_relabel(uint32_t nc, uint64_t Tc, const uint32_t *off,
                                                       const uint32_t *head, const uint32_t *rev, i64 *res,
                                                       i64 *excess, uint32_t *height, uint32_t *fr0, uint32_t *fr1,
                                                       FlowState *st) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / KCDS_WAVE, nwaves = nthreads / KCDS_WAVE;
  const uint32_t N = nc + (uint32_t)Tc + 2, s = N - 2, t = N - 1;
  uint32_t epoch = 0;
  for (uint32_t phase = 0;; ++phase) {
    // ---- global relabel: exact distances to t in the residual graph ----
    for (uint32_t x = tid; x < N; x += nthreads) height[x] = (x == t) ? 0u : N;
    if (tid == 0) {
      fr0[0] = t;
      st->fsize[0] = 1;
      st->fsize[1] = 0;
      st->active = 0;
      st->changed[0] = 0;
      st->changed[1] = 0;
    }
    grid_barrier(&st->barrier, gridDim.x, epoch);
    int cur = 0;
    for (uint32_t dist = 0;; ++dist) {
      const uint32_t fs = ld32(&st->fsize[cur]);
      if (fs == 0) break;
      uint32_t *fr = cur ? fr1 : fr0, *nx = cur ? fr0 : fr1;
      for (uint32_t i = wid; i < fs; i += nwaves) {
        const uint32_t y = fr[i];
        for (uint32_t a = off[y] + lane; a < off[y + 1]; a += KCDS_WAVE) {
          const uint32_t x = head[a];
          if (x == s || ld64(&res[rev[a]]) <= 0) continue;  // x -> y has no residual capacity
          if (ld32(&height[x]) != N) continue;
          if (atomicCAS(&height[x], N, dist + 1) == N) nx[atomicAdd(&st->fsize[cur ^ 1], 1u)] = x;
        }
      }
      grid_barrier(&st->barrier, gridDim.x, epoch);
      if (tid == 0) st->fsize[cur] = 0;
      grid_barrier(&st->barrier, gridDim.x, epoch);
      cur ^= 1;
    }
    // ---- any node with excess that can still reach t? ----
    uint32_t act = 0;
    for (uint32_t x = tid; x < s; x += nthreads) act += (ld64(&excess[x]) > 0 && height[x] < N);
    act = wave_sum_u32(act);
    if (lane == 0 && act) atomicAdd(&st->active, act);
    grid_barrier(&st->barrier, gridDim.x, epoch);
    const uint32_t active = ld32(&st->active);
    if (active == 0 || phase + 1 >= MAX_PHASES) {
      if (tid == 0) {
        st->status = active == 0 ? 1u : 2u;
        st->phases = phase + 1;
      }
      break;
    }
    // ---- push / relabel sweeps ----
    for (uint32_t sweep = 0; sweep < SWEEPS; ++sweep) {
      const uint32_t slot = sweep & 1;
      uint32_t changed = 0;
      // vertex nodes: one wave each (their arc lists can be long)
      for (uint32_t u = wid; u < nc; u += nwaves) {
        if (ld64(&excess[u]) <= 0 || ld32(&height[u]) >= N) continue;
        ull best = ~0ull;
        const uint32_t b = off[u], e = off[u + 1];
        for (uint32_t a = b + lane; a < e; a += KCDS_WAVE)
          if (ld64(&res[a]) > 0) {
            const ull cand = ((ull)ld32(&height[head[a]]) << 32) | (ull)(a - b);
            best = cand < best ? cand : best;
          }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const ull other = __shfl_xor(best, o, 64);
          best = other < best ? other : best;
        }
        if (lane == 0) changed += discharge_step(u, best, off, head, rev, res, excess, height);
      }
      // clique nodes: k arcs, one thread each
      for (uint32_t u = nc + tid; u < s; u += nthreads) {
        if (ld64(&excess[u]) <= 0 || ld32(&height[u]) >= N) continue;
        ull best = ~0ull;
        const uint32_t b = off[u], e = off[u + 1];
        for (uint32_t a = b; a < e; ++a)
          if (ld64(&res[a]) > 0) {
            const ull cand = ((ull)ld32(&height[head[a]]) << 32) | (ull)(a - b);
            best = cand < best ? cand : best;
          }
        changed += discharge_step(u, best, off, head, rev, res, excess, height);
      }
      changed = wave_sum_u32(changed);
      if (lane == 0 && changed) atomicAdd(&st->changed[slot], changed);
      grid_barrier(&st->barrier, gridDim.x, epoch);
      // every thread reads this sweep's slot after the barrier; the other slot
      // was last read before it, so it can be cleared for the next sweep now
      const uint32_t any = ld32(&st->changed[slot]);
      if (tid == 0) st->changed[slot ^ 1] = 0;
      if (any == 0) break;
    }
  }
}

// source side of the maximal min cut: vertices whose height stayed N
__global__ void k_flow_side(uint32_t nc, const uint32_t *height, uint32_t N, uint8_t *side, ull *cnt) {
  ull s = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x) {
    const uint8_t in = height[v] >= N ? 1 : 0;
    side[v] = in;
    s += in;
  }
  s = wave_sum_u64(s);
  if (lane_id() == 0 && s) atomicAdd(&cnt[0], s);
}

__global__ void k_flow_side_cliques(uint64_t Tc, uint32_t k, const uint32_t *ccl, const uint8_t *side, ull *cnt) {
  ull s = 0;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < Tc;
       c += (uint64_t)gridDim.x * blockDim.x) {
    bool all = true;
    for (uint32_t j = 0; j < k; ++j) all = all && side[ccl[c * k + j]];
    s += all;
  }
  s = wave_sum_u64(s);
  if (lane_id() == 0 && s) atomicAdd(&cnt[1], s);
}

__global__ void k_flow_mark(uint32_t nc, const uint32_t *verts, const uint8_t *side, uint8_t *inset) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x)
    if (side[v]) inset[verts[v]] = 1;
}

}  // namespace

FlowOut flow_component(Ctx &c, uint32_t d, uint64_t p, uint64_t q) {
  CompData &g = c.cd;
  FlowBufs &f = c.fb;
  const uint32_t k = c.k;
  const uint32_t nc = g.h_voff[d + 1] - g.h_voff[d];
  const uint64_t Tc = g.h_coff[d + 1] - g.h_coff[d];
  const uint64_t N64 = (uint64_t)nc + Tc + 2, A64 = 4ull * nc + 2ull * k * Tc;
  if (N64 >= 0x7FFFFFFFull || A64 >= 0xFFFFFFFFull)
    fail(KCDS_E_LIMIT, "flow network of component %u too large (%llu nodes, %llu arcs)", d,
         (unsigned long long)N64, (unsigned long long)A64);
  // int64 capacities: q*deg(v) and the total q*k*Tc must not overflow
  if (q > 0 && (double)q * (double)k * (double)(Tc + 1) > 9.0e18)
    fail(KCDS_E_LIMIT, "alpha denominator %llu too large for int64 capacities", (unsigned long long)q);
  const uint32_t N = (uint32_t)N64, A = (uint32_t)A64;
  const uint32_t *ccl = g.ccl.as<uint32_t>() + g.h_coff[d] * k;
  buf_reserve(c, f.off, sizeof(uint32_t) * ((size_t)N + 2));
  buf_reserve(c, f.head, sizeof(uint32_t) * ((size_t)A + 1));
  buf_reserve(c, f.rev, sizeof(uint32_t) * ((size_t)A + 1));
  buf_reserve(c, f.res, sizeof(i64) * ((size_t)A + 1));
  buf_reserve(c, f.excess, sizeof(i64) * ((size_t)N + 1));
  buf_reserve(c, f.height, sizeof(uint32_t) * ((size_t)N + 1));
  buf_reserve(c, f.deg, sizeof(uint32_t) * ((size_t)nc + 1) * 2);
  buf_reserve(c, f.front, sizeof(uint32_t) * ((size_t)N + 1) * 2);
  buf_reserve(c, f.side, (size_t)nc + 1);
  buf_reserve(c, f.scal, 256);
  uint32_t *deg = f.deg.as<uint32_t>(), *cntv = deg + nc + 1;
  HIP_CHECK(hipMemsetAsync(deg, 0, sizeof(uint32_t) * ((size_t)nc + 1) * 2, c.stream));
  if (Tc) {
    k_flow_deg<<<grid1d(Tc * k), 256, 0, c.stream>>>(Tc, k, ccl, deg);
    KCDS_LAUNCH_CHECK();
  }
  if (nc) {
    k_flow_varcs<<<grid1d(nc), 256, 0, c.stream>>>(nc, deg, cntv);
    KCDS_LAUNCH_CHECK();
  }
  scan_u32_to_u32(c, cntv, f.off.as<uint32_t>(), nc);
  k_flow_fixed<<<grid1d(std::max<uint64_t>(nc, Tc + 2)), 256, 0, c.stream>>>(
      nc, Tc, k, (i64)p, (i64)q, f.off.as<uint32_t>(), deg, f.head.as<uint32_t>(), f.rev.as<uint32_t>(),
      f.res.as<i64>(), f.excess.as<i64>());
  KCDS_LAUNCH_CHECK();
  // per-vertex fill counters reuse cntv
  HIP_CHECK(hipMemsetAsync(cntv, 0, sizeof(uint32_t) * ((size_t)nc + 1), c.stream));
  if (Tc) {
    k_flow_cliques<<<grid1d(Tc * k), 256, 0, c.stream>>>(nc, Tc, k, (i64)q, ccl, f.off.as<uint32_t>(), cntv,
                                                         f.head.as<uint32_t>(), f.rev.as<uint32_t>(),
                                                         f.res.as<i64>());
    KCDS_LAUNCH_CHECK();
  }
  FlowState *st = (FlowState *)f.scal.p;
  HIP_CHECK(hipMemsetAsync(st, 0, sizeof(FlowState), c.stream));
  int per_cu = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_push_relabel, 1024, 0));
  if (per_cu < 1) fail(KCDS_E_HIP, "k_push_relabel cannot be resident");
  // small networks use a proportionally small grid (fewer blocks in each barrier)
  uint32_t want = (uint32_t)std::min<uint64_t>((N64 + 1023) / 1024, (uint64_t)c.num_cus * std::min(per_cu, 2));
  uint32_t grid = std::max<uint32_t>(1, want);
  const uint32_t *off = f.off.as<uint32_t>(), *head = f.head.as<uint32_t>(), *rev = f.rev.as<uint32_t>();
  i64 *res = f.res.as<i64>(), *excess = f.excess.as<i64>();
  uint32_t *height = f.height.as<uint32_t>(), *fr0 = f.front.as<uint32_t>(), *fr1 = fr0 + N + 1;
  uint32_t ncc = nc;
  uint64_t Tcc = Tc;
  void *args[] = {(void *)&ncc, (void *)&Tcc, (void *)&off, (void *)&head, (void *)&rev, (void *)&res,
                  (void *)&excess, (void *)&height, (void *)&fr0, (void *)&fr1, (void *)&st};
  HIP_CHECK(hipLaunchCooperativeKernel((const void *)k_push_relabel, dim3(grid), dim3(1024), args, 0, c.stream));
  ull *cnt = (ull *)((char *)f.scal.p + 128);
  HIP_CHECK(hipMemsetAsync(cnt, 0, 2 * sizeof(ull), c.stream));
  if (nc) {
    k_flow_side<<<grid1d(nc), 256, 0, c.stream>>>(nc, height, N, f.side.as<uint8_t>(), cnt);
    KCDS_LAUNCH_CHECK();
  }
  if (Tc) {
    k_flow_side_cliques<<<grid1d(Tc), 256, 0, c.stream>>>(Tc, k, ccl, f.side.as<uint8_t>(), cnt);
    KCDS_LAUNCH_CHECK();
  }
  struct H {
    FlowState st;
    ull cnt[2];
    i64 flow;
  };
  H *h = (H *)c.host_scratch(sizeof(H));
  HIP_CHECK(hipMemcpyAsync(&h->st, st, sizeof(FlowState), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(h->cnt, cnt, 2 * sizeof(ull), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipMemcpyAsync(&h->flow, excess + (N - 1), sizeof(i64), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (h->st.status != 1) fail(KCDS_E_LIMIT, "push-relabel did not converge within %u phases", MAX_PHASES);
  FlowOut o;
  o.value = h->flow;
  o.side_n = h->cnt[0];
  o.side_T = h->cnt[1];
  o.phases = h->st.phases;
  f.last_comp = d;
  f.last_p = p;
  f.last_q = q;
  return o;
}

void flow_mark_side(Ctx &c, uint32_t d) {
  CompData &g = c.cd;
  const uint32_t nc = g.h_voff[d + 1] - g.h_voff[d];
  if (!nc) return;
  k_flow_mark<<<grid1d(nc), 256, 0, c.stream>>>(nc, g.verts.as<uint32_t>() + g.h_voff[d], c.fb.side.as<uint8_t>(),
                                                c.inset.as<uint8_t>());
  KCDS_LAUNCH_CHECK();
}

}  // namespace kcds
