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

// Max-preflow phase of push-relabel over threads [tid, nthreads) of either
// the whole grid (GRID: cooperative launch, software grid barrier) or one
// workgroup (block barrier).  On return st->status = 1 (done) or 2 (phase cap)
// and heights hold exact distances to t (N for nodes that cannot reach t).
template <bool GRID>
__device__ void max_preflow(uint32_t nc, uint64_t Tc, const uint32_t *off, const uint32_t *head, const uint32_t *rev,
                            i64 *res, i64 *excess, uint32_t *height, uint32_t *fr0, uint32_t *fr1, FlowState *st,
                            const uint32_t tid, const uint32_t nthreads, uint32_t &epoch) {
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / KCDS_WAVE, nwaves = nthreads / KCDS_WAVE;
  const uint32_t N = nc + (uint32_t)Tc + 2, s = N - 2, t = N - 1;
  auto sync = [&]() {
    if constexpr (GRID) grid_barrier(&st->barrier, gridDim.x, epoch);
    else {
      __threadfence_block();
      __syncthreads();
    }
  };
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
    sync();
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
      sync();
      if (tid == 0) st->fsize[cur] = 0;
      sync();
      cur ^= 1;
    }
    // ---- any node with excess that can still reach t? ----
    uint32_t act = 0;
    for (uint32_t x = tid; x < s; x += nthreads) act += (ld64(&excess[x]) > 0 && height[x] < N);
    act = wave_sum_u32(act);
    if (lane == 0 && act) atomicAdd(&st->active, act);
    sync();
    const uint32_t active = ld32(&st->active);
    if (active == 0 || phase + 1 >= MAX_PHASES) {
      if (tid == 0) {
        st->status = active == 0 ? 1u : 2u;
        st->phases = phase + 1;
      }
      sync();
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
      sync();
      // every thread reads this sweep's slot after the barrier; the other slot
      // was last read before it, so it can be cleared for the next sweep now
      const uint32_t any = ld32(&st->changed[slot]);
      if (tid == 0) st->changed[slot ^ 1] = 0;
      if (any == 0) break;
    }
    sync();  // the next phase's resets follow every thread's last read
  }
}

__global__ __launch_bounds__(1024) void k_push_relabel(uint32_t nc, uint64_t Tc, const uint32_t *off,
                                                       const uint32_t *head, const uint32_t *rev, i64 *res,
                                                       i64 *excess, uint32_t *height, uint32_t *fr0, uint32_t *fr1,
                                                       FlowState *st) {
  uint32_t epoch = 0;
  max_preflow<true>(nc, Tc, off, head, rev, res, excess, height, fr0, fr1, st,
                    blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, epoch);
}

// ---------------------------------------------------------------------------
// Block-per-component exact search (batched solves).  A persistent grid of
// 1024-thread workgroups pops components from a queue; each workgroup builds
// its component's network in a private slice of the scratch arrays and runs
// the whole Dinkelbach iteration on the device:
//   alpha = p0/q0; repeat { max-flow at alpha; if flow < q*k*Tc then
//   alpha := density of the maximal min-cut source side } until no set is
//   denser than alpha.
// The result is the component's best density reachable from alpha0 and the
// maximal min-cut side at that alpha (empty when no set reaches alpha0).

struct CompTask {
  ull Tc;         // cliques of the component
  ull coff;       // first clique row in ccl
  ull node_base;  // slice of the node arrays
  ull arc_base;   // slice of the arc arrays
  ull p0, q0;     // starting alpha
  uint32_t nc;    // vertices of the component
  uint32_t voff;  // first vertex in verts / cside
};

struct CompOut {
  ull p, q;            // final alpha, reduced
  ull side_n, side_T;  // maximal min-cut side at the final alpha
  uint32_t flows;
  uint32_t status;     // 1 ok, 2 push-relabel phase cap, 3 iteration cap
};

constexpr uint32_t MAX_DINKELBACH = 64;

__device__ ull gcd_u64(ull a, ull b) {
  while (b) {
    const ull t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// exclusive scan of one value per thread over a 1024-thread block
__device__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t *total, uint32_t *wsum) {
  const uint32_t lane = threadIdx.x % KCDS_WAVE, w = threadIdx.x / KCDS_WAVE;
  const uint32_t incl = wave_incl_scan_u32(v);
  if (lane == KCDS_WAVE - 1) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t i = 0; i < blockDim.x / KCDS_WAVE; ++i) {
    if (i < w) before += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(1024) void k_dinkelbach_block(const CompTask *tasks, uint32_t ntasks, uint32_t *qctr,
                                                           uint32_t k, const uint32_t *ccl_all, uint32_t *noff_all,
                                                           i64 *excess_all, uint32_t *height_all, uint32_t *fr0_all,
                                                           uint32_t *fr1_all, uint32_t *head_all, uint32_t *rev_all,
                                                           i64 *res_all, uint8_t *cside, CompOut *outs) {
  __shared__ FlowState st;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t task_id;
  __shared__ ull red[2];
  __shared__ ull alpha[2];
  __shared__ uint32_t done;
  const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
  uint32_t epoch = 0;
  while (true) {
    if (tid == 0) task_id = atomicAdd(qctr, 1u);
    __syncthreads();
    const uint32_t ti = task_id;
    __syncthreads();
    if (ti >= ntasks) break;
    const CompTask tk = tasks[ti];
    const uint32_t nc = tk.nc;
    const ull Tc = tk.Tc;
    const uint32_t N = nc + (uint32_t)Tc + 2, s = N - 2, t = N - 1;
    const uint32_t *ccl = ccl_all + tk.coff * k;
    uint32_t *off = noff_all + tk.node_base, *height = height_all + tk.node_base;
    uint32_t *fr0 = fr0_all + tk.node_base, *fr1 = fr1_all + tk.node_base;
    i64 *excess = excess_all + tk.node_base;
    uint32_t *head = head_all + tk.arc_base, *rev = rev_all + tk.arc_base;
    i64 *res = res_all + tk.arc_base;
    // ---- network topology ----
    for (uint32_t v = tid; v < nc; v += nthreads) fr0[v] = 0;
    __syncthreads();
    for (ull i = tid; i < Tc * k; i += nthreads) atomicAdd(&fr0[ccl[i]], 1u);
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nc; b0 += nthreads) {
      const uint32_t v = b0 + tid;
      const uint32_t cntv = v < nc ? 2 + fr0[v] : 0;
      uint32_t tot;
      const uint32_t ex = block_excl_scan_1024(cntv, &tot, wsum);
      if (v < nc) off[v] = carry + ex;
      carry += tot;
    }
    const uint32_t CA = carry;  // = 2*nc + k*Tc
    const uint32_t SA = CA + (uint32_t)(Tc * k), TA = SA + nc;
    for (ull c = tid; c <= Tc; c += nthreads) off[nc + c] = CA + (uint32_t)(c * k);
    if (tid == 0) {
      off[s] = SA;
      off[t] = TA;
      off[t + 1] = TA + nc;
    }
    for (uint32_t v = tid; v < nc; v += nthreads) {
      const uint32_t a = off[v];
      head[a] = t;
      rev[a] = TA + v;
      head[TA + v] = v;
      rev[TA + v] = a;
      head[a + 1] = s;
      rev[a + 1] = SA + v;
      head[SA + v] = v;
      rev[SA + v] = a + 1;
      fr1[v] = 0;
    }
    __syncthreads();
    for (ull i = tid; i < Tc * k; i += nthreads) {
      const uint32_t v = ccl[i];
      const uint32_t slot = atomicAdd(&fr1[v], 1u);
      const uint32_t a = off[v] + 2 + slot, b = CA + (uint32_t)i;
      head[a] = nc + (uint32_t)(i / k);
      rev[a] = b;
      head[b] = v;
      rev[b] = a;
    }
    if (tid == 0) {
      alpha[0] = tk.p0;
      alpha[1] = tk.q0;
      done = 0;
    }
    __syncthreads();
    uint32_t flows = 0, status = 1;
    for (uint32_t it = 0;; ++it) {
      const i64 p = (i64)alpha[0], q = (i64)alpha[1];
      // residual capacities for alpha = p/q after the preflow saturates s->v
      for (uint32_t v = tid; v < nc; v += nthreads) {
        const uint32_t a = off[v];
        const i64 deg = (i64)(off[v + 1] - a - 2);
        res[a] = (i64)k * p;
        res[TA + v] = 0;
        res[a + 1] = q * deg;
        res[SA + v] = 0;
        excess[v] = q * deg;
      }
      for (ull i = tid; i < Tc * k; i += nthreads) {
        const uint32_t b = CA + (uint32_t)i;
        res[b] = q * (i64)(k - 1);
        res[rev[b]] = q;
      }
      for (ull c = nc + tid; c < N; c += nthreads) excess[c] = 0;
      if (tid == 0) {
        red[0] = 0;
        red[1] = 0;
        st.status = 0;
      }
      __syncthreads();
      max_preflow<false>(nc, Tc, off, head, rev, res, excess, height, fr0, fr1, &st, tid, nthreads, epoch);
      ++flows;
      if (st.status != 1) {
        status = 2;
        break;
      }
      // maximal min-cut side: vertices that cannot reach t
      ull sn = 0, sT = 0;
      for (uint32_t v = tid; v < nc; v += nthreads) sn += height[v] >= N;
      for (ull c = tid; c < Tc; c += nthreads) {
        bool all = true;
        for (uint32_t j = 0; j < k; ++j) all = all && height[ccl[c * k + j]] >= N;
        sT += all;
      }
      sn = wave_sum_u64(sn);
      sT = wave_sum_u64(sT);
      if (lane_id() == 0) {
        if (sn) atomicAdd(&red[0], sn);
        if (sT) atomicAdd(&red[1], sT);
      }
      __syncthreads();
      if (tid == 0) {
        const i64 flow = excess[t];
        const ull full = (ull)q * k * Tc;
        if ((ull)flow >= full) {
          done = 1;  // no set denser than alpha: this side is the maximal tight set
        } else if (red[0] == 0 || it + 1 >= MAX_DINKELBACH) {
          done = 2;
        } else {
          const ull g = gcd_u64(red[1], red[0]);
          alpha[0] = red[1] / g;
          alpha[1] = red[0] / g;
        }
      }
      __syncthreads();
      if (done) {
        if (done == 2) status = 3;
        break;
      }
    }
    for (uint32_t v = tid; v < nc; v += nthreads) cside[tk.voff + v] = (status == 1 && height[v] >= N) ? 1 : 0;
    if (tid == 0) {
      CompOut o;
      o.p = alpha[0];
      o.q = alpha[1];
      o.side_n = status == 1 ? red[0] : 0;
      o.side_T = status == 1 ? red[1] : 0;
      o.flows = flows;
      o.status = status;
      outs[ti] = o;
    }
    __syncthreads();
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


namespace {
__global__ void k_copy_side(uint32_t nc, const uint8_t *side, uint8_t *dst) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x) dst[v] = side[v];
}
__global__ void k_mark_cside(uint32_t nc, const uint32_t *verts, const uint8_t *cside, uint8_t *inset) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x)
    if (cside[v]) inset[verts[v]] = 1;
}
}  // namespace

void flow_store_side(Ctx &c, uint32_t d) {
  CompData &g = c.cd;
  const uint32_t nc = g.h_voff[d + 1] - g.h_voff[d];
  if (!nc) return;
  k_copy_side<<<grid1d(nc), 256, 0, c.stream>>>(nc, c.fb.side.as<uint8_t>(), c.fb.cside.as<uint8_t>() + g.h_voff[d]);
  KCDS_LAUNCH_CHECK();
}

void mark_component_side(Ctx &c, uint32_t d) {
  CompData &g = c.cd;
  const uint32_t nc = g.h_voff[d + 1] - g.h_voff[d];
  if (!nc) return;
  k_mark_cside<<<grid1d(nc), 256, 0, c.stream>>>(nc, g.verts.as<uint32_t>() + g.h_voff[d],
                                                 c.fb.cside.as<uint8_t>() + g.h_voff[d], c.inset.as<uint8_t>());
  KCDS_LAUNCH_CHECK();
}

void dinkelbach_blocks(Ctx &c, const std::vector<BlockTask> &in, std::vector<BlockResult> &out) {
  out.assign(in.size(), BlockResult{});
  if (in.empty()) return;
  CompData &g = c.cd;
  FlowBufs &f = c.fb;
  const uint32_t k = c.k;
  std::vector<CompTask> tasks(in.size());
  ull nodes = 0, arcs = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    const uint32_t d = in[i].d;
    CompTask &t = tasks[i];
    t.nc = g.h_voff[d + 1] - g.h_voff[d];
    t.voff = g.h_voff[d];
    t.Tc = g.h_coff[d + 1] - g.h_coff[d];
    t.coff = g.h_coff[d];
    t.p0 = in[i].p0;
    t.q0 = in[i].q0;
    t.node_base = nodes;
    t.arc_base = arcs;
    nodes += (ull)t.nc + t.Tc + 4;
    arcs += 4ull * t.nc + 2ull * k * t.Tc;
    if ((ull)t.nc + t.Tc + 2 >= 0x7FFFFFFFull || 4ull * t.nc + 2ull * k * t.Tc >= 0xFFFFFFFFull)
      fail(KCDS_E_LIMIT, "component %u too large for the block solver", d);
  }
  buf_reserve(c, f.b_off, sizeof(uint32_t) * (nodes + 1));
  buf_reserve(c, f.b_excess, sizeof(i64) * (nodes + 1));
  buf_reserve(c, f.b_height, sizeof(uint32_t) * (nodes + 1));
  buf_reserve(c, f.b_fr0, sizeof(uint32_t) * (nodes + 1));
  buf_reserve(c, f.b_fr1, sizeof(uint32_t) * (nodes + 1));
  buf_reserve(c, f.b_head, sizeof(uint32_t) * (arcs + 1));
  buf_reserve(c, f.b_rev, sizeof(uint32_t) * (arcs + 1));
  buf_reserve(c, f.b_res, sizeof(i64) * (arcs + 1));
  buf_reserve(c, f.b_tasks, sizeof(CompTask) * tasks.size() + sizeof(CompOut) * tasks.size() + 256);
  CompTask *dt = f.b_tasks.as<CompTask>();
  CompOut *dout = (CompOut *)(((uintptr_t)(dt + tasks.size()) + 63) & ~(uintptr_t)63);
  HIP_CHECK(hipMemcpyAsync(dt, tasks.data(), sizeof(CompTask) * tasks.size(), hipMemcpyHostToDevice, c.stream));
  buf_reserve(c, f.scal, 256);
  uint32_t *qctr = (uint32_t *)((char *)f.scal.p + 192);
  HIP_CHECK(hipMemsetAsync(qctr, 0, sizeof(uint32_t), c.stream));
  int per_cu = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_dinkelbach_block, 1024, 0));
  if (per_cu < 1) fail(KCDS_E_HIP, "k_dinkelbach_block cannot be resident");
  const uint32_t grid =
      (uint32_t)std::min<uint64_t>(tasks.size(), (uint64_t)c.num_cus * (uint64_t)std::min(per_cu, 2));
  k_dinkelbach_block<<<grid, 1024, 0, c.stream>>>(
      dt, (uint32_t)tasks.size(), qctr, k, g.ccl.as<uint32_t>(), f.b_off.as<uint32_t>(), f.b_excess.as<i64>(),
      f.b_height.as<uint32_t>(), f.b_fr0.as<uint32_t>(), f.b_fr1.as<uint32_t>(), f.b_head.as<uint32_t>(),
      f.b_rev.as<uint32_t>(), f.b_res.as<i64>(), f.cside.as<uint8_t>(), dout);
  KCDS_LAUNCH_CHECK();
  std::vector<CompOut> h(tasks.size());
  HIP_CHECK(hipMemcpyAsync(h.data(), dout, sizeof(CompOut) * tasks.size(), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  for (size_t i = 0; i < h.size(); ++i) {
    if (h[i].status != 1)
      fail(KCDS_E_LIMIT, "block solver did not converge on component %u (status %u)", in[i].d, h[i].status);
    out[i].p = h[i].p;
    out[i].q = h[i].q;
    out[i].side_n = h[i].side_n;
    out[i].side_T = h[i].side_T;
    out[i].flows = h[i].flows;
  }
}

}  // namespace kcds
