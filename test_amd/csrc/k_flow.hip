// Clique flow network and exact integer push-relabel max-flow, one workgroup
// per component.
//
// Reference: createFlowNetworkOffset / createFlowNetwork
// (src/helpers.cu:1148-1264) build, per component, the network with nodes
// [vertices | cliques | s | t] and arcs s->v (cap clique-degree), v->t
// (cap k*alpha), v->c (cap 1), c->v (cap k-1), each paired with its reverse
// (flowIndex); preFlow / pushRelabel / globalRelabel / updateFlownetwork
// (src/helpers.cu:1266-1509) run a tile-per-node push-relabel in double
// precision and the host bisects alpha component after component
// (main.cu:903-1268).
//
// Here capacities are int64 and scaled by q for alpha = p/q:
//   s->v q*deg(v), v->t k*p, v->c q, c->v q*(k-1),
// so min cut < q*k*Tc iff some vertex set is denser than alpha.  The arc
// layout needs only one scan (of the vertex arc counts):
//   vertex v : [v->t, v->s, v->c ...]   (2 + deg(v) arcs)
//   clique c : [c->member_0 .. c->member_{k-1}]
//   s        : [s->v for every v],  t : [t->v for every v]
// A persistent grid of 1024-thread workgroups pops components from a queue;
// each builds its component's network in a private slice of the scratch
// arrays and runs the whole exact search on the device:
//   alpha = p0/q0; repeat { max-preflow at alpha; if flow < q*k*Tc then
//   alpha := density of the maximal min-cut source side } until no set is
//   denser than alpha.
// The max-preflow phase alternates a global relabel (level-synchronous BFS
// from t over residual arcs: exact distance labels, height N for nodes that
// cannot reach t) with synchronous discharge sweeps (wave per vertex node,
// thread per clique node; a node pushes along all of its residual arcs to
// the lowest neighbour height at once, or relabels) until a global relabel
// finds no node with excess that can still reach t.  The excess at t is then
// the max-flow value and the vertices that cannot reach t form the source
// side of the maximal minimum cut.  A component's arrays belong to one
// workgroup, so its memory operations use workgroup scope; workgroup
// barriers wait for each wave's own memory operations first.
#include <algorithm>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

typedef long long i64;

struct FlowState {
  uint32_t fsize[2];
  uint32_t active;
  uint32_t changed[3];  // per-sweep change counters in rotation
  uint32_t phases;
  uint32_t status;      // 0 running, 1 done, 2 phase cap reached
};

constexpr uint32_t SWEEPS = 24;
constexpr uint32_t MAX_PHASES = 1u << 14;

uint32_t grid1d(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 16));
}

// Every array of a component's network is private to the one workgroup that
// solves it, so memory operations use workgroup scope: all its waves share
// one CU and its L1, and the atomics stay in the cache hierarchy instead of
// going to the device-wide coherence point.
__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ i64 ld64(const i64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void cst32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void add64(i64 *p, i64 v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t add32(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ bool cas32(uint32_t *p, uint32_t expect, uint32_t v) {
  return __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
}

constexpr uint32_t NO_ARC = 0xFFFFFFFFu;

// Discharge of vertex node u by its wave: push the excess along every
// residual arc whose head has the lowest height, split in arc order, or
// relabel u when no neighbour is lower.  Only u's wave lowers excess[u] and
// the residuals of u's arcs, so reading them first and pushing at most that
// much keeps every value non-negative.  Returns whether anything changed.
__device__ bool discharge_vertex(uint32_t u, const uint32_t *off, const uint32_t *head, const uint32_t *rev, i64 *res,
                                 i64 *excess, uint32_t *height) {
  const uint32_t lane = lane_id();
  const i64 e = ld64(&excess[u]);
  const uint32_t b = ld32(&off[u]), en = ld32(&off[u + 1]);
  uint32_t hmin = NO_ARC;
  for (uint32_t a = b + lane; a < en; a += KCDS_WAVE)
    if (ld64(&res[a]) > 0) hmin = min(hmin, ld32(&height[ld32(&head[a])]));
  hmin = wave_min_u32(hmin);
  if (hmin == NO_ARC) return false;
  if (ld32(&height[u]) <= hmin) {
    if (lane == 0) cst32(&height[u], hmin + 1);
    return true;
  }
  i64 left = e;
  for (uint32_t a0 = b; a0 < en && left > 0; a0 += KCDS_WAVE) {
    const uint32_t a = a0 + lane;
    i64 r = 0;
    uint32_t x = 0;
    if (a < en) {
      r = ld64(&res[a]);
      if (r > 0) {
        x = ld32(&head[a]);
        if (ld32(&height[x]) != hmin) r = 0;
      } else {
        r = 0;
      }
    }
    const ull incl = wave_incl_scan_u64((ull)r);
    const i64 before = (i64)incl - r;
    const i64 d = r > 0 && left > before ? min(r, left - before) : 0;
    if (d > 0) {
      add64(&res[a], -d);
      add64(&res[ld32(&rev[a])], d);
      add64(&excess[x], d);
    }
    const i64 tot = (i64)__shfl(incl, 63, 64);
    left -= min(tot, left);
  }
  if (lane == 0 && e > left) add64(&excess[u], -(e - left));
  return e > left;
}

// Discharge of clique node u (k arcs) by one thread, with the same rule.
__device__ bool discharge_clique(uint32_t u, const uint32_t *off, const uint32_t *head, const uint32_t *rev, i64 *res,
                                 i64 *excess, uint32_t *height) {
  const i64 e = ld64(&excess[u]);
  const uint32_t b = ld32(&off[u]), en = ld32(&off[u + 1]);
  uint32_t hmin = NO_ARC;
  for (uint32_t a = b; a < en; ++a)
    if (ld64(&res[a]) > 0) hmin = min(hmin, ld32(&height[ld32(&head[a])]));
  if (hmin == NO_ARC) return false;
  if (ld32(&height[u]) <= hmin) {
    cst32(&height[u], hmin + 1);
    return true;
  }
  i64 left = e;
  for (uint32_t a = b; a < en && left > 0; ++a) {
    const i64 r = ld64(&res[a]);
    if (r <= 0) continue;
    const uint32_t x = ld32(&head[a]);
    if (ld32(&height[x]) != hmin) continue;
    const i64 d = min(r, left);
    add64(&res[a], -d);
    add64(&res[ld32(&rev[a])], d);
    add64(&excess[x], d);
    left -= d;
  }
  if (e > left) add64(&excess[u], -(e - left));
  return e > left;
}

// Max-preflow phase of push-relabel by one workgroup.  On return
// st->status = 1 (done) or 2 (phase cap) and heights hold exact distances to
// t (N for nodes that cannot reach t).
__device__ void max_preflow(uint32_t nc, uint64_t Tc, const uint32_t *off, const uint32_t *head, const uint32_t *rev,
                            i64 *res, i64 *excess, uint32_t *height, uint32_t *fr0, uint32_t *fr1, FlowState *st,
                            const uint32_t tid, const uint32_t nthreads) {
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / KCDS_WAVE, nwaves = nthreads / KCDS_WAVE;
  const uint32_t N = nc + (uint32_t)Tc + 2, s = N - 2, t = N - 1;
  auto sync = [&]() {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores and atomics are done
    __syncthreads();
  };
  for (uint32_t phase = 0;; ++phase) {
    // ---- global relabel: exact distances to t in the residual graph ----
    for (uint32_t x = tid; x < N; x += nthreads) cst32(&height[x], (x == t) ? 0u : N);
    if (tid == 0) {
      cst32(&fr0[0], t);
      cst32(&st->fsize[0], 1u);
      cst32(&st->fsize[1], 0u);
      cst32(&st->active, 0u);
      cst32(&st->changed[0], 0u);
      cst32(&st->changed[1], 0u);
      cst32(&st->changed[2], 0u);
    }
    sync();
    int cur = 0;
    for (uint32_t dist = 0;; ++dist) {
      const uint32_t fs = ld32(&st->fsize[cur]);
      if (fs == 0) break;
      uint32_t *fr = cur ? fr1 : fr0, *nx = cur ? fr0 : fr1;
      for (uint32_t i = wid; i < fs; i += nwaves) {
        const uint32_t y = ld32(&fr[i]);
        const uint32_t ye = ld32(&off[y + 1]);
        for (uint32_t a = ld32(&off[y]) + lane; a < ye; a += KCDS_WAVE) {
          const uint32_t x = ld32(&head[a]);
          if (x == s || ld64(&res[ld32(&rev[a])]) <= 0) continue;  // x -> y has no residual capacity
          if (ld32(&height[x]) != N) continue;
          if (cas32(&height[x], N, dist + 1)) cst32(&nx[atomicAdd(&st->fsize[cur ^ 1], 1u)], x);
        }
      }
      sync();
      if (tid == 0) cst32(&st->fsize[cur], 0u);
      sync();
      cur ^= 1;
    }
    // ---- any node with excess that can still reach t? ----
    uint32_t act = 0;
    for (uint32_t x = tid; x < s; x += nthreads) act += (ld64(&excess[x]) > 0 && ld32(&height[x]) < N);
    act = wave_sum_u32(act);
    if (lane == 0 && act) atomicAdd(&st->active, act);
    sync();
    const uint32_t active = ld32(&st->active);
    if (active == 0 || phase + 1 >= MAX_PHASES) {
      if (tid == 0) {
        cst32(&st->status, active == 0 ? 1u : 2u);
        cst32(&st->phases, phase + 1);
      }
      sync();
      break;
    }
    // ---- push / relabel sweeps ----
    for (uint32_t sweep = 0; sweep < SWEEPS; ++sweep) {
      // three counters in rotation: sweep i adds to slot i%3; slot (i+1)%3 is
      // cleared during sweep i, after every thread read it (sweep i-2) and
      // before anyone adds to it (sweep i+1, after the barrier of sweep i)
      const uint32_t slot = sweep % 3;
      if (tid == 0) cst32(&st->changed[(sweep + 1) % 3], 0u);
      uint32_t changed = 0;
      // vertex nodes: one wave each (their arc lists can be long)
      for (uint32_t u = wid; u < nc; u += nwaves) {
        if (ld64(&excess[u]) <= 0 || ld32(&height[u]) >= N) continue;
        const bool ch = discharge_vertex(u, off, head, rev, res, excess, height);
        if (lane == 0) changed += ch;
      }
      // clique nodes: k arcs, one thread each
      for (uint32_t u = nc + tid; u < s; u += nthreads) {
        if (ld64(&excess[u]) <= 0 || ld32(&height[u]) >= N) continue;
        changed += discharge_clique(u, off, head, rev, res, excess, height);
      }
      changed = wave_sum_u32(changed);
      if (lane == 0 && changed) atomicAdd(&st->changed[slot], changed);
      sync();
      const uint32_t any = ld32(&st->changed[slot]);
      if (any == 0) break;
    }
    sync();  // the next phase's resets follow every thread's last read
  }
}

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
  i64 flow;            // value of the last max-flow
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

__global__ __launch_bounds__(1024) void k_dinkelbach_block(const CompTask *tasks, uint32_t ntasks, uint32_t *qctr,
                                                           uint32_t k, uint32_t max_iters, const uint32_t *ccl_all, uint32_t *noff_all,
                                                           i64 *excess_all, uint32_t *height_all, uint32_t *fr0_all,
                                                           uint32_t *fr1_all, uint32_t *head_all, uint32_t *rev_all,
                                                           i64 *res_all, uint8_t *cside, CompOut *outs) {
  __shared__ FlowState st;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t task_id;
  __shared__ ull red[2];
  __shared__ ull alpha[2];
  __shared__ uint32_t done;
  __shared__ i64 last_flow;
  const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
  QueueCursor qc = queue_start(blockIdx.x);
  while (true) {
    if (tid == 0) task_id = queue_next(qctr, ntasks, qc);
    block_sync();
    const uint32_t ti = task_id;
    block_sync();
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
    block_sync();
    for (ull i = tid; i < Tc * k; i += nthreads) add32(&fr0[ccl[i]], 1u);
    block_sync();
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nc; b0 += nthreads) {
      const uint32_t v = b0 + tid;
      const uint32_t cntv = v < nc ? 2 + ld32(&fr0[v]) : 0;
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
      const uint32_t a = ld32(&off[v]);
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
    block_sync();
    for (ull i = tid; i < Tc * k; i += nthreads) {
      const uint32_t v = ccl[i];
      const uint32_t slot = add32(&fr1[v], 1u);
      const uint32_t a = ld32(&off[v]) + 2 + slot, b = CA + (uint32_t)i;
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
    block_sync();
    uint32_t flows = 0, status = 1;
    for (uint32_t it = 0;; ++it) {
      const i64 p = (i64)alpha[0], q = (i64)alpha[1];
      // residual capacities for alpha = p/q after the preflow saturates s->v
      for (uint32_t v = tid; v < nc; v += nthreads) {
        const uint32_t a = ld32(&off[v]);
        const i64 deg = (i64)(ld32(&off[v + 1]) - a - 2);
        res[a] = (i64)k * p;
        res[TA + v] = 0;
        res[a + 1] = q * deg;
        res[SA + v] = 0;
        excess[v] = q * deg;
      }
      for (ull i = tid; i < Tc * k; i += nthreads) {
        const uint32_t b = CA + (uint32_t)i;
        res[b] = q * (i64)(k - 1);
        res[ld32(&rev[b])] = q;
      }
      for (ull c = nc + tid; c < N; c += nthreads) excess[c] = 0;
      if (tid == 0) {
        red[0] = 0;
        red[1] = 0;
        cst32(&st.status, 0u);
      }
      block_sync();
      max_preflow(nc, Tc, off, head, rev, res, excess, height, fr0, fr1, &st, tid, nthreads);
      ++flows;
      if (ld32(&st.status) != 1) {
        status = 2;
        break;
      }
      // maximal min-cut side: vertices that cannot reach t
      ull sn = 0, sT = 0;
      for (uint32_t v = tid; v < nc; v += nthreads) sn += ld32(&height[v]) >= N;
      for (ull c = tid; c < Tc; c += nthreads) {
        bool all = true;
        for (uint32_t j = 0; j < k; ++j) all = all && ld32(&height[ccl[c * k + j]]) >= N;
        sT += all;
      }
      sn = wave_sum_u64(sn);
      sT = wave_sum_u64(sT);
      if (lane_id() == 0) {
        if (sn) atomicAdd(&red[0], sn);
        if (sT) atomicAdd(&red[1], sT);
      }
      block_sync();
      if (tid == 0) {
        const i64 flow = ld64(&excess[t]);
        const ull full = (ull)q * k * Tc;
        last_flow = flow;
        if ((ull)flow >= full || it + 1 >= max_iters) {
          done = 1;  // no set denser than alpha (or a single max-flow was asked for)
        } else if (red[0] == 0 || it + 1 >= MAX_DINKELBACH) {
          done = 2;
        } else {
          const ull g = gcd_u64(red[1], red[0]);
          alpha[0] = red[1] / g;
          alpha[1] = red[0] / g;
        }
      }
      block_sync();
      if (done) {
        if (done == 2) status = 3;
        break;
      }
    }
    for (uint32_t v = tid; v < nc; v += nthreads) cside[tk.voff + v] = (status == 1 && ld32(&height[v]) >= N) ? 1 : 0;
    if (tid == 0) {
      CompOut o;
      o.p = alpha[0];
      o.q = alpha[1];
      o.side_n = status == 1 ? red[0] : 0;
      o.side_T = status == 1 ? red[1] : 0;
      o.flow = last_flow;
      o.flows = flows;
      o.status = status;
      outs[ti] = o;
    }
    block_sync();
  }
}

// source side of the maximal min cut: vertices whose height stayed N
__global__ void k_mark_cside(uint32_t nc, const uint32_t *verts, const uint8_t *cside, uint8_t *inset) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nc; v += gridDim.x * blockDim.x)
    if (cside[v]) inset[verts[v]] = 1;
}

// the same for every core vertex of the components flagged in sel
__global__ void k_mark_csides(uint32_t nv, const uint32_t *verts, const uint32_t *cid, const uint8_t *sel,
                              const uint8_t *cside, uint8_t *inset) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x)
    if (cside[i] && sel[cid[verts[i]]]) inset[verts[i]] = 1;
}

// cside = 1 for every core vertex of the components flagged in all
__global__ void k_set_csides(uint32_t nv, const uint32_t *verts, const uint32_t *cid, const uint8_t *all,
                             uint8_t *cside) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x)
    if (all[cid[verts[i]]]) cside[i] = 1;
}

}  // namespace

void mark_component_sides(Ctx &c, const std::vector<uint8_t> &sel) {
  CompData &g = c.cd;
  const uint32_t nv = g.h_voff.back();
  if (!nv) return;
  buf_reserve(c, c.tmp3, sel.size() + 8);
  HIP_CHECK(hipMemcpyAsync(c.tmp3.p, sel.data(), sel.size(), hipMemcpyHostToDevice, c.stream));
  k_mark_csides<<<grid1d(nv), 256, 0, c.stream>>>(nv, g.verts.as<uint32_t>(), g.cid.as<uint32_t>(),
                                                  c.tmp3.as<uint8_t>(), c.fb.cside.as<uint8_t>(),
                                                  c.inset.as<uint8_t>());
  KCDS_LAUNCH_CHECK();
  c.sync();  // sel lives on the host
}

void mark_component_sides_all(Ctx &c, const std::vector<uint8_t> &all) {
  CompData &g = c.cd;
  const uint32_t nv = g.h_voff.back();
  if (!nv) return;
  buf_reserve(c, c.tmp3, all.size() + 8);
  HIP_CHECK(hipMemcpyAsync(c.tmp3.p, all.data(), all.size(), hipMemcpyHostToDevice, c.stream));
  k_set_csides<<<grid1d(nv), 256, 0, c.stream>>>(nv, g.verts.as<uint32_t>(), g.cid.as<uint32_t>(),
                                                 c.tmp3.as<uint8_t>(), c.fb.cside.as<uint8_t>());
  KCDS_LAUNCH_CHECK();
  c.sync();  // all lives on the host
}

void mark_component_side(Ctx &c, uint32_t d) {
  CompData &g = c.cd;
  const uint32_t nc = g.h_voff[d + 1] - g.h_voff[d];
  if (!nc) return;
  k_mark_cside<<<grid1d(nc), 256, 0, c.stream>>>(nc, g.verts.as<uint32_t>() + g.h_voff[d],
                                                 c.fb.cside.as<uint8_t>() + g.h_voff[d], c.inset.as<uint8_t>());
  KCDS_LAUNCH_CHECK();
}

// Enqueues the workgroup-per-component launch on st (after the context
// stream's work so far); block_finish waits for it and fills the results.
void block_launch(Ctx &c, const std::vector<BlockTask> &in, uint32_t max_iters, hipStream_t st, BlockRun &run) {
  run.order.clear();
  run.dout = nullptr;
  run.st = st;
  if (in.empty()) return;
  CompData &g = c.cd;
  FlowBufs &f = c.fb;
  const uint32_t k = c.k;
  if (f.cside.bytes < (size_t)g.h_voff.back() + 8) {
    buf_reserve(c, f.cside, (size_t)g.h_voff.back() + 8);
    HIP_CHECK(hipMemsetAsync(f.cside.p, 0, (size_t)g.h_voff.back() + 8, c.stream));
  }
  // largest components first: the launch lasts as long as its last task, and
  // a big component taken late would end far behind the rest (C4: 126 ms for
  // 4,096 components in component order)
  std::vector<uint32_t> &order = run.order;
  order.resize(in.size());
  for (size_t i = 0; i < in.size(); ++i) order[i] = (uint32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return g.h_coff[in[a].d + 1] - g.h_coff[in[a].d] > g.h_coff[in[b].d + 1] - g.h_coff[in[b].d];
  });
  std::vector<CompTask> tasks(in.size());
  ull nodes = 0, arcs = 0;
  for (size_t j = 0; j < in.size(); ++j) {
    const size_t i = order[j];
    const uint32_t d = in[i].d;
    CompTask &t = tasks[j];
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
      fail(KCDS_E_LIMIT, "flow network of component %u too large (%llu cliques)", d, (unsigned long long)t.Tc);
    // int64 capacities: the largest total flow is q * k * Tc
    if ((double)in[i].q0 * k * (double)(t.Tc + 1) > 9.0e18 || (double)t.nc * k * (double)(t.Tc + 1) > 9.0e18)
      fail(KCDS_E_LIMIT, "capacities of component %u exceed int64", d);
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
  buf_reserve(c, f.scal, QUEUE_WORDS * sizeof(uint32_t));
  CompTask *dt = f.b_tasks.as<CompTask>();
  CompOut *dout = (CompOut *)(((uintptr_t)(dt + tasks.size()) + 63) & ~(uintptr_t)63);
  run.dout = dout;
  if (st != c.stream) {  // after everything enqueued on the context stream so far
    if (!c.ffork) HIP_CHECK(hipEventCreateWithFlags(&c.ffork, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(c.ffork, c.stream));
    HIP_CHECK(hipStreamWaitEvent(st, c.ffork, 0));
  }
  // tasks live on the host until the copy completes
  HIP_CHECK(hipMemcpyAsync(dt, tasks.data(), sizeof(CompTask) * tasks.size(), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipStreamSynchronize(st));
  uint32_t *qctr = (uint32_t *)f.scal.p;
  HIP_CHECK(hipMemsetAsync(qctr, 0, QUEUE_WORDS * sizeof(uint32_t), st));
  int per_cu = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_dinkelbach_block, 1024, 0));
  if (per_cu < 1) fail(KCDS_E_HIP, "k_dinkelbach_block cannot be resident");
  const uint32_t grid =
      (uint32_t)std::min<uint64_t>(tasks.size(), (uint64_t)c.num_cus * (uint64_t)std::min(per_cu, 2));
  k_dinkelbach_block<<<grid, 1024, 0, st>>>(
      dt, (uint32_t)tasks.size(), qctr, k, max_iters, g.ccl.as<uint32_t>(), f.b_off.as<uint32_t>(),
      f.b_excess.as<i64>(), f.b_height.as<uint32_t>(), f.b_fr0.as<uint32_t>(), f.b_fr1.as<uint32_t>(),
      f.b_head.as<uint32_t>(), f.b_rev.as<uint32_t>(), f.b_res.as<i64>(), f.cside.as<uint8_t>(), dout);
  KCDS_LAUNCH_CHECK();
}

void block_finish(Ctx &c, const std::vector<BlockTask> &in, BlockRun &run, std::vector<BlockResult> &out) {
  out.assign(in.size(), BlockResult{});
  if (in.empty()) return;
  const std::vector<uint32_t> &order = run.order;
  std::vector<CompOut> h(in.size());
  HIP_CHECK(hipMemcpyAsync(h.data(), run.dout, sizeof(CompOut) * in.size(), hipMemcpyDeviceToHost, run.st));
  HIP_CHECK(hipStreamSynchronize(run.st));
  for (size_t j = 0; j < h.size(); ++j) {
    const size_t i = order[j];
    if (h[j].status != 1)
      fail(KCDS_E_LIMIT, "block solver did not converge on component %u (status %u)", in[i].d, h[j].status);
    out[i].p = h[j].p;
    out[i].q = h[j].q;
    out[i].side_n = h[j].side_n;
    out[i].side_T = h[j].side_T;
    out[i].flow = h[j].flow;
    out[i].flows = h[j].flows;
  }
}

void dinkelbach_blocks(Ctx &c, const std::vector<BlockTask> &in, std::vector<BlockResult> &out, uint32_t max_iters) {
  BlockRun run;
  block_launch(c, in, max_iters, c.stream, run);
  block_finish(c, in, run, out);
}

}  // namespace kcds
