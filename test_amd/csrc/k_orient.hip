// Degeneracy ordering and DAG orientation.
//
// Reference: src/graph.cpp:53-113 computes a Batagelj-Zaversnik peel on the
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
#include <cstring>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t SORT_WAVE = 1024;  // longest out-list a wave sorts in LDS

__global__ void k_rank_from_lo(const uint32_t *lo, uint32_t n, uint32_t *rank, uint32_t *inv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    uint32_t r = lo[v] - 1;  // lo in 1..n, checked on the host to be a permutation; arcs go up in lo
    rank[v] = r;
    inv[r] = v;
  }
}

// Position of x among the distinct entries a[0..cnt) (number of smaller
// entries); a is 16-byte aligned and padded with ~0 up to a multiple of 4, so
// it is read 4 entries per LDS access with two reads in flight.
__device__ __forceinline__ uint32_t rank_in(const uint32_t *a, uint32_t cnt, uint32_t x) {
  const uint4 *a4 = reinterpret_cast<const uint4 *>(a);
  const uint32_t n4 = (cnt + 3) / 4;
  uint32_t p0 = 0, p1 = 0, j = 0;
  for (; j + 2 <= n4; j += 2) {
    const uint4 u = a4[j], w = a4[j + 1];
    p0 += (u.x < x) + (u.y < x) + (u.z < x) + (u.w < x);
    p1 += (w.x < x) + (w.y < x) + (w.z < x) + (w.w < x);
  }
  if (j < n4) {
    const uint4 u = a4[j];
    p0 += (u.x < x) + (u.y < x) + (u.z < x) + (u.w < x);
  }
  return p0 + p1;
}

constexpr uint32_t HEAVY_DEG = 1024;   // vertices of higher degree are gathered by a whole workgroup
constexpr uint32_t HEAVY_SORT = LIST_WIDE_MAX;  // longest out-list the workgroup gather sorts
constexpr uint32_t HEAVY_THREADS = 1024;
constexpr uint32_t SMALL_DEG = 32;     // vertices of degree <= SMALL_DEG are gathered 64 to a wave
constexpr uint32_t TILE_CAP = 1024;    // arcs a tile pass holds in LDS (>= SMALL_DEG)
constexpr uint32_t GUNROLL = 4;        // 64-arc chunks whose loads are in flight together

// Appends v (where want) to list[] with one counter atomic per wave.
__device__ __forceinline__ void wave_push(bool want, uint32_t v, uint32_t *list, uint32_t *len) {
  const ull bal = __ballot(want);
  if (!bal) return;
  const uint32_t lead = (uint32_t)__builtin_ctzll(bal);
  uint32_t at = 0;
  if (lane_id() == lead) at = atomicAdd(len, (uint32_t)__popcll(bal));
  at = __shfl(at, lead, 64);
  if (want) list[at + lanes_below(bal)] = v;
}

struct TileLds {
  uint32_t buf[TILE_CAP];  // kept new ids of the pass, in arc order
  uint8_t sg[TILE_CAP];    // lane (vertex of the tile) each kept id belongs to
  uint32_t excl[64], b[64], r[64], kst[64], ken[64];
};

// Wave per tile of 64 consecutive vertices, one per lane.  The arcs of the
// tile's vertices of degree <= SMALL_DEG form one flat run (row after row);
// the wave walks it 64 arcs at a time (GUNROLL chunks of loads in flight),
// finds each arc's row by binary search over the row starts, keeps the new
// ids above the row's own and compacts them in LDS, which keeps rows
// contiguous.  A kept id's place in its sorted out-list is the number of
// smaller kept ids of the same row; it is written to tmp[off[v] ..), the
// start of v's undirected list, and dout[rank[v]] = the row's count.  Runs
// longer than TILE_CAP are split at row boundaries into passes.  Vertices of
// higher degree go to mid[] (<= HEAVY_DEG) or heavy[].
__global__ __launch_bounds__(256) void k_dag_gather_tiles(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ rank, uint32_t n,
    uint32_t *__restrict__ dout, uint32_t *__restrict__ tmp, uint32_t *__restrict__ mid, uint32_t *nmid,
    uint32_t *__restrict__ heavy, uint32_t *nheavy) {
  __shared__ TileLds sl[4];
  const uint32_t lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  TileLds &W = sl[wv];
  const uint32_t ntiles = (n + 63) / 64;
  for (uint32_t tile = blockIdx.x * 4 + wv; tile < ntiles; tile += gridDim.x * 4) {
    const uint32_t v = tile * 64 + lane;
    uint32_t b = 0, e = 0, r = 0;
    if (v < n) {
      b = off[v];
      e = off[v + 1];
      r = rank[v];
    }
    const uint32_t d = e - b;
    const bool small = v < n && d <= SMALL_DEG;
    wave_push(v < n && d > SMALL_DEG && d <= HEAVY_DEG, v, mid, nmid);
    wave_push(v < n && d > HEAVY_DEG, v, heavy, nheavy);
    const uint32_t ds = small ? d : 0u;
    const uint32_t incl = wave_incl_scan_u32(ds), excl = incl - ds;
    W.excl[lane] = excl;
    W.b[lane] = b;
    W.r[lane] = r;
    wave_sync();
    const uint32_t nl = min(64u, n - tile * 64);
    uint32_t a = 0;
    while (a < nl) {
      // pass: rows [a, bnd), whose arcs [F0, F1) fit TILE_CAP (row a alone always does)
      const uint32_t F0 = __shfl(excl, a, 64);
      const uint32_t bnd = (uint32_t)__popcll(__ballot(incl <= F0 + TILE_CAP));
      const uint32_t F1 = __shfl(incl, bnd - 1, 64);
      if (lane >= a && lane < bnd) {
        W.kst[lane] = 0;
        W.ken[lane] = 0;
      }
      uint32_t cnt = 0;
      for (uint32_t f0 = F0; f0 < F1; f0 += GUNROLL * KCDS_WAVE) {
        uint32_t u[GUNROLL], sr[GUNROLL];
#pragma unroll
        for (uint32_t q = 0; q < GUNROLL; ++q) {
          const uint32_t f = f0 + q * KCDS_WAVE + lane;
          uint32_t row = a;
          if (f < F1) {
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1)
              if (row + step < bnd && W.excl[row + step] <= f) row += step;
          }
          sr[q] = row;
          u[q] = f < F1 ? nbr[W.b[row] + (f - W.excl[row])] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < GUNROLL; ++q) u[q] = (f0 + q * KCDS_WAVE + lane < F1) ? rank[u[q]] : 0u;
#pragma unroll
        for (uint32_t q = 0; q < GUNROLL; ++q) {
          const bool keep = (f0 + q * KCDS_WAVE + lane < F1) && u[q] > W.r[sr[q]];
          const ull bal = __ballot(keep);
          if (keep) {
            const uint32_t at = cnt + lanes_below(bal);
            W.buf[at] = u[q];
            W.sg[at] = (uint8_t)sr[q];
          }
          cnt += (uint32_t)__popcll(bal);
        }
      }
      wave_sync();
      for (uint32_t i = lane; i < cnt; i += KCDS_WAVE) {  // row ranges of the kept list
        const uint32_t s = W.sg[i];
        if (i == 0 || W.sg[i - 1] != s) W.kst[s] = i;
        if (i + 1 == cnt || W.sg[i + 1] != s) W.ken[s] = i + 1;
      }
      wave_sync();
      for (uint32_t i = lane; i < cnt; i += KCDS_WAVE) {
        const uint32_t s = W.sg[i], x = W.buf[i], en = W.ken[s];
        uint32_t j = W.kst[s], p0 = 0, p1 = 0, p2 = 0, p3 = 0;
        for (; j + 4 <= en; j += 4) {
          p0 += W.buf[j] < x;
          p1 += W.buf[j + 1] < x;
          p2 += W.buf[j + 2] < x;
          p3 += W.buf[j + 3] < x;
        }
        for (; j < en; ++j) p0 += W.buf[j] < x;
        tmp[W.b[s] + p0 + p1 + p2 + p3] = x;
      }
      if (small && lane >= a && lane < bnd) dout[r] = W.ken[lane] - W.kst[lane];
      wave_sync();
      a = bnd;
    }
  }
}

// Sorts the distinct entries buf[0..cnt) (16-byte aligned, padded with ~0 to
// a multiple of 4) into dst[0..cnt): each lane holds M of them in registers
// and counts the smaller entries of the whole list, read once by broadcast.
template <uint32_t M>
__device__ __forceinline__ void wave_rank_sort(const uint32_t *buf, uint32_t cnt, uint32_t *__restrict__ dst) {
  const uint32_t lane = lane_id();
  uint32_t x[M], p[M];
#pragma unroll
  for (uint32_t q = 0; q < M; ++q) {
    const uint32_t i = lane + q * KCDS_WAVE;
    x[q] = i < cnt ? buf[i] : 0xFFFFFFFFu;
    p[q] = 0;
  }
  const uint4 *a4 = reinterpret_cast<const uint4 *>(buf);
  const uint32_t n4 = (cnt + 3) / 4;
  for (uint32_t j = 0; j < n4; ++j) {
    const uint4 u = a4[j];
#pragma unroll
    for (uint32_t q = 0; q < M; ++q) p[q] += (u.x < x[q]) + (u.y < x[q]) + (u.z < x[q]) + (u.w < x[q]);
  }
#pragma unroll
  for (uint32_t q = 0; q < M; ++q)
    if (lane + q * KCDS_WAVE < cnt) dst[p[q]] = x[q];
}

constexpr uint32_t MUNROLL = 8;  // 64-arc chunks in flight per mid vertex

// Wave per vertex of mid[] (degree in (SMALL_DEG, HEAVY_DEG]; 4 waves per
// block, 4 KB of LDS each): gathers the new ids of its later neighbours with
// MUNROLL chunks of loads in flight, sorts them (wave_rank_sort) and writes
// them to tmp[off[v] ..); dout[rank[v]] = count.  A list longer than
// SORT_WAVE is counted but not written (it cannot happen: a mid vertex's
// degree is at most HEAVY_DEG = SORT_WAVE).
__global__ __launch_bounds__(256) void k_dag_gather(const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ rank,
                                                    const uint32_t *__restrict__ mid, const uint32_t *nmid,
                                                    uint32_t *__restrict__ dout, uint32_t *__restrict__ tmp) {
  __shared__ __align__(16) uint32_t sbuf[4][SORT_WAVE + 4];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  uint32_t *buf = sbuf[wv];
  const uint32_t gw = blockIdx.x * 4 + wv, nwv = gridDim.x * 4, nm = *nmid;
  for (uint32_t h = gw; h < nm; h += nwv) {
    const uint32_t v = mid[h];
    const uint32_t b = off[v], e = off[v + 1], r = rank[v];
    uint32_t cnt = 0;
    for (uint32_t i0 = b; i0 < e; i0 += MUNROLL * KCDS_WAVE) {
      uint32_t x[MUNROLL];
#pragma unroll
      for (uint32_t q = 0; q < MUNROLL; ++q) {
        const uint32_t i = i0 + q * KCDS_WAVE + lane;
        x[q] = i < e ? nbr[i] : 0u;
      }
#pragma unroll
      for (uint32_t q = 0; q < MUNROLL; ++q) x[q] = (i0 + q * KCDS_WAVE + lane < e) ? rank[x[q]] : 0u;
#pragma unroll
      for (uint32_t q = 0; q < MUNROLL; ++q) {
        const bool keep = (i0 + q * KCDS_WAVE + lane < e) && x[q] > r;
        const ull bal = __ballot(keep);
        if (keep && cnt + lanes_below(bal) < SORT_WAVE) buf[cnt + lanes_below(bal)] = x[q];
        cnt += (uint32_t)__popcll(bal);
      }
    }
    if (lane == 0) dout[r] = cnt;
    if (lane < 4 && cnt + lane < SORT_WAVE) buf[cnt + lane] = 0xFFFFFFFFu;  // pad to whole 4-entry reads
    wave_sync();
    uint32_t *dst = tmp + b;
    if (cnt <= 64) wave_rank_sort<1>(buf, cnt, dst);
    else if (cnt <= 128) wave_rank_sort<2>(buf, cnt, dst);
    else if (cnt <= 256) wave_rank_sort<4>(buf, cnt, dst);
    else if (cnt <= 512) wave_rank_sort<8>(buf, cnt, dst);
    else if (cnt <= SORT_WAVE) wave_rank_sort<16>(buf, cnt, dst);
    wave_sync();
  }
}

// Workgroup per heavy vertex: the same gather with 16 waves sharing one LDS
// list (slots taken with one LDS atomic per wave and chunk).  Out-lists of up
// to HEAVY_SORT ids are sorted and written; a longer one is counted but not
// written (the caller fails).
__global__ __launch_bounds__(HEAVY_THREADS) void k_dag_gather_heavy(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ rank,
    const uint32_t *__restrict__ heavy, const uint32_t *nheavy, uint32_t *__restrict__ dout,
    uint32_t *__restrict__ tmp) {
  __shared__ __align__(16) uint32_t buf[HEAVY_SORT + 4];
  __shared__ uint32_t cnt;
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const uint32_t nh = *nheavy;
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint32_t v = heavy[h];
    const uint32_t b = off[v], e = off[v + 1], r = rank[v];
    if (tid == 0) cnt = 0;
    block_sync();
    for (uint32_t i0 = b; i0 < e; i0 += HEAVY_THREADS) {
      const uint32_t i = i0 + tid;
      const uint32_t x = (i < e) ? rank[nbr[i]] : 0u;
      const bool keep = (i < e) && x > r;
      const ull bal = __ballot(keep);
      if (bal) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(bal);
        uint32_t at = 0;
        if (lane == lead) at = atomicAdd(&cnt, (uint32_t)__popcll(bal));
        at = __shfl(at, lead, 64) + lanes_below(bal);
        if (keep && at < HEAVY_SORT) buf[at] = x;
      }
    }
    block_sync();
    const uint32_t c = cnt;
    if (tid == 0) dout[r] = c;
    if (tid < 4 && c + tid < HEAVY_SORT) buf[c + tid] = 0xFFFFFFFFu;  // pad to whole 4-entry reads
    block_sync();
    if (c <= HEAVY_SORT) {
      for (uint32_t i = tid; i < c; i += HEAVY_THREADS) tmp[b + rank_in(buf, c, buf[i])] = buf[i];
    }
    block_sync();
  }
}

// 16 lanes per new id r (mean out-degree is small): copies the sorted
// out-list of v = inv[r] into the DAG.
__global__ __launch_bounds__(256) void k_dag_copy(const uint32_t *__restrict__ off, const uint32_t *__restrict__ inv,
                                                  uint32_t n, const uint32_t *__restrict__ doff,
                                                  const uint32_t *__restrict__ tmp, uint32_t *__restrict__ dnbr) {
  const uint32_t sub = threadIdx.x & 15;
  const uint32_t gq = (blockIdx.x * blockDim.x + threadIdx.x) >> 4, nq = (gridDim.x * blockDim.x) >> 4;
  for (uint32_t r = gq; r < n; r += nq) {
    const uint32_t src = off[inv[r]], dst = doff[r], d = doff[r + 1] - dst;
    for (uint32_t i = sub; i < d; i += 16) dnbr[dst + i] = tmp[src + i];
  }
}

// ---------------------------------------------------------------------------
// GPU degeneracy order: level-synchronous k-core peel as one kernel per
// step driven by device-side control state and enqueued by the host in
// batches.
//   OPEN    collect every live vertex of degree <= target into the frontier
//           and take the minimum degree of the others; the control step
//           either starts level = target or, when nothing was collected,
//           retries with target = that minimum (its slack in the slack
//           order; empty levels cost one pass).
//   PROCESS the frontier takes the next ranks (in frontier order) and leaves
//           the graph, decrementing its live neighbours; a neighbour whose
//           degree falls from L+1 to L joins the next sub-round of level L.
// A frontier carries each vertex's adjacency cut into PIECE-neighbour
// pieces, and the decrements run PLANES lanes per piece (a wave takes
// 64 / PLANES pieces at once), so a hub of degree 25,000 is spread over 49
// waves.
// A vertex leaves by having its degree replaced with DEAD; decrements that
// race with that replacement keep the value >= DEAD_MIN, which every reader
// treats as dead.  Kernel boundaries order the steps (the next step reads
// the frontiers this one wrote); inside a step workgroups share only atomic
// counters, and nothing waits on another workgroup.

constexpr uint32_t NSEG = 8;  // frontier segments / counter shards (one per workgroup residue mod 8)

// Counters are 64 words apart so that no two share a cache line (same-address
// device atomics are serialised).
// The degree counters use the scattered layout of kcds_device.hpp (Scatter):
// R-MAT-like graphs number their hubs close together, and packed, the dense
// core's counters would fill a few cache lines on which the decrements of a
// step that removes much of that core at once would queue.  Padding slots
// (vertex >= n) hold DEAD.
struct OrderCtl {
  uint32_t level, target, next_rank, cur, done, phase;
  uint32_t steps;  // steps taken (sizes the next call's first batch)
  uint32_t pad0[64 - 7];
  uint32_t fsize[2][NSEG][64];  // vertex frontier segments: [frontier][segment][0]
  uint32_t psize[2][NSEG][64];  // adjacency-piece frontier segments
  uint32_t gmin[NSEG][64];      // minimum live degree above target, per segment (OPEN)
  uint32_t gblocks[NSEG][64];   // finished workgroups per segment
  uint32_t blocks[64];          // finished segments
};

constexpr uint32_t DEAD = 0xFFFFFFFFu;
constexpr uint32_t DEAD_MIN = 0x80000000u;
constexpr uint32_t PIECE = 256;   // neighbours per decrement work item
constexpr uint32_t PLANES = 32;   // lanes per piece
constexpr uint32_t PUNROLL = PIECE / PLANES;
constexpr uint32_t SMALL_FRONT = 4;  // frontier vertices of at most this degree decrement lane per vertex
constexpr uint32_t DEC_TB = 4096;     // LDS table of a workgroup's pending decrements
constexpr uint32_t DEC_ROUND = 256 * PUNROLL;              // most new keys one workgroup round inserts
constexpr uint32_t DEC_FLUSH = DEC_TB * 3 / 4 - DEC_ROUND;  // flush above this many keys (load stays <= 3/4)
static_assert(DEC_TB * 3 / 4 > DEC_ROUND, "a round must fit the decrement table");
constexpr uint32_t AGG_MIN_PIECES = 2048;  // steps with fewer pieces decrement directly

// Appends v (where want) to a vertex frontier segment fv and, when its
// degree exceeds SMALL_FRONT, its adjacency cut in PIECE-neighbour pieces
// (start << 32 | end) to the piece segment fp; one counter atomic of each
// kind per wave.  Every lane calls it.
__device__ __forceinline__ void wave_append(bool want, uint32_t v, const uint32_t *__restrict__ off, uint32_t *fv,
                                            uint32_t *fvsize, ull *fp, uint32_t *fpsize) {
  const ull bal = __ballot(want);
  if (!bal) return;
  const uint32_t lane = lane_id(), lead = (uint32_t)__builtin_ctzll(bal);
  uint32_t b = 0, e = 0, np = 0;
  if (want) {
    b = off[v];
    e = off[v + 1];
    np = e - b > SMALL_FRONT ? (e - b + PIECE - 1) / PIECE : 0u;
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
    const uint32_t pe = np ? e : b;  // no pieces for a vertex of degree <= SMALL_FRONT
    for (uint32_t st = b; st < pe; st += PIECE) fp[p++] = ((ull)st << 32) | min(st + PIECE, e);
  }
}

// Control step, run by one thread of the last workgroup to finish a step.
// Every workgroup's counter atomics completed before it took its ticket; the
// counters are read back with read-modify-write atomics, which are performed
// where the other workgroups' atomics were.
// Control step, run by wave 0 of the last workgroup: lanes g < NSEG read
// and reset segment g's counters together (one round trip, not one per
// segment); the header values are those this step started from (only the
// control step writes them).
// Threshold from a minimum live degree x: x itself for the exact order,
// x + x>>sh with slack (clamped below the dead marks).
__device__ __forceinline__ uint32_t with_slack(uint32_t x, uint32_t sh) {
  return sh < 32 ? (uint32_t)min<ull>((ull)x + (x >> sh), DEAD_MIN - 1) : x;
}

__device__ void order_control(OrderCtl *ctl, uint32_t n, uint32_t level, uint32_t target, uint32_t next_rank,
                              uint32_t cur, uint32_t phase, uint32_t sh) {
  const uint32_t lane = lane_id();
  uint32_t fc = 0, fn = 0, mn = DEAD;
  if (lane < NSEG) {
    fc = ldc_u32(&ctl->fsize[cur][lane][0]);
    fn = ldc_u32(&ctl->fsize[cur ^ 1][lane][0]);
    mn = atomicExch(&ctl->gmin[lane][0], DEAD);
    ctl->gblocks[lane][0] = 0;
    if (phase == 1) {
      ctl->fsize[cur][lane][0] = 0;
      ctl->psize[cur][lane][0] = 0;
    }
  }
  const uint32_t fcur = wave_sum_u32(fc), fnext = wave_sum_u32(fn);
  mn = wave_min_u32(mn);
  if (lane != 0) return;
  ctl->blocks[0] = 0;
  ++ctl->steps;
  if (phase == 1) {
    ctl->next_rank = next_rank + fcur;
    ctl->cur = cur ^ 1;
    if (next_rank + fcur >= n) {
      ctl->done = 1;
    } else if (fnext == 0) {
      ctl->phase = 0;
      // every live degree now exceeds level.  Exact: try level + 1.  Slack:
      // try the slack of level + 1, which is at most the slack of the true
      // minimum m >= level + 1, so every threshold stays within the slack of a
      // minimum live degree (which never exceeds the degeneracy: out-degrees
      // stay <= kmax + kmax/2); when it collects nothing, that OPEN finds m
      // (C2: 53 steps instead of the 60 of always finding m first)
      ctl->target = sh < 32 ? with_slack(level + 1, sh) : level + 1;
    }
  } else if (fcur > 0) {
    ctl->level = target;
    ctl->phase = 1;
  } else if (mn == DEAD) {
    ctl->done = 1;
  } else {
    ctl->target = with_slack(mn, sh);
  }
}

// One step (OPEN or PROCESS); the last workgroup to finish runs the control
// step, so consecutive steps need no separate control launch.  Frontiers are
// NSEG segments of capacity seg_v (vertices) and seg_p (pieces); workgroup b
// appends to segment b % NSEG.
__global__ __launch_bounds__(256) void k_ko_step(OrderCtl *ctl, uint32_t n, const uint32_t *__restrict__ off,
                                                 const uint32_t *__restrict__ nbr, uint32_t *deg, uint32_t *fv0,
                                                 uint32_t *fv1, ull *fp0, ull *fp1, uint32_t seg_v, uint32_t seg_p,
                                                 uint32_t *rank, uint32_t *inv, uint32_t sh, Scatter sc,
                                                 uint32_t *gone) {
  __shared__ uint32_t s_f[NSEG + 1], s_p[NSEG + 1], s_min, s_last;
  __shared__ uint32_t s_hdr[6], s_sz[2][2][NSEG];
  __shared__ uint32_t t_key[DEC_TB], t_cnt[DEC_TB], t_used[DEC_TB], t_nused[2];
  const uint32_t tid = threadIdx.x, lane = lane_id(), seg = blockIdx.x % NSEG;
  // the header and both frontiers' segment sizes in one round trip
  if (tid < 6) s_hdr[tid] = (&ctl->level)[tid];
  if (tid >= 64 && tid < 64 + 4 * NSEG) {
    const uint32_t i = tid - 64, f = i / (2 * NSEG), kind = (i / NSEG) & 1, g = i % NSEG;
    s_sz[f][kind][g] = kind ? ctl->psize[f][g][0] : ctl->fsize[f][g][0];
  }
  lds_sync();
  const uint32_t level = s_hdr[0], target = s_hdr[1], base = s_hdr[2], cur = s_hdr[3], done = s_hdr[4],
                 phase = s_hdr[5];
  if (done) return;
  uint32_t *fv = cur ? fv1 : fv0;
  uint32_t *nv = cur ? fv0 : fv1;
  ull *fp = cur ? fp1 : fp0;
  ull *np = cur ? fp0 : fp1;
  const uint32_t gtid = blockIdx.x * blockDim.x + tid, gstride = gridDim.x * blockDim.x;
  if (phase == 0) {
    if (tid == 0) s_min = DEAD;
    lds_sync();
    uint32_t m = DEAD;
    const uint32_t K = sc.mask + 1;
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < K; p0 += gstride) {  // slots in order: reads coalesce
      const uint32_t p = p0 + tid;
      const uint32_t d = p < K ? deg[p] : DEAD;
      const uint32_t v = vertex_of(sc, p);
      const bool live = d < DEAD_MIN;
      wave_append(live && d <= target, v, off, fv + (size_t)seg * seg_v, &ctl->fsize[cur][seg][0],
                  fp + (size_t)seg * seg_p, &ctl->psize[cur][seg][0]);
      if (live && d > target) m = min(m, d);
    }
    m = wave_min_u32(m);
    if (lane == 0 && m != DEAD) atomicMin(&s_min, m);
    lds_sync();
    if (tid == 0 && s_min != DEAD) atomicMin(&ctl->gmin[seg][0], s_min);
  } else {
    // segment sizes and prefixes of the current frontier
    if (tid == 0) {
      s_f[0] = 0;
      s_p[0] = 0;
      for (uint32_t g = 0; g < NSEG; ++g) {
        s_f[g + 1] = s_f[g] + s_sz[cur][0][g];
        s_p[g + 1] = s_p[g] + s_sz[cur][1][g];
      }
    }
    lds_sync();
    const uint32_t fs = s_f[NSEG], ps = s_p[NSEG];
    // The frontier takes the next ranks and leaves; a vertex of degree <=
    // SMALL_FRONT (many of an early level) decrements its neighbours from its
    // own lane, its arcs' loads and atomics in flight together.
    for (uint32_t i0 = gtid - lane; i0 < fs; i0 += gstride) {
      const uint32_t i = i0 + lane;
      uint32_t b = 0, e = 0;
      if (i < fs) {
        uint32_t g = 0;
        while (i >= s_f[g + 1]) ++g;
        const uint32_t v = fv[(size_t)g * seg_v + (i - s_f[g])];
        rank[v] = base + i;
        inv[base + i] = v;
        atomicExch(&deg[slot_of(sc, v)], DEAD);
        atomicOr(&gone[v >> 5], 1u << (v & 31));
        b = off[v];
        e = off[v + 1];
        if (e - b > SMALL_FRONT) e = b;  // its pieces are in fp
      }
      const uint32_t dmax = wave_max_u32(e - b);
      uint32_t u[SMALL_FRONT], dv[SMALL_FRONT];
#pragma unroll
      for (uint32_t x = 0; x < SMALL_FRONT; ++x)
        if (x < dmax) u[x] = b + x < e ? nbr[b + x] : 0u;
#pragma unroll
      for (uint32_t x = 0; x < SMALL_FRONT; ++x)
        if (x < dmax) dv[x] = b + x < e ? atomicSub(&deg[slot_of(sc, u[x])], 1u) : DEAD;
#pragma unroll
      for (uint32_t x = 0; x < SMALL_FRONT; ++x)
        if (x < dmax)
          wave_append(dv[x] == level + 1, u[x], off, nv + (size_t)seg * seg_v, &ctl->fsize[cur ^ 1][seg][0],
                      np + (size_t)seg * seg_p, &ctl->psize[cur ^ 1][seg][0]);
    }
    // PLANES lanes per adjacency piece of the frontier.  Decrementing a vertex
    // that already left keeps its value >= DEAD_MIN (it loses at most its
    // degree).  The wide vertices a late level removes have mostly departed
    // neighbours: the departed bitmap (n/8 bytes, cache resident; bits are
    // only ever set, so a stale word only costs an atomic) screens them out.
    // A large level's decrements pile onto few counters (at a mid level of
    // an R-MAT graph, 700,000 decrements onto 15,000 vertices, thousands onto
    // each hub), and same-address device atomics serialise.  In a step of at
    // least AGG_MIN_PIECES pieces a workgroup therefore sums its decrements
    // per counter in an LDS table, over its rounds of 4 waves x 64 / PLANES
    // pieces, and applies each sum with one atomic: the counter crossed from
    // above L to L or below exactly when old > L >= old - sum, which one
    // atomic observes whatever the sum covers (departed counters, >=
    // DEAD_MIN, never satisfy it).
    constexpr uint32_t PW = KCDS_WAVE / PLANES;  // pieces per wave
    constexpr uint32_t BPW = 4 * PW;             // pieces per workgroup round (256 threads)
    const uint32_t sub = lane % PLANES, wv = tid / KCDS_WAVE;
    const bool agg = ps >= AGG_MIN_PIECES;       // uniform over the grid
    if (agg && ps > blockIdx.x * BPW) {
      for (uint32_t t = tid; t < DEC_TB; t += 256) {
        t_key[t] = 0xFFFFFFFFu;
        t_cnt[t] = 0;
      }
      if (tid < 2) t_nused[tid] = 0;
      lds_sync();
    }
    uint32_t par = 0;
    // a lane's piece record for a round is loaded during the round before
    auto piece_at = [&](uint32_t it) -> ull {
      if (it >= ps) return 0ull;
      uint32_t g = 0;
      while (it >= s_p[g + 1]) ++g;
      return fp[(size_t)g * seg_p + (it - s_p[g])];
    };
    const uint32_t my = wv * PW + lane / PLANES;
    ull item_next = piece_at(blockIdx.x * BPW + my);
    for (uint32_t it0 = blockIdx.x * BPW; it0 < ps; it0 += gridDim.x * BPW) {
      const ull item = item_next;
      item_next = piece_at(it0 + gridDim.x * BPW + my);
      const uint32_t st = (uint32_t)(item >> 32), e = (uint32_t)item;
      uint32_t u[PUNROLL], gw[PUNROLL];
#pragma unroll
      for (uint32_t x = 0; x < PUNROLL; ++x) {
        const uint32_t j = st + x * PLANES + sub;
        u[x] = j < e ? nbr[j] : 0u;
      }
#pragma unroll
      for (uint32_t x = 0; x < PUNROLL; ++x) {
        const uint32_t j = st + x * PLANES + sub;
        gw[x] = j < e ? gone[u[x] >> 5] : 0xFFFFFFFFu;
      }
      if (!agg) {
        uint32_t dv[PUNROLL];
#pragma unroll
        for (uint32_t x = 0; x < PUNROLL; ++x)
          dv[x] = (gw[x] >> (u[x] & 31)) & 1u ? DEAD : atomicSub(&deg[slot_of(sc, u[x])], 1u);
#pragma unroll
        for (uint32_t x = 0; x < PUNROLL; ++x)
          wave_append(dv[x] == level + 1, u[x], off, nv + (size_t)seg * seg_v, &ctl->fsize[cur ^ 1][seg][0],
                      np + (size_t)seg * seg_p, &ctl->psize[cur ^ 1][seg][0]);
        continue;
      }
#pragma unroll
      for (uint32_t x = 0; x < PUNROLL; ++x) {
        if ((gw[x] >> (u[x] & 31)) & 1u) continue;
        const uint32_t key = slot_of(sc, u[x]);
        uint32_t h = (key * 0x9E3779B1u) >> (32 - 12);
        while (true) {
          const uint32_t prev = atomicCAS(&t_key[h], 0xFFFFFFFFu, key);
          if (prev == 0xFFFFFFFFu) t_used[atomicAdd(&t_nused[par], 1u)] = h;
          if (prev == 0xFFFFFFFFu || prev == key) break;
          h = (h + 1) & (DEC_TB - 1);
        }
        atomicAdd(&t_cnt[h], 1u);
      }
      lds_sync();
      // The sums stay in the table across rounds (a hub counter then takes one
      // atomic per workgroup, not one per workgroup round) until the next
      // round could fill it past 3/4, or this was the workgroup's last round.
      // Then the used entries are applied (one atomic each) and cleared; the
      // other parity's key counter, last read by the previous flush, is free
      // again for the keys that follow.
      const uint32_t nu = t_nused[par];
      lds_sync();  // every wave has read nu before any inserts more keys
      if (nu <= DEC_FLUSH && it0 + gridDim.x * BPW < ps) continue;  // uniform over the workgroup
      if (tid == 0) t_nused[par ^ 1] = 0;
      for (uint32_t i0 = 0; i0 < nu; i0 += 4 * 256) {
        uint32_t ck[4], cc[4], old[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
          const uint32_t i = i0 + q * 256 + tid;
          ck[q] = 0xFFFFFFFFu;
          cc[q] = 0;
          if (i < nu) {
            const uint32_t h = t_used[i];
            ck[q] = t_key[h];
            cc[q] = t_cnt[h];
            t_key[h] = 0xFFFFFFFFu;
            t_cnt[h] = 0;
          }
        }
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) old[q] = cc[q] ? atomicSub(&deg[ck[q]], cc[q]) : 0u;
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q)
          wave_append(cc[q] && old[q] > level && old[q] - cc[q] <= level, vertex_of(sc, ck[q]), off,
                      nv + (size_t)seg * seg_v, &ctl->fsize[cur ^ 1][seg][0], np + (size_t)seg * seg_p,
                      &ctl->psize[cur ^ 1][seg][0]);
      }
      lds_sync();
      par ^= 1;
    }
  }
  // two-level ticket: the last workgroup of a segment, then the last segment.
  // What the control step reads of this workgroup is complete by now: the
  // frontier-size atomics returned their values before their appends, and
  // thread 0 waits for its own minimum atomic below.  Stores and the
  // decrements' side effects are read only by later launches.
  lds_sync();
  if (tid == 0) {
    wave_drain();
    const uint32_t in_seg = (gridDim.x - seg + NSEG - 1) / NSEG;  // workgroups with this residue
    uint32_t last = 0;
    if (atomicAdd(&ctl->gblocks[seg][0], 1u) == in_seg - 1) {
      const uint32_t segs = min(gridDim.x, NSEG);
      last = atomicAdd(&ctl->blocks[0], 1u) == segs - 1;
    }
    s_last = last;
  }
  lds_sync();
  if (s_last && tid < KCDS_WAVE) order_control(ctl, n, level, target, base, cur, phase, sh);
}

__global__ void k_init_deg(const uint32_t *off, uint32_t n, Scatter sc, uint32_t *deg, uint32_t *gone) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= sc.mask; p += gridDim.x * blockDim.x) {
    const uint32_t v = vertex_of(sc, p);
    deg[p] = v < n ? off[v + 1] - off[v] : DEAD;
  }
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < (n + 31) / 32; w += gridDim.x * blockDim.x) gone[w] = 0;
}

// Degree order: rank[v] = position of v among the vertices sorted by (degree,
// id).  Arcs then point from the lower-degree end to the higher (ties by id);
// the widest out-list is usually near the degeneracy (C2: 423 against a
// degeneracy of 373; C3: 294 against 273) and the out-wedge sum W2 within
// 1 % of the degeneracy order's, but the order is one sort instead of a peel
// of some fifty level-synchronous steps.  The sort is a stable LSD radix
// sort of the ids by degree, 8 bits per pass: ids start in increasing order,
// so equal degrees keep id order.  Each pass is a per-tile digit histogram,
// a scan over (digit, tile), and a scatter in which one wave walks its tile
// in order and places each element after the earlier ones of its digit.
constexpr uint32_t RS_TILE = 1024;  // elements per tile (a scatter wave walks 16 chunks of 64)
constexpr uint32_t RS_CHUNKS = RS_TILE / 64;

__device__ __forceinline__ uint32_t rs_key(const uint32_t *off, const uint32_t *key_in, uint32_t i) {
  return key_in ? key_in[i] : off[i + 1] - off[i];  // first pass: the degree, read from the CSR
}

__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t *off, const uint32_t *key_in, uint32_t n,
                                                 uint32_t shift, uint32_t ntile, uint32_t *hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RS_TILE, end = min(n, base + RS_TILE);
  for (uint32_t i = base + threadIdx.x; i < end; i += blockDim.x)
    atomicAdd(&h[(rs_key(off, key_in, i) >> shift) & 255u], 1u);
  __syncthreads();
  hist[threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];  // digit-major: the scan orders (digit, tile)
}

// One wave per tile.  hoff = exclusive scan of hist.  The last pass writes
// rank/inv instead of the next pass's keys and ids.
__global__ __launch_bounds__(64) void k_rs_scatter(const uint32_t *off, const uint32_t *key_in, const uint32_t *id_in,
                                                   uint32_t n, uint32_t shift, uint32_t ntile, const uint32_t *hoff,
                                                   uint32_t *key_out, uint32_t *id_out, uint32_t *rank,
                                                   uint32_t *inv) {
  __shared__ uint32_t run[256];
  const uint32_t lane = threadIdx.x, base = blockIdx.x * RS_TILE;
#pragma unroll
  for (uint32_t t = 0; t < 4; ++t) run[lane + 64 * t] = hoff[(lane + 64 * t) * ntile + blockIdx.x];
  uint32_t key[RS_CHUNKS], id[RS_CHUNKS];
#pragma unroll
  for (uint32_t ch = 0; ch < RS_CHUNKS; ++ch) {  // every load of the tile in flight at once
    const uint32_t i = base + ch * 64 + lane;
    key[ch] = i < n ? rs_key(off, key_in, i) : 0u;
    id[ch] = i < n ? (id_in ? id_in[i] : i) : 0u;
  }
  __syncthreads();
  const ull below_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (uint32_t ch = 0; ch < RS_CHUNKS; ++ch) {
    const uint32_t i = base + ch * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = (key[ch] >> shift) & 255u;
    ull peers = __ballot(valid);  // lanes of this chunk with the same digit
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const ull m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t before = (uint32_t)__popcll(peers & below_mask), cnt = (uint32_t)__popcll(peers);
    const uint32_t pos = valid ? run[d] + before : 0u;
    __syncthreads();  // every lane has read run[] before the groups advance it
    if (valid && before + 1 == cnt) run[d] += cnt;
    __syncthreads();
    if (valid) {
      if (rank) {
        rank[id[ch]] = pos;
        inv[pos] = id[ch];
      } else {
        key_out[pos] = key[ch];
        id_out[pos] = id[ch];
      }
    }
  }
}

}  // namespace

void degree_rank_gpu(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.rank, sizeof(uint32_t) * (size_t)n + 4);
  buf_reserve(c, c.inv, sizeof(uint32_t) * (size_t)n + 4);
  if (n == 0) return;
  uint32_t bits = 0;
  while (bits < 32 && (c.maxdeg >> bits)) ++bits;
  const uint32_t passes = std::max(1u, (bits + 7) / 8);
  const uint32_t ntile = (n + RS_TILE - 1) / RS_TILE;
  const size_t nh = (size_t)256 * ntile;
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * (4 * ((size_t)n + 64) + 2 * (nh + 64)));
  uint32_t *k0 = c.tmp0.as<uint32_t>(), *i0 = k0 + n + 64, *k1 = i0 + n + 64, *i1 = k1 + n + 64;
  uint32_t *hist = i1 + n + 64, *hoff = hist + nh + 64;
  const uint32_t *kin = nullptr, *iin = nullptr;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = 8 * p;
    const bool last = p + 1 == passes;
    k_rs_hist<<<ntile, 256, 0, c.stream>>>(c.off.as<uint32_t>(), kin, n, shift, ntile, hist);
    KCDS_LAUNCH_CHECK();
    scan_u32_to_u32(c, hist, hoff, nh);
    uint32_t *kout = (p & 1) ? k1 : k0, *iout = (p & 1) ? i1 : i0;
    k_rs_scatter<<<ntile, 64, 0, c.stream>>>(c.off.as<uint32_t>(), kin, iin, n, shift, ntile, hoff, kout, iout,
                                             last ? c.rank.as<uint32_t>() : nullptr,
                                             last ? c.inv.as<uint32_t>() : nullptr);
    KCDS_LAUNCH_CHECK();
    kin = kout;
    iin = iout;
  }
}

void degeneracy_rank_gpu(Ctx &c, uint32_t slack_shift) {
  const uint32_t n = c.n;
  buf_reserve(c, c.rank, sizeof(uint32_t) * (size_t)n + 4);
  buf_reserve(c, c.inv, sizeof(uint32_t) * (size_t)n + 4);
  if (n == 0) return;
  // every segment can hold any frontier: n vertices, their pieces
  const uint32_t seg_v = ((n + 63) / 64) * 64;
  const uint64_t seg_p64 = (((uint64_t)n + c.nnz / PIECE + 64 + 63) / 64) * 64;
  if (seg_p64 * NSEG >= (1ull << 32)) fail(KCDS_E_LIMIT, "graph too large for the ordering frontiers");
  const uint32_t seg_p = (uint32_t)seg_p64;
  const Scatter sc = make_scatter(n);
  const uint32_t K = sc.mask + 1;
  const size_t kslots = ((size_t)K + 63) / 64 * 64;
  const size_t bytes = sizeof(uint32_t) * (kslots + 2 * NSEG * (size_t)seg_v) +
                       sizeof(ull) * 2 * NSEG * (size_t)seg_p + sizeof(OrderCtl) + 1024 +
                       sizeof(uint32_t) * (((size_t)n + 31) / 32) + 256;
  buf_reserve(c, c.tmp0, bytes);
  uint32_t *deg = c.tmp0.as<uint32_t>();
  uint32_t *fv0 = deg + kslots, *fv1 = fv0 + (size_t)NSEG * seg_v;
  ull *fp0 = (ull *)(((uintptr_t)(fv1 + (size_t)NSEG * seg_v) + 255) & ~(uintptr_t)255);
  ull *fp1 = fp0 + (size_t)NSEG * seg_p;
  OrderCtl *ctl = (OrderCtl *)(((uintptr_t)(fp1 + (size_t)NSEG * seg_p) + 255) & ~(uintptr_t)255);
  uint32_t *gone = (uint32_t *)(((uintptr_t)(ctl + 1) + 255) & ~(uintptr_t)255);
  const uint32_t vgrid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)c.num_cus * 2));
  k_init_deg<<<(uint32_t)std::min<uint64_t>((K + 255) / 256, (uint64_t)c.num_cus * 16), 256, 0, c.stream>>>(
      c.off.as<uint32_t>(), n, sc, deg, gone);
  KCDS_LAUNCH_CHECK();
  OrderCtl *h0 = (OrderCtl *)c.host_scratch(sizeof(OrderCtl));
  memset(h0, 0, sizeof(OrderCtl));
  for (uint32_t g = 0; g < NSEG; ++g) h0->gmin[g][0] = DEAD;
  HIP_CHECK(hipMemcpyAsync(ctl, h0, sizeof(OrderCtl), hipMemcpyHostToDevice, c.stream));
  c.sync();  // h0 is reused below
  const uint32_t *off = c.off.as<uint32_t>(), *nbr = c.nbr.as<uint32_t>();
  uint32_t *rank = c.rank.as<uint32_t>(), *inv = c.inv.as<uint32_t>();
  uint32_t *hc = (uint32_t *)h0;
  // the number of steps is a property of the graph: start from the last count
  uint32_t batch = c.order_steps ? (uint32_t)std::min<uint64_t>(c.order_steps, 1u << 20) : 32;
  uint64_t rounds = 0;
  while (true) {
    for (uint32_t r = 0; r < batch; ++r) {
      k_ko_step<<<vgrid, 256, 0, c.stream>>>(ctl, n, off, nbr, deg, fv0, fv1, fp0, fp1, seg_v, seg_p, rank, inv,
                                             slack_shift, sc, gone);
      KCDS_LAUNCH_CHECK();
    }
    rounds += batch;
    HIP_CHECK(hipMemcpyAsync(hc, ctl, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (((OrderCtl *)hc)->done) break;
    // every PROCESS step ranks at least one vertex; an OPEN step is followed
    // by a PROCESS step or by one OPEN step that succeeds
    if (rounds > 3ull * n + 8) fail(KCDS_E_HIP, "ordering did not finish within %llu sub-rounds",
                                    (unsigned long long)rounds);
    batch = std::min<uint32_t>(std::max<uint32_t>(batch, 32) * 2, 128);
  }
  c.order_steps = ((OrderCtl *)hc)->steps;
}

bool orient_from_rank(Ctx &c, bool approx) {
  const uint32_t n = c.n;
  buf_reserve(c, c.doff, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * ((size_t)c.nnz + 1));
  buf_reserve(c, c.tmp2, sizeof(uint32_t) * (2 * (size_t)n + 128));
  uint32_t *dout = c.tmp1.as<uint32_t>(), *tmp = c.tmp0.as<uint32_t>();
  uint32_t *nheavy = c.tmp2.as<uint32_t>(), *nmid = nheavy + 1, *heavy = nheavy + 64, *mid = heavy + n + 64;
  if (n) {
    fill_u32(c, nheavy, 0, 2);
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t tgrid = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, (uint64_t)c.num_cus * 8);
    k_dag_gather_tiles<<<tgrid, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(),
                                                    c.rank.as<uint32_t>(), n, dout, tmp, mid, nmid, heavy, nheavy);
    KCDS_LAUNCH_CHECK();
    k_dag_gather<<<c.num_cus * 8, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(),
                                                      c.rank.as<uint32_t>(), mid, nmid, dout, tmp);
    KCDS_LAUNCH_CHECK();
    k_dag_gather_heavy<<<c.num_cus, HEAVY_THREADS, 0, c.stream>>>(c.off.as<uint32_t>(), c.nbr.as<uint32_t>(),
                                                                   c.rank.as<uint32_t>(), heavy, nheavy, dout, tmp);
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
  if (approx && c.dmax > LIST_MAX_OUTDEG) return false;  // the caller orders exactly instead
  if (c.dmax > LIST_WIDE_MAX)
    fail(KCDS_E_LIMIT, "DAG out-degree %u exceeds %u (orientation not degeneracy-bounded)", c.dmax, LIST_WIDE_MAX);
  buf_reserve(c, c.dnbr, sizeof(uint32_t) * (c.darcs + 1));
  if (n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 15) / 16, (uint64_t)c.num_cus * 40);
    k_dag_copy<<<grid, 256, 0, c.stream>>>(c.off.as<uint32_t>(), c.inv.as<uint32_t>(), n, c.doff.as<uint32_t>(),
                                           tmp, c.dnbr.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.w2 = 0;  // the listing pass reports the out-wedge sum it stages
  return true;
}

// GPU orientation: the slack order (each threshold is the minimum live
// degree rounded up by a half, so out-degrees stay within 3/2 of the
// degeneracy, in a fraction of the steps: C2 53 instead of 609), or the
// exact degeneracy order when KCDS_TUNE_EXACT_ORDER is set or the slack order
// leaves an out-list wider than the listing's LDS rows.
void orient_gpu(Ctx &c) {
  const bool exact = c.tune_flags & KCDS_TUNE_EXACT_ORDER;
  if (!exact && !(c.tune_flags & KCDS_TUNE_DEGENERACY_ORDER)) {
    degree_rank_gpu(c);
    c.order_exact = false;
    c.order_kind = 0;
    if (orient_from_rank(c, true)) return;
  }
  degeneracy_rank_gpu(c, exact ? NO_SLACK : ORDER_SLACK_SHIFT);
  c.order_exact = exact;
  c.order_kind = exact ? 2 : 1;
  if (orient_from_rank(c, !exact)) return;
  degeneracy_rank_gpu(c, NO_SLACK);
  c.order_exact = true;
  c.order_kind = 2;
  orient_from_rank(c, false);
}

void rank_from_listing_order(Ctx &c, const uint32_t *lo_dev) {
  const uint32_t n = c.n;
  if (!n) return;
  k_rank_from_lo<<<(n + 255) / 256, 256, 0, c.stream>>>(lo_dev, n, c.rank.as<uint32_t>(), c.inv.as<uint32_t>());
  KCDS_LAUNCH_CHECK();
}

}  // namespace kcds
