// Wave64 k-clique listing with LDS bitmask rows.
//
// Reference: listAllCliques (main.cu:93-327) lists cliques level by level
// with per-warp partitions of partial cliques, 32-bit neighbour masks
// (listIntialCliques / listMidCliques, src/helpers.cu:175-488), counts them
// (countCliques, src/helpers.cu:1074-1113) and writes them with per-vertex
// clique degrees (writeFinalCliques, src/helpers.cu:489-550).
//
// Here one workgroup of nw wavefronts owns one root r of the DAG (new ids,
// lists ascending) at a time:
//   1. Stage N+(r) = L[0..d) in LDS with an open-addressing table
//      id -> local index.  The waves take the out-lists N+(L[i]) in turn and
//      read each in 64-element chunks, four chunks in flight per wave (the
//      out-wedge sum W2 of the roofline is these reads); every element w is
//      looked up in the table.  A hit j sets bit j of row i and bit i of
//      row j: rows are the symmetric adjacency of the subgraph induced by
//      N+(r), d rows of W 64-bit words stored word-major (A[w*D + j]) so
//      that 64 lanes reading 64 consecutive rows hit 64 distinct banks.
//   2. Because ids increase along every DAG arc, the out-neighbours of L[c]
//      inside N+(r) are row c masked to bits above c, so recursion on
//      candidate sets S (W words, wave-uniform) needs no further global
//      reads: S' = S & A[c] & above(c).  The waves split the first level.
//   3. At two vertices left, the cliques extending the prefix are the edges
//      inside S: lane j (owner of rows j, j+64, ...) adds popc(A[j] & S) to
//      its own counter; the wave sum / 2 is credited to every prefix vertex.
//      Per-vertex counts are exact without touching individual cliques.
//   4. Counters flush to HBM with one 64-bit atomic per (wave, root, member).
// The workgroup exchanges data only through LDS, so its barriers wait for LDS
// operations alone: the flush atomics and the next root's queue ticket stay
// in flight across them.
// The materialising variant (one wave per root) enumerates the same cliques
// in a fixed order and writes them row-major (k new ids per clique) at the
// root's offset.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr int NCLASS = 4;  // W = 1, 2, 4, 8 words per row (out-degree <= 64, 128, 256, 512)
constexpr uint32_t MAX_LDS_DEG = LIST_MAX_OUTDEG;
constexpr uint32_t MAX_WAVES = 8;  // waves per workgroup (launch bound 512)
constexpr uint32_t EMPTY_KEY = 0xFFFFFFFFu;

template <int W> struct Frame {
  static constexpr uint32_t D = 64u * W;  // row capacity
  static constexpr uint32_t HB = W == 1 ? 8 : W == 2 ? 9 : W == 4 ? 10 : 11;
  static constexpr uint32_t H = 1u << HB;  // table slots (load <= 1/4)
  uint32_t L[D];     // N+(r), ascending
  uint2 sinfo[D];    // (doff[L[i]], outdeg(L[i])); length 0 for the last (it has no later candidate)
  ull hent[H];       // table entry: low word L[i] (EMPTY_KEY when free), high word i
  ull A[W * D];      // A[w*D + j]: bits 64w..64w+63 of row j
  ull rt;            // cliques of the current root (summed over waves)
  uint32_t root;
  uint32_t skip;     // the root's graph is already over its admission cap
};

constexpr uint32_t UNROLL = 8;  // 64-element chunks in flight per wave (from one or more out-lists)

// Fibonacci hashing on a 24-bit multiply (full rate; a 32-bit multiply is
// quarter rate): the top bits of the product's low word mix every key bit.
__device__ __forceinline__ uint32_t hslot(uint32_t key, uint32_t hb) {
  return (uint32_t)__umul24(key ^ (key >> 24), 0x9E3779u) >> (32 - hb);
}

__device__ __forceinline__ ull above_mask(int q, uint32_t c) {
  const int cq = (int)(c >> 6), cb = (int)(c & 63);
  if (q < cq) return 0ull;
  if (q > cq) return ~0ull;
  return cb == 63 ? 0ull : (~0ull << (cb + 1));
}

template <int W> __device__ __forceinline__ ull row(const Frame<W> &f, uint32_t j, int w) {
  return f.A[(uint32_t)w * Frame<W>::D + j];
}

// Builds the frame of the root whose out-list is dnbr[base, base + d), d >= 2.
// Every thread of the workgroup calls it.  Returns this lane's share of
// sum over u in N+(r) of outdeg(u).
template <int W>
__device__ __forceinline__ uint32_t build_rows(Frame<W> &f, const uint32_t *__restrict__ doff,
                                               const uint32_t *__restrict__ dnbr, uint32_t base, uint32_t d) {
  constexpr uint32_t D = Frame<W>::D, H = Frame<W>::H, HB = Frame<W>::HB;
  const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, nw = nt >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar loops
  for (uint32_t t = tid; t < H; t += nt) f.hent[t] = ~0ull;
  for (uint32_t j = tid; j < d; j += nt) {
#pragma unroll
    for (int w = 0; w < W; ++w) f.A[(uint32_t)w * D + j] = 0ull;
  }
  lds_sync();
  // stage N+(r), the out-list extents and the id -> index table
  uint32_t *hk = reinterpret_cast<uint32_t *>(f.hent);
  uint32_t raw = 0;
  for (uint32_t t = tid; t < d; t += nt) {
    const uint32_t u = dnbr[base + t];
    const uint32_t ou = doff[u], du = doff[u + 1] - ou;
    f.L[t] = u;
    f.sinfo[t] = make_uint2(ou, t + 1 < d ? du : 0u);
    raw += du;
    uint32_t h = hslot(u, HB);
    while (atomicCAS(&hk[2 * h], EMPTY_KEY, u) != EMPTY_KEY) h = (h + 1) & (H - 1);
    hk[2 * h + 1] = t;
  }
  lds_sync();
  // row i collects N+(L[i]) inside N+(r).  The waves take out-lists in turn
  // (i = wv, wv + nw, ...) and walk them in 64-element chunks; a step takes
  // the next UNROLL chunks, across list boundaries, and issues every load,
  // then every first probe, before any is compared (collisions, rare at load
  // <= 1/4, probe on in a loop afterwards).  Out-lists average a few chunks,
  // so one step overlaps the loads of several lists.
  const uint32_t last = f.L[d - 1];
  // cursor: list ci (its (start, length) in cur, the next list's already
  // requested in nxt) and the offset ct in it
  // (cur is wave-uniform: scalar registers and branches; nxt is made uniform
  // only when it is used, so its LDS read is not waited for at once)
  auto info = [&](uint32_t i) -> uint2 { return i + 1 < d ? f.sinfo[i] : make_uint2(0u, 0u); };
  auto uniform = [](uint2 v) {
    return make_uint2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
  };
  uint32_t ci = wv, ct = 0;
  uint2 cur = uniform(info(ci)), nxt = info(ci + nw);
  auto advance = [&]() {  // to the next non-empty list of this wave
    do {
      ci += nw;
      cur = uniform(nxt);
      nxt = info(ci + nw);
    } while (ci + 1 < d && cur.y == 0);
  };
  if (ci + 1 < d && cur.y == 0) advance();
  while (ci + 1 < d) {
    uint32_t w[UNROLL], h[UNROLL], li[UNROLL];
    ull e[UNROLL];
#pragma unroll
    for (uint32_t x = 0; x < UNROLL; ++x) {
      li[x] = ci;
      w[x] = 0xFFFFFFFFu;
      if (ci + 1 < d) {
        const uint32_t t = ct + lane;
        if (t < cur.y) w[x] = dnbr[cur.x + t];
        ct += 64;
        if (ct >= cur.y) {
          ct = 0;
          advance();
        }
      }
    }
#pragma unroll
    for (uint32_t x = 0; x < UNROLL; ++x) {
      h[x] = hslot(w[x], HB);
      e[x] = f.hent[h[x]];
    }
#pragma unroll
    for (uint32_t x = 0; x < UNROLL; ++x) {
      const bool valid = w[x] <= last;  // also excludes the padding
      uint32_t key = (uint32_t)e[x];
      bool more = valid && key != w[x] && key != EMPTY_KEY;
      while (more) {
        h[x] = (h[x] + 1) & (H - 1);
        e[x] = f.hent[h[x]];
        key = (uint32_t)e[x];
        more = key != w[x] && key != EMPTY_KEY;
      }
      if (valid && key == w[x]) {
        const uint32_t i = li[x];
        const uint32_t j = (uint32_t)(e[x] >> 32);  // j > i: ids ascend along arcs and L is ascending
        atomicOr(&f.A[(j >> 6) * D + i], 1ull << (j & 63));
        atomicOr(&f.A[(i >> 6) * D + j], 1ull << (i & 63));
      }
    }
  }
  lds_sync();
  return raw;
}

// Credits E to local vertex c (owner lane c & 63, slot c >> 6).
template <int W> __device__ __forceinline__ void credit(ull *cnt, uint32_t c, ull E) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int q = 0; q < W; ++q)
    if ((uint32_t)q == (c >> 6) && lane == (c & 63)) cnt[q] += E;
}

// Bits above position b of one word (none for b = 63).
__device__ __forceinline__ ull above_bit(uint32_t b) { return b == 63 ? 0ull : (~0ull << (b + 1)); }

// Counting recursion: REM vertices still to choose from candidate set S.
// Returns the number of REM-cliques inside S (wave-uniform); every member
// vertex's counter is credited with the cliques it belongs to.  Words of S
// below LO are zero (LO is the word of the last chosen vertex: candidates
// only go up), so they are skipped at compile time; words from hi (the
// root's row count in words) up are zero too.  Every row word a step needs
// is requested before any is used, so a step waits for the LDS once, not
// once per word.
template <int W, int REM, int LO> struct Count;

// The candidates of S in word Q and above (Q is a compile-time word index,
// so the words below it are known to be zero).
template <int W, int REM, int Q> struct CountFrom {
  static __device__ ull run(const Frame<W> &f, const ull *S, ull *cnt, uint32_t hi) {
    if constexpr (Q >= W) {
      return 0;
    } else {
      if ((uint32_t)Q >= hi) return 0;
      ull total = 0;
      ull word = S[Q];
      while (word) {
        const uint32_t b = (uint32_t)__builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = (uint32_t)Q * 64 + b;
        ull S2[W];
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) S2[w] = w < Q ? 0ull : row(f, c, w);
#pragma unroll
        for (int w = Q; w < W; ++w) {
          S2[w] = S[w] & uniform_u64(S2[w]) & (w == Q ? above_bit(b) : ~0ull);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        const ull E = Count<W, REM - 1, Q>::run(f, S2, cnt, hi);
        if (lane_id() == b) cnt[Q] += E;
        total += E;
      }
      return total + CountFrom<W, REM, Q + 1>::run(f, S, cnt, hi);
    }
  }
};

template <int W, int REM, int LO> struct Count {
  static __device__ ull run(const Frame<W> &f, const ull *S, ull *cnt, uint32_t hi) {
    return CountFrom<W, REM, LO>::run(f, S, cnt, hi);
  }
};

// Edges inside S (words below LO and from hi up zero), credited to both
// endpoints; rows q of S handled by this wave are those with q % nw == wv
// (all rows when nw == 1).  Returns the wave's share of 2 * edges.  The
// whole wave is active here (control flow above is wave-uniform).
template <int W, int LO>
__device__ __forceinline__ uint32_t edges_in(const Frame<W> &f, const ull *S, ull *cnt, uint32_t hi, uint32_t wv,
                                             uint32_t nw) {
  const uint32_t lane = lane_id();
  uint32_t xs = 0;
#pragma unroll
  for (int q = LO; q < W; ++q) {
    if ((uint32_t)q >= hi) break;
    if (!S[q] || (uint32_t)q % nw != wv) continue;
    if ((S[q] >> lane) & 1ull) {
      const uint32_t j = (uint32_t)q * 64 + lane;
      ull r[W];
#pragma unroll
      for (int w = LO; w < W; ++w) r[w] = row(f, j, w);
      uint32_t x = 0;
#pragma unroll
      for (int w = LO; w < W; ++w) x += (uint32_t)__popcll(r[w] & S[w]);
      cnt[q] += x;
      xs += x;
    }
  }
  return wave_sum_u32_full(xs);
}

template <int W, int LO> struct Count<W, 2, LO> {
  static __device__ ull run(const Frame<W> &f, const ull *S, ull *cnt, uint32_t hi) {
    return (ull)(edges_in<W, LO>(f, S, cnt, hi, 0, 1) >> 1);
  }
};

// Count<W, REM, cq> for a first-level vertex in word cq (wave-uniform).
template <int W, int REM, int Q = 0>
__device__ __forceinline__ ull count_from(uint32_t cq, const Frame<W> &f, const ull *S, ull *cnt, uint32_t hi) {
  if constexpr (Q + 1 < W) {
    if (cq != (uint32_t)Q) return count_from<W, REM, Q + 1>(cq, f, S, cnt, hi);
  }
  return Count<W, REM, Q>::run(f, S, cnt, hi);
}

// Materialising recursion: same traversal, writes each clique as K new ids.
template <int W, int K, int REM> struct Emit {
  static __device__ void run(const Frame<W> &f, const ull *S, uint32_t *pfx, uint32_t *out, ull &pos,
                             ull lim) {
#pragma unroll
    for (int q = 0; q < W; ++q) {
      ull word = S[q];
      while (word) {
        const int b = __builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = (uint32_t)(q * 64 + b);
        ull S2[W];
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          S2[w] = S[w] & row(f, c, w) & above_mask(w, c);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        pfx[K - REM] = f.L[c];
        Emit<W, K, REM - 1>::run(f, S2, pfx, out, pos, lim);
      }
    }
  }
};

template <int W, int K> struct Emit<W, K, 2> {
  static __device__ void run(const Frame<W> &f, const ull *S, uint32_t *pfx, uint32_t *out, ull &pos,
                             ull lim) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const uint32_t j = (uint32_t)q * 64 + lane;
      ull H[W];
      uint32_t x = 0;
      const bool in = (S[q] >> lane) & 1ull;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        H[w] = in ? (row(f, j, w) & S[w] & above_mask(w, j)) : 0ull;
        x += (uint32_t)__popcll(H[w]);
      }
      const uint32_t incl = wave_incl_scan_u32(x);
      const uint32_t tot = __shfl(incl, 63, 64);
      ull at = pos + (incl - x);
      if (x) {
        const uint32_t uj = f.L[j];
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ull h = H[w];
          while (h) {
            const int bb = __builtin_ctzll(h);
            h &= h - 1;
            if (at >= lim) break;  // never taken when count and emit agree
            uint32_t *rw = out + at * K;
#pragma unroll
            for (int t = 0; t < K - 2; ++t) rw[t] = pfx[t];
            rw[K - 2] = uj;
            rw[K - 1] = f.L[w * 64 + bb];
            ++at;
          }
        }
      }
      pos += tot;
    }
  }
};

template <int W> __device__ __forceinline__ void all_bits(ull *S, uint32_t d) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t lo = (uint32_t)w * 64;
    S[w] = d >= lo + 64 ? ~0ull : (d > lo ? ((1ull << (d - lo)) - 1ull) : 0ull);
  }
}

// Hands the root index held by thread 0 to the workgroup.
template <int W> __device__ __forceinline__ uint32_t share_root(Frame<W> &f, uint32_t idx0) {
  if (threadIdx.x == 0) f.root = idx0;
  lds_sync();
  const uint32_t idx = f.root;
  lds_sync();  // every thread has read f.root before it is reused
  return idx;
}

// Adds E (wave-uniform) to the wave's pending count of graph g; at theta the
// pending count goes to gsum[g].  True when gsum[g] has passed cap.
__device__ __forceinline__ bool admit(ull *gsum, uint32_t g, ull E, ull &pend, ull cap, ull theta) {
  pend += E;
  if (pend < theta) return false;
  ull old = 0;
  if (lane_id() == 0) old = atomicAdd(&gsum[g], pend);
  old = __shfl(old, 0, 64);
  const bool over = old + pend > cap;
  pend = 0;
  return over;
}

template <int W, int K>
__global__ __launch_bounds__(64 * MAX_WAVES) void k_list_count(const uint32_t *__restrict__ doff,
                                                               const uint32_t *__restrict__ dnbr,
                                                               const uint4 *__restrict__ roots, uint32_t nroots,
                                                               uint32_t *queue, ull *gcnt, uint64_t cstride,
                                                               uint32_t ncopy, ull *rootT, ull *gtotal,
                                                               const uint32_t *gid, ull *gsum, ull cap, ull theta) {
  __shared__ Frame<W> f;
  const uint32_t tid = threadIdx.x, lane = tid & 63, nw = blockDim.x >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar loops
  ull tot = 0, w2 = 0;
  if (tid == 0) f.rt = 0;
  // per-vertex counts go to one of ncopy copies (summed afterwards), so a
  // hub's counter is not one serialised address for the whole launch
  gcnt += (uint64_t)(blockIdx.x % ncopy) * cstride;
  // Admission cap (cap > 0): gsum[g] gathers a lower bound of graph g's
  // cliques (each wave adds its pending count once it reaches theta); once
  // it passes cap the graph is over budget and its remaining work is
  // skipped (its counts are then partial and the caller drops it).
  ull pend = 0;
  uint32_t pg = 0xFFFFFFFFu;
  QueueCursor qc = queue_start(blockIdx.x);
  uint32_t pre = tid == 0 ? queue_next(queue, nroots, qc) : 0u;  // next root (thread 0)
  while (true) {
    if (tid == 0) {
      f.root = pre;
      uint32_t sk = 0;
      if (cap && pre < nroots) sk = ldc_u64(&gsum[gid ? gid[roots[pre].x] : 0u]) > cap;
      f.skip = sk;
    }
    lds_sync();
    const uint32_t idx = f.root, skip = f.skip;
    lds_sync();  // every thread has read the frame header before it is reused
    if (idx >= nroots) break;
    if (tid == 0) pre = queue_next(queue, nroots, qc);  // in flight while this root is processed
    if (skip) continue;
    const uint4 rec = roots[idx];  // root, its out-list start and out-degree: one round trip
    const uint32_t r = rec.x, base = rec.y, d = rec.z;
    uint32_t g = 0;
    if (cap) {
      g = gid ? gid[r] : 0u;
      if (g != pg) {
        pend = 0;
        pg = g;
      }
    }
    w2 += build_rows<W>(f, doff, dnbr, base, d);
    ull cnt[W];
#pragma unroll
    for (int q = 0; q < W; ++q) cnt[q] = 0;
    ull S[W];
    all_bits<W>(S, d);
    const uint32_t hi = (d + 63) >> 6;  // row words in use
    ull rt = 0;
    if constexpr (K == 3) {
      rt = edges_in<W, 0>(f, S, cnt, hi, wv, nw);  // twice the wave's share of edges
      if (cap) admit(gsum, g, rt >> 1, pend, cap, theta);
    } else {
      for (uint32_t c = wv; c < d; c += nw) {
        const uint32_t cq = c >> 6;
        ull S2[W];
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) S2[w] = row(f, c, w);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          S2[w] = uniform_u64(S2[w]) & above_mask(w, c);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(K - 2)) continue;
        const ull E = count_from<W, K - 2>(cq, f, S2, cnt, hi);
        credit<W>(cnt, c, E);
        rt += E;
        if (cap && admit(gsum, g, E, pend, cap, theta)) break;
      }
    }
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const uint32_t j = (uint32_t)q * 64 + lane;
      if (j < d && cnt[q]) atomicAdd(&gcnt[f.L[j]], cnt[q]);
    }
    if (lane == 0 && rt) atomicAdd(&f.rt, rt);
    lds_sync();
    if (tid == 0) {
      const ull all = K == 3 ? (f.rt >> 1) : f.rt;
      f.rt = 0;
      if (all) atomicAdd(&gcnt[r], all);
      rootT[r] = all;
      tot += all;
    }
  }
  // totals per shard: [s*8] = cliques, [s*8+1] = staged out-wedge sum
  ull *mine = gtotal + (blockIdx.x % QSHARDS) * 8;
  w2 = wave_sum_u64(w2);
  if (lane == 0 && w2) atomicAdd(mine + 1, w2);
  if (tid == 0 && tot) atomicAdd(mine, tot);
}

// gT[gid[r]] += rootT[r], reduced per wave when neighbours share a graph
__global__ void k_graph_T(const ull *rootT, const uint32_t *gid, uint32_t n, ull *gT) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    if (rootT[r]) atomicAdd(&gT[gid[r]], rootT[r]);
}

__global__ void k_drop_graphs(ull *rootT, ull *cnt, const uint32_t *gid, const uint8_t *drop, uint32_t n) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    if (drop[gid[r]]) {
      rootT[r] = 0;
      cnt[r] = 0;
    }
}

// cnt[v] = sum of the ncopy copies (copy 0 is cnt itself)
__global__ void k_sum_copies(ull *cnt, uint64_t n, uint64_t cstride, uint32_t ncopy) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    ull s = cnt[v];
    for (uint32_t c = 1; c < ncopy; ++c) s += cnt[c * cstride + v];
    cnt[v] = s;
  }
}

template <int W, int K>
__global__ __launch_bounds__(64) void k_list_emit(const uint32_t *__restrict__ doff,
                                                  const uint32_t *__restrict__ dnbr,
                                                  const uint4 *__restrict__ roots, uint32_t nroots,
                                                  uint32_t *queue, const ull *rootOff, uint32_t *out) {
  __shared__ Frame<W> f;
  QueueCursor qc = queue_start(blockIdx.x);
  uint32_t pre = threadIdx.x == 0 ? queue_next(queue, nroots, qc) : 0u;
  while (true) {
    const uint32_t idx = share_root(f, pre);
    if (idx >= nroots) break;
    if (threadIdx.x == 0) pre = queue_next(queue, nroots, qc);
    const uint4 rec = roots[idx];
    const uint32_t r = rec.x, base = rec.y, d = rec.z;
    ull pos = rootOff[r];
    const ull lim = rootOff[r + 1];
    if (lim == pos) continue;
    build_rows<W>(f, doff, dnbr, base, d);
    ull S[W];
    all_bits<W>(S, d);
    uint32_t pfx[K];
    pfx[0] = r;
    Emit<W, K, K - 1>::run(f, S, pfx, out, pos, lim);
  }
}

// ---------------------------------------------------------------------------
// Wide roots: out-degree in (512, LIST_WIDE_MAX].  The reference sizes its
// neighbour masks from the widest out-list (src/gpuMemoryAllocation.cu:58),
// so any out-degree lists there.  A root too wide for an LDS frame keeps its
// bitmask rows in HBM instead; one workgroup of WIDE_THREADS takes one root
// at a time:
//   - L = N+(r) and an id -> index table (load <= 1/2) in LDS;
//   - rows G[j*Wd + w] (Wd = ceil(d/64) words, row-major) in the
//     workgroup's own slice of HBM, built with global atomics from the
//     members' out-lists (bit j of row i and bit i of row j per hit);
//   - candidate sets are spread over the lanes of a wave (lane w holds word
//     w; Wd <= 64), so S & row(c) is one coalesced row read and a set's size
//     one DPP sum;
//   - per-member counts in LDS, flushed with one atomic per member.
// Only graphs whose orientation leaves out-lists wider than 512 reach this
// path (a degeneracy above 512), so it favours simplicity over speed.
constexpr uint32_t WIDE_THREADS = 1024;
constexpr uint32_t WIDE_HB = 13, WIDE_H = 1u << WIDE_HB;  // 8,192 table slots for up to 4,096 members

struct WideFrame {
  ull hent[WIDE_H];          // low word member id (EMPTY_KEY when free), high word its index
  ull cnt[LIST_WIDE_MAX];    // cliques of the current root containing member j
  uint32_t L[LIST_WIDE_MAX];  // N+(r), ascending
  ull rt;
  uint32_t root, skip;
};

__device__ __forceinline__ uint32_t whslot(uint32_t key) {
  return (uint32_t)__umul24(key ^ (key >> 24), 0x9E3779u) >> (32 - WIDE_HB);
}

// Workgroup barrier that also orders this workgroup's global stores and
// atomics (the rows) with its later global loads: release, barrier, acquire
// at agent scope (the acquire drops stale L1 lines).
__device__ __forceinline__ void wide_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Builds the rows of the root whose out-list is dnbr[base, base + d) in G.
// Returns this thread's share of the staged out-wedge sum.
__device__ uint32_t wide_build(WideFrame &f, ull *G, const uint32_t *__restrict__ doff,
                               const uint32_t *__restrict__ dnbr, uint32_t base, uint32_t d) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, nw = nt >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t Wd = (d + 63) >> 6;
  for (uint32_t t = tid; t < WIDE_H; t += nt) f.hent[t] = ~0ull;
  for (uint32_t t = tid; t < d; t += nt) f.cnt[t] = 0;
  for (uint64_t e = tid; e < (uint64_t)d * Wd; e += nt) G[e] = 0;
  wide_sync();
  uint32_t *hk = reinterpret_cast<uint32_t *>(f.hent);
  for (uint32_t t = tid; t < d; t += nt) {
    const uint32_t u = dnbr[base + t];
    f.L[t] = u;
    uint32_t h = whslot(u);
    while (atomicCAS(&hk[2 * h], EMPTY_KEY, u) != EMPTY_KEY) h = (h + 1) & (WIDE_H - 1);
    hk[2 * h + 1] = t;
  }
  wide_sync();
  uint32_t raw = 0;
  const uint32_t last = f.L[d - 1];
  for (uint32_t i = wv; i < d; i += nw) {
    const uint32_t u = f.L[i];
    const uint32_t ou = doff[u], du = doff[u + 1] - ou;
    if (lane == 0) raw += du;
    if (i + 1 == d) continue;  // the last member has no later member
    for (uint32_t c0 = 0; c0 < du; c0 += 64) {
      const uint32_t t = c0 + lane;
      const uint32_t w = t < du ? dnbr[ou + t] : 0xFFFFFFFFu;
      if (w > last) continue;
      uint32_t h = whslot(w);
      ull e = f.hent[h];
      while ((uint32_t)e != w && (uint32_t)e != EMPTY_KEY) {
        h = (h + 1) & (WIDE_H - 1);
        e = f.hent[h];
      }
      if ((uint32_t)e == w) {
        const uint32_t j = (uint32_t)(e >> 32);  // j > i: ids ascend along arcs and L is ascending
        atomicOr(&G[(uint64_t)i * Wd + (j >> 6)], 1ull << (j & 63));
        atomicOr(&G[(uint64_t)j * Wd + (i >> 6)], 1ull << (i & 63));
      }
    }
  }
  wide_sync();
  return raw;
}

// Word `lane` of row j (0 past the row's Wd words).
__device__ __forceinline__ ull wide_word(const ull *G, uint32_t Wd, uint32_t j, uint32_t lane) {
  return lane < Wd ? G[(uint64_t)j * Wd + lane] : 0ull;
}

// Lane-distributed bits above vertex c = 64q + b.
__device__ __forceinline__ ull wide_above(uint32_t lane, uint32_t q, uint32_t b) {
  return lane < q ? 0ull : lane > q ? ~0ull : (b == 63 ? 0ull : (~0ull << (b + 1)));
}

__device__ __forceinline__ ull wave_word(ull S, uint32_t q) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)S, q), hi = __builtin_amdgcn_readlane((uint32_t)(S >> 32), q);
  return ((ull)hi << 32) | lo;
}

// REM-cliques inside the lane-distributed set S, every member credited in
// f.cnt; returns the count (wave-uniform).  The whole wave is active.
template <int REM> struct WCount {
  static __device__ ull run(WideFrame &f, const ull *G, uint32_t Wd, ull S) {
    const uint32_t lane = lane_id();
    ull total = 0;
    ull nz = __ballot(S != 0ull);
    while (nz) {
      const uint32_t q = (uint32_t)__builtin_ctzll(nz);
      nz &= nz - 1;
      ull word = wave_word(S, q);
      while (word) {
        const uint32_t b = (uint32_t)__builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = q * 64 + b;
        const ull row = wide_word(G, Wd, c, lane);
        if constexpr (REM == 2) {
          const uint32_t x = wave_sum_u32_full((uint32_t)__popcll(row & S));
          if (lane == 0 && x) atomicAdd(&f.cnt[c], (ull)x);
          total += x;
        } else {
          const ull S2 = S & row & wide_above(lane, q, b);
          const uint32_t pc = wave_sum_u32_full((uint32_t)__popcll(S2));
          if (pc < (uint32_t)(REM - 1)) continue;
          const ull E = WCount<REM - 1>::run(f, G, Wd, S2);
          if (lane == 0 && E) atomicAdd(&f.cnt[c], E);
          total += E;
        }
      }
    }
    return REM == 2 ? total / 2 : total;
  }
};

template <int K>
__global__ __launch_bounds__(WIDE_THREADS) void k_list_count_wide(
    const uint32_t *__restrict__ doff, const uint32_t *__restrict__ dnbr, const uint4 *__restrict__ roots,
    uint32_t nroots, uint32_t *queue, ull *gcnt, uint64_t cstride, uint32_t ncopy, ull *rootT, ull *gtotal,
    const uint32_t *gid, ull *gsum, ull cap, ull *rowbuf, uint64_t rowstride) {
  __shared__ WideFrame f;
  const uint32_t tid = threadIdx.x, lane = tid & 63, nw = blockDim.x >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  ull *G = rowbuf + (uint64_t)blockIdx.x * rowstride;
  gcnt += (uint64_t)(blockIdx.x % ncopy) * cstride;
  ull tot = 0, w2 = 0;
  QueueCursor qc = queue_start(blockIdx.x);
  uint32_t pre = tid == 0 ? queue_next(queue, nroots, qc) : 0u;
  while (true) {
    if (tid == 0) {
      f.root = pre;
      f.rt = 0;
      f.skip = cap && pre < nroots && ldc_u64(&gsum[gid ? gid[roots[pre].x] : 0u]) > cap;
    }
    lds_sync();
    const uint32_t idx = f.root, skip = f.skip;
    lds_sync();
    if (idx >= nroots) break;
    if (tid == 0) pre = queue_next(queue, nroots, qc);
    if (skip) continue;
    const uint4 rec = roots[idx];
    const uint32_t r = rec.x, base = rec.y, d = rec.z, Wd = (d + 63) >> 6;
    w2 += wide_build(f, G, doff, dnbr, base, d);
    ull rt = 0;
    if constexpr (K == 3) {
      const ull Sall = lane < Wd ? (lane + 1 < Wd || (d & 63) == 0 ? ~0ull : ((1ull << (d & 63)) - 1ull)) : 0ull;
      for (uint32_t j = wv; j < d; j += nw) {
        const uint32_t x = wave_sum_u32_full((uint32_t)__popcll(wide_word(G, Wd, j, lane) & Sall));
        if (lane == 0 && x) atomicAdd(&f.cnt[j], (ull)x);
        rt += x;  // twice the wave's share of edges
      }
    } else {
      for (uint32_t c = wv; c < d; c += nw) {
        const ull S = wide_word(G, Wd, c, lane) & wide_above(lane, c >> 6, c & 63);
        const uint32_t pc = wave_sum_u32_full((uint32_t)__popcll(S));
        if (pc < (uint32_t)(K - 2)) continue;
        const ull E = WCount<K - 2>::run(f, G, Wd, S);
        if (lane == 0 && E) atomicAdd(&f.cnt[c], E);
        rt += E;
      }
    }
    if (lane == 0 && rt) atomicAdd(&f.rt, rt);
    lds_sync();
    const ull all = K == 3 ? (f.rt >> 1) : f.rt;
    for (uint32_t t = tid; t < d; t += blockDim.x)
      if (f.cnt[t]) atomicAdd(&gcnt[f.L[t]], f.cnt[t]);
    if (tid == 0) {
      if (all) atomicAdd(&gcnt[r], all);
      rootT[r] = all;
      if (cap && all) atomicAdd(&gsum[gid ? gid[r] : 0u], all);
      tot += all;
    }
    lds_sync();  // the counts are read before the next root clears them
  }
  ull *mine = gtotal + (blockIdx.x % QSHARDS) * 8;
  w2 = wave_sum_u64(w2);
  if (lane == 0 && w2) atomicAdd(mine + 1, w2);
  if (tid == 0 && tot) atomicAdd(mine, tot);
}

// Materialises the cliques of wide roots: the workgroup builds the rows and
// wave 0 writes the root's cliques in a fixed order (REM-cliques of the
// lane-distributed set S below the prefix pfx) from rootOff[r] on.
template <int K, int REM> struct WEmit {
  static __device__ void run(const WideFrame &f, const ull *G, uint32_t Wd, ull S, uint32_t *pfx, uint32_t *out,
                             ull &pos, ull lim) {
    const uint32_t lane = lane_id();
    ull nz = __ballot(S != 0ull);
    while (nz) {
      const uint32_t q = (uint32_t)__builtin_ctzll(nz);
      nz &= nz - 1;
      ull word = wave_word(S, q);
      while (word) {
        const uint32_t b = (uint32_t)__builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = q * 64 + b;
        const ull H = S & wide_word(G, Wd, c, lane) & wide_above(lane, q, b);
        if constexpr (REM == 2) {
          const uint32_t x = (uint32_t)__popcll(H);
          const uint32_t incl = wave_incl_scan_u32(x);
          const uint32_t tot = __shfl(incl, 63, 64);
          ull at = pos + (incl - x);
          ull h = H;
          while (h) {
            const uint32_t bb = (uint32_t)__builtin_ctzll(h);
            h &= h - 1;
            if (at >= lim) break;  // never taken when count and emit agree
            uint32_t *rw = out + at * K;
#pragma unroll
            for (int t = 0; t < K - 2; ++t) rw[t] = pfx[t];
            rw[K - 2] = f.L[c];
            rw[K - 1] = f.L[lane * 64 + bb];
            ++at;
          }
          pos += tot;
        } else {
          const uint32_t pc = wave_sum_u32_full((uint32_t)__popcll(H));
          if (pc < (uint32_t)(REM - 1)) continue;
          pfx[K - REM] = f.L[c];
          WEmit<K, REM - 1>::run(f, G, Wd, H, pfx, out, pos, lim);
        }
      }
    }
  }
};

template <int K>
__global__ __launch_bounds__(WIDE_THREADS) void k_list_emit_wide(
    const uint32_t *__restrict__ doff, const uint32_t *__restrict__ dnbr, const uint4 *__restrict__ roots,
    uint32_t nroots, uint32_t *queue, const ull *rootOff, uint32_t *out, ull *rowbuf, uint64_t rowstride) {
  __shared__ WideFrame f;
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  ull *G = rowbuf + (uint64_t)blockIdx.x * rowstride;
  QueueCursor qc = queue_start(blockIdx.x);
  uint32_t pre = tid == 0 ? queue_next(queue, nroots, qc) : 0u;
  while (true) {
    if (tid == 0) f.root = pre;
    lds_sync();
    const uint32_t idx = f.root;
    lds_sync();
    if (idx >= nroots) break;
    if (tid == 0) pre = queue_next(queue, nroots, qc);
    const uint4 rec = roots[idx];
    const uint32_t r = rec.x, base = rec.y, d = rec.z, Wd = (d + 63) >> 6;
    ull pos = rootOff[r];
    const ull lim = rootOff[r + 1];
    if (lim == pos) continue;
    wide_build(f, G, doff, dnbr, base, d);
    if (tid < 64) {  // wave 0: a fixed order
      const ull Sall = lane < Wd ? (lane + 1 < Wd || (d & 63) == 0 ? ~0ull : ((1ull << (d & 63)) - 1ull)) : 0ull;
      uint32_t pfx[K];
      pfx[0] = r;
      WEmit<K, K - 1>::run(f, G, Wd, Sall, pfx, out, pos, lim);
    }
    lds_sync();  // wave 0 is done with the rows before the next root rebuilds them
  }
}

// Root classification: class c holds roots with 64*2^(c-1) < outdeg <= 64*2^c
// and outdeg >= k-1 (smaller roots have no k-clique); class 4 = the wide
// roots (rows in HBM).
__device__ __forceinline__ int width_class(uint32_t d) {
  return d <= 64 ? 0 : d <= 128 ? 1 : d <= 256 ? 2 : d <= 512 ? 3 : 4;
}

__global__ __launch_bounds__(256) void k_class_count(const uint32_t *doff, uint32_t n, uint32_t k,
                                                     uint32_t *ccount) {
  __shared__ uint32_t sc[8];
  if (threadIdx.x < 8) sc[threadIdx.x] = 0;
  __syncthreads();
  uint32_t mine[5] = {0, 0, 0, 0, 0};
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t d = doff[r + 1] - doff[r];
    if (d + 1 < k) continue;
    const int c = width_class(d);
#pragma unroll
    for (int i = 0; i < 5; ++i) mine[i] += (c == i);
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const uint32_t v = wave_sum_u32(mine[i]);
    if (lane_id() == 0 && v) atomicAdd(&sc[i], v);
  }
  __syncthreads();
  if (threadIdx.x < 5 && sc[threadIdx.x]) atomicAdd(&ccount[threadIdx.x], sc[threadIdx.x]);
}

// Writes each root into its class's range (classes in order, each starting
// where the counts of the earlier ones end): one global atomic per class per
// 256 roots.
constexpr int NCLASS_ALL = NCLASS + 1;  // the LDS classes, then the wide roots

__global__ __launch_bounds__(256) void k_class_fill(const uint32_t *doff, uint32_t n, uint32_t k,
                                                    const uint32_t *ccount, uint32_t *cfill, uint4 *roots) {
  __shared__ uint32_t wc[4][NCLASS_ALL], bbase[NCLASS_ALL];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  for (uint32_t r0 = blockIdx.x * 256; r0 < n; r0 += gridDim.x * 256) {
    const uint32_t r = r0 + threadIdx.x;
    int c = -1;
    uint32_t b = 0, d = 0;
    if (r < n) {
      b = doff[r];
      d = doff[r + 1] - b;
      if (d + 1 >= k) c = width_class(d);
    }
    ull bal[NCLASS_ALL];
#pragma unroll
    for (int i = 0; i < NCLASS_ALL; ++i) {
      bal[i] = __ballot(c == i);
      if (lane == 0) wc[wv][i] = (uint32_t)__popcll(bal[i]);
    }
    __syncthreads();
    if (threadIdx.x < NCLASS_ALL) {
      const uint32_t i = threadIdx.x;
      const uint32_t tot = wc[0][i] + wc[1][i] + wc[2][i] + wc[3][i];
      uint32_t start = 0;
      for (uint32_t j = 0; j < i; ++j) start += ccount[j];
      bbase[i] = tot ? start + atomicAdd(&cfill[i], tot) : 0u;
    }
    __syncthreads();
    if (c >= 0 && c < NCLASS_ALL) {
      uint32_t at = bbase[c] + lanes_below(bal[c]);
      for (uint32_t v = 0; v < wv; ++v) at += wc[v][c];
      roots[at] = make_uint4(r, b, d, 0u);
    }
    __syncthreads();
  }
}

// incidence index: inc[incoff[v] + i] = id of the i-th clique containing v.
// A round of the workgroup takes INC_E * 256 consecutive clique entries and
// reserves each vertex's slots with one atomic per round (hub vertices repeat
// across thousands of consecutive cliques).
constexpr int INC_E = 8, INC_TB = 4096;

__global__ __launch_bounds__(256) void k_incidence(const uint32_t *cl, uint64_t T, uint32_t k, const ull *incoff,
                                                   uint32_t *fillc, uint32_t *inc) {
  __shared__ KeyTable<INC_TB> tb;
  const uint64_t Tk = T * k, per = (uint64_t)INC_E * 256;
  for (uint64_t r0 = (uint64_t)blockIdx.x * per; r0 < Tk; r0 += (uint64_t)gridDim.x * per) {
    bool want[INC_E];
    uint32_t v[INC_E], slot[INC_E];
#pragma unroll
    for (int e = 0; e < INC_E; ++e) {
      const uint64_t i = r0 + (uint64_t)e * 256 + threadIdx.x;
      want[e] = i < Tk;
      v[e] = want[e] ? cl[i] : 0u;
    }
    block_slot_by_key<INC_E, INC_TB, uint32_t>(tb, want, v, slot, fillc);
#pragma unroll
    for (int e = 0; e < INC_E; ++e) {
      const uint64_t i = r0 + (uint64_t)e * 256 + threadIdx.x;
      if (want[e] && incoff[v[e]] + slot[e] < incoff[v[e] + 1]) inc[incoff[v[e]] + slot[e]] = (uint32_t)(i / k);
    }
  }
}

// k = 2 (edge density: main.cu:1365-1423 with writeEdgeCliques,
// helpers.cu:552-585): the 2-cliques are the DAG arcs.  cnt[r] = degree,
// rootT[r] = out-degree, cliques row a = (r, dnbr[a]) for arc a of root r.
__global__ void k_edge_out(const uint32_t *doff, uint32_t n, ull *cnt, ull *rootT) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const ull d = doff[r + 1] - doff[r];
    cnt[r] = d;
    rootT[r] = d;
  }
}

// in-degrees: rounds of INC_E * 256 arcs, one atomic per distinct head per round
__global__ __launch_bounds__(256) void k_edge_in(const uint32_t *dnbr, uint64_t arcs, ull *cnt) {
  __shared__ KeyTable<INC_TB> tb;
  const uint64_t per = (uint64_t)INC_E * 256;
  for (uint64_t r0 = (uint64_t)blockIdx.x * per; r0 < arcs; r0 += (uint64_t)gridDim.x * per) {
    bool want[INC_E];
    uint32_t v[INC_E];
    ull slot[INC_E];
#pragma unroll
    for (int e = 0; e < INC_E; ++e) {
      const uint64_t i = r0 + (uint64_t)e * 256 + threadIdx.x;
      want[e] = i < arcs;
      v[e] = want[e] ? dnbr[i] : 0u;
    }
    block_slot_by_key<INC_E, INC_TB, ull>(tb, want, v, slot, cnt);
  }
}

__global__ __launch_bounds__(256) void k_edge_emit(const uint32_t *doff, const uint32_t *dnbr, uint32_t n,
                                                   uint32_t *cl) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwv = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t r = gw; r < n; r += nwv)
    for (uint32_t a = doff[r] + lane; a < doff[r + 1]; a += KCDS_WAVE) {
      cl[2ull * a] = r;
      cl[2ull * a + 1] = dnbr[a];
    }
}

typedef void (*CountFn)(const uint32_t *, const uint32_t *, const uint4 *, uint32_t, uint32_t *, ull *, uint64_t,
                        uint32_t, ull *, ull *, const uint32_t *, ull *, ull, ull);
typedef void (*EmitFn)(const uint32_t *, const uint32_t *, const uint4 *, uint32_t, uint32_t *, const ull *,
                       uint32_t *);

template <int W> CountFn count_fn(uint32_t k) {
  switch (k) {
    case 3: return k_list_count<W, 3>;
    case 4: return k_list_count<W, 4>;
    case 5: return k_list_count<W, 5>;
    case 6: return k_list_count<W, 6>;
    case 7: return k_list_count<W, 7>;
    case 8: return k_list_count<W, 8>;
  }
  return nullptr;
}

template <int W> EmitFn emit_fn(uint32_t k) {
  switch (k) {
    case 3: return k_list_emit<W, 3>;
    case 4: return k_list_emit<W, 4>;
    case 5: return k_list_emit<W, 5>;
    case 6: return k_list_emit<W, 6>;
    case 7: return k_list_emit<W, 7>;
    case 8: return k_list_emit<W, 8>;
  }
  return nullptr;
}

CountFn class_count_fn(int cls, uint32_t k) {
  return cls == 0 ? count_fn<1>(k) : cls == 1 ? count_fn<2>(k) : cls == 2 ? count_fn<4>(k) : count_fn<8>(k);
}

// Waves per root by class when not tuned: heavy classes split each root.
const uint32_t kWavesPerRoot[NCLASS] = {1, 2, 4, 8};
constexpr uint32_t WAVE_SLOTS_PER_CU = 32;

uint32_t lds_bytes(int cls) {
  switch (cls) {
    case 0: return sizeof(Frame<1>);
    case 1: return sizeof(Frame<2>);
    case 2: return sizeof(Frame<4>);
    default: return sizeof(Frame<8>);
  }
}


uint32_t wg_grid(const Ctx &c, int cls, uint32_t nroots, uint32_t nw) {
  uint32_t slots = c.list_waves_per_cu ? c.list_waves_per_cu : WAVE_SLOTS_PER_CU;
  uint32_t per_cu = std::max<uint32_t>(1, slots / nw);
  per_cu = std::min<uint32_t>(per_cu, std::max<uint32_t>(1, (160u * 1024u) / lds_bytes(cls)));
  return std::max<uint32_t>(1, std::min<uint32_t>(nroots, (uint32_t)c.num_cus * per_cu));
}

// c.qctr layout (32-bit words): [0, 64) classification counters, then one
// sharded queue per class for counting, one per class for materialising and
// one for the autotuner.
constexpr uint32_t QBASE = 256;
constexpr uint32_t QCTR_WORDS = QBASE + (2 * NCLASS + 3) * QUEUE_WORDS;
uint32_t *count_queue(Ctx &c, int cls) { return c.qctr.as<uint32_t>() + QBASE + cls * QUEUE_WORDS; }
uint32_t *emit_queue(Ctx &c, int cls) { return c.qctr.as<uint32_t>() + QBASE + (NCLASS + cls) * QUEUE_WORDS; }
uint32_t *tune_queue(Ctx &c) { return c.qctr.as<uint32_t>() + QBASE + 2 * NCLASS * QUEUE_WORDS; }
uint32_t *wide_count_queue(Ctx &c) { return c.qctr.as<uint32_t>() + QBASE + (2 * NCLASS + 1) * QUEUE_WORDS; }
uint32_t *wide_emit_queue(Ctx &c) { return c.qctr.as<uint32_t>() + QBASE + (2 * NCLASS + 2) * QUEUE_WORDS; }

typedef void (*WideCountFn)(const uint32_t *, const uint32_t *, const uint4 *, uint32_t, uint32_t *, ull *, uint64_t,
                            uint32_t, ull *, ull *, const uint32_t *, ull *, ull, ull *, uint64_t);
typedef void (*WideEmitFn)(const uint32_t *, const uint32_t *, const uint4 *, uint32_t, uint32_t *, const ull *,
                           uint32_t *, ull *, uint64_t);

WideCountFn wide_count_fn(uint32_t k) {
  switch (k) {
    case 3: return k_list_count_wide<3>;
    case 4: return k_list_count_wide<4>;
    case 5: return k_list_count_wide<5>;
    case 6: return k_list_count_wide<6>;
    case 7: return k_list_count_wide<7>;
    case 8: return k_list_count_wide<8>;
  }
  return nullptr;
}

WideEmitFn wide_emit_fn(uint32_t k) {
  switch (k) {
    case 3: return k_list_emit_wide<3>;
    case 4: return k_list_emit_wide<4>;
    case 5: return k_list_emit_wide<5>;
    case 6: return k_list_emit_wide<6>;
    case 7: return k_list_emit_wide<7>;
    case 8: return k_list_emit_wide<8>;
  }
  return nullptr;
}

// Workgroups for the wide roots and the HBM rows they need: one slice of
// dmax x ceil(dmax/64) words per workgroup.
uint32_t wide_grid(Ctx &c, uint64_t &rowstride) {
  const uint32_t nr = c.class_count[NCLASS];
  rowstride = (uint64_t)c.dmax * ((c.dmax + 63) / 64);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(nr, (uint32_t)c.num_cus));
  buf_reserve(c, c.wide, sizeof(ull) * (rowstride * grid + 64));
  return grid;
}

// Copies of the per-vertex counters (8 while they fit comfortably).
uint32_t count_copies(uint32_t n) { return n <= (1u << 26) ? 8u : 1u; }
uint64_t copy_stride(uint32_t n) { return (((uint64_t)n + 1 + 63) / 64) * 64; }

void classify_roots(Ctx &c, uint32_t k) {
  const uint32_t n = c.n;
  buf_reserve(c, c.roots, sizeof(uint4) * ((size_t)n + 1));
  buf_reserve(c, c.qctr, QCTR_WORDS * sizeof(uint32_t));
  uint32_t *q = c.qctr.as<uint32_t>();
  HIP_CHECK(hipMemsetAsync(q, 0, QCTR_WORDS * sizeof(uint32_t), c.stream));
  uint32_t *ccount = q + 16, *cfill = q + 32;
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 256));
  if (n) {
    k_class_count<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), n, k, ccount);
    KCDS_LAUNCH_CHECK();
  }
  uint32_t *h = (uint32_t *)c.host_scratch(64);
  HIP_CHECK(hipMemcpyAsync(h, ccount, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (h[4] && c.dmax > LIST_WIDE_MAX)
    fail(KCDS_E_LIMIT, "out-degree %u above %u (the widest out-list listed)", c.dmax, LIST_WIDE_MAX);
  uint32_t s = 0;
  for (int i = 0; i < NCLASS_ALL; ++i) {
    c.class_count[i] = h[i];
    c.class_start[i] = s;
    s += h[i];
  }
  if (n && s) {
    k_class_fill<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), n, k, ccount, cfill, c.roots.as<uint4>());
    KCDS_LAUNCH_CHECK();
  }
}

// One launch of class cls over its roots; counters accumulate into the
// ncopy copies of gcnt, rootT and the sharded totals.
void launch_class(Ctx &c, int cls, uint32_t k, uint32_t nw, uint32_t *queue, ull *gcnt, ull *rootT, ull *totals,
                  ull cap = 0, hipStream_t st = nullptr) {
  if (!st) st = c.stream;
  const uint32_t nr = c.class_count[cls];
  const uint4 *roots = c.roots.as<uint4>() + c.class_start[cls];
  const uint32_t *gid = c.ngraphs > 1 ? c.gid.as<uint32_t>() : nullptr;
  const ull theta = cap / 64 + 1;
  hipLaunchKernelGGL(class_count_fn(cls, k), dim3(wg_grid(c, cls, nr, nw)), dim3(64 * nw), 0, st,
                     c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(), roots, nr, queue, gcnt, copy_stride(c.n),
                     count_copies(c.n), rootT, totals, gid, cap ? c.gsum.as<ull>() : nullptr, cap, theta);
  KCDS_LAUNCH_CHECK();
}

// Sums the sharded totals on the host: (cliques, staged out-wedge sum).
void read_totals(Ctx &c, const ull *totals, ull &T, ull &w2) {
  ull *h = (ull *)c.host_scratch(sizeof(ull) * QSHARDS * 8);
  HIP_CHECK(hipMemcpyAsync(h, totals, sizeof(ull) * QSHARDS * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  T = 0;
  w2 = 0;
  for (uint32_t i = 0; i < QSHARDS; ++i) {
    T += h[i * 8];
    w2 += h[i * 8 + 1];
  }
}

// The width classes with waves per root nw[cls], all at once: the widest on
// the context stream, the others on three more streams (default priority),
// each class filling the others' tails (they share only atomic counters).
// Returns the number of launches; the context stream waits for all.
uint32_t run_classes(Ctx &c, uint32_t k, const uint32_t nw[NCLASS], ull *gcnt, ull *rootT, ull *totals, ull cap) {
  if (!c.lfork) {
    HIP_CHECK(hipEventCreateWithFlags(&c.lfork, hipEventDisableTiming));
    for (int i = 0; i < 3; ++i) {
      if (!c.lstream[i]) HIP_CHECK(hipStreamCreateWithFlags(&c.lstream[i], hipStreamNonBlocking));
      HIP_CHECK(hipEventCreateWithFlags(&c.ljoin[i], hipEventDisableTiming));
    }
  }
  uint32_t launches = 0;
  if (c.class_count[NCLASS]) {  // the wide roots first, on the context stream (their rows buffer is reserved here)
    uint64_t rowstride = 0;
    const uint32_t grid = wide_grid(c, rowstride);
    const uint32_t *gid = c.ngraphs > 1 ? c.gid.as<uint32_t>() : nullptr;
    hipLaunchKernelGGL(wide_count_fn(k), dim3(grid), dim3(WIDE_THREADS), 0, c.stream, c.doff.as<uint32_t>(),
                       c.dnbr.as<uint32_t>(), c.roots.as<uint4>() + c.class_start[NCLASS], c.class_count[NCLASS],
                       wide_count_queue(c), gcnt, copy_stride(c.n), count_copies(c.n), rootT, totals, gid,
                       cap ? c.gsum.as<ull>() : nullptr, cap, c.wide.as<ull>(), rowstride);
    KCDS_LAUNCH_CHECK();
    ++launches;
  }
  HIP_CHECK(hipEventRecord(c.lfork, c.stream));
  for (int cls = NCLASS - 1; cls >= 0; --cls) {
    if (!c.class_count[cls]) continue;
    hipStream_t st = cls == NCLASS - 1 ? c.stream : c.lstream[cls];
    if (st != c.stream) HIP_CHECK(hipStreamWaitEvent(st, c.lfork, 0));
    launch_class(c, cls, k, nw[cls], count_queue(c, cls), gcnt, rootT, totals, cap, st);
    if (st != c.stream) {
      HIP_CHECK(hipEventRecord(c.ljoin[cls], st));
      HIP_CHECK(hipStreamWaitEvent(c.stream, c.ljoin[cls], 0));
    }
    ++launches;
  }
  return launches;
}

}  // namespace

uint32_t list_waves_per_root(const Ctx &c, int cls, uint32_t k) {
  uint32_t nw = c.list_waves_per_root[cls] ? c.list_waves_per_root[cls] : kWavesPerRoot[cls];
  nw = std::min<uint32_t>(std::max<uint32_t>(nw, 1), MAX_WAVES);
  if (k == 3) nw = std::min<uint32_t>(nw, 1u << cls);  // K = 3 splits rows by 64-row word
  return nw;
}

// Picks the waves per root of every class by timing each candidate on the
// current graph (results of the trial runs go to scratch buffers).
void list_autotune(Ctx &c, uint32_t k) {
  if (!c.oriented) fail(KCDS_E_STATE, "autotune before kcds_orient");
  if (k < 2 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
  if (k == 2) {  // no listing kernels to shape
    for (int cls = 0; cls < NCLASS; ++cls) c.tune_ms[cls] = 0;
    return;
  }
  const uint32_t n = c.n;
  classify_roots(c, k);
  const uint64_t cw = copy_stride(n) * count_copies(n);
  buf_reserve(c, c.tmp2, sizeof(ull) * cw);
  buf_reserve(c, c.tmp3, sizeof(ull) * ((size_t)n + 2));
  buf_reserve(c, c.lsum, sizeof(ull) * QSHARDS * 8 * 2);
  ull *gcnt = c.tmp2.as<ull>(), *rootT = c.tmp3.as<ull>(), *totals = c.lsum.as<ull>() + QSHARDS * 8;
  uint32_t *q = tune_queue(c);
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  auto allowed = [&](int cls, uint32_t nw) { return !(k == 3 && nw > (1u << cls)); };
  // 1. each class alone: the best waves per root as a starting point (and
  //    the isolated times reported)
  float iso[NCLASS][4] = {};
  uint32_t shape[NCLASS];
  for (int cls = 0; cls < NCLASS; ++cls) {
    shape[cls] = kWavesPerRoot[cls];
    if (!c.class_count[cls]) continue;
    float best = 1e30f;
    for (uint32_t nw = 1, i = 0; nw <= MAX_WAVES; nw *= 2, ++i) {
      if (!allowed(cls, nw)) break;
      float tsum = 0;
      for (int rep = 0; rep < 2; ++rep) {
        HIP_CHECK(hipMemsetAsync(q, 0, QUEUE_WORDS * sizeof(uint32_t), c.stream));
        HIP_CHECK(hipEventRecord(e0, c.stream));
        launch_class(c, cls, k, nw, q, gcnt, rootT, totals);
        HIP_CHECK(hipEventRecord(e1, c.stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (rep) tsum += ms;  // first run warms caches
      }
      iso[cls][i] = tsum;
      if (tsum < best) {
        best = tsum;
        shape[cls] = nw;
      }
    }
  }
  // 2. the classes run at once and share the GPU, so the best shape of the
  //    pass is not the classes' best alone (C2: [4,4,8,8] from the isolated
  //    times, 1.86 ms; [2,2,4,8], 1.66 ms): coordinate descent on the time
  //    of the whole concurrent pass, two rounds over the classes
  auto pass_ms = [&](const uint32_t *sh) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      HIP_CHECK(hipMemsetAsync(count_queue(c, 0), 0, NCLASS * QUEUE_WORDS * sizeof(uint32_t), c.stream));
      HIP_CHECK(hipEventRecord(e0, c.stream));
      run_classes(c, k, sh, gcnt, rootT, totals, 0);
      HIP_CHECK(hipEventRecord(e1, c.stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (rep) best = std::min(best, ms);  // first run warms caches
    }
    return best;
  };
  float cur = pass_ms(shape);
  for (int round = 0; round < 2; ++round) {
    for (int cls = 0; cls < NCLASS; ++cls) {
      if (!c.class_count[cls]) continue;
      for (uint32_t nw = 1; nw <= MAX_WAVES; nw *= 2) {
        if (!allowed(cls, nw)) break;
        if (nw == shape[cls]) continue;
        uint32_t trial[NCLASS];
        for (int i = 0; i < NCLASS; ++i) trial[i] = shape[i];
        trial[cls] = nw;
        const float t = pass_ms(trial);
        if (t < cur) {
          cur = t;
          shape[cls] = nw;
        }
      }
    }
  }
  for (int cls = 0; cls < NCLASS; ++cls) {
    c.list_waves_per_root[cls] = shape[cls];
    uint32_t i = 0;
    while ((1u << i) < shape[cls]) ++i;
    c.tune_ms[cls] = c.class_count[cls] ? iso[cls][i] : 0.f;
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
}

// k = 2: T = arcs, per-vertex counts = degrees (no listing kernels)
void list_count_edges(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.cnt, sizeof(ull) * ((size_t)n + 1));
  buf_reserve(c, c.rootT, sizeof(ull) * ((size_t)n + 2));
  HIP_CHECK(hipMemsetAsync(c.rootT.p, 0, sizeof(ull) * ((size_t)n + 2), c.stream));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, c.stream));
  if (n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c.num_cus * 8);
    k_edge_out<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), n, c.cnt.as<ull>(), c.rootT.as<ull>());
    KCDS_LAUNCH_CHECK();
  }
  if (c.darcs) {
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((c.darcs + INC_E * 256 - 1) / (INC_E * 256), (uint64_t)c.num_cus * 2));
    k_edge_in<<<grid, 256, 0, c.stream>>>(c.dnbr.as<uint32_t>(), c.darcs, c.cnt.as<ull>());
    KCDS_LAUNCH_CHECK();
  }
  HIP_CHECK(hipEventRecord(e1, c.stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  for (int i = 0; i < NCLASS; ++i) c.class_count[i] = 0;
  c.list_kernel_ms = ms;
  c.list_kernel_launches = 2;
  c.count_cap = 0;
  c.T = c.darcs;
  c.w2 = 0;
  c.k = 2;
  c.counted = true;
  c.listed = false;
  c.peeled = false;
  c.have_comp = false;
}

void list_count(Ctx &c, uint32_t k, ull cap) {
  if (!c.oriented) fail(KCDS_E_STATE, "kcds_count/kcds_list before kcds_orient");
  if (k < 2 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
  if (k == 2) {
    list_count_edges(c);
    return;
  }
  const uint32_t n = c.n;
  classify_roots(c, k);
  const uint32_t ncopy = count_copies(n);
  const uint64_t cstride = copy_stride(n);
  buf_reserve(c, c.cnt, sizeof(ull) * cstride * ncopy);
  buf_reserve(c, c.rootT, sizeof(ull) * ((size_t)n + 2));
  buf_reserve(c, c.lsum, sizeof(ull) * QSHARDS * 8 * 2);
  ull *totals = c.lsum.as<ull>();
  if (cap) buf_reserve(c, c.gsum, sizeof(ull) * ((size_t)c.ngraphs + 1));
  zero_u64(c, c.cnt.as<ull>(), cstride * ncopy, c.rootT.as<ull>(), (uint64_t)n + 2, totals, (uint64_t)QSHARDS * 8,
           cap ? c.gsum.as<ull>() : nullptr, cap ? c.ngraphs : 0);
  c.count_cap = cap;
  if (!c.ltime[0]) {
    HIP_CHECK(hipEventCreate(&c.ltime[0]));
    HIP_CHECK(hipEventCreate(&c.ltime[1]));
  }
  hipEvent_t e0 = c.ltime[0], e1 = c.ltime[1];
  HIP_CHECK(hipEventRecord(e0, c.stream));
  uint32_t shape[NCLASS];
  for (int cls = 0; cls < NCLASS; ++cls) shape[cls] = list_waves_per_root(c, cls, k);
  uint32_t launches = run_classes(c, k, shape, c.cnt.as<ull>(), c.rootT.as<ull>(), totals, cap);
  if (ncopy > 1 && n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c.num_cus * 8);
    k_sum_copies<<<grid, 256, 0, c.stream>>>(c.cnt.as<ull>(), n, cstride, ncopy);
    KCDS_LAUNCH_CHECK();
    ++launches;
  }
  HIP_CHECK(hipEventRecord(e1, c.stream));
  ull T = 0, w2 = 0;
  read_totals(c, totals, T, w2);
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  c.list_kernel_ms = ms;
  c.list_kernel_launches = launches;
  c.T = T;
  c.w2 = w2;
  c.k = k;
  c.counted = true;
  c.listed = false;
  c.peeled = false;
  c.have_comp = false;
}

void graph_counts(Ctx &c, std::vector<ull> &gT) {
  const uint32_t G = c.ngraphs, n = c.n;
  gT.assign(G, 0);
  if (G == 1) {
    gT[0] = c.T;
  } else {
    buf_reserve(c, c.tmp3, sizeof(ull) * ((size_t)G + 1));
    ull *d = c.tmp3.as<ull>();
    HIP_CHECK(hipMemsetAsync(d, 0, sizeof(ull) * G, c.stream));
    if (n) {
      const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c.num_cus * 8);
      k_graph_T<<<grid, 256, 0, c.stream>>>(c.rootT.as<ull>(), c.gid.as<uint32_t>(), n, d);
      KCDS_LAUNCH_CHECK();
    }
    HIP_CHECK(hipMemcpyAsync(gT.data(), d, sizeof(ull) * G, hipMemcpyDeviceToHost, c.stream));
  }
  if (c.count_cap) {  // graphs stopped at the cap: their lower bound (counts are partial)
    std::vector<ull> gs(G);
    HIP_CHECK(hipMemcpyAsync(gs.data(), c.gsum.p, sizeof(ull) * G, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (uint32_t g = 0; g < G; ++g) gT[g] = std::max(gT[g], gs[g]);
  }
  c.sync();
}

ull drop_graphs(Ctx &c, const std::vector<uint8_t> &drop, const std::vector<ull> &gT) {
  const uint32_t G = c.ngraphs, n = c.n;
  ull keep = 0;
  for (uint32_t g = 0; g < G; ++g)
    if (!drop[g]) keep += gT[g];
  buf_reserve(c, c.tmp3, (size_t)G + 8);
  HIP_CHECK(hipMemcpyAsync(c.tmp3.p, drop.data(), G, hipMemcpyHostToDevice, c.stream));
  if (n) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c.num_cus * 8);
    k_drop_graphs<<<grid, 256, 0, c.stream>>>(c.rootT.as<ull>(), c.cnt.as<ull>(), c.gid.as<uint32_t>(),
                                              c.tmp3.as<uint8_t>(), n);
    KCDS_LAUNCH_CHECK();
  }
  c.sync();  // drop lives on the host
  c.T = keep;
  return keep;
}

void list_materialize(Ctx &c, uint32_t k) {
  if (!c.counted || c.k != k) list_count(c, k);
  const uint32_t n = c.n;
  const uint64_t T = c.T;
  if (T >= 0xFFFFFFFFull) fail(KCDS_E_LIMIT, "T = %llu cliques exceeds the 32-bit clique id space",
                               (unsigned long long)T);
  // per-root offsets (new ids) and vertex->clique incidence offsets
  buf_reserve(c, c.tmp2, sizeof(ull) * ((size_t)n + 2));
  ull *rootOff = c.tmp2.as<ull>();
  scan_u64_to_u64(c, c.rootT.as<ull>(), rootOff, n);
  buf_reserve(c, c.incoff, sizeof(ull) * ((size_t)n + 1));
  scan_u64_to_u64(c, c.cnt.as<ull>(), c.incoff.as<ull>(), n);
  buf_reserve(c, c.cliques, sizeof(uint32_t) * (T * k + 1));
  buf_reserve(c, c.inc, sizeof(uint32_t) * (T * k + 1));
  if (k == 2 && n) {  // the arcs themselves, at rootOff[r] = doff[r]
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 3) / 4, (uint64_t)c.num_cus * 16);
    k_edge_emit<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(), n,
                                            c.cliques.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  if (k > 2) HIP_CHECK(hipMemsetAsync(emit_queue(c, 0), 0, NCLASS * QUEUE_WORDS * sizeof(uint32_t), c.stream));
  if (k > 2 && c.class_count[NCLASS]) {
    HIP_CHECK(hipMemsetAsync(wide_emit_queue(c), 0, QUEUE_WORDS * sizeof(uint32_t), c.stream));
    uint64_t rowstride = 0;
    const uint32_t grid = wide_grid(c, rowstride);
    hipLaunchKernelGGL(wide_emit_fn(k), dim3(grid), dim3(WIDE_THREADS), 0, c.stream, c.doff.as<uint32_t>(),
                       c.dnbr.as<uint32_t>(), c.roots.as<uint4>() + c.class_start[NCLASS], c.class_count[NCLASS],
                       wide_emit_queue(c), rootOff, c.cliques.as<uint32_t>(), c.wide.as<ull>(), rowstride);
    KCDS_LAUNCH_CHECK();
  }
  for (int cls = NCLASS - 1; cls >= 0 && k > 2; --cls) {
    const uint32_t nr = c.class_count[cls];
    if (!nr) continue;
    EmitFn fn = cls == 0 ? emit_fn<1>(k) : cls == 1 ? emit_fn<2>(k) : cls == 2 ? emit_fn<4>(k) : emit_fn<8>(k);
    const uint4 *roots = c.roots.as<uint4>() + c.class_start[cls];
    hipLaunchKernelGGL(fn, dim3(wg_grid(c, cls, nr, 1)), dim3(64), 0, c.stream, c.doff.as<uint32_t>(),
                       c.dnbr.as<uint32_t>(), roots, nr, emit_queue(c, cls), rootOff, c.cliques.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  buf_reserve(c, c.tmp3, sizeof(uint32_t) * ((size_t)n + 1));
  HIP_CHECK(hipMemsetAsync(c.tmp3.p, 0, sizeof(uint32_t) * ((size_t)n + 1), c.stream));
  if (T) {
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((T * k + INC_E * 256 - 1) / (INC_E * 256), (uint64_t)c.num_cus * 2));
    k_incidence<<<grid, 256, 0, c.stream>>>(c.cliques.as<uint32_t>(), T, k, c.incoff.as<ull>(),
                                            c.tmp3.as<uint32_t>(), c.inc.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.sync();
  c.listed = true;
}

// ---------------------------------------------------------------------------
// Allocation-free count (kcds_count_dag_noalloc): the counting kernels above
// over a DAG the caller holds in device memory, every temporary in a static
// device array.  Layout of the array (bytes, each part 256-aligned):
//   [0, NA_Q)            classification counters and the class queues
//                        (QCTR_WORDS words; word 0 = DAG-form error flag)
//   totals               QSHARDS * 8 sharded (cliques, wedges) pairs
//   rootT                KCDS_NOALLOC_MAX_N + 2 per-root counts
//   roots                KCDS_NOALLOC_MAX_N + 1 root records (uint4)
// The first three parts are zeroed by a fill kernel at every call.
namespace {

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_Q = na_align(sizeof(uint32_t) * QCTR_WORDS);
constexpr size_t NA_TOT = na_align(sizeof(ull) * QSHARDS * 8);
constexpr size_t NA_ROOTT = na_align(sizeof(ull) * ((size_t)KCDS_NOALLOC_MAX_N + 2));
constexpr size_t NA_ROOTS = na_align(sizeof(uint4) * ((size_t)KCDS_NOALLOC_MAX_N + 1));
constexpr size_t NA_BYTES = NA_Q + NA_TOT + NA_ROOTT + NA_ROOTS;
static_assert(NA_BYTES <= (size_t)KCDS_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds its stated size");

static __device__ __align__(256) unsigned char scratch[KCDS_NOALLOC_SCRATCH_BYTES];

// per device: the array's device address (hipGetSymbolAddress) and CU count
constexpr int NA_MAX_DEV = 64;
std::mutex na_mu;
void *na_ptr[NA_MAX_DEV];
int na_cus[NA_MAX_DEV];

__global__ void k_na_zero(uint32_t *p, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ void k_na_zero64(ull *p, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// listing form: offsets non-decreasing, every arc r -> w with r < w < n, each
// out-list strictly ascending; any violation sets *err
__global__ void k_na_check(const uint32_t *doff, const uint32_t *dnbr, uint32_t n, uint32_t *err) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t b = doff[r], e = doff[r + 1];
    bool bad = e < b;
    uint32_t prev = r;
    for (uint32_t a = b; !bad && a < e; ++a) {
      const uint32_t w = dnbr[a];
      bad = w <= prev || w >= n;
      prev = w;
    }
    if (bad) atomicOr(err, 1u);
  }
}

uint32_t na_grid(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 4096));
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int d) {
    HIP_CHECK(hipGetDevice(&prev));
    if (prev != d) HIP_CHECK(hipSetDevice(d));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

void na_count(int device, uint32_t n, const uint32_t *doff, const uint32_t *dnbr, uint32_t k, ull *per_vertex,
              ull *T_out, hipStream_t st) {
  if (!doff || !T_out || (n && !per_vertex)) fail(KCDS_E_INVALID, "null argument");
  if (k < 2 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
  if (n > KCDS_NOALLOC_MAX_N) fail(KCDS_E_LIMIT, "n = %u above %u", n, KCDS_NOALLOC_MAX_N);
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev || device >= NA_MAX_DEV) fail(KCDS_E_INVALID, "device %d", device);
  DeviceGuard dg(device);
  unsigned char *base;
  int cus;
  {
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[device]) {  // once per device, with the device current
      void *p = nullptr;
      HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      int nc = 0;
      HIP_CHECK(hipDeviceGetAttribute(&nc, hipDeviceAttributeMultiprocessorCount, device));
      na_cus[device] = nc;
      na_ptr[device] = p;
    }
    base = (unsigned char *)na_ptr[device];
    cus = na_cus[device];
  }
  uint32_t *q = (uint32_t *)base;
  ull *totals = (ull *)(base + NA_Q), *rootT = (ull *)(base + NA_Q + NA_TOT);
  uint4 *roots = (uint4 *)(base + NA_Q + NA_TOT + NA_ROOTT);
  uint32_t *err = q, *ccount = q + 16, *cfill = q + 32;
  const uint64_t zw = (NA_Q + NA_TOT + sizeof(ull) * ((uint64_t)n + 2)) / 4;
  k_na_zero<<<na_grid(zw), 256, 0, st>>>(q, zw);
  KCDS_LAUNCH_CHECK();
  if (n) {
    k_na_zero64<<<na_grid(n), 256, 0, st>>>(per_vertex, n);
    KCDS_LAUNCH_CHECK();
    k_na_check<<<na_grid(n), 256, 0, st>>>(doff, dnbr, n, err);
    KCDS_LAUNCH_CHECK();
    if (k > 2) {
      const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 256));
      k_class_count<<<grid, 256, 0, st>>>(doff, n, k, ccount);
      KCDS_LAUNCH_CHECK();
    }
  }
  uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0}, arcs = 0, e = 0;
  HIP_CHECK(hipMemcpyAsync(&e, err, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(h, ccount, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(&arcs, doff + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (e) fail(KCDS_E_INVALID, "DAG not in listing form (arcs must go to higher ids, out-lists ascending)");
  if (h[4]) fail(KCDS_E_LIMIT, "%u roots have out-degree above %u (the LDS row limit)", h[4], MAX_LDS_DEG);
  if (k == 2) {
    if (n) {
      k_edge_out<<<na_grid(n), 256, 0, st>>>(doff, n, per_vertex, rootT);
      KCDS_LAUNCH_CHECK();
    }
    if (arcs) {
      const uint32_t grid = (uint32_t)std::max<uint64_t>(
          1, std::min<uint64_t>((arcs + INC_E * 256 - 1) / (INC_E * 256), (uint64_t)cus * 2));
      k_edge_in<<<grid, 256, 0, st>>>(dnbr, arcs, per_vertex);
      KCDS_LAUNCH_CHECK();
    }
    HIP_CHECK(hipStreamSynchronize(st));
    *T_out = arcs;
    return;
  }
  uint32_t start[NCLASS], s = 0;
  for (int i = 0; i < NCLASS; ++i) {
    start[i] = s;
    s += h[i];
  }
  if (s) {
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 256));
    k_class_fill<<<grid, 256, 0, st>>>(doff, n, k, ccount, cfill, roots);
    KCDS_LAUNCH_CHECK();
  }
  for (int cls = NCLASS - 1; cls >= 0; --cls) {
    const uint32_t nr = h[cls];
    if (!nr) continue;
    uint32_t nw = kWavesPerRoot[cls];
    if (k == 3) nw = std::min<uint32_t>(nw, 1u << cls);
    uint32_t per_cu = std::max<uint32_t>(1, WAVE_SLOTS_PER_CU / nw);
    per_cu = std::min<uint32_t>(per_cu, std::max<uint32_t>(1, (160u * 1024u) / lds_bytes(cls)));
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(nr, (uint32_t)cus * per_cu));
    hipLaunchKernelGGL(class_count_fn(cls, k), dim3(grid), dim3(64 * nw), 0, st, doff, dnbr, roots + start[cls], nr,
                       q + QBASE + cls * QUEUE_WORDS, per_vertex, copy_stride(n), 1u, rootT, totals, nullptr,
                       nullptr, 0ull, 1ull);
    KCDS_LAUNCH_CHECK();
  }
  ull ht[QSHARDS * 8];
  HIP_CHECK(hipMemcpyAsync(ht, totals, sizeof(ht), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  ull T = 0;
  for (uint32_t i = 0; i < QSHARDS; ++i) T += ht[i * 8];
  *T_out = T;
}

}  // namespace

}  // namespace kcds

extern "C" int kcds_count_dag_noalloc(int device, uint32_t n, const uint32_t *doff, const uint32_t *dnbr, uint32_t k,
                                      uint64_t *per_vertex, uint64_t *T, void *stream) {
  try {
    kcds::na_count(device, n, doff, dnbr, k, (unsigned long long *)per_vertex, (unsigned long long *)T,
                   (hipStream_t)stream);
    return KCDS_OK;
  } catch (const kcds::Error &e) {
    if (getenv("KCDS_LOG")) fprintf(stderr, "kcds_count_dag_noalloc: %s\n", e.msg.c_str());
    return e.code;
  }
}
