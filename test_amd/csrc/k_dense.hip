// Exact densest-set search on a k'-core that is small in vertices but huge in
// cliques (10^9..10^11): no clique is stored and no clique flow network is
// built.
//
// Reference: dynamicExactAlgo (main.cu:729-1270) bisects alpha with a
// push-relabel max-flow on the clique network of each component (a node per
// clique, helpers.cu:1148-1264).  At 10^9 cliques and more that network does
// not fit in any GPU's memory.
//
// Here the same optimum is found through the dual of that network.  Split
// every clique's unit of weight among its members; the largest vertex load
// of a split bounds every set's density from above (a set's cliques are
// split among its own members), and the best split meets rho* (max-flow /
// min-cut).  The splits form a polytope whose extreme points are the greedy
// splits -- each clique wholly to its member that comes first in a vertex
// order -- and a greedy split is one listing pass on the GPU: orient by the
// order and count the cliques rooted at each vertex (the k'-core, at most
// 1,024 vertices, is held as a bitmask matrix in LDS; a wave per (root,
// 64-candidate word) task).  Wolfe's minimum-norm-point algorithm (a corral
// of greedy splits, each new one the greedy split of the current point's
// order) drives a convex combination of greedy splits toward the balanced
// split; its level sets give the candidate S (the best prefix of the vertices
// ordered by load, counted exactly in one more pass).
//
// Certificate (exact, integer).  With integer weights mu_t on the greedy
// splits q_t, the pair holdings H(u, w) = sum_t mu_t P_t(u, w) (P_t(u, w) =
// cliques held by u that contain w, an output of the same pass) say how
// much weight u can hand to w: moving the fraction theta_uw <= 1/(k-1) of
// every such clique's share from u to w keeps every share non-negative
// (a clique has k-1 other members), so any flow with f(u->w) <=
// H(u, w)/(k-1) is a valid re-split.  If an integer max-flow finds a
// re-split that leaves every vertex of S at exactly rho_S = c_k(S)/|S| and
// every other vertex strictly below it, then every set A has c_k(A) <=
// rho_S |A| with equality only inside S: rho* = rho_S and S is the unique
// maximal densest set.  All quantities are exact integers (128-bit).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

typedef __int128 i128;
typedef unsigned __int128 u128;

constexpr uint32_t DN_MAX = 1024;     // largest core the LDS matrix holds
constexpr uint32_t DN_THREADS = 512;

__device__ __forceinline__ ull dabove(int q, uint32_t c) {
  const int cq = (int)(c >> 6), cb = (int)(c & 63);
  if (q < cq) return 0ull;
  if (q > cq) return ~0ull;
  return cb == 63 ? 0ull : (~0ull << (cb + 1));
}

__device__ __forceinline__ ull dbelow(int q, uint32_t c) {
  const int cq = (int)(c >> 6), cb = (int)(c & 63);
  if (q < cq) return ~0ull;
  if (q > cq) return 0ull;
  return cb == 0 ? 0ull : ((1ull << cb) - 1ull);
}

// Rows in LDS, word-major: A[w * 64W + j] holds bits 64w..64w+63 of row j.
template <int W> __device__ __forceinline__ ull drow(const ull *A, uint32_t j, int w) {
  return A[(uint32_t)w * (64u * W) + j];
}

template <int W, bool WP> __device__ __forceinline__ void dcredit(ull *cr, uint32_t c, ull E) {
  if constexpr (WP) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int q = 0; q < W; ++q)
      if ((uint32_t)q == (c >> 6) && lane == (c & 63)) cr[q] += E;
  }
}

// REM-cliques inside S (one wave); with WP every member is credited U per
// clique.  Every row word of a step is requested before the first is used
// (one LDS wait per step), and candidate sets live in scalar registers.
template <int W, int REM, bool WP, ull U = 1> struct DCount {
  static __device__ ull run(const ull *A, const ull *S, ull *cr) {
    ull total = 0;
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
        for (int w = 0; w < W; ++w) S2[w] = drow<W>(A, c, w);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          S2[w] = S[w] & uniform_u64(S2[w]) & dabove(w, c);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        const ull E = DCount<W, REM - 1, WP, U>::run(A, S2, cr);
        dcredit<W, WP>(cr, c, E * U);
        total += E;
      }
    }
    return total;
  }
};

// Row words [LO, W) of lane j against S, all requested before any is used.
template <int W, int LO>
__device__ __forceinline__ uint32_t drow_popc(const ull *A, uint32_t j, const ull *S) {
  ull r[W];
#pragma unroll
  for (int w = LO; w < W; ++w) r[w] = drow<W>(A, j, w);
  uint32_t x = 0;
#pragma unroll
  for (int w = LO; w < W; ++w) x += (uint32_t)__popcll(r[w] & S[w]);
  return x;
}

// Edges inside S.  When the lower half of S is empty (candidates only go up,
// so that is common) only the upper half of each row is read.  The whole
// wave is active here (control flow above is wave-uniform).
template <int W, bool WP, ull U> struct DCount<W, 2, WP, U> {
  static __device__ ull run(const ull *A, const ull *S, ull *cr) {
    const uint32_t lane = lane_id();
    constexpr int H = W >= 8 ? W / 2 : 0;
    bool low = false;
#pragma unroll
    for (int w = 0; w < H; ++w) low |= S[w] != 0ull;
    uint32_t xs = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if (!S[q]) continue;
      if ((S[q] >> lane) & 1ull) {
        const uint32_t j = (uint32_t)q * 64 + lane;
        const uint32_t x = low ? drow_popc<W, 0>(A, j, S) : drow_popc<W, H>(A, j, S);
        if constexpr (WP) cr[q] += (ull)x * U;  // modulo 2^64: U may stand for -U
        xs += x;
      }
    }
    return (ull)(wave_sum_u32_full(xs) >> 1);
  }
};

template <int W, bool WP, ull U> struct DCount<W, 1, WP, U> {
  static __device__ ull run(const ull *, const ull *S, ull *cr) {
    const uint32_t lane = lane_id();
    ull tot = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if constexpr (WP)
        if ((S[q] >> lane) & 1ull) cr[q] += U;
      tot += (ull)__popcll(S[q]);
    }
    return tot;
  }
};

// Pair holdings with a critical set C.  A clique's share may move from its
// holder to its other members in any proportions that sum to at most 1; the
// rule used here: a clique with nc >= 1 critical members besides its holder
// offers 1/nc to each of them and nothing to the rest, one with none offers
// 1/(k-1) to each.  Holdings are kept in units of 1/L, L = lcm(1..k-1), so
// the plain pass credits U = L/(k-1) per clique and k_dn_crit adds, for the
// cliques that meet C, the difference between the two rules.
template <int K> struct Units {
  static constexpr uint32_t L = K <= 3 ? 2u : K == 4 ? 6u : K == 5 ? 12u : K <= 7 ? 60u : 420u;
  static constexpr uint32_t U = L / (K - 1);
  static constexpr ull NEG_U = 0ull - (ull)U;  // -U modulo 2^64
};

// REM-cliques inside S, each completing a clique that already has pc >= 1
// critical members besides its holder: members are credited the rule's
// share minus U (modulo 2^64: a negative difference wraps, and every total
// it joins is non-negative); h[t] += REM-cliques with t critical members.
template <int W, int K, int REM> struct DCountD {
  static __device__ void run(const ull *A, const ull *S, const ull *Cm, uint32_t pc, ull *cr, ull *h) {
    constexpr uint32_t L = Units<K>::L, U = Units<K>::U;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      ull word = S[q];
      while (word) {
        const int b = __builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = (uint32_t)(q * 64 + b);
        ull S2[W];
        uint32_t pcnt = 0;
        bool anyc = false;
#pragma unroll
        for (int w = 0; w < W; ++w) S2[w] = drow<W>(A, c, w);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          S2[w] = S[w] & uniform_u64(S2[w]) & dabove(w, c);
          pcnt += (uint32_t)__popcll(S2[w]);
          anyc |= (S2[w] & Cm[w]) != 0ull;
        }
        if (pcnt < (uint32_t)(REM - 1)) continue;
        const uint32_t cc = (uint32_t)((Cm[q] >> b) & 1ull), pc2 = pc + cc;
        ull hz[REM];
#pragma unroll
        for (int t = 0; t < REM; ++t) hz[t] = 0;
        if (anyc)
          DCountD<W, K, REM - 1>::run(A, S2, Cm, pc2, cr, hz);
        else
          hz[0] = DCount<W, REM - 1, true, Units<K>::NEG_U>::run(A, S2, cr);
        ull credit = 0;
#pragma unroll
        for (int t = 0; t < REM; ++t) credit += hz[t] * (cc ? (ull)(L / (pc2 + (uint32_t)t)) - U : Units<K>::NEG_U);
        dcredit<W, true>(cr, c, credit);
#pragma unroll
        for (int t = 0; t < REM; ++t) h[t + (int)cc] += hz[t];
      }
    }
  }
};

template <int W, int K> struct DCountD<W, K, 2> {
  static __device__ void run(const ull *A, const ull *S, const ull *Cm, uint32_t pc, ull *cr, ull *h) {
    constexpr uint32_t L = Units<K>::L, U = Units<K>::U;
    const uint32_t lane = lane_id();
    // a critical lane's cliques have pc + 1 or pc + 2 critical members (pc +
    // 2 <= k - 1 at this level)
    const ull da = (ull)(L / (pc + 1)) - U, db = (ull)(L / (pc + 2)) - U;
    uint32_t xs = 0, x0s = 0, x2s = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if (!S[q]) continue;
      if ((S[q] >> lane) & 1ull) {
        const uint32_t j = (uint32_t)q * 64 + lane;
        ull r[W];
#pragma unroll
        for (int w = 0; w < W; ++w) r[w] = drow<W>(A, j, w);
        uint32_t x = 0, x1 = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          x += (uint32_t)__popcll(r[w] & S[w]);
          x1 += (uint32_t)__popcll(r[w] & S[w] & Cm[w]);
        }
        if ((Cm[q] >> lane) & 1ull) {
          cr[q] += (ull)(x - x1) * da + (ull)x1 * db;
          x2s += x1;
        } else {
          cr[q] += (ull)x * Units<K>::NEG_U;
          x0s += x - x1;
        }
        xs += x;
      }
    }
    // every edge is seen from both ends
    const ull e = wave_sum_u32_full(xs) >> 1, e0 = wave_sum_u32_full(x0s) >> 1, e2 = wave_sum_u32_full(x2s) >> 1;
    h[0] += e0;
    h[1] += e - e0 - e2;
    h[2] += e2;
  }
};

template <int W, int K> struct DCountD<W, K, 1> {
  static __device__ void run(const ull *, const ull *S, const ull *Cm, uint32_t pc, ull *cr, ull *h) {
    constexpr uint32_t L = Units<K>::L, U = Units<K>::U;
    const uint32_t lane = lane_id();
    const ull da = (ull)(L / (pc + 1)) - U;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if ((S[q] >> lane) & 1ull) cr[q] += ((Cm[q] >> lane) & 1ull) ? da : Units<K>::NEG_U;
      h[0] += (ull)__popcll(S[q] & ~Cm[q]);
      h[1] += (ull)__popcll(S[q] & Cm[q]);
    }
  }
};

// One greedy pass over the rank-ordered core matrix Mr (n rows of W words,
// row-major; vertex i = rank i, arcs point up in rank).  q[i] = cliques whose
// lowest-rank member is i; with WP, P[i*n + j] = U times those of them
// containing j (the 1/(k-1) rule in units of 1/L).  A wave takes (root i,
// candidate word w0) tasks from a ticket counter.
template <int W, int K, bool WP>
__global__ __launch_bounds__(DN_THREADS) void k_dn_pass(const ull *__restrict__ Mr, uint32_t n, uint32_t split,
                                                        uint32_t *ticket, ull *q, ull *P) {
  constexpr uint32_t NP = 64u * W;
  constexpr ull U = Units<K>::U;
  __shared__ ull A[W * NP];
  for (uint32_t e = threadIdx.x; e < W * NP; e += blockDim.x) {
    const uint32_t w = e / NP, j = e % NP;
    A[e] = j < n ? Mr[(size_t)j * W + w] : 0ull;
  }
  __syncthreads();
  const uint32_t lane = lane_id();
  // task (i, w0, sub): root i's candidates in word w0 whose bit index is in
  // the sub-th of `split` equal ranges (small cores need many tasks to fill
  // the GPU: a dense core's low-rank roots carry most of the work)
  const uint32_t ntask = n * W * split, bits = 64 / split;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (t >= ntask) break;
    const uint32_t sub = t % split, iw = t / split;
    const uint32_t i = iw / W, w0 = iw % W;
    if (w0 < (i >> 6)) continue;
    ull Sr[W];
#pragma unroll
    for (int w = 0; w < W; ++w) Sr[w] = uniform_u64(drow<W>(A, i, w)) & dabove(w, i);
    ull cand = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
      if ((uint32_t)w == w0) cand = Sr[w];
    cand &= bits == 64 ? ~0ull : (((1ull << bits) - 1ull) << (sub * bits));
    if (!cand) continue;
    ull cr[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cr[w] = 0;
    ull Et = 0;
    while (cand) {
      const int b = __builtin_ctzll(cand);
      cand &= cand - 1;
      const uint32_t c = w0 * 64 + (uint32_t)b;
      ull S2[W];
      uint32_t pc = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) S2[w] = drow<W>(A, c, w);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        S2[w] = Sr[w] & uniform_u64(S2[w]) & dabove(w, c);
        pc += (uint32_t)__popcll(S2[w]);
      }
      if (pc + 2 < (uint32_t)K) continue;
      const ull E = DCount<W, K - 2, WP, U>::run(A, S2, cr);
      dcredit<W, WP>(cr, c, E * U);
      Et += E;
    }
    if (lane == 0 && Et) atomicAdd(&q[i], Et);
    if constexpr (WP) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t j = (uint32_t)w * 64 + lane;
        if (cr[w] && j < n) atomicAdd(&P[(size_t)i * n + j], cr[w]);
      }
    }
  }
}

// After k_dn_pass with WP: the holdings of the cliques that meet the
// critical set C besides their holder, moved from the 1/(k-1) rule to the
// critical rule.  Each such clique is visited once, from its holder h and
// its lowest-rank critical member v other than h: task (v = crit[ci], h)
// with h < v adjacent to v; the other members lie above h, adjacent to h and
// v, and are not critical below v.  Cm: C as a bitmask in rank order.
template <int W, int K>
__global__ __launch_bounds__(DN_THREADS) void k_dn_crit(const ull *__restrict__ Mr, uint32_t n,
                                                        const uint32_t *__restrict__ crit, uint32_t ncrit,
                                                        const ull *__restrict__ cmask, uint32_t *ticket, ull *P) {
  constexpr uint32_t NP = 64u * W;
  constexpr uint32_t L = Units<K>::L, U = Units<K>::U;
  __shared__ ull A[W * NP];
  for (uint32_t e = threadIdx.x; e < W * NP; e += blockDim.x) {
    const uint32_t w = e / NP, j = e % NP;
    A[e] = j < n ? Mr[(size_t)j * W + w] : 0ull;
  }
  __syncthreads();
  ull Cm[W];
#pragma unroll
  for (int w = 0; w < W; ++w) Cm[w] = uniform_u64(cmask[w]);
  const uint32_t lane = lane_id();
  const uint32_t ntask = ncrit * n;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (t >= ntask) break;
    const uint32_t v = __builtin_amdgcn_readfirstlane(crit[t / n]), h = t % n;
    if (h >= v) continue;
    ull S[W];
    bool adj = false, any = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const ull rv = uniform_u64(drow<W>(A, v, w));
      S[w] = uniform_u64(drow<W>(A, h, w)) & dabove(w, h) & rv & ~(Cm[w] & dbelow(w, v));
      if ((uint32_t)w == (h >> 6)) adj = (rv >> (h & 63)) & 1ull;
      any |= S[w] != 0ull;
    }
    if (!adj || (K > 2 && !any)) continue;
    ull cr[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cr[w] = 0;
    ull hz[K - 1];
#pragma unroll
    for (int j = 0; j < K - 1; ++j) hz[j] = 0;
    DCountD<W, K, K - 2>::run(A, S, Cm, 1, cr, hz);
    ull credit = 0;
#pragma unroll
    for (int j = 0; j < K - 1; ++j) credit += hz[j] * ((ull)(L / (1u + (uint32_t)j)) - U);
    dcredit<W, true>(cr, v, credit);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t j = (uint32_t)w * 64 + lane;
      if (cr[w] && j < n) atomicAdd(&P[(size_t)h * n + j], cr[w]);
    }
  }
}

// core vertices (new ids) of graph 0: core >= k'
__global__ void k_dn_collect(uint32_t n, const long long *core, long long kp, uint32_t *cnt, uint32_t *verts) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    if (core[r] >= kp) {
      const uint32_t at = atomicAdd(cnt, 1u);
      if (at < DN_MAX) verts[at] = r;
    }
}

__global__ void k_dn_loc(uint32_t nc, const uint32_t *verts, uint32_t *loc) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < nc; u += gridDim.x * blockDim.x) loc[verts[u]] = u;
}

__global__ void k_dn_unloc(uint32_t nc, const uint32_t *verts, uint32_t *loc) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < nc; u += gridDim.x * blockDim.x)
    loc[verts[u]] = 0xFFFFFFFFu;
}

// M0 (local ids, row-major n x W words) from the core DAG, a wave per vertex
__global__ void k_dn_adj(uint32_t nc, uint32_t W, const uint32_t *verts, const uint32_t *loc, const uint32_t *kdoff,
                         const uint32_t *kdnbr, ull *M0) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwv = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t u = gw; u < nc; u += nwv) {
    const uint32_t r = verts[u];
    for (uint32_t a = kdoff[r] + lane; a < kdoff[r + 1]; a += KCDS_WAVE) {
      const uint32_t w = loc[kdnbr[a]];
      if (w >= nc) continue;
      atomicOr(&M0[(size_t)u * W + (w >> 6)], 1ull << (w & 63));
      atomicOr(&M0[(size_t)w * W + (u >> 6)], 1ull << (u & 63));
    }
  }
}

// Mr = M0 with vertex u moved to rank[u]
__global__ void k_dn_perm(const ull *M0, uint32_t n, uint32_t W, const uint32_t *rank, ull *Mr) {
  const uint64_t nn = (uint64_t)n * n;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)(e / n), w = (uint32_t)(e % n);
    if ((M0[(size_t)u * W + (w >> 6)] >> (w & 63)) & 1ull) {
      const uint32_t ru = rank[u], rw = rank[w];
      atomicOr(&Mr[(size_t)ru * W + (rw >> 6)], 1ull << (rw & 63));
    }
  }
}

// back to local ids: Pl[u*n + w] = P[rank u * n + rank w], ql[u] = q[rank u]
__global__ void k_dn_unrank(uint32_t n, const uint32_t *rank, const ull *q, const ull *P, ull *ql, ull *Pl) {
  const uint64_t nn = (uint64_t)n * n;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)(e / n), w = (uint32_t)(e % n);
    if (P) Pl[e] = P[(size_t)rank[u] * n + rank[w]];
    if (w == 0) ql[u] = q[rank[u]];
  }
}

// H[e] = sum_t mu[t] * Phist[slot[t]][e] as 128-bit (lo, hi)
__global__ void k_dn_hold(uint64_t nn, uint32_t nt, const ull *mu, const uint32_t *slot, const ull *Phist, ull *Hlo,
                          ull *Hhi) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (uint64_t)gridDim.x * blockDim.x) {
    ull lo = 0, hi = 0;
    for (uint32_t t = 0; t < nt; ++t) {
      const ull m = mu[t];
      if (!m) continue;
      const ull p = Phist[(size_t)slot[t] * nn + e];
      const ull plo = m * p, phi = __umul64hi(m, p);
      const ull nlo = lo + plo;
      hi += phi + (nlo < lo ? 1ull : 0ull);
      lo = nlo;
    }
    Hlo[e] = lo;
    Hhi[e] = hi;
  }
}

__global__ void k_dn_zero(ull *p, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// 64-bit words per matrix row for n vertices (the instantiated widths; the
// count's work per row operation grows with the width)
uint32_t dense_words(uint32_t n) {
  return n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : n <= 640 ? 10 : n <= 768 ? 12 : 16;
}

uint32_t dgrid(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 16));
}

typedef void (*PassFn)(const ull *, uint32_t, uint32_t, uint32_t *, ull *, ull *);

template <int W, bool WP> PassFn pass_fn_k(uint32_t k) {
  switch (k) {
    case 3: return k_dn_pass<W, 3, WP>;
    case 4: return k_dn_pass<W, 4, WP>;
    case 5: return k_dn_pass<W, 5, WP>;
    case 6: return k_dn_pass<W, 6, WP>;
    case 7: return k_dn_pass<W, 7, WP>;
    case 8: return k_dn_pass<W, 8, WP>;
  }
  return nullptr;
}

template <bool WP> PassFn pass_fn(uint32_t W, uint32_t k) {
  switch (W) {
    case 1: return pass_fn_k<1, WP>(k);
    case 2: return pass_fn_k<2, WP>(k);
    case 4: return pass_fn_k<4, WP>(k);
    case 8: return pass_fn_k<8, WP>(k);
    case 10: return pass_fn_k<10, WP>(k);
    case 12: return pass_fn_k<12, WP>(k);
    case 16: return pass_fn_k<16, WP>(k);
  }
  return nullptr;
}

typedef void (*CritFn)(const ull *, uint32_t, const uint32_t *, uint32_t, const ull *, uint32_t *, ull *);

template <int W> CritFn crit_fn_k(uint32_t k) {
  switch (k) {
    case 3: return k_dn_crit<W, 3>;
    case 4: return k_dn_crit<W, 4>;
    case 5: return k_dn_crit<W, 5>;
    case 6: return k_dn_crit<W, 6>;
    case 7: return k_dn_crit<W, 7>;
    case 8: return k_dn_crit<W, 8>;
  }
  return nullptr;
}

CritFn crit_fn(uint32_t W, uint32_t k) {
  switch (W) {
    case 1: return crit_fn_k<1>(k);
    case 2: return crit_fn_k<2>(k);
    case 4: return crit_fn_k<4>(k);
    case 8: return crit_fn_k<8>(k);
    case 10: return crit_fn_k<10>(k);
    case 12: return crit_fn_k<12>(k);
    case 16: return crit_fn_k<16>(k);
  }
  return nullptr;
}

// Integer max-flow (Dinic) with 128-bit capacities.
struct Dinic {
  struct Arc {
    uint32_t to, rev;
    i128 cap;
  };
  std::vector<std::vector<Arc>> g;
  std::vector<int> level;
  std::vector<size_t> it;
  explicit Dinic(uint32_t n) : g(n), level(n), it(n) {}
  void add(uint32_t a, uint32_t b, i128 cap) {
    if (cap <= 0) return;
    g[a].push_back(Arc{b, (uint32_t)g[b].size(), cap});
    g[b].push_back(Arc{a, (uint32_t)g[a].size() - 1, 0});
  }
  bool bfs(uint32_t s, uint32_t t) {
    std::fill(level.begin(), level.end(), -1);
    std::vector<uint32_t> qu{s};
    level[s] = 0;
    for (size_t h = 0; h < qu.size(); ++h)
      for (const Arc &a : g[qu[h]])
        if (a.cap > 0 && level[a.to] < 0) {
          level[a.to] = level[qu[h]] + 1;
          qu.push_back(a.to);
        }
    return level[t] >= 0;
  }
  i128 dfs(uint32_t v, uint32_t t, i128 f) {
    if (v == t) return f;
    for (size_t &i = it[v]; i < g[v].size(); ++i) {
      Arc &a = g[v][i];
      if (a.cap <= 0 || level[a.to] != level[v] + 1) continue;
      const i128 d = dfs(a.to, t, std::min(f, a.cap));
      if (d > 0) {
        a.cap -= d;
        g[a.to][a.rev].cap += d;
        return d;
      }
    }
    return 0;
  }
  i128 run(uint32_t s, uint32_t t) {
    i128 flow = 0;
    while (bfs(s, t)) {
      std::fill(it.begin(), it.end(), 0);
      while (true) {
        const i128 f = dfs(s, t, ((i128)1) << 120);
        if (f <= 0) break;
        flow += f;
      }
    }
    return flow;
  }
};

// Wolfe's minimum-norm-point algorithm over the polytope of clique splits:
// a corral of affinely independent greedy splits (atoms) with convex
// weights, the Cholesky factor R of (Q^T Q + 1 1^T) updated as atoms enter
// and leave, so every affine minimisation costs O(m^2).
struct Corral {
  uint32_t n = 0, cap = 0, m = 0;
  std::vector<std::vector<ull>> q;   // exact atoms (local ids)
  std::vector<std::vector<double>> qs;  // scaled atoms
  std::vector<uint32_t> slot;        // pair-holding slot of each atom
  std::vector<double> lam;
  std::vector<double> R;             // row-major cap x cap, upper triangular
  double &r(uint32_t i, uint32_t j) { return R[(size_t)i * cap + j]; }

  // adds atom a (scaled copy sa); false when it is affinely dependent
  bool add(std::vector<ull> &&a, std::vector<double> &&sa, uint32_t sl) {
    std::vector<double> v(m + 1);
    for (uint32_t i = 0; i < m; ++i) {
      double d = 1.0;
      for (uint32_t u = 0; u < n; ++u) d += qs[i][u] * sa[u];
      v[i] = d;
    }
    double ss = 1.0;
    for (uint32_t u = 0; u < n; ++u) ss += sa[u] * sa[u];
    double rr = 0;
    for (uint32_t i = 0; i < m; ++i) {
      double t = v[i];
      for (uint32_t j = 0; j < i; ++j) t -= r(j, i) * v[j];
      v[i] = t / r(i, i);
      rr += v[i] * v[i];
    }
    const double d2 = ss - rr;
    if (!(d2 > 1e-13 * ss)) return false;
    for (uint32_t i = 0; i < m; ++i) r(i, m) = v[i];
    r(m, m) = std::sqrt(d2);
    q.push_back(std::move(a));
    qs.push_back(std::move(sa));
    slot.push_back(sl);
    lam.push_back(0.0);
    ++m;
    return true;
  }

  // removes atom p (its slot is returned)
  uint32_t remove(uint32_t p) {
    const uint32_t sl = slot[p];
    for (uint32_t j = p; j + 1 < m; ++j)
      for (uint32_t i = 0; i <= j + 1 && i < m; ++i) r(i, j) = r(i, j + 1);
    // the columns from p on are upper Hessenberg: Givens rotations on rows (j, j+1)
    for (uint32_t j = p; j + 1 < m; ++j) {
      const double a = r(j, j), b = r(j + 1, j);
      const double h = std::hypot(a, b);
      const double cs = h > 0 ? a / h : 1.0, sn = h > 0 ? b / h : 0.0;
      for (uint32_t col = j; col + 1 < m; ++col) {
        const double x1 = r(j, col), x2 = r(j + 1, col);
        r(j, col) = cs * x1 + sn * x2;
        r(j + 1, col) = -sn * x1 + cs * x2;
      }
    }
    for (uint32_t i = 0; i < m; ++i) r(i, m - 1) = 0;
    for (uint32_t j = 0; j < m; ++j) r(m - 1, j) = 0;
    q.erase(q.begin() + p);
    qs.erase(qs.begin() + p);
    slot.erase(slot.begin() + p);
    lam.erase(lam.begin() + p);
    --m;
    return sl;
  }

  // weights of the point of least norm in the affine hull of the corral
  std::vector<double> affine_min() {
    std::vector<double> y(m, 1.0);
    for (uint32_t i = 0; i < m; ++i) {
      double t = y[i];
      for (uint32_t j = 0; j < i; ++j) t -= r(j, i) * y[j];
      y[i] = t / r(i, i);
    }
    for (uint32_t i = m; i-- > 0;) {
      double t = y[i];
      for (uint32_t j = i + 1; j < m; ++j) t -= r(i, j) * y[j];
      y[i] = t / r(i, i);
    }
    double s = 0;
    for (double v : y) s += v;
    for (double &v : y) v /= s;
    return y;
  }

  void point(const std::vector<double> &w, std::vector<double> &x) const {
    std::fill(x.begin(), x.end(), 0.0);
    for (uint32_t i = 0; i < m; ++i)
      for (uint32_t u = 0; u < n; ++u) x[u] += w[i] * qs[i][u];
  }
};

struct DenseSolver {
  Ctx &c;
  uint32_t k, n, W;
  DevBuf M0, Mr, rankd, qd, Pd, ql, Phist, mu, slotd, Hlo, Hhi, tick;
  uint32_t hist_cap = 0;
  uint32_t passes = 0;
  double pass_ms = 0;  // wall time inside pass()
  // the critical set (local ids) and, per history slot, the order and the
  // critical-set size its holdings were computed with
  std::vector<uint8_t> crit;
  uint32_t ncrit = 0;
  std::vector<std::vector<uint32_t>> slot_order;
  std::vector<uint32_t> slot_ncrit;

  DenseSolver(Ctx &cc, uint32_t kk, uint32_t nn) : c(cc), k(kk), n(nn), crit(nn, 0) {
    W = dense_words(n);
  }
  // upload layout: ranks (n), the critical set as a rank-order bitmask (W
  // words), the critical vertices' ranks (at most CRIT_MAX)
  static constexpr uint32_t CRIT_MAX = 64;
  size_t cm_at() const { return (sizeof(uint32_t) * n + 7) & ~(size_t)7; }
  size_t cl_at() const { return cm_at() + sizeof(ull) * W; }
  size_t rank_bytes() const { return cl_at() + sizeof(uint32_t) * CRIT_MAX; }
  ~DenseSolver() {
    for (DevBuf *b : {&M0, &Mr, &rankd, &qd, &Pd, &ql, &Phist, &mu, &slotd, &Hlo, &Hhi, &tick}) buf_free(c, *b);
  }

  // q (local ids) of the greedy split by order (order[0] first); with want_P
  // the pair holdings go to history slot `slot`.
  std::vector<ull> pass(const std::vector<uint32_t> &order, bool want_P, uint32_t slot) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t nn = (uint64_t)n * n;
    // ranks and the critical set, in one upload
    std::vector<uint8_t> up(rank_bytes(), 0);
    uint32_t *rank = (uint32_t *)up.data(), *cl = (uint32_t *)(up.data() + cl_at());
    ull *cm = (ull *)(up.data() + cm_at());
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n; ++i) {
      rank[order[i]] = i;
      if (crit[order[i]]) {
        cm[i >> 6] |= 1ull << (i & 63);
        cl[nc++] = i;
      }
    }
    if (nc != ncrit || nc > CRIT_MAX) fail(KCDS_E_INVALID, "dense pass: critical set of %u", nc);
    HIP_CHECK(hipMemcpyAsync(rankd.p, up.data(), up.size(), hipMemcpyHostToDevice, c.stream));
    if (want_P) {
      slot_order[slot] = order;
      slot_ncrit[slot] = ncrit;
    }
    k_dn_zero<<<dgrid((uint64_t)n * W), 256, 0, c.stream>>>(Mr.as<ull>(), (uint64_t)n * W);
    k_dn_zero<<<dgrid(n), 256, 0, c.stream>>>(qd.as<ull>(), n);
    if (want_P) k_dn_zero<<<dgrid(nn), 256, 0, c.stream>>>(Pd.as<ull>(), nn);
    HIP_CHECK(hipMemsetAsync(tick.p, 0, 8, c.stream));
    k_dn_perm<<<dgrid(nn), 256, 0, c.stream>>>(M0.as<ull>(), n, W, rankd.as<uint32_t>(), Mr.as<ull>());
    KCDS_LAUNCH_CHECK();
    PassFn fn = want_P ? pass_fn<true>(W, k) : pass_fn<false>(W, k);
    if (!fn) fail(KCDS_E_INVALID, "dense pass: k = %u, %u words", k, W);
    // at least ~16 tasks per wave of the launch
    uint32_t split = 1;
    while (split < 64 && (uint64_t)n * W * split < 16ull * c.num_cus * (DN_THREADS / 64)) split *= 2;
    hipLaunchKernelGGL(fn, dim3(c.num_cus), dim3(DN_THREADS), 0, c.stream, Mr.as<ull>(), n, split,
                       tick.as<uint32_t>(), qd.as<ull>(), want_P ? Pd.as<ull>() : nullptr);
    KCDS_LAUNCH_CHECK();
    if (want_P && nc) {
      CritFn cf = crit_fn(W, k);
      if (!cf) fail(KCDS_E_INVALID, "dense critical pass: k = %u, %u words", k, W);
      hipLaunchKernelGGL(cf, dim3(c.num_cus), dim3(DN_THREADS), 0, c.stream, Mr.as<ull>(), n,
                         (const uint32_t *)(rankd.as<uint8_t>() + cl_at()), nc,
                         (const ull *)(rankd.as<uint8_t>() + cm_at()), tick.as<uint32_t>() + 1, Pd.as<ull>());
      KCDS_LAUNCH_CHECK();
    }
    ull *Pl = want_P ? Phist.as<ull>() + (size_t)slot * nn : nullptr;
    k_dn_unrank<<<dgrid(nn), 256, 0, c.stream>>>(n, rankd.as<uint32_t>(), qd.as<ull>(),
                                                 want_P ? Pd.as<ull>() : nullptr, ql.as<ull>(), Pl);
    KCDS_LAUNCH_CHECK();
    std::vector<ull> out(n);
    HIP_CHECK(hipMemcpyAsync(out.data(), ql.p, sizeof(ull) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    ++passes;
    pass_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return out;
  }
};



// One sub-round of the dense-mode peel: the live cliques that meet the
// frontier F, each counted once, from its smallest frontier member v: the
// cliques v forms inside X_v = live vertices minus the frontier members below
// v.  Task (F[fi], candidate word w0, sub-range) as in k_dn_pass.
// dec[j] += destroyed cliques containing j (frontier members included: the
// caller ignores them); *total += destroyed cliques.  K >= 3.
template <int W, int K>
__global__ __launch_bounds__(DN_THREADS) void k_dp_pass(const ull *__restrict__ M, uint32_t n,
                                                        const uint32_t *__restrict__ F, uint32_t nf,
                                                        const ull *__restrict__ live, const ull *__restrict__ fmask,
                                                        uint32_t split, uint32_t *ticket, ull *dec, ull *total) {
  constexpr uint32_t NP = 64u * W;
  __shared__ ull A[W * NP];
  __shared__ ull s_live[W], s_f[W];
  for (uint32_t e = threadIdx.x; e < W * NP; e += blockDim.x) {
    const uint32_t w = e / NP, j = e % NP;
    A[e] = j < n ? M[(size_t)j * W + w] : 0ull;
  }
  if (threadIdx.x < W) {
    s_live[threadIdx.x] = live[threadIdx.x];
    s_f[threadIdx.x] = fmask[threadIdx.x];
  }
  __syncthreads();
  const uint32_t lane = lane_id();
  const uint32_t ntask = nf * W * split, bits = 64 / split;
  ull tot = 0;
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (t >= ntask) break;
    const uint32_t sub = t % split, iw = t / split;
    const uint32_t fi = iw / W, w0 = iw % W;
    const uint32_t v = F[fi];
    ull S[W];
#pragma unroll
    for (int w = 0; w < W; ++w) S[w] = uniform_u64(drow<W>(A, v, w) & s_live[w] & ~(s_f[w] & dbelow(w, v)));
    ull cand = 0;
#pragma unroll
    for (int w = 0; w < W; ++w)
      if ((uint32_t)w == w0) cand = S[w];
    cand &= bits == 64 ? ~0ull : (((1ull << bits) - 1ull) << (sub * bits));
    if (!cand) continue;
    ull cr[W];
#pragma unroll
    for (int w = 0; w < W; ++w) cr[w] = 0;
    ull Et = 0;
    while (cand) {
      const int b = __builtin_ctzll(cand);
      cand &= cand - 1;
      const uint32_t c = w0 * 64 + (uint32_t)b;
      ull S2[W];
      uint32_t pc = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) S2[w] = drow<W>(A, c, w);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        S2[w] = S[w] & uniform_u64(S2[w]) & dabove(w, c);
        pc += (uint32_t)__popcll(S2[w]);
      }
      if (pc + 2 < (uint32_t)K) continue;
      const ull E = DCount<W, K - 2, true>::run(A, S2, cr);
      dcredit<W, true>(cr, c, E);
      Et += E;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t j = (uint32_t)w * 64 + lane;
      if (cr[w] && j < n) atomicAdd(&dec[j], cr[w]);
    }
    tot += Et;
  }
  if (lane == 0 && tot) atomicAdd(total, tot);
}

// live vertices (core = -1) of the enumerating peel
__global__ void k_dp_collect(uint32_t n, const long long *core, uint32_t *cnt, uint32_t *verts) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    if (core[r] == -1) {
      const uint32_t at = atomicAdd(cnt, 1u);
      if (at < DN_MAX) verts[at] = r;
    }
}

__global__ void k_dp_gather(uint32_t nd, const uint32_t *verts, const ull *cnt, Scatter sc, ull *out) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < nd; u += gridDim.x * blockDim.x)
    out[u] = cnt[slot_of(sc, verts[u])];
}

__global__ void k_dp_write_core(uint32_t nd, const uint32_t *verts, const long long *core_l, long long *core) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < nd; u += gridDim.x * blockDim.x)
    core[verts[u]] = core_l[u];
}

typedef void (*DpFn)(const ull *, uint32_t, const uint32_t *, uint32_t, const ull *, const ull *, uint32_t,
                     uint32_t *, ull *, ull *);

template <int W> DpFn dp_fn_k(uint32_t k) {
  switch (k) {
    case 3: return k_dp_pass<W, 3>;
    case 4: return k_dp_pass<W, 4>;
    case 5: return k_dp_pass<W, 5>;
    case 6: return k_dp_pass<W, 6>;
    case 7: return k_dp_pass<W, 7>;
    case 8: return k_dp_pass<W, 8>;
  }
  return nullptr;
}

DpFn dp_fn(uint32_t W, uint32_t k) {
  switch (W) {
    case 1: return dp_fn_k<1>(k);
    case 2: return dp_fn_k<2>(k);
    case 4: return dp_fn_k<4>(k);
    case 8: return dp_fn_k<8>(k);
    case 10: return dp_fn_k<10>(k);
    case 12: return dp_fn_k<12>(k);
    case 16: return dp_fn_k<16>(k);
  }
  return nullptr;
}

}  // namespace

// Dense exact search over graph 0's k'-core (c.kdoff/c.kdnbr hold its DAG).
// Returns false when the core has more than 1,024 vertices or the
// certificate was not reached within max_iters major cycles of the
// minimum-norm-point search.  On
// success: num/den = cliques in / size of the maximal densest set, which
// c.inset (new ids) marks; ncomp = components of the core's clique graph;
// passes = greedy passes run.
bool dense_core_solve(Ctx &c, uint32_t k, ull &num, ull &den, uint32_t &ncomp, uint32_t &passes, uint32_t &critical,
                      uint32_t max_iters) {
  const uint32_t nall = c.n;
  const long long kp = (long long)c.h_gpeel[0].k_prime;
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * (DN_MAX + 64));
  uint32_t *dcnt = c.tmp1.as<uint32_t>(), *dverts = dcnt + 64;
  HIP_CHECK(hipMemsetAsync(dcnt, 0, 4, c.stream));
  k_dn_collect<<<dgrid(nall), 256, 0, c.stream>>>(nall, c.core.as<long long>(), kp, dcnt, dverts);
  KCDS_LAUNCH_CHECK();
  uint32_t nc = 0;
  HIP_CHECK(hipMemcpyAsync(&nc, dcnt, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (nc == 0 || nc > DN_MAX || k < 3) return false;
  std::vector<uint32_t> verts(nc);
  HIP_CHECK(hipMemcpyAsync(verts.data(), dverts, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  std::sort(verts.begin(), verts.end());
  DenseSolver D(c, k, nc);
  const uint32_t n = nc, W = D.W;
  // holdings are in units of 1/L, L = lcm(1..k-1); unit = L/(k-1)
  const uint32_t L = k <= 3 ? 2u : k == 4 ? 6u : k == 5 ? 12u : k <= 7 ? 60u : 420u, unit = L / (k - 1);
  const uint64_t nn = (uint64_t)n * n;
  buf_reserve(c, D.M0, sizeof(ull) * (size_t)n * W);
  buf_reserve(c, D.Mr, sizeof(ull) * (size_t)n * W);
  buf_reserve(c, D.rankd, D.rank_bytes());
  buf_reserve(c, D.qd, sizeof(ull) * (n + 1));
  buf_reserve(c, D.ql, sizeof(ull) * (n + 1));
  buf_reserve(c, D.Pd, sizeof(ull) * nn);
  buf_reserve(c, D.tick, 256);
  // pair holdings of every atom in the corral (at most n + 1) and of the
  // newest split
  D.hist_cap = n + 3;
  buf_reserve(c, D.Phist, sizeof(ull) * nn * D.hist_cap);
  D.slot_order.assign(D.hist_cap, {});
  D.slot_ncrit.assign(D.hist_cap, 0);
  // local ids and the core matrix
  buf_reserve(c, c.tmp2, sizeof(uint32_t) * ((size_t)nall + 1));
  uint32_t *loc = c.tmp2.as<uint32_t>();
  HIP_CHECK(hipMemsetAsync(loc, 0xFF, sizeof(uint32_t) * (size_t)nall, c.stream));
  HIP_CHECK(hipMemcpyAsync(dverts, verts.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
  k_dn_loc<<<dgrid(n), 256, 0, c.stream>>>(n, dverts, loc);
  k_dn_zero<<<dgrid((uint64_t)n * W), 256, 0, c.stream>>>(D.M0.as<ull>(), (uint64_t)n * W);
  k_dn_adj<<<dgrid((uint64_t)n * 64), 256, 0, c.stream>>>(n, W, dverts, loc, c.kdoff.as<uint32_t>(),
                                                          c.kdnbr.as<uint32_t>(), D.M0.as<ull>());
  KCDS_LAUNCH_CHECK();
  // the split by local id: per-vertex clique counts (held + contained) and
  // the clique graph's components (members joined to their clique's root)
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::vector<ull> deg(n, 0);
  {
    std::vector<ull> q0 = D.pass(order, true, 0);
    std::vector<ull> P0(nn);
    HIP_CHECK(hipMemcpyAsync(P0.data(), D.Phist.p, sizeof(ull) * nn, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::vector<uint32_t> par(n);
    std::iota(par.begin(), par.end(), 0u);
    std::function<uint32_t(uint32_t)> find = [&](uint32_t a) { return par[a] == a ? a : (par[a] = find(par[a])); };
    std::vector<uint8_t> any(n, 0);
    for (uint32_t u = 0; u < n; ++u) {
      deg[u] += q0[u];
      for (uint32_t w = 0; w < n; ++w)
        if (P0[(size_t)u * n + w]) {
          deg[w] += P0[(size_t)u * n + w] / unit;  // no critical set yet: unit per clique
          any[u] = any[w] = 1;
          par[find(u)] = find(w);
        }
    }
    ncomp = 0;
    for (uint32_t u = 0; u < n; ++u) ncomp += any[u] && find(u) == u;
  }
  // start: each clique to its member of smallest clique-degree
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return deg[a] < deg[b]; });
  // the critical set starts with the core's members of least slack: those
  // whose clique-degree is within 10 % of the core's density c(core)/n (the
  // certificate's usual bottleneck when the answer is the whole core), at
  // most SEED_CRIT of them, so the atoms carry the rule from the first one
  {
    constexpr uint32_t SEED_CRIT = 8;
    u128 tot = 0;
    for (uint32_t u = 0; u < n; ++u) tot += deg[u];
    const u128 cc = tot / k;  // every clique is counted at its k members
    for (uint32_t i = 0; i < n && D.ncrit < SEED_CRIT; ++i) {
      const uint32_t u = order[i];
      // deg[u] <= 1.1 * cc / n, in integers
      if ((u128)deg[u] * 10 * n > cc * 11) break;
      D.crit[u] = 1;
      ++D.ncrit;
    }
  }
  Corral Q;
  Q.n = n;
  Q.cap = n + 3;
  Q.R.assign((size_t)Q.cap * Q.cap, 0.0);
  std::vector<uint32_t> free_slots;
  for (uint32_t sl = D.hist_cap; sl-- > 0;) free_slots.push_back(sl);
  double scale = 1.0;
  auto scaled = [&](const std::vector<ull> &a) {
    std::vector<double> v(n);
    for (uint32_t u = 0; u < n; ++u) v[u] = (double)a[u] / scale;
    return v;
  };
  std::vector<double> x(n);
  {
    const uint32_t sl = free_slots.back();
    free_slots.pop_back();
    std::vector<ull> q0 = D.pass(order, true, sl);
    for (ull v : q0) scale = std::max(scale, (double)v);
    std::vector<double> s0 = scaled(q0);
    Q.add(std::move(q0), std::move(s0), sl);
    Q.lam[0] = 1.0;
    Q.point(Q.lam, x);
  }
  // the count of every prefix of the vertices ordered by load (descending):
  // a greedy pass of the reversed order credits each clique to its last member
  auto candidate = [&](std::vector<uint32_t> &S, ull &cS) {
    std::vector<uint32_t> desc(n);
    std::iota(desc.begin(), desc.end(), 0u);
    std::stable_sort(desc.begin(), desc.end(), [&](uint32_t a, uint32_t b) { return x[a] > x[b]; });
    std::vector<uint32_t> rev(desc.rbegin(), desc.rend());
    std::vector<ull> last = D.pass(rev, false, 0);
    ull acc = 0, bn = 0, bd = 1;
    uint32_t bj = 0;
    for (uint32_t j = 0; j < n; ++j) {
      acc += last[desc[j]];
      if ((u128)acc * bd >= (u128)bn * (j + 1)) {  // ties keep the larger prefix
        bn = acc;
        bd = j + 1;
        bj = j + 1;
      }
    }
    S.assign(desc.begin(), desc.begin() + bj);
    cS = bn;
  };
  const bool log = getenv("KCDS_LOG") != nullptr;  // progress lines on stderr
  // the certificate for S (see the file comment); true when it holds.  When
  // it fails, unsat = the vertices whose shortfall or room the flow could not
  // fill, the largest first.
  auto certify = [&](const std::vector<uint32_t> &S, ull cS, std::vector<uint32_t> &unsat) -> bool {
    unsat.clear();
    const uint32_t T = Q.m;
    double ls = 0;
    for (double l : Q.lam) ls += std::max(0.0, l);
    std::vector<ull> mu(T);
    for (uint32_t t = 0; t < T; ++t) mu[t] = (ull)std::floor(std::max(0.0, Q.lam[t]) / ls * 1125899906842624.0);  // 2^50
    i128 Dsum = 0;
    for (ull m : mu) Dsum += m;
    if (Dsum == 0) return false;
    buf_reserve(c, D.mu, sizeof(ull) * (T + 1));
    buf_reserve(c, D.slotd, sizeof(uint32_t) * (T + 1));
    buf_reserve(c, D.Hlo, sizeof(ull) * nn);
    buf_reserve(c, D.Hhi, sizeof(ull) * nn);
    HIP_CHECK(hipMemcpyAsync(D.mu.p, mu.data(), sizeof(ull) * T, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(D.slotd.p, Q.slot.data(), sizeof(uint32_t) * T, hipMemcpyHostToDevice, c.stream));
    k_dn_hold<<<dgrid(nn), 256, 0, c.stream>>>(nn, T, D.mu.as<ull>(), D.slotd.as<uint32_t>(), D.Phist.as<ull>(),
                                               D.Hlo.as<ull>(), D.Hhi.as<ull>());
    KCDS_LAUNCH_CHECK();
    std::vector<ull> hlo(nn), hhi(nn);
    HIP_CHECK(hipMemcpyAsync(hlo.data(), D.Hlo.p, sizeof(ull) * nn, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(hhi.data(), D.Hhi.p, sizeof(ull) * nn, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    const i128 s = (i128)S.size();
    std::vector<uint8_t> inS(n, 0);
    for (uint32_t v : S) inS[v] = 1;
    // loads and target in units of 1/(D |S|): X'_v = |S| sum mu q, tau = D c_S
    const i128 tau = Dsum * (i128)cS;
    std::vector<i128> X(n, 0);
    for (uint32_t t = 0; t < T; ++t)
      for (uint32_t u = 0; u < n; ++u) X[u] += (i128)mu[t] * (i128)Q.q[t][u];
    for (uint32_t u = 0; u < n; ++u) X[u] *= s;
    const uint32_t SRC = n, SNK = n + 1, OPT = n + 2;
    Dinic F(n + 3);
    i128 sup = 0, mand = 0;
    for (uint32_t u = 0; u < n; ++u) {
      if (inS[u]) {
        if (X[u] > tau) {
          F.add(SRC, u, X[u] - tau);
          sup += X[u] - tau;
        } else if (X[u] < tau) {
          F.add(u, SNK, tau - X[u]);
          mand += tau - X[u];
        }
      } else if (X[u] >= tau) {
        F.add(SRC, u, X[u] - tau + 1);
        sup += X[u] - tau + 1;
      } else if (X[u] < tau - 1) {
        F.add(u, OPT, tau - 1 - X[u]);
      }
    }
    if (sup < mand) return false;
    F.add(OPT, SNK, sup - mand);
    // u may hand w the part of its holdings that w may take: H(u, w) / L
    for (uint32_t u = 0; u < n; ++u)
      for (uint32_t w = 0; w < n; ++w) {
        if (u == w) continue;
        const u128 h = ((u128)hhi[(size_t)u * n + w] << 64) | hlo[(size_t)u * n + w];
        if (!h) continue;
        F.add(u, w, (i128)((h * (u128)s) / (u128)L));
      }
    const i128 fl = F.run(SRC, SNK);
    if (fl == sup) return true;
    // the vertices the flow could not fill: members of S short of tau and
    // outsiders with room left (the excess of S is mostly the shares S holds
    // in cliques with outsiders, which have to pass to those outsiders)
    std::vector<std::pair<i128, uint32_t>> un;
    for (uint32_t u = 0; u < n; ++u)
      for (const Dinic::Arc &a : F.g[u])
        if ((a.to == SNK || a.to == OPT) && a.cap > 0) un.push_back({a.cap, u});
    std::sort(un.begin(), un.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (const auto &e : un) unsat.push_back(e.second);
    if (log)
      fprintf(stderr, "kcds dense: certificate short: flow/supply %.6f, %zu vertices unfilled, %u critical\n",
              (double)fl / (double)sup, un.size(), D.ncrit);
    return false;
  };
  std::vector<uint32_t> S;
  ull cS = 0;
  bool ok = false;
  uint32_t attempts = 0;
  double cert_ms = 0;
  // A member of S whose clique-degree in S barely exceeds rho_S must end
  // with almost all of its cliques, and an outsider must end with almost all
  // the shares of its cliques with S; under the 1/(k-1) rule a holder can
  // hand such a vertex only a (k-1)-th of its share, so the certificate
  // would wait for x to converge far.  The vertices the flow left unfilled
  // join the critical set, the corral's holdings are recomputed under it,
  // and the certificate is tried again at once.
  constexpr uint32_t CRIT_STEP = 8, CRIT_MAX = DenseSolver::CRIT_MAX;
  auto attempt = [&]() {
    ++attempts;
    candidate(S, cS);
    if (S.empty()) return false;
    const auto t0 = std::chrono::steady_clock::now();
    bool r = false;
    std::vector<uint32_t> unsat;
    while (!(r = certify(S, cS, unsat))) {
      uint32_t added = 0;
      for (uint32_t u : unsat)
        if (!D.crit[u] && added < CRIT_STEP && D.ncrit < CRIT_MAX) {
          D.crit[u] = 1;
          ++D.ncrit;
          ++added;
        }
      if (!added) break;
      for (uint32_t t = 0; t < Q.m; ++t)
        if (D.slot_ncrit[Q.slot[t]] != D.ncrit) (void)D.pass(D.slot_order[Q.slot[t]], true, Q.slot[t]);
    }
    cert_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return r;
  };
  const auto t_start = std::chrono::steady_clock::now();
  // the certificate has slack, so it often holds well before x has
  // converged: it is tried from cycle 16 or a gap of 3e-3 on (C2 19 passes,
  // C3 19; from cycle 32 or 1e-3, 30 and 28), then each time the gap has
  // fallen tenfold or the cycle count has grown by half
  double next_try_gap = 3e-3;
  uint32_t next_try_major = 16;
  for (uint32_t major = 1; major <= max_iters; ++major) {
    // linear oracle: the greedy split minimising <x, q> (each clique to its
    // member of smallest load)
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return x[a] < x[b]; });
    const uint32_t sl = free_slots.back();
    std::vector<ull> q = D.pass(order, true, sl);
    std::vector<double> sq = scaled(q);
    double xx = 0, xq = 0;
    for (uint32_t u = 0; u < n; ++u) {
      xx += x[u] * x[u];
      xq += x[u] * sq[u];
    }
    bool again = false;
    for (uint32_t i = 0; i < Q.m && !again; ++i) again = Q.q[i] == q;
    if (log && (major % 100 == 0 || again))
      fprintf(stderr, "kcds dense: n=%u major=%u corral=%u gap=%.3e\n", n, major, Q.m, (xx - xq) / xx);
    const double gap = (xx - xq) / xx;
    if (again || gap <= 1e-12 || !Q.add(std::move(q), std::move(sq), sl)) {
      ok = attempt();  // x is the minimum-norm point to working precision
      break;
    }
    if (gap <= next_try_gap || major >= next_try_major) {
      next_try_gap = gap * 0.1;
      next_try_major = major + std::max(32u, major / 2);
      if (log) fprintf(stderr, "kcds dense: attempt at major=%u gap=%.3e corral=%u\n", major, gap, Q.m);
      if ((ok = attempt())) break;
    }
    free_slots.pop_back();
    // minor cycles: move to the affine minimiser, or as far toward it as the
    // weights stay non-negative, dropping the atoms whose weight reaches 0
    for (uint32_t minor = 0; minor < 4 * Q.cap; ++minor) {
      std::vector<double> a = Q.affine_min();
      double theta = 1.0;
      bool inside = true;
      for (uint32_t i = 0; i < Q.m; ++i)
        if (a[i] <= 1e-14) {
          inside = false;
          if (Q.lam[i] - a[i] > 0) theta = std::min(theta, Q.lam[i] / (Q.lam[i] - a[i]));
        }
      if (inside) {
        Q.lam = a;
        break;
      }
      theta = std::max(0.0, theta);
      for (uint32_t i = 0; i < Q.m; ++i) Q.lam[i] += theta * (a[i] - Q.lam[i]);
      // drop the smallest weight and every vanished one
      uint32_t arg = 0;
      for (uint32_t i = 1; i < Q.m; ++i)
        if (Q.lam[i] < Q.lam[arg]) arg = i;
      for (uint32_t i = Q.m; i-- > 0;)
        if (Q.m > 1 && (i == arg || Q.lam[i] <= 1e-15)) free_slots.push_back(Q.remove(i));
      double s = 0;
      for (double l : Q.lam) s += l;
      for (double &l : Q.lam) l /= s;
    }
    Q.point(Q.lam, x);
  }
  passes = D.passes;
  critical = D.ncrit;
  if (log)
    fprintf(stderr, "kcds dense: n=%u k=%u passes=%u attempts=%u critical=%u ok=%d total %.1f ms, passes %.1f ms, "
            "certify %.1f ms\n", n, k, D.passes, attempts, D.ncrit, (int)ok,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count(), D.pass_ms,
            cert_ms);
  if (!ok) return false;
  // mark S (new ids) in c.inset
  buf_reserve(c, c.inset, (size_t)nall + 8);
  HIP_CHECK(hipMemsetAsync(c.inset.p, 0, nall, c.stream));
  std::vector<uint8_t> ins(nall, 0);
  HIP_CHECK(hipMemcpyAsync(ins.data(), c.inset.p, nall, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  for (uint32_t u : S) ins[verts[u]] = 1;
  HIP_CHECK(hipMemcpyAsync(c.inset.p, ins.data(), nall, hipMemcpyHostToDevice, c.stream));
  c.sync();
  num = cS;
  den = (ull)S.size();
  return true;
}


// Dense mode of the enumerating peel for a single graph whose live vertices
// (core = -1) number at most 1,024: the live subgraph as an LDS bitmask
// matrix, every sub-round's destroyed cliques counted in one pass rooted at
// the frontier (k_dp_pass), levels and densities kept on the host with the
// rules of the step kernels.  cnt/sc: the working counts (scattered layout);
// gp: graph 0's device record, brought up to date; last_level: the level
// that just ended (-1: none yet).  Returns the levels peeled here, or -1
// (nothing done) when more than 1,024 vertices are live.
long long dense_peel_finish(Ctx &c, const ull *cnt, Scatter sc, GraphPeel *gp, long long last_level) {
  const uint32_t nall = c.n, k = c.k;
  if (k < 3) return -1;
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * (DN_MAX + 64));
  uint32_t *dcnt = c.tmp1.as<uint32_t>(), *dverts = dcnt + 64;
  HIP_CHECK(hipMemsetAsync(dcnt, 0, 4, c.stream));
  k_dp_collect<<<dgrid(nall), 256, 0, c.stream>>>(nall, c.core.as<long long>(), dcnt, dverts);
  KCDS_LAUNCH_CHECK();
  uint32_t nd = 0;
  HIP_CHECK(hipMemcpyAsync(&nd, dcnt, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (nd > DN_MAX) return -1;
  GraphPeel P;
  HIP_CHECK(hipMemcpyAsync(&P, gp, sizeof(GraphPeel), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  // the level that just ended has not had its density recorded yet
  if (last_level >= 0 && P.touch && P.rem > 0 &&
      (u128)P.live * P.max_den >= (u128)P.max_num * P.rem) {
    P.max_num = P.live;
    P.max_den = P.rem;
  }
  P.touch = 0;
  long long levels = 0;
  if (nd > 0) {
    std::vector<uint32_t> verts(nd);
    HIP_CHECK(hipMemcpyAsync(verts.data(), dverts, sizeof(uint32_t) * nd, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::sort(verts.begin(), verts.end());
    const uint32_t n = nd, W = dense_words(n);
    DpFn fn = dp_fn(W, k);
    if (!fn) fail(KCDS_E_INVALID, "dense peel: k = %u", k);
    DevBuf M, loc, Fd, masks, dec, cl, tick, cstore;
    struct Free {
      Ctx &c;
      std::vector<DevBuf *> b;
      ~Free() {
        for (DevBuf *x : b) buf_free(c, *x);
      }
    } fr{c, {&M, &loc, &Fd, &masks, &dec, &cl, &tick, &cstore}};
    buf_reserve(c, M, sizeof(ull) * (size_t)n * W);
    buf_reserve(c, loc, sizeof(uint32_t) * ((size_t)nall + 1));
    buf_reserve(c, Fd, sizeof(uint32_t) * (n + 1));
    buf_reserve(c, masks, sizeof(ull) * 2 * 16);
    buf_reserve(c, dec, sizeof(ull) * (n + 1) + 64);
    buf_reserve(c, cl, sizeof(long long) * (n + 1));
    buf_reserve(c, tick, 256);
    buf_reserve(c, cstore, sizeof(ull) * (n + 1));
    HIP_CHECK(hipMemcpyAsync(dverts, verts.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemsetAsync(loc.p, 0xFF, sizeof(uint32_t) * (size_t)nall, c.stream));
    k_dn_loc<<<dgrid(n), 256, 0, c.stream>>>(n, dverts, loc.as<uint32_t>());
    k_dn_zero<<<dgrid((uint64_t)n * W), 256, 0, c.stream>>>(M.as<ull>(), (uint64_t)n * W);
    k_dn_adj<<<dgrid((uint64_t)n * 64), 256, 0, c.stream>>>(n, W, dverts, loc.as<uint32_t>(), c.doff.as<uint32_t>(),
                                                            c.dnbr.as<uint32_t>(), M.as<ull>());
    k_dp_gather<<<dgrid(n), 256, 0, c.stream>>>(n, dverts, cnt, sc, cstore.as<ull>());
    KCDS_LAUNCH_CHECK();
    std::vector<ull> cv(n);
    HIP_CHECK(hipMemcpyAsync(cv.data(), cstore.p, sizeof(ull) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::vector<uint8_t> alive(n, 1);
    std::vector<long long> core(n, -1);
    std::vector<ull> hmask(32, 0), dech(n + 1);
    uint32_t left = n;
    long long L = last_level;
    std::vector<uint32_t> F;
    const bool log = getenv("KCDS_LOG") != nullptr;  // progress lines on stderr
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t subrounds = 0;
    while (left) {
      ull m = ~0ull;
      for (uint32_t u = 0; u < n; ++u)
        if (alive[u]) m = std::min(m, cv[u]);
      L = std::max<long long>(L + 1, (long long)m);  // every live count exceeds the last level
      ++levels;
      while (true) {  // the level's sub-rounds
        F.clear();
        for (uint32_t u = 0; u < n; ++u)
          if (alive[u] && cv[u] <= (ull)L) F.push_back(u);
        if (F.empty()) break;
        ++subrounds;
        std::fill(hmask.begin(), hmask.end(), 0ull);
        for (uint32_t u = 0; u < n; ++u)
          if (alive[u]) hmask[u >> 6] |= 1ull << (u & 63);
        for (uint32_t u : F) hmask[16 + (u >> 6)] |= 1ull << (u & 63);
        HIP_CHECK(hipMemcpyAsync(masks.p, hmask.data(), sizeof(ull) * 32, hipMemcpyHostToDevice, c.stream));
        HIP_CHECK(hipMemcpyAsync(Fd.p, F.data(), sizeof(uint32_t) * F.size(), hipMemcpyHostToDevice, c.stream));
        HIP_CHECK(hipMemsetAsync(dec.p, 0, sizeof(ull) * (n + 1), c.stream));
        HIP_CHECK(hipMemsetAsync(tick.p, 0, 4, c.stream));
        const uint32_t nf = (uint32_t)F.size();
        uint32_t split = 1;
        while (split < 64 && (uint64_t)nf * W * split < 16ull * c.num_cus * (DN_THREADS / 64)) split *= 2;
        hipLaunchKernelGGL(fn, dim3(c.num_cus), dim3(DN_THREADS), 0, c.stream, M.as<ull>(), n, Fd.as<uint32_t>(), nf,
                           masks.as<ull>(), masks.as<ull>() + 16, split, tick.as<uint32_t>(), dec.as<ull>(),
                           dec.as<ull>() + n);
        KCDS_LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(dech.data(), dec.p, sizeof(ull) * (n + 1), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        for (uint32_t u : F) {
          alive[u] = 0;
          core[u] = L;
        }
        left -= nf;
        P.rem -= nf;
        P.live -= dech[n];
        for (uint32_t u = 0; u < n; ++u)
          if (alive[u]) cv[u] -= dech[u];
      }
      // density of the level just completed (ties keep the later core)
      if (P.rem > 0 && (u128)P.live * P.max_den >= (u128)P.max_num * P.rem) {
        P.max_num = P.live;
        P.max_den = P.rem;
      }
    }
    if (log)
      fprintf(stderr, "kcds dense peel: %u live vertices from level %lld: %lld levels, %llu sub-rounds, %.1f ms\n", n,
              last_level, levels, (unsigned long long)subrounds,
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    HIP_CHECK(hipMemcpyAsync(cl.p, core.data(), sizeof(long long) * n, hipMemcpyHostToDevice, c.stream));
    k_dp_write_core<<<dgrid(n), 256, 0, c.stream>>>(n, dverts, cl.as<long long>(), c.core.as<long long>());
    KCDS_LAUNCH_CHECK();
  }
  HIP_CHECK(hipMemcpyAsync(gp, &P, sizeof(GraphPeel), hipMemcpyHostToDevice, c.stream));
  c.sync();
  return levels;
}

}  // namespace kcds
