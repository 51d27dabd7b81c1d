// k-clique-core peeling without a stored clique list.
//
// Reference: cliqueCoreDecompose (main.cu:329-420) drives selectNodes and
// processNodesByWarp (src/helpers.cu:587-726) over the materialised clique
// list (trie, status): processNodesByWarp scans all T cliques for every
// removed vertex, marks its live cliques destroyed at level L and decrements
// the other members.  The level rule, the density bookkeeping and the k'
// rule are those of k_peel.hip.
//
// Here nothing per clique is stored, so graphs whose clique list would not
// fit in HBM (T of 10^9..10^11) peel too.  A sub-round removes a frontier F
// (live vertices of count <= L).  Every live clique that meets F is
// destroyed exactly once, by its smallest frontier member v (new ids): v
// enumerates the cliques it forms inside X_v = (live vertices outside F) +
// (frontier vertices above v) and decrements their members outside F.
// Frontier members are never decremented, so a frontier vertex's count stays
// the number of its live cliques for the whole sub-round; a count of 0 skips
// the vertex.  Two forms of the enumeration, one workgroup per work item:
//   narrow v (degree <= 512): stage N(v) & X_v as an LDS frame (bitmask rows
//     of the induced subgraph, built from the DAG out-lists exactly as the
//     listing kernel builds N+(r)'s frame) and count its (k-1)-cliques with
//     per-member credits;
//   wide v: for each candidate c in N(v) & X_v, the cliques whose smallest
//     member besides v is c.  The frame is N+(c) & X_v (at most the DAG's
//     out-degree bound), the anchor set is its members adjacent to v (v's
//     own row when v > c, else N+(v) looked up in the frame table), and the
//     (k-2)-cliques inside the anchor are counted; c is credited with their
//     number.  A work item carries 16 candidates.
// Levels and sub-rounds run as in the grid peel: one fused kernel per step
// (OPEN a level or PROCESS a sub-round) whose last workgroup runs the
// control step; the host enqueues steps in batches.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t ESEG = 4;                    // frontier segments / counter shards
constexpr uint32_t ED = 512;                    // frame rows (the DAG out-degree bound)
constexpr uint32_t EHB = 11, EH = 1u << EHB;    // frame table slots (load <= 1/4)
constexpr uint32_t ETHREADS = 512;
constexpr uint32_t EPIECE = 16;                 // candidates per wide work item
constexpr uint32_t ECHUNK_LOG = 16;             // a narrow frame takes a work item per 2^16 of its cliques
constexpr uint32_t EMAXP = 64;                  //   up to 64 (k >= 4)
constexpr uint32_t EUNROLL = 8;                 // 64-element chunks in flight per wave
constexpr uint32_t EMPTY_ID = 0xFFFFFFFFu;
constexpr uint32_t FMAX = 4096;                 // widest stored frame (members; see k_ep_frames_wide)
constexpr uint32_t FWORDS = FMAX / 64;
constexpr uint32_t FW_HB = 13, FW_H = 1u << FW_HB;  // k_ep_frames_wide's table slots (load <= 1/2)
constexpr uint32_t FW_ROWS = 128;                // rows k_ep_frames_wide builds per pass
constexpr uint32_t LIVE = 0xFFFFFFFFu;          // erem of a vertex still in the graph

// Counters 64 words apart (same-address device atomics are serialised).
struct EpCtl {
  ull level, target;
  uint32_t cur, done, phase, steps, levels, handoff, pad0[64 - 10];
  uint32_t fsize[2][ESEG][64];  // vertex frontier segments
  uint32_t isize[2][ESEG][64];  // work-item segments
  ull gmin[ESEG][8];            // minimum live count above target (OPEN)
  uint32_t gblocks[ESEG][64];
  uint32_t blocks[64];
  uint32_t ntouch[64];          // graphs touched by the current level
  ull dmin[ESEG][8];            // lowest count above the level left by decrements (PROCESS)
  ull nextmin[8];               // minimum live count above the level when it opened
  uint32_t ticket[64];          // next work item of the current PROCESS step
};

struct EFrame {
  uint32_t L[ED];     // members (new ids), in staging order
  uint2 sinfo[ED];    // (doff, out-degree) of each member
  ull hent[EH];       // table entry: low word member id (EMPTY_ID when free), high word index
  ull A[8 * ED];      // A[w*ED + j]: bits 64w..64w+63 of row j (symmetric adjacency)
  ull fr[8];          // bit j: member j is in the current frontier (never decremented)
  ull anc[8];         // anchor set of a wide task
  ull rt;             // 2 x cliques of the current task, summed over the waves
  uint32_t d, iv;     // members; index of v among them (EMPTY_ID: not staged)
  uint32_t item_idx, skip;
  uint4 item;
  uint32_t cand[EPIECE], crem[EPIECE];
  uint32_t wcnt[ETHREADS / 64];  // members kept by each wave in the current staging chunk
  ull smask[FWORDS];             // e_load: members of a stored frame that lie in X_v
  ull fmask[FWORDS];             //   those of them in the current frontier
  ull cmv[FWORDS * 6];           //   compress round masks of smask's words
  uint32_t sbase[FWORDS];        //   kept members before each word
};
static_assert(ETHREADS >= ED, "e_load takes a stored frame's members one per thread");

__device__ __forceinline__ uint32_t ehslot(uint32_t key) {
  return (uint32_t)__umul24(key ^ (key >> 24), 0x9E3779u) >> (32 - EHB);
}

__device__ __forceinline__ ull above_bits(int q, uint32_t c) {
  const int cq = (int)(c >> 6), cb = (int)(c & 63);
  if (q < cq) return 0ull;
  if (q > cq) return ~0ull;
  return cb == 63 ? 0ull : (~0ull << (cb + 1));
}

__device__ __forceinline__ ull erow(const EFrame &f, uint32_t j, int w) { return f.A[(uint32_t)w * ED + j]; }

__device__ __forceinline__ uint32_t words_of(uint32_t d) { return d <= 64 ? 1u : d <= 128 ? 2u : d <= 256 ? 4u : 8u; }
// 64-bit words per row of a stored frame of d members
__host__ __device__ __forceinline__ uint32_t stored_words(uint32_t d) {
  return d <= 64 ? 1u : d <= 128 ? 2u : d <= 256 ? 4u : d <= 512 ? 8u : (d + 63) / 64;
}

template <int W> __device__ __forceinline__ void ecredit(ull *cr, uint32_t c, ull E) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int q = 0; q < W; ++q)
    if ((uint32_t)q == (c >> 6) && lane == (c & 63)) cr[q] += E;
}

// Bits above position b of one word (none for b = 63).
__device__ __forceinline__ ull above_bit1(uint32_t b) { return b == 63 ? 0ull : (~0ull << (b + 1)); }

// Edges inside S (words below LO zero), both endpoints credited; rows q with
// q % nw == wv belong to this wave.  Returns the wave's share of 2 x edges.
// Each lane requests all its row words before using any; the whole wave is
// active here (the callers' control flow is workgroup-uniform).
template <int W, int LO>
__device__ __forceinline__ uint32_t e_edges(const EFrame &f, const ull *S, ull *cr, uint32_t wv, uint32_t nw) {
  const uint32_t lane = lane_id();
  uint32_t xs = 0;
#pragma unroll
  for (int q = LO; q < W; ++q) {
    if (!S[q] || (uint32_t)q % nw != wv) continue;
    if ((S[q] >> lane) & 1ull) {
      const uint32_t j = (uint32_t)q * 64 + lane;
      ull r[W];
#pragma unroll
      for (int w = LO; w < W; ++w) r[w] = erow(f, j, w);
      uint32_t x = 0;
#pragma unroll
      for (int w = LO; w < W; ++w) x += (uint32_t)__popcll(r[w] & S[w]);
      cr[q] += x;
      xs += x;
    }
  }
  return wave_sum_u32_full(xs);
}

template <int W, int REM, int LO> struct ECount;

// REM-cliques inside S whose first member lies in word Q or above (Q is a
// compile-time word index; the words of S below LO <= Q are zero).  The row
// words of a step are all requested before the first is used, and S stays in
// scalar registers.
template <int W, int REM, int Q> struct ECountFrom {
  static __device__ ull run(const EFrame &f, const ull *S, ull *cr) {
    if constexpr (Q >= W) {
      return 0;
    } else {
      ull total = 0;
      ull word = S[Q];
      while (word) {
        const uint32_t b = (uint32_t)__builtin_ctzll(word);
        word &= word - 1;
        const uint32_t c = (uint32_t)Q * 64 + b;
        ull S2[W];
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) S2[w] = w < Q ? 0ull : erow(f, c, w);
#pragma unroll
        for (int w = Q; w < W; ++w) {
          S2[w] = S[w] & uniform_u64(S2[w]) & (w == Q ? above_bit1(b) : ~0ull);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        const ull E = ECount<W, REM - 1, Q>::run(f, S2, cr);
        if (lane_id() == b) cr[Q] += E;
        total += E;
      }
      return total + ECountFrom<W, REM, Q + 1>::run(f, S, cr);
    }
  }
};

// REM-cliques inside S (one wave; words of S below LO are zero), every member
// credited; returns the count.
template <int W, int REM, int LO> struct ECount {
  static __device__ ull run(const EFrame &f, const ull *S, ull *cr) { return ECountFrom<W, REM, LO>::run(f, S, cr); }
};

template <int W, int LO> struct ECount<W, 2, LO> {
  static __device__ ull run(const EFrame &f, const ull *S, ull *cr) {
    return (ull)(e_edges<W, LO>(f, S, cr, 0, 1) >> 1);
  }
};

// ECount<W, REM, cq> for a first-level vertex in word cq (wave-uniform).
template <int W, int REM, int Q = 0>
__device__ __forceinline__ ull ecount_from(uint32_t cq, const EFrame &f, const ull *S, ull *cr) {
  if constexpr (Q + 1 < W) {
    if (cq != (uint32_t)Q) return ecount_from<W, REM, Q + 1>(cq, f, S, cr);
  }
  return ECount<W, REM, Q>::run(f, S, cr);
}

// REM-cliques inside S with per-member credits, the first level split over
// the workgroup's waves.  Returns this wave's share of 2 x the count.
template <int W, int REM>
__device__ ull e_count_set(const EFrame &f, const ull *S, ull *cr, uint32_t wv, uint32_t nw) {
  if constexpr (REM == 1) {
    if (wv != 0) return 0;
    const uint32_t lane = lane_id();
    ull tot = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if ((S[q] >> lane) & 1ull) cr[q] += 1;
      tot += (ull)__popcll(S[q]);
    }
    return 2 * tot;
  } else if constexpr (REM == 2) {
    return e_edges<W, 0>(f, S, cr, wv, nw);
  } else {
    ull tot = 0;
    uint32_t idx = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      ull word = S[q];
      while (word) {
        const int b = __builtin_ctzll(word);
        word &= word - 1;
        if ((idx++ % nw) != wv) continue;
        const uint32_t c = (uint32_t)(q * 64 + b);
        ull S2[W];
        uint32_t pc = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) S2[w] = erow(f, c, w);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          S2[w] = S[w] & uniform_u64(S2[w]) & above_bits(w, c);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        const ull E = ecount_from<W, REM - 1>(c >> 6, f, S2, cr);
        ecredit<W>(cr, c, E);
        tot += E;
      }
    }
    return 2 * tot;
  }
}

template <int REM>
__device__ __forceinline__ ull e_count(const EFrame &f, uint32_t Wd, const ull *S, ull *cr, uint32_t wv, uint32_t nw) {
  switch (Wd) {
    case 1: return e_count_set<1, REM>(f, S, cr, wv, nw);
    case 2: return e_count_set<2, REM>(f, S, cr, wv, nw);
    case 4: return e_count_set<4, REM>(f, S, cr, wv, nw);
    default: return e_count_set<8, REM>(f, S, cr, wv, nw);
  }
}

// Appends u (where want) to the next frontier: its vertex entry and its work
// items (none when its count is 0; one for a narrow vertex, or one per 2^16
// of its cliques, at most 64 and one per 16 neighbours, each counting a
// share of the frame's first-level members, when splits is set; one per 16
// neighbours for a wide one), and stamps erem[u] = the step that removes it.
// A vertex is narrow when at most ED of its neighbours can lie in its frame
// X_v at that step: neighbours stamped for an earlier step never come back,
// so those not stamped below it bound the frame from above (the wave counts
// them for a vertex of degree above ED, stopping past ED).
// Every lane of the wave calls it.
struct EAppend {
  const uint32_t *foff, *fnbr;
  uint32_t *fv, *fvsize;
  uint4 *fi;
  uint32_t *fisize;
  uint32_t *erem;
  uint32_t snext, splits;
};

__device__ __forceinline__ void e_append(bool want, uint32_t u, ull cu, const EAppend &a) {
  const ull bal = __ballot(want);
  if (!bal) return;
  const uint32_t lane = lane_id(), lead = (uint32_t)__builtin_ctzll(bal);
  uint32_t ni = 0, deg = 0;
  if (want) {
    deg = a.foff[u + 1] - a.foff[u];
    stc_u32(&a.erem[u], a.snext);
  }
  bool wide = false;
  ull big = __ballot(want && cu != 0 && deg > ED);
  while (big) {
    const uint32_t l = (uint32_t)__builtin_ctzll(big);
    big &= big - 1;
    const uint32_t x = __shfl(u, l, 64), xd = __shfl(deg, l, 64), xb = a.foff[x];
    uint32_t in_frame = 0;
    for (uint32_t t0 = 0; t0 < xd && in_frame <= ED; t0 += KCDS_WAVE) {
      const uint32_t t = t0 + lane;
      const bool in = t < xd && ldc_u32(&a.erem[a.fnbr[xb + t]]) >= a.snext;
      in_frame += (uint32_t)__popcll(__ballot(in));
    }
    if (lane == l) wide = in_frame > ED;
  }
  uint32_t pieces = 1;
  if (want && !wide && a.splits && (cu >> ECHUNK_LOG) > 0)
    pieces = (uint32_t)min(min((cu >> ECHUNK_LOG) + 1, (ull)EMAXP), (ull)((deg + EPIECE - 1) / EPIECE));
  if (want) ni = cu == 0 ? 0u : !wide ? pieces : (deg + EPIECE - 1) / EPIECE;
  const uint32_t incl = wave_incl_scan_u32(ni);
  const uint32_t tot = __shfl(incl, 63, 64);
  uint32_t vb = 0, ib = 0;
  if (lane == lead) {
    vb = atomicAdd(a.fvsize, (uint32_t)__popcll(bal));
    if (tot) ib = atomicAdd(a.fisize, tot);
  }
  vb = __shfl(vb, lead, 64);
  ib = __shfl(ib, lead, 64);
  if (want) {
    a.fv[vb + lanes_below(bal)] = u;
    const uint32_t p = ib + incl - ni;
    if (!wide) {
      if (ni == 1) a.fi[p] = make_uint4(u, 0u, 0u, 0u);
      else
        for (uint32_t j = 0; j < ni; ++j) a.fi[p + j] = make_uint4(u, j, ni, 2u);
    } else {
      for (uint32_t j = 0; j < ni; ++j) a.fi[p + j] = make_uint4(u, j * EPIECE, min((j + 1) * EPIECE, deg), 1u);
    }
  }
}

// Decrements from one wave's credits: members outside the frontier lose
// their credited cliques; one whose count falls from above L to L or below
// joins the next sub-round; dm tracks the lowest count left above L.
template <int W>
__device__ __forceinline__ void e_flush_w(const EFrame &f, uint32_t d, const ull *cr, ull L, ull *cnt, Scatter sc,
                                          const EAppend &a, ull &dm) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int q = 0; q < W; ++q) {
    const uint32_t j = (uint32_t)q * 64 + lane;
    const ull x = j < d ? cr[q] : 0ull;
    const bool dec = x != 0 && !((f.fr[q] >> lane) & 1ull);
    const uint32_t u = dec ? f.L[j] : 0u;
    ull old = 0;
    if (dec) old = atomicSub(&cnt[slot_of(sc, u)], x);
    const ull now = old - x;
    if (dec && now > L) dm = min(dm, now);
    e_append(dec && old > L && now <= L, u, now, a);
  }
}

__device__ __forceinline__ void e_flush(const EFrame &f, uint32_t Wd, uint32_t d, const ull *cr, ull L, ull *cnt,
                                        Scatter sc, const EAppend &a, ull &dm) {
  switch (Wd) {
    case 1: e_flush_w<1>(f, d, cr, L, cnt, sc, a, dm); break;
    case 2: e_flush_w<2>(f, d, cr, L, cnt, sc, a, dm); break;
    case 4: e_flush_w<4>(f, d, cr, L, cnt, sc, a, dm); break;
    default: e_flush_w<8>(f, d, cr, L, cnt, sc, a, dm); break;
  }
}

// Stages the members of src[0, len) that lie in X_v (and v itself when
// keep_v), with their out-list extents, frontier bits and the id -> index
// table; clears their rows.  Every thread calls it.  At most ED of them are
// kept (len may be larger).  Members keep the order of src, so every
// workgroup that stages the same frame numbers its members alike (the
// pieces of a split frame partition its cliques by first member).
__device__ void e_stage(EFrame &f, const uint32_t *__restrict__ src, uint32_t len, uint32_t v, uint32_t s,
                        const uint32_t *__restrict__ erem, const uint32_t *__restrict__ doff, bool keep_v) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid & 63;
  for (uint32_t t = tid; t < EH; t += nt) f.hent[t] = ~0ull;
  if (tid < 8) {
    f.fr[tid] = 0;
    f.anc[tid] = 0;
  }
  if (tid == 0) {
    f.d = 0;
    f.iv = EMPTY_ID;
  }
  lds_sync();
  const uint32_t wv = tid >> 6, nw = nt >> 6;
  uint32_t run = 0;
  for (uint32_t t0 = 0; t0 < len; t0 += nt) {
    const uint32_t t = t0 + tid;
    const uint32_t u = t < len ? src[t] : EMPTY_ID;
    uint32_t r = 0, ou = 0, ou1 = 0;
    if (u != EMPTY_ID) {
      r = erem[u];
      ou = doff[u];
      ou1 = doff[u + 1];
    }
    const bool isv = keep_v && u == v;
    const bool keep = u != EMPTY_ID && (isv || r > s || (r == s && u > v));
    const ull bal = __ballot(keep);
    if (lane == 0) f.wcnt[wv] = (uint32_t)__popcll(bal);
    lds_sync();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < nw; ++w) {
      const uint32_t cw = f.wcnt[w];
      before += w < wv ? cw : 0u;
      all += cw;
    }
    if (keep) {
      const uint32_t pos = run + before + lanes_below(bal);
      f.L[pos] = u;
      f.sinfo[pos] = make_uint2(ou, ou1 - ou);
      if (!isv && r == s) atomicOr(&f.fr[pos >> 6], 1ull << (pos & 63));
      if (isv) f.iv = pos;
    }
    run += all;
    lds_sync();  // wcnt is rewritten by the next chunk
  }
  if (tid == 0) f.d = run;
  lds_sync();
  const uint32_t d = f.d;
  uint32_t *hk = reinterpret_cast<uint32_t *>(f.hent);
  for (uint32_t t = tid; t < d; t += nt) {
    const uint32_t u = f.L[t];
    uint32_t h = ehslot(u);
    while (atomicCAS(&hk[2 * h], EMPTY_ID, u) != EMPTY_ID) h = (h + 1) & (EH - 1);
    hk[2 * h + 1] = t;
#pragma unroll
    for (int w = 0; w < 8; ++w) f.A[(uint32_t)w * ED + t] = 0ull;
  }
  lds_sync();
}

// Looks w up in the frame table: its index, or EMPTY_ID.
__device__ __forceinline__ uint32_t e_find(const EFrame &f, uint32_t w) {
  uint32_t h = ehslot(w);
  while (true) {
    const ull e = f.hent[h];
    const uint32_t key = (uint32_t)e;
    if (key == w) return (uint32_t)(e >> 32);
    if (key == EMPTY_ID) return EMPTY_ID;
    h = (h + 1) & (EH - 1);
  }
}

// Stages a stored frame (k_ep_frames, k_ep_frames_wide): the rows of the
// subgraph induced by all of N(v) = src[0, len) (len <= FMAX), word-major,
// row i's word w at rows[w * len + i], stored_words(len) words.  The members
// in X_v (at most ED of them: the frontier admits v as narrow only then)
// are kept in src order, as e_stage keeps them, with their frontier bits;
// each kept row is restricted to the kept members and compressed to the new
// numbering, so the frame equals the one e_stage + e_build would build,
// without reading any out-list.
__device__ void e_load(EFrame &f, const uint32_t *__restrict__ src, uint32_t len, uint32_t v, uint32_t s,
                       const uint32_t *__restrict__ erem, const ull *__restrict__ rows) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
  const uint32_t Wf = stored_words(len);
  if (tid < 8) f.fr[tid] = 0;
  // members in X_v, and those in the frontier: one mask word per wave and chunk
  for (uint32_t t0 = 0; t0 < len; t0 += nt) {
    const uint32_t i = t0 + tid;
    const uint32_t u = i < len ? src[i] : EMPTY_ID;
    const uint32_t r = u != EMPTY_ID ? erem[u] : 0u;
    const bool keep = u != EMPTY_ID && (r > s || (r == s && u > v));
    const ull bal = __ballot(keep), fb = __ballot(keep && r == s);
    const uint32_t q = (t0 >> 6) + wv;
    if (lane == 0 && q < Wf) {
      f.smask[q] = bal;
      f.fmask[q] = fb;
    }
  }
  lds_sync();
  if (tid < Wf) compress_masks(f.smask[tid], &f.cmv[tid * 6]);
  if (wv == 0) {
    const uint32_t x = lane < Wf ? (uint32_t)__popcll(f.smask[lane]) : 0u;
    const uint32_t incl = wave_incl_scan_u32(x);
    f.sbase[lane] = incl - x;
    if (lane == 63) f.d = min(incl, ED);
  }
  lds_sync();
  for (uint32_t i = tid; i < len; i += nt) {
    const uint32_t q = i >> 6, bi = i & 63;
    const ull m = f.smask[q];
    if (!((m >> bi) & 1ull)) continue;
    const uint32_t pos = f.sbase[q] + (uint32_t)__popcll(m & ((1ull << bi) - 1ull));
    if (pos >= ED) continue;
    f.L[pos] = src[i];
    if ((f.fmask[q] >> bi) & 1ull) atomicOr(&f.fr[pos >> 6], 1ull << (pos & 63));
    // the kept bits of each stored word, packed; words fill in order
    ull acc = 0;
    uint32_t fill = 0, ow = 0;
    for (uint32_t w = 0; w < Wf; ++w) {
      const ull mw = f.smask[w];
      const uint32_t pcw = (uint32_t)__popcll(mw);
      if (!pcw) continue;
      const ull y = compress_apply(rows[(size_t)w * len + i] & mw, &f.cmv[w * 6]);
      acc |= y << fill;
      const uint32_t nf = fill + pcw;
      if (nf >= 64) {
        if (ow < 8) f.A[ow * ED + pos] = acc;
        ++ow;
        acc = fill ? (y >> (64 - fill)) : 0ull;
        fill = nf - 64;
      } else {
        fill = nf;
      }
    }
    if (fill && ow < 8) f.A[ow * ED + pos] = acc;
    if (fill) ++ow;
    for (; ow < 8; ++ow) f.A[ow * ED + pos] = 0ull;
  }
  lds_sync();
}

// Rows of the staged members: every member's DAG out-list is read (64-element
// chunks, EUNROLL in flight per wave, across list boundaries) and looked up
// in the table; a hit j in member i's list sets bit j of row i and bit i of
// row j (each edge lies in exactly one out-list).
__device__ void e_build(EFrame &f, uint32_t d, const uint32_t *__restrict__ dnbr) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, nw = blockDim.x >> 6;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  auto info = [&](uint32_t i) -> uint2 { return i < d ? f.sinfo[i] : make_uint2(0u, 0u); };
  auto uniform = [](uint2 x) {
    return make_uint2(__builtin_amdgcn_readfirstlane(x.x), __builtin_amdgcn_readfirstlane(x.y));
  };
  uint32_t ci = wv, ct = 0;
  uint2 cur = uniform(info(ci)), nxt = info(ci + nw);
  auto advance = [&]() {
    do {
      ci += nw;
      cur = uniform(nxt);
      nxt = info(ci + nw);
    } while (ci < d && cur.y == 0);
  };
  if (ci < d && cur.y == 0) advance();
  while (ci < d) {
    uint32_t w[EUNROLL], h[EUNROLL], li[EUNROLL];
    ull e[EUNROLL];
#pragma unroll
    for (uint32_t x = 0; x < EUNROLL; ++x) {
      li[x] = ci;
      w[x] = EMPTY_ID;
      if (ci < d) {
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
    for (uint32_t x = 0; x < EUNROLL; ++x) {
      h[x] = ehslot(w[x]);
      e[x] = f.hent[h[x]];
    }
#pragma unroll
    for (uint32_t x = 0; x < EUNROLL; ++x) {
      const bool valid = w[x] != EMPTY_ID;
      uint32_t key = (uint32_t)e[x];
      bool more = valid && key != w[x] && key != EMPTY_ID;
      while (more) {
        h[x] = (h[x] + 1) & (EH - 1);
        e[x] = f.hent[h[x]];
        key = (uint32_t)e[x];
        more = key != w[x] && key != EMPTY_ID;
      }
      if (valid && key == w[x]) {
        const uint32_t i = li[x], j = (uint32_t)(e[x] >> 32);
        atomicOr(&f.A[(j >> 6) * ED + i], 1ull << (j & 63));
        atomicOr(&f.A[(i >> 6) * ED + j], 1ull << (i & 63));
      }
    }
  }
  lds_sync();
}

// Sums the waves' 2 x clique shares of a task; thread 0 gets the count.
__device__ __forceinline__ ull e_task_total(EFrame &f, ull share) {
  if (lane_id() == 0 && share) atomicAdd(&f.rt, share);
  lds_sync();
  const ull E = f.rt >> 1;
  lds_sync();
  if (threadIdx.x == 0) f.rt = 0;
  return E;
}

struct EArgs {
  const uint32_t *foff, *fnbr, *doff, *dnbr;
  const ull *fo, *fstore;  // stored frames (k_ep_frames), or null
  ull *cnt;
  uint32_t *erem;
  GraphPeel *gp;
  Scatter sc;
  ull L;
  uint32_t s;
};

// Narrow v: one frame of N(v) & X_v, its (K-1)-cliques; piece j of P counts
// those whose first member is the j-th of every P in turn (K >= 4 when P > 1).
template <int K>
__device__ void e_narrow(EFrame &f, uint32_t v, uint32_t g, uint32_t j, uint32_t P, const EArgs &x,
                         const EAppend &a, ull &dm) {
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
  const uint32_t b = x.foff[v], len = x.foff[v + 1] - b;
  const bool stored = x.fstore && x.fo[v + 1] > x.fo[v];
  if (stored) e_load(f, x.fnbr + b, len, v, x.s, x.erem, x.fstore + x.fo[v]);
  else e_stage(f, x.fnbr + b, len, v, x.s, x.erem, x.doff, false);
  const uint32_t d = f.d;
  if (d + 1 < (uint32_t)K) {
    lds_sync();
    return;
  }
  if (!stored) e_build(f, d, x.dnbr);
  const uint32_t Wd = words_of(d);
  ull S[8], cr[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    const uint32_t lo = (uint32_t)w * 64;
    S[w] = d >= lo + 64 ? ~0ull : (d > lo ? ((1ull << (d - lo)) - 1ull) : 0ull);
    cr[w] = 0;
  }
  ull share;
  if constexpr (K >= 4) share = e_count<K - 1>(f, Wd, S, cr, wv + nw * j, nw * P);
  else share = e_count<K - 1>(f, Wd, S, cr, wv, nw);
  e_flush(f, Wd, d, cr, x.L, x.cnt, x.sc, a, dm);
  const ull E = e_task_total(f, share);
  if (threadIdx.x == 0 && E) atomicSub(&x.gp[g].live, E);
  lds_sync();
}

// Wide v: candidates c = N(v)[lo, hi) that lie in X_v, each with the frame
// N+(c) & X_v and the anchor set of members adjacent to v.
template <int K>
__device__ void e_wide(EFrame &f, uint32_t v, uint32_t g, uint32_t lo, uint32_t hi, const EArgs &x,
                       const EAppend &a, ull &dm) {
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = blockDim.x >> 6;
  const uint32_t nc = hi - lo, b = x.foff[v];
  if (tid < nc) {
    const uint32_t c = x.fnbr[b + lo + tid];
    f.cand[tid] = c;
    f.crem[tid] = x.erem[c];
  }
  lds_sync();
  ull Etot = 0;
  for (uint32_t j = 0; j < nc; ++j) {
    const uint32_t c = f.cand[j], rc = f.crem[j];
    if (!(rc > x.s || (rc == x.s && c > v))) continue;
    const uint32_t cb = x.doff[c], clen = x.doff[c + 1] - cb;
    if (clen + 2 < (uint32_t)K) continue;  // the anchor holds at most clen members
    e_stage(f, x.dnbr + cb, clen, v, x.s, x.erem, x.doff, true);
    const uint32_t d = f.d, iv = f.iv;
    if (iv != EMPTY_ID) {
      e_build(f, d, x.dnbr);
      if (tid < 8) f.anc[tid] = erow(f, iv, (int)tid) & ~((tid == (iv >> 6)) ? (1ull << (iv & 63)) : 0ull);
      lds_sync();
    } else {
      // v < c: members adjacent to v are those in N+(v)
      const uint32_t vb = x.doff[v], vlen = x.doff[v + 1] - vb;
      for (uint32_t t = tid; t < vlen; t += blockDim.x) {
        const uint32_t i = e_find(f, x.dnbr[vb + t]);
        if (i != EMPTY_ID) atomicOr(&f.anc[i >> 6], 1ull << (i & 63));
      }
      lds_sync();
      uint32_t pc = 0;
#pragma unroll
      for (int w = 0; w < 8; ++w) pc += (uint32_t)__popcll(f.anc[w]);
      if (pc + 2 < (uint32_t)K) {
        lds_sync();
        continue;
      }
      if (K > 3) e_build(f, d, x.dnbr);  // K = 3 counts anchor members alone
    }
    const uint32_t Wd = words_of(d);
    ull S[8], cr[8];
    uint32_t pc = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      S[w] = (uint32_t)w < Wd ? f.anc[w] : 0ull;
      pc += (uint32_t)__popcll(S[w]);
      cr[w] = 0;
    }
    ull share = 0;
    if (pc + 2 >= (uint32_t)K) {
      share = e_count<K - 2>(f, Wd, S, cr, wv, nw);
      e_flush(f, Wd, d, cr, x.L, x.cnt, x.sc, a, dm);
    }
    const ull E = e_task_total(f, share);
    Etot += E;
    // c lies in every one of these cliques
    if (tid < 64) {
      const bool dec = lane == 0 && E && rc > x.s;
      ull old = 0;
      if (dec) old = atomicSub(&x.cnt[slot_of(x.sc, c)], E);
      const ull now = old - E;
      if (dec && now > x.L) dm = min(dm, now);
      e_append(dec && old > x.L && now <= x.L, c, now, a);
    }
    lds_sync();
  }
  if (tid == 0 && Etot) atomicSub(&x.gp[g].live, Etot);
}

// k = 2: the cliques of v are its edges into X_v.
__device__ void e_edges_item(EFrame &f, uint32_t v, uint32_t g, uint32_t lo, uint32_t hi, const EArgs &x,
                             const EAppend &a, ull &dm) {
  const uint32_t b = x.foff[v];
  ull mine = 0;
  for (uint32_t t0 = lo; t0 < hi; t0 += blockDim.x) {
    const uint32_t t = t0 + threadIdx.x;
    const uint32_t u = t < hi ? x.fnbr[b + t] : 0u;
    const uint32_t r = t < hi ? x.erem[u] : 0u;
    const bool in = t < hi && (r > x.s || (r == x.s && u > v));
    const bool dec = in && r > x.s;
    mine += in;
    ull old = 0;
    if (dec) old = atomicSub(&x.cnt[slot_of(x.sc, u)], 1ull);
    const ull now = old - 1;
    if (dec && now > x.L) dm = min(dm, now);
    e_append(dec && old > x.L && now <= x.L, u, now, a);
  }
  const ull E = e_task_total(f, 2 * wave_sum_u64(mine));
  if (threadIdx.x == 0 && E) atomicSub(&x.gp[g].live, E);
  lds_sync();
}

__device__ __forceinline__ bool ratio_ge(ull a, ull b, ull c, ull d) {
  ull hi1 = __umul64hi(a, d), lo1 = a * d;
  ull hi2 = __umul64hi(c, b), lo2 = c * b;
  return hi1 > hi2 || (hi1 == hi2 && lo1 >= lo2);
}

// Lanes with want grouped by key (first eight groups; later lanes alone):
// the group size on the leader lane, 0 elsewhere.
__device__ __forceinline__ uint32_t e_group(bool want, uint32_t key) {
  const uint32_t lane = lane_id();
  uint32_t size = 0;
  ull pending = __ballot(want);
#pragma unroll 1
  for (int round = 0; round < 8 && pending; ++round) {
    const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
    const uint32_t kl = __shfl(key, lead, 64);
    const ull same = __ballot(want && key == kl) & pending;
    if (lane == lead) size = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  if ((pending >> lane) & 1ull) size = 1;
  return size;
}

// The last workgroup of a launch (two-level ticket) returns true.
__device__ bool ep_last_block(EpCtl *ctl, uint32_t *s_last) {
  lds_sync();
  if (threadIdx.x == 0) {
    wave_drain();
    const uint32_t seg = blockIdx.x % ESEG;
    const uint32_t in_seg = (gridDim.x - seg + ESEG - 1) / ESEG;
    uint32_t last = 0;
    if (atomicAdd(&ctl->gblocks[seg][0], 1u) == in_seg - 1)
      last = atomicAdd(&ctl->blocks[0], 1u) == min(gridDim.x, ESEG) - 1;
    *s_last = last;
  }
  lds_sync();
  return *s_last != 0;
}

// One step of the enumerating peel.
//   OPEN    records the density of every graph the completed level touched,
//           then collects every live vertex with count <= target (and its
//           work items) and the minimum count of the others; the control
//           step starts level = target, retries at that minimum, or ends.
//   PROCESS marks the frontier (core = L), then the workgroups take its work
//           items from a ticket counter and enumerate / decrement.
template <int K>
__global__ __launch_bounds__(ETHREADS) void k_ep_step(EpCtl *ctl, uint32_t n, const uint32_t *__restrict__ foff,
                                                     const uint32_t *__restrict__ fnbr,
                                                     const uint32_t *__restrict__ doff,
                                                     const uint32_t *__restrict__ dnbr, ull *cnt, long long *core,
                                                     uint32_t *erem, uint32_t *fv0, uint32_t *fv1, uint4 *fi0,
                                                     uint4 *fi1, uint32_t seg_v, uint32_t seg_i, GraphPeel *gp,
                                                     const uint32_t *__restrict__ gid, uint32_t *tlist, Scatter sc,
                                                     uint8_t *alive, const ull *__restrict__ fo,
                                                     const ull *__restrict__ fstore) {
  __shared__ EFrame f;
  __shared__ ull s_hdr[2];
  __shared__ uint32_t s_hdr2[4], s_f[ESEG + 1], s_i[ESEG + 1], s_last, s_sz[2][2][ESEG];
  __shared__ ull s_min;
  const uint32_t tid = threadIdx.x, lane = lane_id(), seg = blockIdx.x % ESEG;
  if (tid < 2) s_hdr[tid] = (&ctl->level)[tid];
  if (tid < 4) s_hdr2[tid] = (&ctl->cur)[tid];
  if (tid >= 64 && tid < 64 + 4 * ESEG) {
    const uint32_t i = tid - 64, fb = i / (2 * ESEG), kind = (i / ESEG) & 1, gs = i % ESEG;
    s_sz[fb][kind][gs] = kind ? ctl->isize[fb][gs][0] : ctl->fsize[fb][gs][0];
  }
  if (tid == 0) f.rt = 0;
  lds_sync();
  const ull L = s_hdr[0], target = s_hdr[1];
  const uint32_t cur = s_hdr2[0], done = s_hdr2[1], phase = s_hdr2[2], s = s_hdr2[3];
  if (done) return;
  uint32_t *fv = cur ? fv1 : fv0, *nv = cur ? fv0 : fv1;
  uint4 *fi = cur ? fi1 : fi0, *ni = cur ? fi0 : fi1;
  const uint32_t gtid = blockIdx.x * blockDim.x + tid, gstride = gridDim.x * blockDim.x;
  if (phase == 0) {
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
    const EAppend a{foff, fnbr, fv + (size_t)seg * seg_v, &ctl->fsize[cur][seg][0], fi + (size_t)seg * seg_i,
                    &ctl->isize[cur][seg][0], erem, s + 1, K >= 4};
    ull m = ~0ull;
    const uint32_t Kc = sc.mask + 1;
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < Kc; p0 += gstride) {
      const uint32_t p = p0 + tid, v = vertex_of(sc, p);
      const bool live = p < Kc && alive[p];  // read in slot order: no gather of core[v]
      const ull cv = live ? cnt[p] : ~0ull;
      e_append(live && cv <= target, v, cv, a);
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
      s_i[0] = 0;
      for (uint32_t gs = 0; gs < ESEG; ++gs) {
        s_f[gs + 1] = s_f[gs] + s_sz[cur][0][gs];
        s_i[gs + 1] = s_i[gs] + s_sz[cur][1][gs];
      }
    }
    lds_sync();
    const uint32_t fs = s_f[ESEG], is = s_i[ESEG];
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
        alive[slot_of(sc, v)] = 0;
        g = gid[v];
      }
      const uint32_t gsize = e_group(valid, g);
      if (gsize) {
        atomicSub(&gp[g].rem, gsize);
        if (atomicExch(&gp[g].touch, 1u) == 0u) tlist[atomicAdd(&ctl->ntouch[0], 1u)] = g;
      }
    }
    const EArgs x{foff, fnbr, doff, dnbr, fo, fstore, cnt, erem, gp, sc, L, s};
    const EAppend a{foff, fnbr, nv + (size_t)seg * seg_v, &ctl->fsize[cur ^ 1][seg][0], ni + (size_t)seg * seg_i,
                    &ctl->isize[cur ^ 1][seg][0], erem, s + 1, K >= 4};
    ull dm = ~0ull;
    // item b goes to workgroup b; the items past the grid are taken from the
    // ticket counter by the workgroups that finish first (so a step with
    // fewer items than workgroups touches no shared counter)
    bool first = true;
    while (true) {
      if (tid == 0) {
        const uint32_t it = first ? blockIdx.x : gridDim.x + atomicAdd(&ctl->ticket[0], 1u);
        f.item_idx = it;
        if (it < is) {
          uint32_t sg = 0;
          while (it >= s_i[sg + 1]) ++sg;
          const uint4 item = fi[(size_t)sg * seg_i + (it - s_i[sg])];
          f.item = item;
          f.skip = ldc_u64(&cnt[slot_of(sc, item.x)]) == 0;
        }
      }
      lds_sync();
      const uint32_t it = f.item_idx;
      const uint4 item = f.item;
      const uint32_t skip = f.skip;
      lds_sync();  // every thread has read the item before it is replaced
      first = false;
      // no own item (blockIdx.x >= is: then none past the grid either), or
      // the counter ran past the last item
      if (it >= is) break;
      if (skip) continue;
      const uint32_t v = item.x, g = gid[v];
      if constexpr (K == 2) {
        const uint32_t deg = foff[v + 1] - foff[v];
        e_edges_item(f, v, g, item.w ? item.y : 0u, item.w ? item.z : deg, x, a, dm);
      } else {
        if (item.w == 0) e_narrow<K>(f, v, g, 0, 1, x, a, dm);
        else if (item.w == 2) e_narrow<K>(f, v, g, item.y, item.z, x, a, dm);
        else e_wide<K>(f, v, g, item.y, item.z, x, a, dm);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dm = min(dm, (ull)__shfl_xor(dm, o, 64));
    if (tid == 0) s_min = ~0ull;
    lds_sync();
    if (lane == 0 && dm != ~0ull) atomicMin(&s_min, dm);
    lds_sync();
    if (tid == 0 && s_min != ~0ull) atomicMin(&ctl->dmin[seg][0], s_min);
  }
  // control step: wave 0 of the last workgroup
  if (ep_last_block(ctl, &s_last) && tid < KCDS_WAVE) {
    uint32_t fc = 0, fn = 0;
    ull mn = ~0ull, dn = ~0ull;
    if (lane < ESEG) {
      fc = ldc_u32(&ctl->fsize[cur][lane][0]);
      fn = ldc_u32(&ctl->fsize[cur ^ 1][lane][0]);
      if (phase == 1) dn = ldc_u64(&ctl->dmin[lane][0]);
      mn = atomicExch(&ctl->gmin[lane][0], ~0ull);
      ctl->gblocks[lane][0] = 0;
      if (phase == 1) {
        ctl->fsize[cur][lane][0] = 0;
        ctl->isize[cur][lane][0] = 0;
      }
    }
    const uint32_t fcur = wave_sum_u32(fc), fnext = wave_sum_u32(fn);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, (ull)__shfl_xor(mn, o, 64));
    if (!(phase == 1 && fnext == 0)) dn = ~0ull;
    else if (lane < ESEG) stc_u64(&ctl->dmin[lane][0], ~0ull);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dn = min(dn, (ull)__shfl_xor(dn, o, 64));
    if (lane == 0) {
      ctl->blocks[0] = 0;
      ctl->ticket[0] = 0;
      ++ctl->steps;
      if (phase == 1) {
        ctl->cur = cur ^ 1;
        if (fnext == 0) {
          const ull nm = min(ctl->nextmin[0], dn);
          ctl->phase = 0;
          ctl->target = nm == ~0ull ? L + 1 : max(nm, L + 1);
          // a single graph small enough for the dense mode: hand over
          if (ctl->handoff && ldc_u32(&gp[0].rem) <= ctl->handoff) ctl->done = 2;
        }
      } else {
        ctl->ntouch[0] = 0;
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

// Words of v's stored frame: stored_words(d) * d for a vertex of degree
// d <= maxd that lies in a clique (its count is positive), else none; above
// degree ED only for a vertex whose frame may be split into pieces (count
// from 2^ECHUNK_LOG up), each of which would build it again.
__global__ void k_ep_fsize(uint32_t n, const uint32_t *foff, const ull *cnt0, uint32_t maxd, ull *words) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t d = foff[v + 1] - foff[v];
    const ull c = cnt0[v];
    words[v] = d <= maxd && c && (d <= ED || (c >> ECHUNK_LOG) > 0) ? (ull)stored_words(d) * d : 0ull;
  }
}

// The stored frames, one workgroup per vertex at a time: N(v) staged whole
// (every vertex is still live) and its rows built from the out-lists, then
// written word-major to fstore + fo[v].  Runs after k_ep_init.
__global__ __launch_bounds__(ETHREADS) void k_ep_frames(uint32_t n, const uint32_t *__restrict__ foff,
                                                       const uint32_t *__restrict__ fnbr,
                                                       const uint32_t *__restrict__ doff,
                                                       const uint32_t *__restrict__ dnbr,
                                                       const uint32_t *__restrict__ erem, const ull *__restrict__ fo,
                                                       ull *fstore, uint32_t *ticket) {
  __shared__ EFrame f;
  __shared__ uint32_t s_v;
  while (true) {
    if (threadIdx.x == 0) s_v = atomicAdd(ticket, 1u);
    lds_sync();
    const uint32_t v = s_v;
    lds_sync();  // every thread has read s_v
    if (v >= n) break;
    const uint32_t b = foff[v], d = foff[v + 1] - b;
    if (fo[v + 1] == fo[v] || d > ED) continue;
    e_stage(f, fnbr + b, d, v, 0u, erem, doff, false);
    e_build(f, d, dnbr);
    const uint32_t W = words_of(d);
    ull *dst = fstore + fo[v];
    for (uint32_t t = threadIdx.x; t < W * d; t += blockDim.x) dst[t] = f.A[(t / d) * ED + t % d];
    lds_sync();
  }
}

__device__ __forceinline__ uint32_t fwslot(uint32_t key) {
  return (uint32_t)__umul24(key ^ (key >> 24), 0x9E3779u) >> (32 - FW_HB);
}

// The stored frames of vertices of degree ED+1..FMAX, one workgroup per
// vertex at a time: N(v) in a table (id -> index), then FW_ROWS rows at a
// time, row i = the members adjacent to member i, found by looking up each
// entry of member i's full adjacency list (so no row needs another's
// out-list), written word-major to fstore + fo[v].
__global__ __launch_bounds__(ETHREADS) void k_ep_frames_wide(uint32_t n, const uint32_t *__restrict__ foff,
                                                            const uint32_t *__restrict__ fnbr,
                                                            const ull *__restrict__ fo, ull *fstore,
                                                            uint32_t *ticket) {
  __shared__ ull hent[FW_H];
  __shared__ ull blk[FW_ROWS * FWORDS];
  __shared__ uint32_t s_v;
  const uint32_t tid = threadIdx.x, lane = tid & 63, nt = blockDim.x;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nt >> 6;
  uint32_t *hk = reinterpret_cast<uint32_t *>(hent);
  while (true) {
    if (tid == 0) s_v = atomicAdd(ticket, 1u);
    lds_sync();
    const uint32_t v = s_v;
    lds_sync();  // every thread has read s_v
    if (v >= n) break;
    const uint32_t b = foff[v], d = foff[v + 1] - b;
    if (fo[v + 1] == fo[v] || d <= ED || d > FMAX) continue;
    for (uint32_t t = tid; t < FW_H; t += nt) hent[t] = ~0ull;
    lds_sync();
    for (uint32_t t = tid; t < d; t += nt) {
      const uint32_t u = fnbr[b + t];
      uint32_t h = fwslot(u);
      while (atomicCAS(&hk[2 * h], EMPTY_ID, u) != EMPTY_ID) h = (h + 1) & (FW_H - 1);
      hk[2 * h + 1] = t;
    }
    lds_sync();
    const uint32_t Wf = stored_words(d);
    ull *dst = fstore + fo[v];
    for (uint32_t r0 = 0; r0 < d; r0 += FW_ROWS) {
      const uint32_t nr = min(FW_ROWS, d - r0);
      for (uint32_t t = tid; t < nr * Wf; t += nt) blk[t] = 0ull;
      lds_sync();
      for (uint32_t row = wv; row < nr; row += nw) {
        const uint32_t u = fnbr[b + r0 + row], ub = foff[u], ue = foff[u + 1];
        for (uint32_t e = ub + lane; e < ue; e += KCDS_WAVE) {
          const uint32_t w = fnbr[e];
          uint32_t h = fwslot(w);
          while (true) {
            const ull he = hent[h];
            const uint32_t key = (uint32_t)he;
            if (key == w) {
              const uint32_t j = (uint32_t)(he >> 32);
              atomicOr(&blk[row * Wf + (j >> 6)], 1ull << (j & 63));
              break;
            }
            if (key == EMPTY_ID) break;
            h = (h + 1) & (FW_H - 1);
          }
        }
      }
      lds_sync();
      for (uint32_t t = tid; t < nr * Wf; t += nt) {
        const uint32_t w = t / nr, row = t % nr;
        dst[(size_t)w * d + r0 + row] = blk[row * Wf + w];
      }
      lds_sync();
    }
  }
}

// working counts (scattered layout), core = -1, every vertex live (alive is
// the slot-ordered live flag that OPEN scans; slots past n stay 0)
__global__ void k_ep_init(uint32_t n, const ull *cnt0, ull *cnt, long long *core, uint32_t *erem, Scatter sc,
                          uint8_t *alive) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = slot_of(sc, i);
    cnt[p] = cnt0[i];
    alive[p] = 1;
    core[i] = -1;
    erem[i] = LIVE;
  }
}

// full adjacency over new ids: in-degree counts
__global__ void k_fadj_indeg(const uint32_t *dnbr, uint64_t arcs, uint32_t *indeg) {
  for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < arcs; a += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&indeg[dnbr[a]], 1u);
}

__global__ void k_fadj_deg(uint32_t n, const uint32_t *doff, const uint32_t *indeg, uint32_t *deg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    deg[v] = indeg[v] + (doff[v + 1] - doff[v]);
}

// out-lists after the in-lists; in-list slots taken by atomics (the frame
// does not need the members in order); fill[] counts the in-list entries
__global__ void k_fadj_fill(uint32_t n, const uint32_t *doff, const uint32_t *dnbr, const uint32_t *foff,
                            const uint32_t *indeg, uint32_t *fill, uint32_t *fnbr) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwv = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t r = gw; r < n; r += nwv) {
    const uint32_t ob = foff[r] + indeg[r];
    for (uint32_t a = doff[r] + lane; a < doff[r + 1]; a += KCDS_WAVE) {
      const uint32_t u = dnbr[a];
      fnbr[ob + (a - doff[r])] = u;
      fnbr[foff[u] + atomicAdd(&fill[u], 1u)] = r;
    }
  }
}

// The k'-core's DAG: arcs between vertices whose core reaches their graph's
// k' (both ends lie in one graph), in the same ascending order.
__device__ __forceinline__ bool in_core(uint32_t v, const long long *core, const uint32_t *gid, const GraphPeel *gp) {
  return core[v] >= (long long)gp[gid[v]].k_prime;
}

__global__ void k_core_dag(uint32_t n, const uint32_t *doff, const uint32_t *dnbr, const long long *core,
                           const uint32_t *gid, const GraphPeel *gp, const uint32_t *koff, uint32_t *kdeg,
                           uint32_t *knbr) {
  const uint32_t lane = lane_id();
  const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwv = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t r = gw; r < n; r += nwv) {
    const bool rin = in_core(r, core, gid, gp);
    uint32_t at = 0;
    for (uint32_t a0 = doff[r]; a0 < doff[r + 1]; a0 += KCDS_WAVE) {
      const uint32_t a = a0 + lane;
      const bool keep = rin && a < doff[r + 1] && in_core(dnbr[a], core, gid, gp);
      const ull bal = __ballot(keep);
      if (knbr && keep) knbr[koff[r] + at + lanes_below(bal)] = dnbr[a];
      at += (uint32_t)__popcll(bal);
    }
    if (!knbr && lane == 0) kdeg[r] = at;
  }
}

__global__ void k_fill_i32(int32_t *p, int32_t v, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

typedef void (*EpStepFn)(EpCtl *, uint32_t, const uint32_t *, const uint32_t *, const uint32_t *, const uint32_t *,
                         ull *, long long *, uint32_t *, uint32_t *, uint32_t *, uint4 *, uint4 *, uint32_t, uint32_t,
                         GraphPeel *, const uint32_t *, uint32_t *, Scatter, uint8_t *, const ull *, const ull *);

EpStepFn ep_step_fn(uint32_t k) {
  switch (k) {
    case 2: return k_ep_step<2>;
    case 3: return k_ep_step<3>;
    case 4: return k_ep_step<4>;
    case 5: return k_ep_step<5>;
    case 6: return k_ep_step<6>;
    case 7: return k_ep_step<7>;
    case 8: return k_ep_step<8>;
  }
  return nullptr;
}

uint32_t egrid(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 16));
}

}  // namespace

// foff/fnbr: for new id v, its in-neighbours (any order) then N+(v).
void build_full_adjacency(Ctx &c) {
  const uint32_t n = c.n;
  const uint64_t arcs = c.darcs;
  buf_reserve(c, c.foff, sizeof(uint32_t) * ((size_t)n + 2));
  buf_reserve(c, c.fnbr, sizeof(uint32_t) * (2 * arcs + 1));
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * (3 * (size_t)n + 4));
  uint32_t *indeg = c.tmp1.as<uint32_t>(), *deg = indeg + n + 1, *fill = deg + n + 1;
  HIP_CHECK(hipMemsetAsync(indeg, 0, sizeof(uint32_t) * (3 * (size_t)n + 4), c.stream));
  if (arcs) {
    k_fadj_indeg<<<egrid(arcs), 256, 0, c.stream>>>(c.dnbr.as<uint32_t>(), arcs, indeg);
    KCDS_LAUNCH_CHECK();
  }
  if (n) {
    k_fadj_deg<<<egrid(n), 256, 0, c.stream>>>(n, c.doff.as<uint32_t>(), indeg, deg);
    KCDS_LAUNCH_CHECK();
  }
  scan_u32_to_u32(c, deg, c.foff.as<uint32_t>(), n);
  if (n) {
    k_fadj_fill<<<egrid((uint64_t)n * 64), 256, 0, c.stream>>>(n, c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(),
                                                                c.foff.as<uint32_t>(), indeg, fill,
                                                                c.fnbr.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
}

void epeel_run(Ctx &c) {
  if (!c.counted) fail(KCDS_E_STATE, "peel before the count pass");
  const uint32_t n = c.n, k = c.k, G = c.ngraphs;
  if (c.dmax > ED) fail(KCDS_E_LIMIT, "DAG out-degree %u above %u (the enumerating peel's frame)", c.dmax, ED);
  graph_ids(c);
  build_full_adjacency(c);
  buf_reserve(c, c.core, sizeof(long long) * ((size_t)n + 1));
  buf_reserve(c, c.erem, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.gpeel, sizeof(GraphPeel) * ((size_t)G + 1));
  buf_reserve(c, c.tlist, sizeof(uint32_t) * ((size_t)G + 1));
  GraphPeel *gp = c.gpeel.as<GraphPeel>();
  const uint32_t seg_v = ((n + 63) / 64) * 64;
  const uint64_t seg_i64 = (((uint64_t)n + (2 * c.darcs) / EPIECE + 64 + 63) / 64) * 64;
  if (seg_i64 * ESEG * 2 >= (1ull << 32)) fail(KCDS_E_LIMIT, "graph too large for the peeling frontiers");
  const uint32_t seg_i = (uint32_t)seg_i64;
  const Scatter sc = make_scatter(n);
  const size_t kslots = ((size_t)sc.mask + 1 + 63) / 64 * 64;
  const size_t bytes = sizeof(ull) * kslots + sizeof(uint32_t) * 2 * ESEG * (size_t)seg_v +
                       sizeof(uint4) * 2 * ESEG * (size_t)seg_i + sizeof(EpCtl) + kslots + 1024;
  buf_reserve(c, c.tmp0, bytes);
  ull *cnt = c.tmp0.as<ull>();
  uint32_t *fv0 = (uint32_t *)(cnt + kslots), *fv1 = fv0 + (size_t)ESEG * seg_v;
  uint4 *fi0 = (uint4 *)(((uintptr_t)(fv1 + (size_t)ESEG * seg_v) + 255) & ~(uintptr_t)255);
  uint4 *fi1 = fi0 + (size_t)ESEG * seg_i;
  EpCtl *ctl = (EpCtl *)(((uintptr_t)(fi1 + (size_t)ESEG * seg_i) + 255) & ~(uintptr_t)255);
  uint8_t *alive = (uint8_t *)(ctl + 1);  // kslots bytes, live flag per counter slot
  EpCtl *h0 = (EpCtl *)c.host_scratch(sizeof(EpCtl));
  memset(h0, 0, sizeof(EpCtl));
  for (uint32_t g = 0; g < ESEG; ++g) {
    h0->gmin[g][0] = ~0ull;
    h0->dmin[g][0] = ~0ull;
  }
  h0->nextmin[0] = ~0ull;
  // with KCDS_TUNE_DENSE_PEEL a single graph goes over to the dense mode
  // (dense_peel_finish) once at most 1,024 vertices are live, at a level
  // boundary; from the start when it is that small
  const bool dense_mode = G == 1 && k >= 3 && (c.tune_flags & KCDS_TUNE_DENSE_PEEL);
  h0->handoff = dense_mode ? 1024u : 0u;
  const bool dense_now = dense_mode && n <= 1024u;
  h0->done = n == 0 || dense_now;
  HIP_CHECK(hipMemcpyAsync(ctl, h0, sizeof(EpCtl), hipMemcpyHostToDevice, c.stream));
  c.sync();  // h0 is reused below
  if (n) {
    HIP_CHECK(hipMemsetAsync(alive, 0, kslots, c.stream));
    k_ep_init<<<egrid(n), 256, 0, c.stream>>>(n, c.cnt.as<ull>(), cnt, c.core.as<long long>(), c.erem.as<uint32_t>(),
                                               sc, alive);
    KCDS_LAUNCH_CHECK();
  }
  peel_graphs_begin(c, gp);
  EpStepFn step = ep_step_fn(k);
  if (!step) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
  // one workgroup of ETHREADS threads per CU (a 55 KB frame; its 150-250
  // VGPRs per thread leave room for one 512-thread workgroup per CU)
  const uint32_t sgrid = (uint32_t)c.num_cus * (ETHREADS >= 512 ? 1u : 2u);
  // stored frames: each narrow vertex's rows built once, all vertices in
  // parallel, so that a step only restricts and renumbers them
  const ull *fo = nullptr, *fstore = nullptr;
  if (n && k >= 3 && !(c.tune_flags & KCDS_TUNE_BUILD_FRAMES)) {
    buf_reserve(c, c.fofs, sizeof(ull) * ((size_t)n + 2));
    buf_reserve(c, c.tmp1, sizeof(ull) * ((size_t)n + 64));
    ull *words = c.tmp1.as<ull>();
    uint32_t *ticket = (uint32_t *)(words + n + 32);
    // vertices up to degree FMAX; up to ED alone when those do not fit
    ull total = 0;
    size_t fbytes = 0;
    for (uint32_t maxd : {FMAX, ED}) {
      k_ep_fsize<<<egrid(n), 256, 0, c.stream>>>(n, c.foff.as<uint32_t>(), c.cnt.as<ull>(), maxd, words);
      KCDS_LAUNCH_CHECK();
      scan_u64_to_u64(c, words, c.fofs.as<ull>(), n);
      HIP_CHECK(hipMemcpyAsync(&total, c.fofs.as<ull>() + n, sizeof(ull), hipMemcpyDeviceToHost, c.stream));
      c.sync();
      fbytes = (size_t)total * sizeof(ull);
      // a store kept from an earlier solve is reused where it is large enough
      const size_t grow = fbytes > c.fstore.bytes ? fbytes - c.fstore.bytes : 0;
      if (c.used + grow + ((size_t)64 << 20) <= c.budget) break;
      total = 0;
    }
    if (total) {
      buf_reserve(c, c.fstore, fbytes);
      HIP_CHECK(hipMemsetAsync(ticket, 0, sizeof(uint32_t), c.stream));
      k_ep_frames<<<sgrid, ETHREADS, 0, c.stream>>>(n, c.foff.as<uint32_t>(), c.fnbr.as<uint32_t>(),
                                                    c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(),
                                                    c.erem.as<uint32_t>(), c.fofs.as<ull>(), c.fstore.as<ull>(),
                                                    ticket);
      KCDS_LAUNCH_CHECK();
      if (c.maxdeg > ED) {
        HIP_CHECK(hipMemsetAsync(ticket + 1, 0, sizeof(uint32_t), c.stream));
        k_ep_frames_wide<<<sgrid, ETHREADS, 0, c.stream>>>(n, c.foff.as<uint32_t>(), c.fnbr.as<uint32_t>(),
                                                           c.fofs.as<ull>(), c.fstore.as<ull>(), ticket + 1);
        KCDS_LAUNCH_CHECK();
      }
      fo = c.fofs.as<ull>();
      fstore = c.fstore.as<ull>();
    }
  }
  uint32_t batch = c.peel_steps ? (uint32_t)std::min<uint64_t>(c.peel_steps, 1u << 20) : 32;
  uint64_t rounds = 0;
  const bool log = getenv("KCDS_LOG") != nullptr;  // progress lines on stderr
  auto t_start = std::chrono::steady_clock::now(), t_log = t_start;
  while (n && !dense_now) {
    for (uint32_t r = 0; r < batch; ++r) {
      hipLaunchKernelGGL(step, dim3(sgrid), dim3(ETHREADS), 0, c.stream, ctl, n, c.foff.as<uint32_t>(),
                         c.fnbr.as<uint32_t>(), c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(), cnt,
                         c.core.as<long long>(), c.erem.as<uint32_t>(), fv0, fv1, fi0, fi1, seg_v, seg_i, gp,
                         c.gid.as<uint32_t>(), c.tlist.as<uint32_t>(), sc, alive, fo, fstore);
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
    const auto now = std::chrono::steady_clock::now();
    if (log && std::chrono::duration<double>(now - t_log).count() > 5.0) {
      t_log = now;
      fprintf(stderr, "kcds epeel: n=%u k=%u steps=%u levels=%u level=%llu %.1f s\n", n, k, h0->steps, h0->levels,
              (unsigned long long)h0->level, std::chrono::duration<double>(now - t_start).count());
    }
  }
  c.peel_steps = h0->steps;
  uint32_t levels = h0->levels;
  // the store is kept for the next solve (freeing waits for the whole
  // device, other streams included) unless the context is past half its HBM
  // budget, when the later stages get it back
  if (fstore && c.used > c.budget / 2) {
    buf_free(c, c.fstore);
    buf_free(c, c.fofs);
  }
  if (log)
    fprintf(stderr, "kcds epeel: step kernels ended after %u steps, %u levels, %.1f ms\n", h0->steps, h0->levels,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count());
  if (n && (dense_now || h0->done == 2)) {
    const long long dl = dense_peel_finish(c, cnt, sc, gp, dense_now ? -1 : (long long)h0->level);
    if (dl < 0) fail(KCDS_E_STATE, "dense peel mode: more than 1,024 live vertices");
    levels += (uint32_t)dl;
  }
  peel_graphs_end(c, gp, levels);
  c.enum_peel = true;
  c.core_listed = false;
  c.core_counted = false;
}


// After an enumerating peel: counts the cliques of every graph's k'-core (the
// core cliques, status >= k' in the reference's terms) on the core's DAG and
// records each graph's count; when they fit in HBM, materialises them with
// their incidence and status = INT32_MAX, so that the components and flow
// stages run unchanged.  c.T becomes the number of core cliques.
// The DAG restricted to every graph's k'-core (c.kdoff, c.kdnbr); returns
// its arcs.
uint64_t build_core_dag(Ctx &c) {
  const uint32_t n = c.n;
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * ((size_t)n + 2));
  buf_reserve(c, c.kdoff, sizeof(uint32_t) * ((size_t)n + 2));
  uint32_t *kdeg = c.tmp1.as<uint32_t>();
  const GraphPeel *gp = c.gpeel.as<GraphPeel>();
  if (n) {
    k_core_dag<<<egrid((uint64_t)n * 64), 256, 0, c.stream>>>(n, c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(),
                                                               c.core.as<long long>(), c.gid.as<uint32_t>(), gp,
                                                               nullptr, kdeg, nullptr);
    KCDS_LAUNCH_CHECK();
  }
  scan_u32_to_u32(c, kdeg, c.kdoff.as<uint32_t>(), n);
  uint32_t *h = (uint32_t *)c.host_scratch(16);
  HIP_CHECK(hipMemcpyAsync(h, c.kdoff.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  const uint64_t karcs = h[0];
  buf_reserve(c, c.kdnbr, sizeof(uint32_t) * (karcs + 1));
  if (n && karcs) {
    k_core_dag<<<egrid((uint64_t)n * 64), 256, 0, c.stream>>>(n, c.doff.as<uint32_t>(), c.dnbr.as<uint32_t>(),
                                                               c.core.as<long long>(), c.gid.as<uint32_t>(), gp,
                                                               c.kdoff.as<uint32_t>(), nullptr,
                                                               c.kdnbr.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  return karcs;
}

void core_materialize(Ctx &c, bool materialize) {
  if (!c.peeled || !c.enum_peel) fail(KCDS_E_STATE, "core cliques before the enumerating peel");
  const uint32_t k = c.k, G = c.ngraphs;
  const uint64_t karcs = build_core_dag(c);
  // list on the core DAG (the full DAG is swapped back afterwards)
  const kcds_peel_stats keep_stats = c.peel;
  const uint64_t keep_w2 = c.w2, keep_arcs = c.darcs;
  std::swap(c.doff, c.kdoff);
  std::swap(c.dnbr, c.kdnbr);
  c.darcs = karcs;
  struct Back {
    Ctx &c;
    uint64_t w2, arcs;
    ~Back() {
      std::swap(c.doff, c.kdoff);
      std::swap(c.dnbr, c.kdnbr);
      c.w2 = w2;
      c.darcs = arcs;
    }
  } back{c, keep_w2, keep_arcs};
  list_count(c, k);
  std::vector<ull> gT;
  graph_counts(c, gT);
  const uint64_t Tc = c.T;
  // the core cliques, their incidence and status: Tc * (8k + 4) bytes
  const size_t avail = c.budget > c.used ? c.budget - c.used : 0;
  const bool fits = materialize && Tc < 0xFFFFFFFFull && (double)Tc * (8.0 * k + 4.0) <= 0.85 * (double)avail;
  if (Tc && fits) list_materialize(c, k);
  if (fits) {
    buf_reserve(c, c.status, sizeof(int32_t) * (Tc + 1));
    if (Tc) {
      k_fill_i32<<<egrid(Tc), 256, 0, c.stream>>>(c.status.as<int32_t>(), 0x7FFFFFFF, Tc);
      KCDS_LAUNCH_CHECK();
    }
  }
  for (uint32_t g = 0; g < G; ++g) c.h_gpeel[g].core_cliques = gT[g];
  HIP_CHECK(hipMemcpyAsync(c.gpeel.p, c.h_gpeel.data(), sizeof(GraphPeel) * G, hipMemcpyHostToDevice, c.stream));
  c.sync();
  c.peel = keep_stats;
  c.peel.core_cliques = gT[0];
  c.peeled = true;
  c.core_listed = fits;
  c.core_counted = true;
  c.counted = false;  // c.cnt now holds the core's counts: a new peel needs a new count pass
  c.have_comp = false;
}

}  // namespace kcds
