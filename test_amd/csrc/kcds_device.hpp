// Device-side helpers shared by the k-CDS kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KCDS_WAVE 64

#ifndef KCDS_ULL_DEFINED
#define KCDS_ULL_DEFINED
typedef unsigned long long ull;
#endif

// Orders this wave's LDS accesses: LDS instructions of one wave execute in
// issue order, so a compiler-level fence plus a wave barrier makes lane A's
// LDS write visible to lane B's later read.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Workgroup barrier that also waits for this wave's outstanding memory
// operations, so LDS and global writes made before it are complete.
__device__ __forceinline__ void block_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

// Waits until this wave's memory operations so far have completed (the
// compiler may not move memory operations across it).
__device__ __forceinline__ void wave_drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Scattered layout for per-vertex counters that take many atomics: vertex
// v's counter sits at slot (v * a) mod K (K = n rounded up to a power of two,
// a odd, so the map is a bijection; inverse p -> (p * ainv) mod K).  Hot
// vertices with close ids then get cache lines of their own instead of
// queueing their atomics on a few shared lines.
struct Scatter {
  uint32_t mask, a, ainv;
};
__device__ __forceinline__ uint32_t slot_of(const Scatter &sc, uint32_t v) { return (v * sc.a) & sc.mask; }
__device__ __forceinline__ uint32_t vertex_of(const Scatter &sc, uint32_t p) { return (p * sc.ainv) & sc.mask; }
inline Scatter make_scatter(uint32_t n) {
  Scatter sc;
  uint32_t K = 1;
  while (K < n) K <<= 1;
  sc.mask = K - 1;
  sc.a = 0x9E3779B1u;
  sc.ainv = sc.a;  // Newton iteration for the inverse mod 2^32
  for (int i = 0; i < 5; ++i) sc.ainv *= 2u - sc.a * sc.ainv;
  return sc;
}

// Workgroup barrier for data exchanged through LDS only: it waits for the
// wave's LDS operations but leaves its global loads, stores and atomics in
// flight (a plain __syncthreads also drains those).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t wave_bcast0(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ ull wave_sum_u64(ull v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the 64 lanes of a wave in which every lane is active, as a scalar.
// Data-parallel-primitive (DPP) row shifts build each 16-lane row's inclusive
// prefix sum (lane 15 holds the row's total), two row broadcasts carry the
// totals of rows 0..2 upward, and lane 63 ends with the sum: six VALU adds and
// no LDS traffic (a shuffle round trips through the LDS crossbar six times).
// Lanes switched off by divergence would contribute stale values, so callers
// use it only where control flow has been wave-uniform.
// Bit compress (generalised extract): the bits of x at the set positions of
// a mask m, packed to the low end in order.  Warren, Hacker's Delight, 2nd
// ed., section 7-4: six rounds, each moving the selected bits right by 2^i;
// the round masks depend on m alone, so they are computed once per mask.
__host__ __device__ __forceinline__ void compress_masks(ull m, ull *mv) {
  ull mk = ~m << 1;  // counts the zeros of m to the right of each position
  for (int i = 0; i < 6; ++i) {
    ull mp = mk ^ (mk << 1);  // parallel suffix of mk
    mp ^= mp << 2;
    mp ^= mp << 4;
    mp ^= mp << 8;
    mp ^= mp << 16;
    mp ^= mp << 32;
    const ull v = mp & m;  // bits that move 2^i places this round
    mv[i] = v;
    m = (m ^ v) | (v >> (1u << i));
    mk &= ~mp;
  }
}
// x must already be restricted to the mask the round masks came from.
__host__ __device__ __forceinline__ ull compress_apply(ull x, const ull *mv) {
  for (int i = 0; i < 6; ++i) {
    const ull t = x & mv[i];
    x = (x ^ t) | (t >> (1u << i));
  }
  return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32_full(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// A wave-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ ull uniform_u64(ull v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((ull)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, 64));
  return v;
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

__device__ __forceinline__ ull wave_incl_scan_u64(ull v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    ull t = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// Number of lanes below this one whose predicate is set.
__device__ __forceinline__ uint32_t lanes_below(ull ballot) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(ballot >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)ballot, 0u));
}

// Index of the first element of a[0..len) that is >= key (a sorted ascending).
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t len, uint32_t key) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Index of the first element of a[0..len) that is > key.
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t len, uint32_t key) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Same-address device atomics are serialised (about 11 ns each on MI355X), so
// a work queue touched by thousands of workgroups is split into QSHARDS
// contiguous shards with one counter each, QSTRIDE words apart.  A grabber
// starts at its own shard and moves on when one runs dry; it returns n when
// every shard is empty.
constexpr uint32_t QSHARDS = 32, QSTRIDE = 64;
constexpr uint32_t QUEUE_WORDS = QSHARDS * QSTRIDE;

struct QueueCursor {
  uint32_t shard, tries;
};

__device__ __forceinline__ QueueCursor queue_start(uint32_t id) { return QueueCursor{id % QSHARDS, 0}; }

__device__ __forceinline__ uint32_t queue_next(uint32_t *q, uint32_t n, QueueCursor &cur) {
  const uint32_t per = (n + QSHARDS - 1) / QSHARDS;
  while (cur.tries < QSHARDS) {
    const uint32_t lo = cur.shard * per;
    if (lo < n) {
      const uint32_t i = atomicAdd(&q[cur.shard * QSTRIDE], 1u);
      if (i < min(per, n - lo)) return lo + i;
    }
    cur.shard = cur.shard + 1 == QSHARDS ? 0 : cur.shard + 1;
    ++cur.tries;
  }
  return n;
}

// Exclusive scan of one value per thread over the workgroup (blockDim a
// multiple of 64, at most 1024); wsum is 16 words of LDS.  Every thread gets
// the total.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t v, uint32_t *total, uint32_t *wsum) {
  const uint32_t lane = threadIdx.x % KCDS_WAVE, w = threadIdx.x / KCDS_WAVE;
  const uint32_t incl = wave_incl_scan_u32(v);
  if (lane == KCDS_WAVE - 1) wsum[w] = incl;
  block_sync();
  uint32_t before = 0, all = 0;
  for (uint32_t i = 0; i < blockDim.x / KCDS_WAVE; ++i) {
    if (i < w) before += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__device__ __forceinline__ ull block_excl_scan_u64(ull v, ull *total, ull *wsum) {
  const uint32_t lane = threadIdx.x % KCDS_WAVE, w = threadIdx.x / KCDS_WAVE;
  const ull incl = wave_incl_scan_u64(v);
  if (lane == KCDS_WAVE - 1) wsum[w] = incl;
  block_sync();
  ull before = 0, all = 0;
  for (uint32_t i = 0; i < blockDim.x / KCDS_WAVE; ++i) {
    if (i < w) before += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

// Device-coherent loads: read data that other waves of this launch changed
// with atomics or stores (they bypass this CU's vector L1).
__device__ __forceinline__ uint32_t ldc_u32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ldc_i32(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ldc_i64(const long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stc_i64(long long *p, long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ ull ldc_u64(const ull *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Device-coherent stores, paired with the loads above for data that other
// waves of the same launch read or update with atomics.
__device__ __forceinline__ void stc_u32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stc_i32(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stc_u64(ull *p, ull v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-key slot assignment aggregated over a workgroup round: every thread
// brings E (want, key) pairs, and each distinct key of the round takes its
// slots in base[key] with one device atomic.  A hub key that repeats across
// the whole round (thousands of cliques of one vertex) therefore costs one
// atomic per round instead of one per wave, which keeps it from turning
// into one hot L2 line.  The table (LDS, 2 * blockDim.x * E <= TB slots,
// open addressing) is cleared by the call; every thread of the workgroup
// must call it.  slot[e] is valid where want[e].
template <int TB> struct KeyTable {
  uint32_t key[TB];
  uint32_t cnt[TB];
  ull base[TB];
};

template <int E, int TB, typename T>
__device__ void block_slot_by_key(KeyTable<TB> &tb, const bool *want, const uint32_t *key, T *slot, T *base) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  for (uint32_t t = tid; t < TB; t += nt) {
    tb.key[t] = 0xFFFFFFFFu;
    tb.cnt[t] = 0;
  }
  __syncthreads();
  uint32_t h[E], loc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    h[e] = 0;
    loc[e] = 0;
    if (!want[e]) continue;
    uint32_t x = ((uint32_t)__umul24(key[e] ^ (key[e] >> 24), 0x9E3779u) >> 8) & (TB - 1);
    while (true) {
      const uint32_t prev = atomicCAS(&tb.key[x], 0xFFFFFFFFu, key[e]);
      if (prev == 0xFFFFFFFFu || prev == key[e]) break;
      x = (x + 1) & (TB - 1);
    }
    h[e] = x;
    loc[e] = atomicAdd(&tb.cnt[x], 1u);
  }
  __syncthreads();
  for (uint32_t t = tid; t < TB; t += nt)
    if (tb.cnt[t]) tb.base[t] = (ull)atomicAdd(&base[tb.key[t]], (T)tb.cnt[t]);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (want[e]) slot[e] = (T)tb.base[h[e]] + (T)loc[e];
  __syncthreads();  // the table is reused by the next round
}

// Slot of this lane in base[key] for every lane with want: base[key] grows by
// the number of such lanes.  Lanes that share a key (a hub vertex, the same
// component) are served by one device atomic per group for the first eight
// groups; the remaining lanes use one atomic each.  Every lane of the wave
// must call it.
template <typename T>
__device__ __forceinline__ T wave_slot_by_key(bool want, uint32_t key, T *base) {
  const uint32_t lane = lane_id();
  T slot = 0;
  ull pending = __ballot(want);
#pragma unroll 1
  for (int round = 0; round < 8 && pending; ++round) {
    const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
    const uint32_t kl = __shfl(key, lead, 64);
    const ull same = __ballot(want && key == kl) & pending;
    T b = 0;
    if (lane == lead) b = atomicAdd(&base[kl], (T)__popcll(same));
    b = __shfl(b, lead, 64);
    if ((same >> lane) & 1ull) slot = b + (T)lanes_below(same);
    pending &= ~same;
  }
  if ((pending >> lane) & 1ull) slot = atomicAdd(&base[key], (T)1);
  return slot;
}

