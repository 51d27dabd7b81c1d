// Wave64 k-clique listing with LDS bitmask rows.
//
// Reference: listAllCliques (main.cu:93-327) lists cliques level by level
// with per-warp partitions of partial cliques, 32-bit neighbour masks
// (listIntialCliques / listMidCliques, src/helpers.cu:175-488), counts them
// (countCliques, src/helpers.cu:1074-1113) and writes them with per-vertex
// clique degrees (writeFinalCliques, src/helpers.cu:489-550).
//
// Here one wavefront owns one root r of the DAG (new ids, lists ascending):
//   1. stage N+(r) = L[0..d) in LDS and build the symmetric adjacency of the
//      induced subgraph as d rows of W 64-bit words, A[i] bit j = {L[i],L[j]}
//      is an edge.  Every element of N+(L[i]) is looked up in L by binary
//      search (the out-wedge sum W2 of the roofline is these reads).
//   2. Because ids increase along every DAG arc, the out-neighbours of L[c]
//      inside N+(r) are A[c] masked to bits above c, so recursion on
//      candidate sets S (W words, wave-uniform) needs no further global
//      reads: S' = S & A[c] & above(c).
//   3. At two vertices left, the cliques extending the prefix are the edges
//      inside S: lane j (owner of rows j, j+64, ...) adds popc(A[j] & S) to
//      its own counter; the wave sum / 2 is credited to every prefix vertex.
//      Per-vertex counts are exact without touching individual cliques.
//   4. Counters flush to HBM with one 64-bit atomic per (root, member).
// The materialising variant enumerates the same cliques in the same order and
// writes them row-major (k new ids per clique) at the root's offset.
#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr int NCLASS = 4;  // W = 1, 2, 4, 8 words per row (out-degree <= 64, 128, 256, 512)
constexpr uint32_t MAX_LDS_DEG = 512;

template <int W> struct Frame {
  uint32_t L[64 * W];        // N+(r)
  uint32_t pre[64 * W + 1];  // prefix sums of outdeg(L[i])
  uint32_t st[64 * W];       // doff[L[i]] - pre[i]
  ull A[64 * W * W];         // symmetric adjacency rows
};

__device__ __forceinline__ ull above_mask(int q, uint32_t c) {
  const int cq = (int)(c >> 6), cb = (int)(c & 63);
  if (q < cq) return 0ull;
  if (q > cq) return ~0ull;
  return cb == 63 ? 0ull : (~0ull << (cb + 1));
}

template <int W>
__device__ void build_rows(Frame<W> &f, const uint32_t *__restrict__ doff, const uint32_t *__restrict__ dnbr,
                           uint32_t base, uint32_t d) {
  const uint32_t lane = lane_id();
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < d; t0 += KCDS_WAVE) {
    const uint32_t t = t0 + lane;
    uint32_t du = 0, ou = 0;
    if (t < d) {
      uint32_t u = dnbr[base + t];
      ou = doff[u];
      du = doff[u + 1] - ou;
      f.L[t] = u;
    }
    uint32_t incl = wave_incl_scan_u32(du);
    if (t < d) {
      f.pre[t + 1] = carry + incl;
      f.st[t] = ou - (carry + incl - du);
    }
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) f.pre[0] = 0;
  for (uint32_t i = lane; i < d * W; i += KCDS_WAVE) f.A[i] = 0ull;
  wave_sync();
  const uint32_t E = carry;
  const uint32_t last = f.L[d - 1];
  for (uint32_t e0 = 0; e0 < E; e0 += KCDS_WAVE) {
    const uint32_t e = e0 + lane;
    if (e < E) {
      const uint32_t i = upper_bound_u32(f.pre, d + 1, e) - 1;
      const uint32_t w = dnbr[f.st[i] + e];
      if (w <= last) {
        const uint32_t j = i + 1 + lower_bound_u32(f.L + i + 1, d - i - 1, w);
        if (j < d && f.L[j] == w) {
          atomicOr(&f.A[i * W + (j >> 6)], 1ull << (j & 63));
          atomicOr(&f.A[j * W + (i >> 6)], 1ull << (i & 63));
        }
      }
    }
  }
  wave_sync();
}

// Credits E to local vertex c (owner lane c & 63, slot c >> 6).
template <int W> __device__ __forceinline__ void credit(ull *cnt, uint32_t c, ull E) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int q = 0; q < W; ++q)
    if ((uint32_t)q == (c >> 6) && lane == (c & 63)) cnt[q] += E;
}

// Counting recursion: REM vertices still to choose from candidate set S.
// Returns the number of REM-cliques inside S (wave-uniform); every member
// vertex's counter is credited with the cliques it belongs to.
template <int W, int REM> struct Count {
  static __device__ ull run(const Frame<W> &f, const ull *S, ull *cnt) {
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
        for (int w = 0; w < W; ++w) {
          S2[w] = S[w] & f.A[c * W + w] & above_mask(w, c);
          pc += (uint32_t)__popcll(S2[w]);
        }
        if (pc < (uint32_t)(REM - 1)) continue;
        const ull E = Count<W, REM - 1>::run(f, S2, cnt);
        credit<W>(cnt, c, E);
        total += E;
      }
    }
    return total;
  }
};

template <int W> struct Count<W, 2> {
  static __device__ ull run(const Frame<W> &f, const ull *S, ull *cnt) {
    const uint32_t lane = lane_id();
    uint32_t xs = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      if ((S[q] >> lane) & 1ull) {
        const uint32_t j = (uint32_t)q * 64 + lane;
        uint32_t x = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) x += (uint32_t)__popcll(f.A[j * W + w] & S[w]);
        cnt[q] += x;
        xs += x;
      }
    }
    return (ull)(wave_sum_u32(xs) >> 1);
  }
};

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
          S2[w] = S[w] & f.A[c * W + w] & above_mask(w, c);
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
        H[w] = in ? (f.A[j * W + w] & S[w] & above_mask(w, j)) : 0ull;
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
            uint32_t *row = out + at * K;
#pragma unroll
            for (int t = 0; t < K - 2; ++t) row[t] = pfx[t];
            row[K - 2] = uj;
            row[K - 1] = f.L[w * 64 + bb];
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

template <int W, int K>
__global__ __launch_bounds__(64) void k_list_count(const uint32_t *__restrict__ doff,
                                                   const uint32_t *__restrict__ dnbr,
                                                   const uint32_t *__restrict__ roots, uint32_t nroots,
                                                   uint32_t *qctr, ull *gcnt, ull *rootT, ull *gtotal) {
  __shared__ Frame<W> f;
  const uint32_t lane = lane_id();
  ull tot = 0;
  while (true) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(qctr, 1u);
    idx = wave_bcast0(__shfl(idx, 0, 64));
    if (idx >= nroots) break;
    const uint32_t r = roots[idx];
    const uint32_t base = doff[r], d = doff[r + 1] - base;
    build_rows<W>(f, doff, dnbr, base, d);
    ull cnt[W];
#pragma unroll
    for (int q = 0; q < W; ++q) cnt[q] = 0;
    ull S[W];
    all_bits<W>(S, d);
    const ull rt = Count<W, K - 1>::run(f, S, cnt);
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const uint32_t j = (uint32_t)q * 64 + lane;
      if (j < d && cnt[q]) atomicAdd(&gcnt[f.L[j]], cnt[q]);
    }
    if (lane == 0) {
      if (rt) atomicAdd(&gcnt[r], rt);
      rootT[r] = rt;
    }
    tot += rt;
    wave_sync();
  }
  if (lane == 0 && tot) atomicAdd(gtotal, tot);
}

template <int W, int K>
__global__ __launch_bounds__(64) void k_list_emit(const uint32_t *__restrict__ doff,
                                                  const uint32_t *__restrict__ dnbr,
                                                  const uint32_t *__restrict__ roots, uint32_t nroots,
                                                  uint32_t *qctr, const ull *rootOff, uint32_t *out) {
  __shared__ Frame<W> f;
  const uint32_t lane = lane_id();
  while (true) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(qctr, 1u);
    idx = wave_bcast0(__shfl(idx, 0, 64));
    if (idx >= nroots) break;
    const uint32_t r = roots[idx];
    ull pos = rootOff[r];
    const ull lim = rootOff[r + 1];
    if (lim == pos) continue;
    const uint32_t base = doff[r], d = doff[r + 1] - base;
    build_rows<W>(f, doff, dnbr, base, d);
    ull S[W];
    all_bits<W>(S, d);
    uint32_t pfx[K];
    pfx[0] = r;
    Emit<W, K, K - 1>::run(f, S, pfx, out, pos, lim);
    wave_sync();
  }
}

// Root classification: class c holds roots with 64*2^(c-1) < outdeg <= 64*2^c
// and outdeg >= k-1 (smaller roots have no k-clique).
__device__ __forceinline__ int width_class(uint32_t d) {
  return d <= 64 ? 0 : d <= 128 ? 1 : d <= 256 ? 2 : d <= 512 ? 3 : 4;
}

__global__ void k_class_count(const uint32_t *doff, uint32_t n, uint32_t k, uint32_t *ccount) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    uint32_t d = doff[r + 1] - doff[r];
    if (d + 1 < k) continue;
    atomicAdd(&ccount[width_class(d)], 1u);
  }
}

__global__ void k_class_fill(const uint32_t *doff, uint32_t n, uint32_t k, const uint32_t *cstart,
                             uint32_t *cfill, uint32_t *roots) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    uint32_t d = doff[r + 1] - doff[r];
    if (d + 1 < k) continue;
    int c = width_class(d);
    roots[cstart[c] + atomicAdd(&cfill[c], 1u)] = r;
  }
}

// incidence index: inc[incoff[v] + i] = id of the i-th clique containing v
__global__ void k_incidence(const uint32_t *cl, uint64_t T, uint32_t k, const ull *incoff, uint32_t *fillc,
                            uint32_t *inc) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
       c += (uint64_t)gridDim.x * blockDim.x) {
    for (uint32_t t = 0; t < k; ++t) {
      uint32_t v = cl[c * k + t];
      uint32_t slot = atomicAdd(&fillc[v], 1u);
      if (incoff[v] + slot < incoff[v + 1]) inc[incoff[v] + slot] = (uint32_t)c;
    }
  }
}

typedef void (*CountFn)(const uint32_t *, const uint32_t *, const uint32_t *, uint32_t, uint32_t *, ull *, ull *,
                        ull *);
typedef void (*EmitFn)(const uint32_t *, const uint32_t *, const uint32_t *, uint32_t, uint32_t *, const ull *,
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

const uint32_t kWaveDefault[NCLASS] = {16, 12, 8, 4};

uint32_t lds_bytes(int cls) {
  switch (cls) {
    case 0: return sizeof(Frame<1>);
    case 1: return sizeof(Frame<2>);
    case 2: return sizeof(Frame<4>);
    default: return sizeof(Frame<8>);
  }
}

void classify_roots(Ctx &c, uint32_t k) {
  const uint32_t n = c.n;
  buf_reserve(c, c.roots, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.qctr, 64 * sizeof(uint32_t));
  uint32_t *q = c.qctr.as<uint32_t>();
  HIP_CHECK(hipMemsetAsync(q, 0, 64 * sizeof(uint32_t), c.stream));
  uint32_t *ccount = q + 16, *cstart = q + 24, *cfill = q + 32;
  uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 4096));
  if (n) {
    k_class_count<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), n, k, ccount);
    KCDS_LAUNCH_CHECK();
  }
  uint32_t *h = (uint32_t *)c.host_scratch(64);
  HIP_CHECK(hipMemcpyAsync(h, ccount, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (h[4]) fail(KCDS_E_LIMIT, "%u roots have out-degree above %u (the LDS row limit)", h[4], MAX_LDS_DEG);
  uint32_t s = 0;
  uint32_t hs[8] = {0};
  for (int i = 0; i < NCLASS; ++i) {
    c.class_count[i] = h[i];
    c.class_start[i] = s;
    hs[i] = s;
    s += h[i];
  }
  HIP_CHECK(hipMemcpyAsync(cstart, hs, 8 * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
  if (n && s) {
    k_class_fill<<<grid, 256, 0, c.stream>>>(c.doff.as<uint32_t>(), n, k, cstart, cfill, c.roots.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.sync();  // hs lives on the host stack
}

uint32_t wave_grid(Ctx &c, int cls, uint32_t nroots) {
  uint32_t per_cu = c.list_waves_per_cu ? c.list_waves_per_cu : kWaveDefault[cls];
  uint32_t lds_per_cu = 160u * 1024u;
  per_cu = std::min<uint32_t>(per_cu, std::max<uint32_t>(1, lds_per_cu / lds_bytes(cls)));
  return std::max<uint32_t>(1, std::min<uint32_t>(nroots, (uint32_t)c.num_cus * per_cu));
}

}  // namespace

void list_count(Ctx &c, uint32_t k) {
  if (!c.oriented) fail(KCDS_E_STATE, "kcds_count/kcds_list before kcds_orient");
  if (k < 3 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 3..8", k);
  const uint32_t n = c.n;
  classify_roots(c, k);
  buf_reserve(c, c.cnt, sizeof(ull) * ((size_t)n + 1));
  buf_reserve(c, c.rootT, sizeof(ull) * ((size_t)n + 2));
  buf_reserve(c, c.scal, 64 * sizeof(ull));
  HIP_CHECK(hipMemsetAsync(c.cnt.p, 0, sizeof(ull) * ((size_t)n + 1), c.stream));
  HIP_CHECK(hipMemsetAsync(c.rootT.p, 0, sizeof(ull) * ((size_t)n + 2), c.stream));
  ull *total = c.scal.as<ull>();
  HIP_CHECK(hipMemsetAsync(total, 0, sizeof(ull), c.stream));
  uint32_t *q = c.qctr.as<uint32_t>();
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, c.stream));
  uint32_t launches = 0;
  for (int cls = NCLASS - 1; cls >= 0; --cls) {
    const uint32_t nr = c.class_count[cls];
    if (!nr) continue;
    CountFn fn = cls == 0 ? count_fn<1>(k) : cls == 1 ? count_fn<2>(k) : cls == 2 ? count_fn<4>(k) : count_fn<8>(k);
    const uint32_t *roots = c.roots.as<uint32_t>() + c.class_start[cls];
    hipLaunchKernelGGL(fn, dim3(wave_grid(c, cls, nr)), dim3(64), 0, c.stream, c.doff.as<uint32_t>(),
                       c.dnbr.as<uint32_t>(), roots, nr, q + cls, c.cnt.as<ull>(), c.rootT.as<ull>(), total);
    KCDS_LAUNCH_CHECK();
    ++launches;
  }
  HIP_CHECK(hipEventRecord(e1, c.stream));
  ull *h = (ull *)c.host_scratch(8);
  HIP_CHECK(hipMemcpyAsync(h, total, sizeof(ull), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  c.list_kernel_ms = ms;
  c.list_kernel_launches = launches;
  c.T = *h;
  c.k = k;
  c.counted = true;
  c.listed = false;
  c.peeled = false;
  c.have_comp = false;
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
  uint32_t *q = c.qctr.as<uint32_t>();
  HIP_CHECK(hipMemsetAsync(q, 0, 8 * sizeof(uint32_t), c.stream));
  for (int cls = NCLASS - 1; cls >= 0; --cls) {
    const uint32_t nr = c.class_count[cls];
    if (!nr) continue;
    EmitFn fn = cls == 0 ? emit_fn<1>(k) : cls == 1 ? emit_fn<2>(k) : cls == 2 ? emit_fn<4>(k) : emit_fn<8>(k);
    const uint32_t *roots = c.roots.as<uint32_t>() + c.class_start[cls];
    hipLaunchKernelGGL(fn, dim3(wave_grid(c, cls, nr)), dim3(64), 0, c.stream, c.doff.as<uint32_t>(),
                       c.dnbr.as<uint32_t>(), roots, nr, q + cls, rootOff, c.cliques.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  buf_reserve(c, c.tmp3, sizeof(uint32_t) * ((size_t)n + 1));
  HIP_CHECK(hipMemsetAsync(c.tmp3.p, 0, sizeof(uint32_t) * ((size_t)n + 1), c.stream));
  if (T) {
    uint32_t grid = (uint32_t)std::min<uint64_t>((T + 255) / 256, 65535ull * 16);
    k_incidence<<<grid, 256, 0, c.stream>>>(c.cliques.as<uint32_t>(), T, k, c.incoff.as<ull>(),
                                            c.tmp3.as<uint32_t>(), c.inc.as<uint32_t>());
    KCDS_LAUNCH_CHECK();
  }
  c.sync();
  c.listed = true;
}

}  // namespace kcds
