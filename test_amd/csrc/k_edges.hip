// Edge list -> undirected CSR on the GPU: the input side in front of the
// solve.  The reference reads its graph on the host, one line at a time
// (src/graph.cpp:7-40: an adjacency file, self-loops skipped, the lists kept
// as written); the configurations it is benchmarked on are edge lists of
// 10^6-10^7 pairs, with both directions, repeats and self-loops possible.
// Here the pairs go to HBM as they are and the CSR that every later stage
// reads (neighbour lists ascending, no duplicate, no self-loop) is built by:
//   1. arcs: each pair (u, v) gives u->v and v->u; a self-loop or an id >= n
//      gives two sentinel arcs (source n) that sort last and are dropped, and
//      an id >= n also raises the error flag;
//   2. a stable LSD radix sort of the arcs by (source, target): the 8-bit
//      digits of the target first, then those of the source.  A pass is a
//      per-tile digit histogram, an exclusive scan over (digit, tile) and a
//      scatter in which one wave walks its tile in order, so equal digits keep
//      their order (HBM traffic: 16 bytes read and 16 written per arc a pass);
//   3. unique: an arc is kept when it is no sentinel and differs from the arc
//      before it; the kept targets are scanned into place, and each run of
//      equal sources within a wave adds its length to that source's degree;
//   4. offsets: an exclusive scan of the degrees; the widest list by a max
//      reduction.
#include <algorithm>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr uint32_t ED_TILE = 1024;  // arcs per scatter tile: one wave walks 16 chunks of 64
constexpr uint32_t ED_CHUNKS = ED_TILE / KCDS_WAVE;

// 256-thread workgroups for a grid-stride loop over len items
uint32_t ed_grid(uint64_t len) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((len + 255) / 256, 65535ull * 4));
}

__global__ __launch_bounds__(256) void k_ed_arcs(const uint2 *__restrict__ uv, ull m, uint32_t n,
                                                 uint2 *__restrict__ src, uint2 *__restrict__ dst,
                                                 uint32_t *__restrict__ meta) {
  bool bad = false;
  for (ull i = (ull)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (ull)gridDim.x * blockDim.x) {
    uint2 e = uv[i];
    if (e.x >= n || e.y >= n) {
      bad = true;
      e.x = e.y = n;
    } else if (e.x == e.y) {
      e.x = e.y = n;
    }
    src[i] = make_uint2(e.x, e.y);  // arcs 2i (u -> v) and 2i + 1 (v -> u)
    dst[i] = make_uint2(e.y, e.x);
  }
  if (__ballot(bad) && lane_id() == 0) atomicOr(&meta[0], 1u);
}

// hist[digit * ntile + tile]: digit-major, so the scan orders (digit, tile)
__global__ __launch_bounds__(256) void k_ed_hist(const uint32_t *__restrict__ key, uint32_t len, uint32_t shift,
                                                 uint32_t ntile, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * ED_TILE, end = min(len, base + ED_TILE);
  for (uint32_t i = base + threadIdx.x; i < end; i += blockDim.x) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];
}

// One wave per tile; hoff = exclusive scan of hist.  Within a chunk of 64 the
// lanes holding the same digit find each other with eight ballots; the one
// with the most lanes before it advances the digit's cursor.
__global__ __launch_bounds__(64) void k_ed_scatter(const uint32_t *__restrict__ key_in,
                                                   const uint32_t *__restrict__ val_in, uint32_t len, uint32_t shift,
                                                   uint32_t ntile, const uint32_t *__restrict__ hoff,
                                                   uint32_t *__restrict__ key_out, uint32_t *__restrict__ val_out) {
  __shared__ uint32_t cursor[256];
  const uint32_t lane = threadIdx.x, base = blockIdx.x * ED_TILE;
#pragma unroll
  for (uint32_t t = 0; t < 4; ++t) cursor[lane + 64 * t] = hoff[(lane + 64 * t) * ntile + blockIdx.x];
  uint32_t key[ED_CHUNKS], val[ED_CHUNKS];
#pragma unroll
  for (uint32_t ch = 0; ch < ED_CHUNKS; ++ch) {  // the whole tile's loads in flight together
    const uint32_t i = base + ch * KCDS_WAVE + lane;
    key[ch] = i < len ? key_in[i] : 0u;
    val[ch] = i < len ? val_in[i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t ch = 0; ch < ED_CHUNKS; ++ch) {
    const bool valid = base + ch * KCDS_WAVE + lane < len;
    const uint32_t d = (key[ch] >> shift) & 255u;
    ull same = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < 8; ++b) {
      const ull m = __ballot((d >> b) & 1u);
      same &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t before = lanes_below(same), cnt = (uint32_t)__popcll(same);
    const uint32_t pos = valid ? cursor[d] + before : 0u;
    __syncthreads();  // every lane has its position before any cursor moves
    if (valid && before + 1 == cnt) cursor[d] += cnt;
    __syncthreads();
    if (valid) {
      key_out[pos] = key[ch];
      val_out[pos] = val[ch];
    }
  }
}

// keep[i] = 1 for the first of each run of equal (source, target) that is no
// sentinel; keep[len] = 0 so that the scan's last entry is the kept total
__global__ __launch_bounds__(256) void k_ed_unique(const uint32_t *__restrict__ S, const uint32_t *__restrict__ D,
                                                   uint32_t len, uint32_t n, uint32_t *__restrict__ keep) {
  for (ull i = (ull)blockIdx.x * blockDim.x + threadIdx.x; i <= len; i += (ull)gridDim.x * blockDim.x) {
    uint32_t k = 0;
    if (i < len) k = S[i] < n && (i == 0 || S[i] != S[i - 1] || D[i] != D[i - 1]);
    keep[i] = k;
  }
}

__global__ __launch_bounds__(256) void k_ed_place(const uint32_t *__restrict__ S, const uint32_t *__restrict__ D,
                                                  const uint32_t *__restrict__ keep, const uint32_t *__restrict__ pos,
                                                  uint32_t len, uint32_t *__restrict__ nbr, uint32_t *__restrict__ deg) {
  // every lane of a wave takes part in the grouping: the loop bound is wave-uniform
  const ull stride = (ull)gridDim.x * blockDim.x;
  for (ull i0 = (ull)blockIdx.x * blockDim.x; i0 < len; i0 += stride) {
    const ull i = i0 + threadIdx.x;
    const bool k = i < len && keep[i];
    const uint32_t s = k ? S[i] : 0u;
    if (k) nbr[pos[i]] = D[i];
    // sorted by source, so a wave's kept arcs fall in a few runs of equal source
    const ull kept = __ballot(k);
    ull pending = kept;
    while (pending) {
      const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
      const uint32_t sl = __shfl(s, lead, 64);
      const ull run = __ballot(k && s == sl) & pending;
      if (lane_id() == lead) atomicAdd(&deg[sl], (uint32_t)__popcll(run));
      pending &= ~run;
    }
  }
}

}  // namespace

void build_csr_from_edges(Ctx &c, uint32_t n, uint64_t m, const uint32_t *uv_host) {
  const uint64_t L64 = 2 * m;
  // 32-bit arc positions, with a tile's worth of headroom below 2^32
  if (L64 > 0xFFFF0000ull) fail(KCDS_E_LIMIT, "%llu edges: at most 2^31 - 2^15", (unsigned long long)m);
  if (n == 0xFFFFFFFFu) fail(KCDS_E_LIMIT, "n = %u: at most 2^32 - 2 vertices", n);
  const uint32_t L = (uint32_t)L64;
  const uint32_t ntile = (L + ED_TILE - 1) / ED_TILE;
  const size_t nh = (size_t)256 * ntile;
  // tmp0: the pairs; tmp1: two (source, target) arrays, keep and its scan;
  // tmp2: the digit histogram and its scan; tmp3: meta words and degrees
  buf_reserve(c, c.tmp0, sizeof(uint32_t) * (L64 + 2));
  buf_reserve(c, c.tmp1, sizeof(uint32_t) * (6 * L64 + 16 * 64));
  buf_reserve(c, c.tmp2, sizeof(uint32_t) * (2 * nh + 128));
  buf_reserve(c, c.tmp3, sizeof(uint32_t) * ((size_t)n + 1 + 64));
  buf_reserve(c, c.off, sizeof(uint32_t) * ((size_t)n + 1));
  buf_reserve(c, c.nbr, sizeof(uint32_t) * (L64 + 1));
  const size_t stride = (L64 + 63) & ~(size_t)63;  // keeps each array 256-byte aligned
  uint32_t *S0 = c.tmp1.as<uint32_t>(), *D0 = S0 + stride, *S1 = D0 + stride, *D1 = S1 + stride;
  uint32_t *keep = D1 + stride, *pos = keep + stride + 64;
  uint32_t *hist = c.tmp2.as<uint32_t>(), *hoff = hist + ((nh + 63) & ~(size_t)63);
  uint32_t *meta = c.tmp3.as<uint32_t>(), *deg = meta + 64;
  uint32_t *off = c.off.as<uint32_t>(), *nbr = c.nbr.as<uint32_t>();
  fill_u32(c, meta, 0, 64);
  fill_u32(c, deg, 0, (size_t)n + 1);
  if (L) {
    HIP_CHECK(hipMemcpyAsync(c.tmp0.p, uv_host, sizeof(uint32_t) * L64, hipMemcpyHostToDevice, c.stream));
    k_ed_arcs<<<ed_grid(m), 256, 0, c.stream>>>(c.tmp0.as<uint2>(), m, n, (uint2 *)S0, (uint2 *)D0, meta);
    KCDS_LAUNCH_CHECK();
    // digits of n itself: the sentinel source must sort
    uint32_t bits = 0;
    while (bits < 32 && (n >> bits)) ++bits;
    const uint32_t dpk = std::max(1u, (bits + 7) / 8);
    uint32_t *S = S0, *D = D0, *So = S1, *Do = D1;
    for (uint32_t p = 0; p < 2 * dpk; ++p) {
      const bool by_target = p < dpk;
      const uint32_t shift = 8 * (by_target ? p : p - dpk);
      const uint32_t *kin = by_target ? D : S, *vin = by_target ? S : D;
      uint32_t *kout = by_target ? Do : So, *vout = by_target ? So : Do;
      k_ed_hist<<<ntile, 256, 0, c.stream>>>(kin, L, shift, ntile, hist);
      KCDS_LAUNCH_CHECK();
      scan_u32_to_u32(c, hist, hoff, nh);
      k_ed_scatter<<<ntile, 64, 0, c.stream>>>(kin, vin, L, shift, ntile, hoff, kout, vout);
      KCDS_LAUNCH_CHECK();
      std::swap(S, So);
      std::swap(D, Do);
    }
    k_ed_unique<<<ed_grid((uint64_t)L + 1), 256, 0, c.stream>>>(S, D, L, n, keep);
    KCDS_LAUNCH_CHECK();
    scan_u32_to_u32(c, keep, pos, (uint64_t)L + 1);
    k_ed_place<<<ed_grid(L), 256, 0, c.stream>>>(S, D, keep, pos, L, nbr, deg);
    KCDS_LAUNCH_CHECK();
  }
  scan_u32_to_u32(c, deg, off, (uint64_t)n + 1);
  max_u32(c, deg, n, meta + 1);
  HIP_CHECK(hipMemcpyAsync(meta + 2, off + n, sizeof(uint32_t), hipMemcpyDeviceToDevice, c.stream));
  uint32_t *h = (uint32_t *)c.host_scratch(4 * sizeof(uint32_t));
  HIP_CHECK(hipMemcpyAsync(h, meta, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (h[0]) fail(KCDS_E_INVALID, "an edge endpoint is out of range (n = %u)", n);
  c.nnz = h[2];
  c.maxdeg = h[1];
}

}  // namespace kcds
