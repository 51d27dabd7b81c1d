// Device-wide primitives: exclusive scan, fill, max and sum reductions.
// Written directly for wave64: 256-thread blocks, 16 items per thread.
#include <algorithm>

#include "kcds_device.hpp"
#include "kcds_internal.hpp"

namespace kcds {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;

// Exclusive block scan of one value per thread; returns this thread's prefix,
// *total receives the block total.
template <typename T> __device__ T block_excl_scan(T v, T *total) {
  __shared__ T wsum[SCAN_THREADS / KCDS_WAVE];
  const uint32_t lane = threadIdx.x % KCDS_WAVE, w = threadIdx.x / KCDS_WAVE;
  T incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(incl, o, 64);
    if (lane >= (uint32_t)o) incl += t;
  }
  if (lane == KCDS_WAVE - 1) wsum[w] = incl;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < SCAN_THREADS / KCDS_WAVE; ++i) {
    if (i < (int)w) before += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

template <typename In, typename Out>
__global__ __launch_bounds__(SCAN_THREADS) void k_chunk_sums(const In *in, uint64_t len, Out *bsum) {
  uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK + (uint64_t)threadIdx.x * SCAN_ITEMS;
  Out s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    if (base + i < len) s += (Out)in[base + i];
  Out tot;
  block_excl_scan<Out>(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename In, typename Out>
__global__ __launch_bounds__(SCAN_THREADS) void k_chunk_apply(const In *in, Out *out, uint64_t len,
                                                              const Out *boff) {
  uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK + (uint64_t)threadIdx.x * SCAN_ITEMS;
  In vals[SCAN_ITEMS];
  Out s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    vals[i] = (base + i < len) ? in[base + i] : (In)0;
    s += (Out)vals[i];
  }
  Out tot;
  Out run = block_excl_scan<Out>(s, &tot) + (boff ? boff[blockIdx.x] : (Out)0);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    if (base + i < len) {
      out[base + i] = run;
      run += (Out)vals[i];
    }
}

template <typename Out> __global__ void k_write_total(Out *out, uint64_t len, const Out *boff, uint64_t nb) {
  out[len] = boff[nb];
}

template <typename In, typename Out> __global__ void k_last_total(const In *in, Out *out, uint64_t len) {
  out[len] = out[len - 1] + (Out)in[len - 1];
}

// Words of block sums a scan of len entries needs at every level below it.
uint64_t scan_ws_words(uint64_t len) {
  uint64_t w = 0;
  for (uint64_t nb = (len + SCAN_CHUNK - 1) / SCAN_CHUNK; nb > 1; nb = (nb + SCAN_CHUNK - 1) / SCAN_CHUNK)
    w += 2 * nb + 1;
  return w;
}

// ws: scan_ws_words(len) entries of block sums (bsum, then boff, per level).
template <typename In, typename Out> void scan_impl(Ctx &c, const In *in, Out *out, uint64_t len, Out *ws) {
  if (len == 0) {
    HIP_CHECK(hipMemsetAsync(out, 0, sizeof(Out), c.stream));
    return;
  }
  uint64_t nb = (len + SCAN_CHUNK - 1) / SCAN_CHUNK;
  if (nb == 1) {
    // single chunk: the total is the last prefix plus the last value
    k_chunk_apply<In, Out><<<1, SCAN_THREADS, 0, c.stream>>>(in, out, len, nullptr);
    KCDS_LAUNCH_CHECK();
    k_last_total<In, Out><<<1, 1, 0, c.stream>>>(in, out, len);
    KCDS_LAUNCH_CHECK();
    return;
  }
  Out *bsum = ws, *boff = ws + nb;
  k_chunk_sums<In, Out><<<(uint32_t)nb, SCAN_THREADS, 0, c.stream>>>(in, len, bsum);
  KCDS_LAUNCH_CHECK();
  scan_impl<Out, Out>(c, bsum, boff, nb, ws + 2 * nb + 1);
  k_chunk_apply<In, Out><<<(uint32_t)nb, SCAN_THREADS, 0, c.stream>>>(in, out, len, boff);
  KCDS_LAUNCH_CHECK();
  k_write_total<Out><<<1, 1, 0, c.stream>>>(out, len, boff, nb);
  KCDS_LAUNCH_CHECK();
}

template <typename In, typename Out> void scan_top(Ctx &c, const In *in, Out *out, uint64_t len) {
  const uint64_t w = scan_ws_words(len);
  if (w) buf_reserve(c, c.scan_ws, sizeof(Out) * w);
  scan_impl<In, Out>(c, in, out, len, w ? c.scan_ws.as<Out>() : nullptr);
}

__global__ void k_zero4(ull *p0, uint64_t n0, ull *p1, uint64_t n1, ull *p2, uint64_t n2, ull *p3, uint64_t n3) {
  const uint64_t tot = n0 + n1 + n2 + n3, stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += stride) {
    if (i < n0) p0[i] = 0;
    else if (i < n0 + n1) p1[i - n0] = 0;
    else if (i < n0 + n1 + n2) p2[i - n0 - n1] = 0;
    else p3[i - n0 - n1 - n2] = 0;
  }
}

template <typename T> __global__ void k_fill(T *p, T v, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Reductions: one device atomic per 256-thread block (a block reduces in LDS
// first); launched with at most REDUCE_BLOCKS blocks, since same-address
// device atomics are serialised.
constexpr uint32_t REDUCE_BLOCKS = 256;

__global__ __launch_bounds__(256) void k_max_u32(const uint32_t *in, uint64_t len, uint32_t *out) {
  __shared__ uint32_t s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
  m = wave_max_u32(m);
  if (lane_id() == 0) atomicMax(&s, m);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, s);
}

__global__ __launch_bounds__(256) void k_max_u64(const ull *in, uint64_t len, ull *out) {
  __shared__ ull s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  ull m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, in[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (ull)__shfl_xor(m, o, 64));
  if (lane_id() == 0) atomicMax(&s, m);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, s);
}

__global__ __launch_bounds__(256) void k_sum_u64(const ull *in, uint64_t len, ull *out) {
  __shared__ ull s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  ull t = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    t += in[i];
  t = wave_sum_u64(t);
  if (lane_id() == 0) atomicAdd(&s, t);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, s);
}

uint32_t reduce_grid(uint64_t len) {
  const uint64_t g = (len + 255) / 256;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, REDUCE_BLOCKS));
}

uint32_t grid_for(uint64_t len, uint32_t threads) {
  uint64_t g = (len + threads - 1) / threads;
  if (g > 65535ull * 4) g = 65535ull * 4;
  return (uint32_t)(g ? g : 1);
}

template <typename T>
__global__ void k_gather_by_rank(const T *src_new, const uint32_t *rank, uint32_t n, T *dst) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    dst[v] = src_new[rank[v]];
}

}  // namespace

// dst[v] = src[rank[v]]: new-id space -> original ids
void gather_by_rank_u64(Ctx &c, const ull *src_new, ull *dst_orig) {
  if (!c.n) return;
  k_gather_by_rank<ull><<<grid_for(c.n, 256), 256, 0, c.stream>>>(src_new, c.rank.as<uint32_t>(), c.n, dst_orig);
  KCDS_LAUNCH_CHECK();
}
void gather_by_rank_u32(Ctx &c, const uint32_t *src_new, uint32_t *dst_orig) {
  if (!c.n) return;
  k_gather_by_rank<uint32_t><<<grid_for(c.n, 256), 256, 0, c.stream>>>(src_new, c.rank.as<uint32_t>(), c.n,
                                                                       dst_orig);
  KCDS_LAUNCH_CHECK();
}

void scan_u32_to_u32(Ctx &c, const uint32_t *in, uint32_t *out, uint64_t len) {
  scan_top<uint32_t, uint32_t>(c, in, out, len);
}
void scan_u64_to_u64(Ctx &c, const ull *in, ull *out, uint64_t len) { scan_top<ull, ull>(c, in, out, len); }

void zero_u64(Ctx &c, ull *p0, uint64_t n0, ull *p1, uint64_t n1, ull *p2, uint64_t n2, ull *p3, uint64_t n3) {
  if (!p0) n0 = 0;
  if (!p1) n1 = 0;
  if (!p2) n2 = 0;
  if (!p3) n3 = 0;
  const uint64_t tot = n0 + n1 + n2 + n3;
  if (!tot) return;
  k_zero4<<<grid_for(tot, 256), 256, 0, c.stream>>>(p0, n0, p1, n1, p2, n2, p3, n3);
  KCDS_LAUNCH_CHECK();
}

void fill_u32(Ctx &c, uint32_t *p, uint32_t v, uint64_t len) {
  if (!len) return;
  k_fill<uint32_t><<<grid_for(len, 256), 256, 0, c.stream>>>(p, v, len);
  KCDS_LAUNCH_CHECK();
}
void fill_u64(Ctx &c, ull *p, ull v, uint64_t len) {
  if (!len) return;
  k_fill<ull><<<grid_for(len, 256), 256, 0, c.stream>>>(p, v, len);
  KCDS_LAUNCH_CHECK();
}

void max_u32(Ctx &c, const uint32_t *in, uint64_t len, uint32_t *out_dev) {
  fill_u32(c, out_dev, 0, 1);
  if (!len) return;
  k_max_u32<<<reduce_grid(len), 256, 0, c.stream>>>(in, len, out_dev);
  KCDS_LAUNCH_CHECK();
}

ull max_u64_host(Ctx &c, const ull *in, uint64_t len) {
  buf_reserve(c, c.scal, 64 * sizeof(ull));
  ull *d = c.scal.as<ull>() + 62;
  fill_u64(c, d, 0, 1);
  if (len) {
    k_max_u64<<<reduce_grid(len), 256, 0, c.stream>>>(in, len, d);
    KCDS_LAUNCH_CHECK();
  }
  ull *h = (ull *)c.host_scratch(sizeof(ull));
  HIP_CHECK(hipMemcpyAsync(h, d, sizeof(ull), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return *h;
}

ull sum_u64_host(Ctx &c, const ull *in, uint64_t len) {
  buf_reserve(c, c.scal, 64 * sizeof(ull));
  ull *d = c.scal.as<ull>() + 63;
  fill_u64(c, d, 0, 1);
  if (len) {
    k_sum_u64<<<reduce_grid(len), 256, 0, c.stream>>>(in, len, d);
    KCDS_LAUNCH_CHECK();
  }
  ull *h = (ull *)c.host_scratch(sizeof(ull));
  HIP_CHECK(hipMemcpyAsync(h, d, sizeof(ull), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return *h;
}

}  // namespace kcds
