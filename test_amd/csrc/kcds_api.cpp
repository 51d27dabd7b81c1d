// C ABI of libkcds (include/kcds.h): argument checking, the context and its
// device buffers, and the stage sequence of a solve.  Kernels live in the
// k_*.hip files; this file only launches them through kcds_internal.hpp.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "kcds_internal.hpp"

namespace kcds {

void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error{code, buf};
}

void buf_reserve(Ctx &c, DevBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return;
  bytes = std::max<size_t>(bytes, 256);
  bytes = (bytes + 255) & ~(size_t)255;
  if (b.p) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
    HIP_CHECK(hipFree(b.p));
    c.used -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }
  if (c.used + bytes > c.budget)
    fail(KCDS_E_NOMEM, "HBM budget exceeded: need %zu more bytes, %llu of %llu in use", bytes,
         (unsigned long long)c.used, (unsigned long long)c.budget);
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    fail(KCDS_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  c.used += bytes;
}

void buf_free(Ctx &c, DevBuf &b) {
  if (b.p) {
    (void)hipStreamSynchronize(c.stream);
    (void)hipFree(b.p);
    c.used -= b.bytes;
  }
  b.p = nullptr;
  b.bytes = 0;
}

void *Ctx::host_scratch(size_t bytes) {
  if (bytes > pinned_bytes) {
    if (pinned) HIP_CHECK(hipHostFree(pinned));
    pinned = nullptr;
    pinned_bytes = std::max<size_t>(bytes, 4096);
    HIP_CHECK(hipHostMalloc(&pinned, pinned_bytes, hipHostMallocDefault));
  }
  return pinned;
}

void gather_by_rank_u64(Ctx &c, const ull *src_new, ull *dst_orig);
void gather_by_rank_u32(Ctx &c, const uint32_t *src_new, uint32_t *dst_orig);
void rank_from_listing_order(Ctx &c, const uint32_t *lo_dev);
void maxflow_component(Ctx &c, uint32_t comp_id, uint64_t p, uint64_t q, int64_t *flow, uint8_t *side);

static void reset_after_graph(Ctx &c) {
  c.oriented = false;
  c.counted = false;
  c.listed = false;
  c.peeled = false;
  c.have_comp = false;
  c.T = 0;
}

}  // namespace kcds

using namespace kcds;

template <typename F> static int guarded(kcds_ctx *ctx, F &&f) {
  if (!ctx) return KCDS_E_INVALID;
  try {
    HIP_CHECK(hipSetDevice(ctx->c.device));
    f(ctx->c);
    ctx->c.err.clear();
    return KCDS_OK;
  } catch (const Error &e) {
    ctx->c.err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    ctx->c.err = "host allocation failed";
    return KCDS_E_NOMEM;
  }
}

extern "C" {

int kcds_create(kcds_ctx **out, int device, uint64_t hbm_budget_bytes) {
  if (!out) return KCDS_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KCDS_E_HIP;
  if (device < 0 || device >= ndev) return KCDS_E_INVALID;
  kcds_ctx *ctx = new kcds_ctx();
  Ctx &c = ctx->c;
  c.device = device;
  try {
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    c.num_cus = prop.multiProcessorCount;
    size_t freeb = 0, total = 0;
    HIP_CHECK(hipMemGetInfo(&freeb, &total));
    c.budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(freeb * 0.9);
  } catch (const Error &e) {
    delete ctx;
    return e.code;
  }
  *out = ctx;
  return KCDS_OK;
}

void kcds_destroy(kcds_ctx *ctx) {
  if (!ctx) return;
  Ctx &c = ctx->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  CompData &g = c.cd;
  FlowBufs &f = c.fb;
  DevBuf *bufs[] = {&c.off, &c.nbr, &c.rank, &c.inv, &c.doff, &c.dnbr, &c.roots, &c.qctr, &c.cnt, &c.scal,
                    &c.rootT, &c.lsum, &c.cliques, &c.incoff, &c.inc, &c.core, &c.status, &c.foff, &c.fnbr, &c.erem, &c.kdoff, &c.kdnbr, &c.comp, &c.tmp0,
                    &c.tmp1, &c.tmp2, &c.tmp3, &c.gvoff, &c.gid, &c.gpeel, &c.tlist, &c.inset, &c.wide, &c.fofs, &c.fstore,
                    &g.par, &g.flag, &g.dense, &g.cid, &g.local, &g.vcount, &g.voff, &g.verts, &g.ccount,
                    &g.coff, &g.cfill, &g.ccl, &f.scal, &f.cside, &f.g_off, &f.g_head, &f.g_rev, &f.g_res,
                    &f.g_excess, &f.g_height, &f.g_fr, &f.g_ctl, &c.gsum, &f.g_nch, &f.g_choff, &f.g_chv,
                    &f.g_vmin, &f.g_csum, &f.g_esnap, &f.g_cmin, &f.g_hsnap, &f.b_off, &f.b_excess, &f.b_height, &f.b_fr0, &f.b_fr1, &f.b_head, &f.b_rev, &f.b_res, &f.b_tasks};
  for (DevBuf *b : bufs) buf_free(c, *b);
  if (c.pinned) (void)hipHostFree(c.pinned);
  for (int i = 0; i < 3; ++i) {
    if (c.lstream[i]) (void)hipStreamDestroy(c.lstream[i]);
    if (c.ljoin[i]) (void)hipEventDestroy(c.ljoin[i]);
  }
  if (c.lfork) (void)hipEventDestroy(c.lfork);
  for (int i = 0; i < 2; ++i)
    if (c.ltime[i]) (void)hipEventDestroy(c.ltime[i]);
  if (c.ffork) (void)hipEventDestroy(c.ffork);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
}

const char *kcds_last_error(const kcds_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int kcds_upload_csr(kcds_ctx *ctx, uint32_t n, const uint32_t *offset, const uint32_t *nbr) {
  return guarded(ctx, [&](Ctx &c) {
    if (!offset || (!nbr && offset[n] != 0)) fail(KCDS_E_INVALID, "null CSR array");
    if (offset[0] != 0) fail(KCDS_E_INVALID, "offset[0] must be 0");
    const uint64_t nnz = offset[n];
    // every list must hold ids < n, no self-loop, strictly ascending (the
    // loaders and generators produce this form; duplicates would count
    // cliques twice)
    uint32_t maxdeg = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (offset[v + 1] < offset[v]) fail(KCDS_E_INVALID, "offset not monotone at %u", v);
      maxdeg = std::max(maxdeg, offset[v + 1] - offset[v]);
      for (uint32_t i = offset[v]; i < offset[v + 1]; ++i) {
        uint32_t u = nbr[i];
        if (u >= n) fail(KCDS_E_INVALID, "neighbour id %u of vertex %u out of range", u, v);
        if (u == v) fail(KCDS_E_INVALID, "self-loop at vertex %u", v);
        if (i > offset[v] && nbr[i - 1] >= u)
          fail(KCDS_E_INVALID, "neighbour list of vertex %u not strictly ascending", v);
      }
    }
    reset_after_graph(c);
    c.have_graph = false;
    buf_reserve(c, c.off, sizeof(uint32_t) * ((size_t)n + 1));
    buf_reserve(c, c.nbr, sizeof(uint32_t) * (nnz + 1));
    HIP_CHECK(hipMemcpyAsync(c.off.p, offset, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, c.stream));
    if (nnz)
      HIP_CHECK(hipMemcpyAsync(c.nbr.p, nbr, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, c.stream));
    c.sync();
    c.n = n;
    c.nnz = nnz;
    c.maxdeg = maxdeg;
    c.ngraphs = 1;
    c.h_gvoff.assign({0u, n});
    c.have_graph = true;
  });
}

int kcds_upload_edges(kcds_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *uv, uint64_t *m_out) {
  return guarded(ctx, [&](Ctx &c) {
    if (m && !uv) fail(KCDS_E_INVALID, "null edge array");
    reset_after_graph(c);
    c.have_graph = false;
    build_csr_from_edges(c, n, m, uv);
    c.n = n;
    c.ngraphs = 1;
    c.h_gvoff.assign({0u, n});
    c.have_graph = true;
    if (m_out) *m_out = c.nnz / 2;
  });
}

int kcds_download_csr(kcds_ctx *ctx, uint32_t *offset, uint32_t *nbr, uint64_t cap, uint64_t *nnz) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.have_graph) fail(KCDS_E_STATE, "kcds_download_csr before a graph upload");
    if (nnz) *nnz = c.nnz;
    if (offset)
      HIP_CHECK(hipMemcpyAsync(offset, c.off.p, sizeof(uint32_t) * ((size_t)c.n + 1), hipMemcpyDeviceToHost,
                               c.stream));
    const uint64_t take = std::min<uint64_t>(cap, c.nnz);
    if (nbr && take)
      HIP_CHECK(hipMemcpyAsync(nbr, c.nbr.p, sizeof(uint32_t) * take, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int kcds_orient(kcds_ctx *ctx, const uint32_t *listing_order) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.have_graph) fail(KCDS_E_STATE, "kcds_orient before kcds_upload_csr");
    reset_after_graph(c);
    const uint32_t n = c.n;
    if (listing_order) {
      std::vector<uint8_t> seen(n, 0);
      for (uint32_t v = 0; v < n; ++v) {
        uint32_t x = listing_order[v];
        if (x < 1 || x > n || seen[x - 1]) fail(KCDS_E_INVALID, "listing_order is not a permutation of 1..n");
        seen[x - 1] = 1;
      }
      buf_reserve(c, c.rank, sizeof(uint32_t) * (size_t)n + 4);
      buf_reserve(c, c.inv, sizeof(uint32_t) * (size_t)n + 4);
      buf_reserve(c, c.tmp0, sizeof(uint32_t) * (size_t)n + 4);
      HIP_CHECK(hipMemcpyAsync(c.tmp0.p, listing_order, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c.stream));
      rank_from_listing_order(c, c.tmp0.as<uint32_t>());
      c.order_replaced = !orient_from_rank(c, true);
      c.order_kind = 3;
      if (c.order_replaced) orient_gpu(c);  // an out-list wider than the frames: the GPU order
    } else {
      c.order_replaced = false;
      orient_gpu(c);
    }
    c.oriented = true;
  });
}

int kcds_count(kcds_ctx *ctx, uint32_t k, uint64_t *T, int64_t *per_vertex) {
  return guarded(ctx, [&](Ctx &c) {
    list_count(c, k);
    if (T) *T = c.T;
    if (per_vertex && c.n) {
      buf_reserve(c, c.tmp1, sizeof(ull) * (size_t)c.n);
      gather_by_rank_u64(c, c.cnt.as<ull>(), c.tmp1.as<ull>());
      HIP_CHECK(hipMemcpyAsync(per_vertex, c.tmp1.p, sizeof(int64_t) * c.n, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
  });
}

int kcds_list(kcds_ctx *ctx, uint32_t k, uint64_t *T) {
  return guarded(ctx, [&](Ctx &c) {
    list_materialize(c, k);
    if (T) *T = c.T;
  });
}

int kcds_download_cliques(kcds_ctx *ctx, uint32_t *out, uint64_t cap, uint64_t *T) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.listed) fail(KCDS_E_STATE, "kcds_download_cliques before kcds_list");
    const uint32_t k = c.k;
    if (T) *T = c.T;
    if (!out || !cap || !c.T) return;
    std::vector<uint32_t> cl(c.T * k), inv(c.n);
    HIP_CHECK(hipMemcpyAsync(cl.data(), c.cliques.p, sizeof(uint32_t) * c.T * k, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(inv.data(), c.inv.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (auto &x : cl) x = inv[x];
    for (uint64_t i = 0; i < c.T; ++i) std::sort(cl.begin() + i * k, cl.begin() + (i + 1) * k);
    std::vector<uint64_t> idx(c.T);
    std::iota(idx.begin(), idx.end(), 0ull);
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
      return std::lexicographical_compare(cl.begin() + a * k, cl.begin() + (a + 1) * k, cl.begin() + b * k,
                                          cl.begin() + (b + 1) * k);
    });
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, c.T); ++i)
      memcpy(out + i * k, cl.data() + idx[i] * k, sizeof(uint32_t) * k);
  });
}

int kcds_peel(kcds_ctx *ctx, int64_t *clique_core, kcds_peel_stats *stats) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.counted) fail(KCDS_E_STATE, "kcds_peel before kcds_count or kcds_list");
    if (c.listed && !(c.tune_flags & KCDS_TUNE_ENUM_PEEL)) {
      c.enum_peel = false;
      peel_run(c);
    } else {
      epeel_run(c);
      core_materialize(c);  // the k'-core's cliques (the statistics report their number)
    }
    if (stats) *stats = c.peel;
    if (clique_core && c.n) {
      buf_reserve(c, c.tmp1, sizeof(int64_t) * (size_t)c.n);
      gather_by_rank_u64(c, c.core.as<ull>(), c.tmp1.as<ull>());  // int64 core numbers, copied bitwise
      HIP_CHECK(hipMemcpyAsync(clique_core, c.tmp1.p, sizeof(int64_t) * c.n, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
  });
}

int kcds_components(kcds_ctx *ctx, uint32_t *comp, uint32_t *ncomp) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.peeled) fail(KCDS_E_STATE, "kcds_components before kcds_peel");
    if (c.enum_peel && !c.core_counted) core_materialize(c);
    if (c.enum_peel && !c.core_listed)
      fail(KCDS_E_LIMIT, "the k'-core holds %llu cliques, too many to list in HBM",
           (unsigned long long)c.peel.core_cliques);
    components_run(c);
    if (ncomp) *ncomp = c.ncomp;
    if (comp && c.n) {
      buf_reserve(c, c.tmp1, sizeof(uint32_t) * (size_t)c.n);
      gather_by_rank_u32(c, c.comp.as<uint32_t>(), c.tmp1.as<uint32_t>());
      HIP_CHECK(hipMemcpyAsync(comp, c.tmp1.p, sizeof(uint32_t) * c.n, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
  });
}

int kcds_maxflow(kcds_ctx *ctx, uint32_t comp_id, uint64_t p, uint64_t q, int64_t *flow_value, uint8_t *side) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.have_comp) fail(KCDS_E_STATE, "kcds_maxflow before kcds_components");
    if (q == 0) fail(KCDS_E_INVALID, "alpha denominator q must be positive");
    maxflow_component(c, comp_id, p, q, flow_value, side);
  });
}

int kcds_solve(kcds_ctx *ctx, uint32_t k, kcds_result *res, uint8_t *in_set) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.have_graph) fail(KCDS_E_STATE, "kcds_solve before kcds_upload_csr");
    if (k < 2 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
    c.ngraphs = 1;
    c.h_gvoff.assign({0u, c.n});
    solve_graphs(c, k, res, nullptr, in_set);
  });
}

int kcds_solve_batch(kcds_ctx *ctx, uint32_t k, uint32_t ngraphs, const uint32_t *graph_voff,
                     kcds_graph_result *results, uint8_t *in_set) {
  return guarded(ctx, [&](Ctx &c) {
    if (!c.have_graph) fail(KCDS_E_STATE, "kcds_solve_batch before kcds_upload_csr");
    if (k < 2 || k > 8) fail(KCDS_E_INVALID, "k = %u outside 2..8", k);
    if (!graph_voff || !results || ngraphs == 0) fail(KCDS_E_INVALID, "null batch partition or results");
    if (graph_voff[0] != 0 || graph_voff[ngraphs] != c.n)
      fail(KCDS_E_INVALID, "graph_voff must run from 0 to n");
    for (uint32_t g = 0; g < ngraphs; ++g)
      if (graph_voff[g + 1] < graph_voff[g]) fail(KCDS_E_INVALID, "graph_voff not monotone at %u", g);
    c.ngraphs = ngraphs;
    c.h_gvoff.assign(graph_voff, graph_voff + ngraphs + 1);
    struct Restore {
      Ctx &c;
      ~Restore() {
        c.ngraphs = 1;
        c.h_gvoff.assign({0u, c.n});
      }
    } restore{c};
    solve_graphs(c, k, nullptr, results, in_set);
  });
}

int kcds_sync(kcds_ctx *ctx) {
  return guarded(ctx, [&](Ctx &c) { c.sync(); });
}

int kcds_set_tuning(kcds_ctx *ctx, uint32_t list_waves_per_cu, uint32_t flags) {
  return guarded(ctx, [&](Ctx &c) {
    c.list_waves_per_cu = list_waves_per_cu;
    c.tune_flags = flags;
  });
}

int kcds_get_list_info(kcds_ctx *ctx, kcds_list_info *info) {
  return guarded(ctx, [&](Ctx &c) {
    if (!info) fail(KCDS_E_INVALID, "null info");
    info->n = c.n;
    info->arcs = c.darcs;
    info->w2 = c.w2;
    info->max_outdeg = c.dmax;
    info->kernel_ms = c.list_kernel_ms;
    info->kernel_launches = c.list_kernel_launches;
    info->order_replaced = c.order_replaced ? 1u : 0u;
    info->order_kind = c.order_kind;
    for (int i = 0; i < 4; ++i) {
      info->class_roots[i] = c.class_count[i];
      info->waves_per_root[i] = list_waves_per_root(c, i, c.k ? c.k : 4);
    }
  });
}

int kcds_set_graph_budget(kcds_ctx *ctx, uint64_t bytes) {
  return guarded(ctx, [&](Ctx &c) { c.graph_budget = bytes; });
}

int kcds_set_dense_budget(kcds_ctx *ctx, uint32_t major_cycles) {
  return guarded(ctx, [&](Ctx &c) { c.dense_budget = major_cycles; });
}

int kcds_set_list_shape(kcds_ctx *ctx, const uint32_t *waves_per_root) {
  return guarded(ctx, [&](Ctx &c) {
    for (int i = 0; i < 4; ++i) {
      const uint32_t w = waves_per_root ? waves_per_root[i] : 0;
      if (w > 8) fail(KCDS_E_INVALID, "waves per root %u above 8", w);
      c.list_waves_per_root[i] = w;
    }
  });
}

int kcds_autotune_list(kcds_ctx *ctx, uint32_t k, uint32_t *waves_per_root, float *class_ms) {
  return guarded(ctx, [&](Ctx &c) {
    list_autotune(c, k);
    for (int i = 0; i < 4; ++i) {
      if (waves_per_root) waves_per_root[i] = list_waves_per_root(c, i, k);
      if (class_ms) class_ms[i] = c.tune_ms[i];
    }
  });
}

}  // extern "C"
