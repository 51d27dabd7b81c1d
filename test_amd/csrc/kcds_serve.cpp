// Serving: continuous batching of exact k-CDS requests on one GPU, behind the
// C ABI (kcds_server_create / kcds_submit / kcds_poll / kcds_server_destroy).
//
// No counterpart in the reference, which solves one graph per process
// (main.cu:1272-1522).
//
// Requests enter per-k queues in submission order.  Batch slots each own a
// kcds context (one HIP stream and resident, grow-only device buffers); a free
// batch slot takes the k whose oldest request has waited longest and batches
// that k's queued requests, oldest first, up to an edge budget, into one
// kcds_solve_batch.  Requests submitted while batches run join the next batch
// of the next free slot, and the slots' batches overlap on the device.
//
// Admission: in a batch every graph is counted first and one whose clique
// list would exceed the per-graph budget (T * (8k + 4) bytes) is not solved
// there; it moves to the large lane, whose slots solve one graph at a time
// without a clique list (enumerating peel; dense exact search for a huge
// k'-core).  A graph with more than max_cliques cliques is answered over
// budget.  A batch that exceeds its slot's HBM budget is split in halves.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "kcds_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t ticket = 0;
  uint32_t k = 0;
  std::vector<uint32_t> off, nbr;
  double t_submit = 0, t_done = -1;
  kcds_graph_result res{};
  std::vector<uint8_t> in_set;
  uint32_t batch = 0, lane = 0;
  int code = 0;  // 0, or the kcds_status that ended it
  bool done = false;
};

}  // namespace

struct kcds_server {
  kcds_server_config cfg{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<uint32_t, std::deque<std::shared_ptr<Req>>> queues;  // k -> waiting requests
  std::deque<std::shared_ptr<Req>> large;                       // the large lane
  std::map<uint64_t, std::shared_ptr<Req>> all;
  uint64_t next_ticket = 1;
  bool stopping = false;
  uint32_t busy = 0;
  std::vector<std::thread> threads;
  std::vector<kcds_ctx *> ctxs;
  std::string err;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void finish(kcds_server *s, const std::vector<std::shared_ptr<Req>> &reqs) {
  std::lock_guard<std::mutex> g(s->mu);
  const double t = s->now();
  for (auto &r : reqs) {
    r->t_done = t;
    r->done = true;
  }
  s->cv_done.notify_all();
}

// Solves `batch` (one k) on ctx; over-budget graphs go to the large lane.
void solve_batch(kcds_server *s, kcds_ctx *ctx, std::vector<std::shared_ptr<Req>> batch) {
  const uint32_t k = batch[0]->k, G = (uint32_t)batch.size();
  std::vector<uint32_t> off(1, 0), nbr, voff(1, 0);
  for (auto &r : batch) {
    const uint32_t n = (uint32_t)r->off.size() - 1, base = voff.back(), e0 = off.back();
    for (uint32_t v = 0; v < n; ++v) off.push_back(e0 + r->off[v + 1]);
    for (uint32_t x : r->nbr) nbr.push_back(x + base);
    voff.push_back(base + n);
  }
  const uint32_t n = voff.back();
  std::vector<kcds_graph_result> res(G);
  std::vector<uint8_t> ins(s->cfg.want_sets ? (size_t)n + 1 : 0);
  // a batch graph over graph_budget is not solved here (large lane, or over
  // budget); with no graph_budget the clique cap itself bounds the batch
  const uint64_t cap_bytes = s->cfg.max_cliques ? s->cfg.max_cliques * (8ull * k + 4ull) : 0;
  (void)kcds_set_graph_budget(ctx, s->cfg.graph_budget ? s->cfg.graph_budget : cap_bytes);
  int rc = kcds_upload_csr(ctx, n, off.data(), nbr.data());
  if (rc == KCDS_OK)
    rc = kcds_solve_batch(ctx, k, G, voff.data(), res.data(), s->cfg.want_sets ? ins.data() : nullptr);
  if (rc != KCDS_OK) {
    if (G > 1) {  // over the slot's budget, or one graph beyond a limit: split and retry
      std::vector<std::shared_ptr<Req>> a(batch.begin(), batch.begin() + G / 2), b(batch.begin() + G / 2, batch.end());
      solve_batch(s, ctx, a);
      solve_batch(s, ctx, b);
      return;
    }
    if (rc == KCDS_E_NOMEM && s->cfg.large_slots) {  // a lone graph over the slot's HBM budget: the large lane
      std::lock_guard<std::mutex> g(s->mu);
      s->large.push_back(batch[0]);
      s->cv_work.notify_all();
      return;
    }
    batch[0]->code = rc;
    batch[0]->res.status = KCDS_GRAPH_OVER_BUDGET;
    batch[0]->batch = 1;
    finish(s, batch);
    return;
  }
  std::vector<std::shared_ptr<Req>> done, big;
  for (uint32_t i = 0; i < G; ++i) {
    auto &r = batch[i];
    if (res[i].status == KCDS_GRAPH_OVER_BUDGET && s->cfg.large_slots) {
      big.push_back(r);
      continue;
    }
    if (res[i].status == KCDS_GRAPH_OVER_BUDGET) r->code = KCDS_E_NOMEM;
    r->res = res[i];
    r->batch = G;
    if (s->cfg.want_sets) r->in_set.assign(ins.begin() + voff[i], ins.begin() + voff[i + 1]);
    done.push_back(r);
  }
  if (!done.empty()) finish(s, done);
  if (!big.empty()) {
    std::lock_guard<std::mutex> g(s->mu);
    for (auto &r : big) s->large.push_back(r);
    s->cv_work.notify_all();
  }
}

// One graph alone without a clique list (the solve picks the enumerating
// peel when the list would not fit), capped at max_cliques cliques.
void solve_large(kcds_server *s, kcds_ctx *ctx, const std::shared_ptr<Req> &r) {
  const uint32_t n = (uint32_t)r->off.size() - 1;
  (void)kcds_set_graph_budget(ctx, s->cfg.max_cliques ? s->cfg.max_cliques * (8ull * r->k + 4ull) : 0);
  kcds_result res{};
  std::vector<uint8_t> ins(s->cfg.want_sets ? (size_t)n + 1 : 0);
  int rc = kcds_upload_csr(ctx, n, r->off.data(), r->nbr.data());
  if (rc == KCDS_OK) rc = kcds_solve(ctx, r->k, &res, s->cfg.want_sets ? ins.data() : nullptr);
  r->batch = 1;
  r->lane = 1;
  if (rc == KCDS_OK) {
    r->res.T = res.T;
    r->res.rho_num = res.rho_num;
    r->res.rho_den = res.rho_den;
    r->res.k_prime = res.k_prime;
    r->res.core_vertices = res.core_vertices;
    r->res.ncomp = res.ncomp;
    r->res.flow_calls = res.flow_calls;
    r->res.status = KCDS_GRAPH_SOLVED;
    if (s->cfg.want_sets) {
      ins.resize(n);
      r->in_set = std::move(ins);
    }
  } else {
    r->res.status = KCDS_GRAPH_OVER_BUDGET;
    r->code = rc;
  }
  finish(s, {r});
}

// The oldest k's requests, oldest first, up to the edge budget (at least one).
std::vector<std::shared_ptr<Req>> take(kcds_server *s) {
  uint32_t best = 0;
  double bt = 0;
  bool any = false;
  for (auto &kv : s->queues)
    if (!kv.second.empty() && (!any || kv.second.front()->t_submit < bt)) {
      best = kv.first;
      bt = kv.second.front()->t_submit;
      any = true;
    }
  std::vector<std::shared_ptr<Req>> batch;
  if (!any) return batch;
  auto &q = s->queues[best];
  uint64_t edges = 0;
  while (!q.empty() && batch.size() < s->cfg.max_batch) {
    const uint64_t e = q.front()->nbr.size();
    if (!batch.empty() && edges + e > s->cfg.batch_edges) break;
    batch.push_back(q.front());
    q.pop_front();
    edges += e;
  }
  return batch;
}

void batch_worker(kcds_server *s, kcds_ctx *ctx) {
  while (true) {
    std::vector<std::shared_ptr<Req>> batch;
    {
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_work.wait(g, [&] {
        for (auto &kv : s->queues)
          if (!kv.second.empty()) return true;
        return s->stopping;
      });
      batch = take(s);
      if (batch.empty()) return;  // stopping and drained
      ++s->busy;
    }
    solve_batch(s, ctx, batch);
    std::lock_guard<std::mutex> g(s->mu);
    --s->busy;
    s->cv_work.notify_all();
  }
}

void large_worker(kcds_server *s, kcds_ctx *ctx) {
  while (true) {
    std::shared_ptr<Req> r;
    {
      std::unique_lock<std::mutex> g(s->mu);
      // the large lane ends once nothing can reach it any more
      s->cv_work.wait(g, [&] {
        if (!s->large.empty()) return true;
        if (!s->stopping) return false;
        for (auto &kv : s->queues)
          if (!kv.second.empty()) return false;
        return s->busy == 0;
      });
      if (s->large.empty()) return;
      r = s->large.front();
      s->large.pop_front();
    }
    solve_large(s, ctx, r);
  }
}

}  // namespace

extern "C" {

int kcds_server_create(kcds_server **out, const kcds_server_config *cfg) {
  if (!out || !cfg || cfg->slots == 0) return KCDS_E_INVALID;
  *out = nullptr;
  kcds_server *s = new kcds_server();
  s->cfg = *cfg;
  if (!s->cfg.batch_edges) s->cfg.batch_edges = 4000000;
  if (!s->cfg.max_batch) s->cfg.max_batch = 512;
  s->t0 = Clock::now();
  const uint32_t nctx = cfg->slots + cfg->large_slots;
  for (uint32_t i = 0; i < nctx; ++i) {
    kcds_ctx *c = nullptr;
    const int rc = kcds_create(&c, cfg->device, cfg->hbm_budget ? cfg->hbm_budget / nctx : 0);
    if (rc != KCDS_OK) {
      for (kcds_ctx *x : s->ctxs) kcds_destroy(x);
      delete s;
      return rc;
    }
    s->ctxs.push_back(c);
  }
  for (uint32_t i = 0; i < cfg->slots; ++i) s->threads.emplace_back(batch_worker, s, s->ctxs[i]);
  for (uint32_t i = cfg->slots; i < nctx; ++i) s->threads.emplace_back(large_worker, s, s->ctxs[i]);
  *out = s;
  return KCDS_OK;
}

int kcds_submit(kcds_server *s, uint32_t n, const uint32_t *offset, const uint32_t *nbr, uint32_t k,
                uint64_t *ticket) {
  if (!s || !offset || !ticket || (!nbr && offset[n] != 0)) return KCDS_E_INVALID;
  if (k < 2 || k > 8) return KCDS_E_INVALID;
  auto r = std::make_shared<Req>();
  r->k = k;
  r->off.assign(offset, offset + n + 1);
  r->nbr.assign(nbr, nbr + offset[n]);
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stopping) return KCDS_E_STATE;
  r->ticket = s->next_ticket++;
  r->t_submit = s->now();
  s->all[r->ticket] = r;
  s->queues[k].push_back(r);
  *ticket = r->ticket;
  s->cv_work.notify_all();
  return KCDS_OK;
}

int kcds_poll(kcds_server *s, uint64_t ticket, int32_t wait_ms, kcds_served *out, uint8_t *in_set) {
  if (!s) return KCDS_E_INVALID;
  std::unique_lock<std::mutex> g(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return KCDS_E_INVALID;
  std::shared_ptr<Req> r = it->second;
  if (!r->done) {
    if (wait_ms == 0) return KCDS_E_PENDING;
    auto pred = [&] { return r->done; };
    if (wait_ms < 0) s->cv_done.wait(g, pred);
    else if (!s->cv_done.wait_for(g, std::chrono::milliseconds(wait_ms), pred)) return KCDS_E_PENDING;
  }
  if (out) {
    out->result = r->res;
    out->t_submit = r->t_submit;
    out->t_done = r->t_done;
    out->batch = r->batch;
    out->lane = r->lane;
    out->code = r->code;
  }
  if (in_set && !r->in_set.empty()) memcpy(in_set, r->in_set.data(), r->in_set.size());
  s->all.erase(it);
  return KCDS_OK;
}

double kcds_server_clock(const kcds_server *s) { return s ? s->now() : 0.0; }

int kcds_server_cancel_pending(kcds_server *s) {
  if (!s) return KCDS_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  const double t = s->now();
  int nc = 0;
  auto cancel = [&](std::shared_ptr<Req> &r) {
    r->res.status = KCDS_GRAPH_OVER_BUDGET;
    r->code = KCDS_E_STATE;
    r->t_done = t;
    r->done = true;
    ++nc;
  };
  for (auto &kv : s->queues) {
    for (auto &r : kv.second) cancel(r);
    kv.second.clear();
  }
  for (auto &r : s->large) cancel(r);
  s->large.clear();
  s->cv_done.notify_all();
  s->cv_work.notify_all();
  return nc;
}

void kcds_server_destroy(kcds_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stopping = true;
    s->cv_work.notify_all();
  }
  for (auto &t : s->threads) t.join();
  for (kcds_ctx *c : s->ctxs) kcds_destroy(c);
  delete s;
}

}  // extern "C"
