"""Serving loop: continuous batching of exact k-CDS requests on one GPU.

A request is (arrival time, CSR graph, k).  Requests enter per-k queues at
their arrival times.  S worker slots each own one kcds context (one HIP
stream and its resident, grow-only device buffers: CSR, DAG, cliques, flow
networks), so batches of different slots run concurrently on the device.  A
free slot takes the k whose oldest request has waited longest and batches
every queued request of that k, oldest first, up to an edge budget, into one
kcds_solve_batch (one orientation, listing, peel and flow launch for the
whole batch).  Requests that arrive while batches run wait for the next free
slot and join the next batch: continuous batching, not a queue in front of
one solver.

HBM admission:
  * a slot's context has an HBM budget (its share of the device budget);
    a batch that would exceed it is split in halves and retried;
  * every graph is admitted on its count pass: one that needs more than the
    per-graph budget for its materialised cliques (T * (8k + 4) bytes) is
    answered KCDS_GRAPH_OVER_BUDGET without being solved, and the rest of
    its batch is solved normally.

Latency is measured from a request's scheduled arrival to the moment its
batch's results are back on the host (queueing included).  The plain form
for comparison is the same loop with one slot and batches of one graph.

No counterpart in the reference (it solves one graph per process).
"""
import heapq
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import kcds


@dataclass
class Request:
    rid: int
    t_arrival: float  # scheduled arrival, seconds from the start of the run
    offset: np.ndarray
    nbr: np.ndarray
    k: int
    t_done: float = -1.0
    status: int = -1  # kcds.GRAPH_SOLVED / kcds.GRAPH_OVER_BUDGET; -2 = failed
    T: int = 0
    rho_num: int = 0
    rho_den: int = 0
    in_set: np.ndarray = None
    batch: int = 0  # size of the batch that answered it
    error: str = ""

    @property
    def n(self):
        return len(self.offset) - 1

    @property
    def edges(self):
        return int(self.offset[-1])


@dataclass
class ServeStats:
    completed: int
    solved: int
    over_budget: int
    failed: int
    wall_s: float
    solves_per_s: float
    p50_ms: float
    p99_ms: float
    mean_batch: float
    batches: int
    latencies_ms: list = field(default_factory=list)


class Server:
    """Continuous-batching request loop over `slots` device contexts."""

    def __init__(self, slots=4, batch_edges=4_000_000, max_batch=512, graph_budget=1 << 30, hbm_budget=0,
                 want_sets=False, device=0):
        self.slots = slots
        self.batch_edges = batch_edges
        self.max_batch = max_batch
        self.graph_budget = graph_budget
        self.hbm_budget = hbm_budget
        self.want_sets = want_sets
        self.device = device
        self._solvers = []
        for _ in range(slots):
            s = kcds.Solver(device, hbm_budget // slots if hbm_budget else 0)
            s.set_graph_budget(graph_budget)
            self._solvers.append(s)

    def close(self):
        for s in self._solvers:
            s.close()
        self._solvers = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- one batch on one slot -------------------------------------------------
    def _solve(self, solver, batch):
        k = batch[0].k
        try:
            res, sets = solver.solve_batch([(r.offset, r.nbr) for r in batch], k, want_set=self.want_sets)
        except kcds.KcdsError as e:
            if len(batch) > 1:  # over the slot's budget, or one graph beyond a limit: split and retry
                h = len(batch) // 2
                return self._solve(solver, batch[:h]) + self._solve(solver, batch[h:])
            r = batch[0]
            if e.code == kcds.E_NOMEM:  # a single graph over the slot's budget
                r.status = kcds.GRAPH_OVER_BUDGET
            else:
                r.status = -2
            r.error = str(e)
            r.batch = 1
            return [(r, None)]
        out = []
        for i, r in enumerate(batch):
            g = res[i]
            r.status, r.T, r.rho_num, r.rho_den = int(g.status), int(g.T), int(g.rho_num), int(g.rho_den)
            r.batch = len(batch)
            out.append((r, sets[i] if sets else None))
        return out

    # -- the loop ---------------------------------------------------------------
    def run(self, requests, plain=False):
        """Serves `requests` (iterable of Request, ascending t_arrival) in real time.

        plain=True serves them one graph per call on one slot (the comparison
        form).  Returns (requests, ServeStats)."""
        reqs = list(requests)
        queues = {}  # k -> heap of (t_arrival, rid, request)
        lock = threading.Condition()
        state = {"fed": 0, "done": 0, "batches": 0, "batched": 0}
        nslots = 1 if plain else self.slots
        t0 = time.perf_counter()

        def feeder():
            for r in reqs:
                delay = r.t_arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                with lock:
                    heapq.heappush(queues.setdefault(r.k, []), (r.t_arrival, r.rid, r))
                    state["fed"] += 1
                    lock.notify()

        def take():
            # the k whose oldest request has waited longest; its requests oldest
            # first up to the edge budget (always at least one)
            best = None
            for k, q in queues.items():
                if q and (best is None or q[0][0] < queues[best][0][0]):
                    best = k
            if best is None:
                return None
            q = queues[best]
            batch, edges = [], 0
            limit = 1 if plain else self.max_batch
            while q and len(batch) < limit:
                r = q[0][2]
                if batch and edges + r.edges > self.batch_edges:
                    break
                heapq.heappop(q)
                batch.append(r)
                edges += r.edges
            return batch

        def worker(slot):
            solver = self._solvers[slot]
            while True:
                with lock:
                    while True:
                        batch = take()
                        if batch or (state["fed"] == len(reqs)):
                            break
                        lock.wait(0.05)
                    if not batch:
                        return
                out = self._solve(solver, batch)
                t = time.perf_counter() - t0
                with lock:
                    for r, st in out:
                        r.t_done = t
                        r.in_set = st
                    state["done"] += len(out)
                    state["batches"] += 1
                    state["batched"] += len(out)
                    lock.notify_all()

        threads = [threading.Thread(target=feeder, daemon=True)]
        threads += [threading.Thread(target=worker, args=(i,), daemon=True) for i in range(nslots)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        wall = time.perf_counter() - t0
        lat = [(r.t_done - r.t_arrival) * 1e3 for r in reqs if r.status == kcds.GRAPH_SOLVED]
        solved = len(lat)
        stats = ServeStats(
            completed=len(reqs), solved=solved,
            over_budget=sum(r.status == kcds.GRAPH_OVER_BUDGET for r in reqs),
            failed=sum(r.status == -2 for r in reqs), wall_s=wall,
            solves_per_s=solved / wall if wall > 0 else 0.0,
            p50_ms=float(np.percentile(lat, 50)) if lat else 0.0,
            p99_ms=float(np.percentile(lat, 99)) if lat else 0.0,
            mean_batch=state["batched"] / max(1, state["batches"]), batches=state["batches"],
            latencies_ms=lat)
        return reqs, stats


def c5_requests(count=None, duration=None, time_scale=1.0, max_n=None, threads=8):
    """The C5 request stream (Poisson 200/s, R-MAT, k in 3..6), generated in order.

    count / duration bound the sample (first requests of the stream);
    time_scale stretches arrival times; max_n skips requests above that size.
    Graphs are generated on `threads` host threads (the generator is native
    and releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    from . import gen
    picked = []
    for i, spec in enumerate(gen.request_specs()):
        if count is not None and len(picked) >= count:
            break
        if duration is not None and spec[0] >= duration:
            break
        if max_n is not None and spec[1] > max_n:
            continue
        picked.append((i, spec))

    def make(item):
        i, spec = item
        g, k = gen.request_graph(spec)
        return Request(rid=i, t_arrival=spec[0] * time_scale, offset=np.asarray(g.offset, np.uint32),
                       nbr=np.asarray(g.nbr, np.uint32), k=k)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(make, picked))
