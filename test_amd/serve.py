"""Serving loop: continuous batching of exact k-CDS requests on one GPU.

A thin front end of the native server in libkcds.so (kcds_server_create /
kcds_submit / kcds_poll, include/kcds.h).  The server owns `slots` batch
contexts and `large_slots` large-lane contexts, each with its own host
thread and HIP stream, so batches of different slots overlap on the device.
Requests wait in per-k queues; a free batch slot solves the waiting requests
of the k whose oldest request has waited longest (oldest first, up to an edge
budget) in one kcds_solve_batch.  Requests that arrive while batches run join
the next batch: continuous batching, not a queue in front of one solver.

Admission:
  * a graph whose clique list would exceed `graph_budget` bytes
    (T * (8k + 4)) is not solved in a batch: it moves to the large lane,
    which solves one graph at a time without a clique list (enumerating peel,
    dense exact search for a huge k'-core);
  * a graph with more than `max_cliques` cliques is answered over budget;
  * a batch over its slot's HBM budget is split in halves.

Latency runs from a request's scheduled arrival to the moment the server
finished it (queueing included).  The plain form for comparison is the same
server with one batch slot and batches of one graph.

No counterpart in the reference (it solves one graph per process).
"""
import ctypes
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
    t_submit: float = -1.0
    status: int = -1  # kcds.GRAPH_SOLVED / kcds.GRAPH_OVER_BUDGET; -2 = failed; -3 = not answered
    code: int = 0     # kcds status that left it unsolved (E_NOMEM: over budget, E_STATE: cancelled)
    T: int = 0
    rho_num: int = 0
    rho_den: int = 0
    in_set: np.ndarray = None
    batch: int = 0  # size of the batch that answered it
    lane: int = 0   # 0 batch slot, 1 large lane
    error: str = ""

    @property
    def n(self):
        return len(self.offset) - 1

    @property
    def edges(self):
        return int(self.offset[-1])


@dataclass
class ServeStats:
    offered: int
    solved: int
    over_budget: int
    cancelled: int
    failed: int
    wall_s: float
    solves_per_s: float
    p50_ms: float       # over solved requests
    p99_ms: float
    p50_all_ms: float   # over every offered request (unanswered ones at the end of the run)
    p99_all_ms: float
    mean_batch: float
    batches: int
    large_lane: int
    max_pending: int
    latencies_ms: list = field(default_factory=list)

    @property
    def completed(self):
        return self.solved + self.over_budget + self.failed


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("slots", ctypes.c_uint32), ("large_slots", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("batch_edges", ctypes.c_uint64), ("graph_budget", ctypes.c_uint64),
                ("max_cliques", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64), ("want_sets", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class Served(ctypes.Structure):
    _fields_ = [("result", kcds.GraphResult), ("t_submit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("batch", ctypes.c_uint32), ("lane", ctypes.c_uint32), ("code", ctypes.c_int32),
                ("pad", ctypes.c_uint32)]


def _lib():
    L = kcds.lib()
    if not getattr(L, "_serve_bound", False):
        V = ctypes.c_void_p
        L.kcds_server_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(ServerConfig)]
        L.kcds_submit.argtypes = [V, ctypes.c_uint32, kcds.U32P, kcds.U32P, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_uint64)]
        L.kcds_poll.argtypes = [V, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(Served), kcds.U8P]
        L.kcds_server_clock.argtypes = [V]
        L.kcds_server_clock.restype = ctypes.c_double
        L.kcds_server_cancel_pending.argtypes = [V]
        L.kcds_server_destroy.argtypes = [V]
        L.kcds_server_destroy.restype = None
        L._serve_bound = True
    return L


class Server:
    """Continuous-batching request loop (the native server of libkcds.so)."""

    def __init__(self, slots=4, large_slots=1, batch_edges=4_000_000, max_batch=512, graph_budget=1 << 30,
                 max_cliques=0, hbm_budget=0, want_sets=False, device=0):
        self.cfg = dict(slots=slots, large_slots=large_slots, batch_edges=batch_edges, max_batch=max_batch,
                        graph_budget=graph_budget, max_cliques=max_cliques, hbm_budget=hbm_budget,
                        want_sets=want_sets, device=device)

    def close(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, requests, plain=False, deadline_s=None):
        """Serves `requests` (ascending t_arrival) in real time.

        plain=True: one batch slot, batches of one graph (the comparison form).
        deadline_s: seconds after the last arrival to wait; requests not
        started by then are cancelled and the run waits for those in flight.
        Returns (requests, ServeStats)."""
        L = _lib()
        c = self.cfg
        cfg = ServerConfig(device=c["device"], slots=1 if plain else c["slots"], large_slots=c["large_slots"],
                           max_batch=1 if plain else c["max_batch"], batch_edges=c["batch_edges"],
                           graph_budget=c["graph_budget"], max_cliques=c["max_cliques"],
                           hbm_budget=c["hbm_budget"], want_sets=1 if c["want_sets"] else 0)
        h = ctypes.c_void_p()
        rc = L.kcds_server_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != 0:
            raise kcds.KcdsError(rc, "kcds_server_create failed")
        reqs = list(requests)
        tickets = {}
        try:
            t0 = time.perf_counter()
            c0 = L.kcds_server_clock(h)
            for r in reqs:
                delay = r.t_arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                off = np.ascontiguousarray(r.offset, np.uint32)
                nbr = np.ascontiguousarray(r.nbr, np.uint32)
                t = ctypes.c_uint64()
                rc = L.kcds_submit(h, r.n, off.ctypes.data_as(kcds.U32P), nbr.ctypes.data_as(kcds.U32P), r.k,
                                   ctypes.byref(t))
                if rc != 0:
                    r.status, r.error = -2, f"submit failed ({rc})"
                    continue
                tickets[r.rid] = t.value
            if deadline_s is not None:
                # after the deadline, requests no slot has started are cancelled
                timer = threading.Timer(max(0.0, reqs[-1].t_arrival + deadline_s - (time.perf_counter() - t0))
                                        if reqs else 0.0, lambda: L.kcds_server_cancel_pending(h))
                timer.start()
            out = Served()
            for r in reqs:
                if r.rid not in tickets:
                    continue
                ins = np.zeros(max(r.n, 1), np.uint8) if c["want_sets"] else None
                rc = L.kcds_poll(h, tickets[r.rid], -1, ctypes.byref(out),
                                 ins.ctypes.data_as(kcds.U8P) if ins is not None else None)
                if rc != 0:
                    r.status, r.error = -2, f"poll failed ({rc})"
                    continue
                g = out.result
                r.status, r.T, r.rho_num, r.rho_den = int(g.status), int(g.T), int(g.rho_num), int(g.rho_den)
                r.code, r.batch, r.lane = int(out.code), int(out.batch), int(out.lane)
                r.t_done = out.t_done - c0
                r.t_submit = out.t_submit - c0
                if ins is not None and r.status == kcds.GRAPH_SOLVED:
                    r.in_set = ins[: r.n].astype(bool)
            if deadline_s is not None:
                timer.cancel()
            wall = time.perf_counter() - t0
        finally:
            L.kcds_server_destroy(h)
        solved = [r for r in reqs if r.status == kcds.GRAPH_SOLVED]
        lat = [(r.t_done - r.t_arrival) * 1e3 for r in solved]
        # every offered request: unanswered ones count as waiting until the end of the run
        lat_all = [((r.t_done if r.t_done >= 0 and r.code != kcds.E_STATE else wall) - r.t_arrival) * 1e3
                   for r in reqs]
        # queue depth: requests submitted and not yet answered, at each submission
        ev = sorted([(r.t_submit, 1) for r in reqs if r.t_submit >= 0] +
                    [(r.t_done, -1) for r in reqs if r.t_submit >= 0 and r.t_done >= 0])
        depth = max_depth = 0
        for _, d in ev:
            depth += d
            max_depth = max(max_depth, depth)
        in_batches = [r for r in reqs if r.batch > 0 and r.lane == 0]
        nb = sum(1.0 / r.batch for r in in_batches)
        stats = ServeStats(
            offered=len(reqs), solved=len(solved),
            over_budget=sum(r.status == kcds.GRAPH_OVER_BUDGET and r.code != kcds.E_STATE for r in reqs),
            cancelled=sum(r.code == kcds.E_STATE for r in reqs),
            failed=sum(r.status == -2 for r in reqs), wall_s=wall,
            solves_per_s=len(solved) / wall if wall > 0 else 0.0,
            p50_ms=float(np.percentile(lat, 50)) if lat else 0.0,
            p99_ms=float(np.percentile(lat, 99)) if lat else 0.0,
            p50_all_ms=float(np.percentile(lat_all, 50)) if lat_all else 0.0,
            p99_all_ms=float(np.percentile(lat_all, 99)) if lat_all else 0.0,
            mean_batch=len(in_batches) / max(1.0, nb),
            batches=int(round(nb)), large_lane=sum(r.lane == 1 for r in reqs), max_pending=max_depth,
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
