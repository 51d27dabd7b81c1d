"""Seeded synthetic workloads for k-CDS (ctypes front end of libkcdsgen.so).

Configs (seed = 0x4B434453 + config index, per-graph streams seed ^ i):

  C1  ER G(2000, 0.005) + planted 40-clique, k=3
  C2  R-MAT scale 18, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05),
      + one planted 60-clique and five 30-cliques, k=4
  C3  Chung-Lu n=1e6, Pareto exponent 2.1 weights in [2, 5000] scaled to
      ~10M expected edges, + 80-vertex near-clique of edge density 0.95, k=5
  C4  4096 ER graphs, n log-uniform in [50, 5000], average degree
      log-uniform in [4, 40], + one planted clique of size U{8..30}, k=4
  C5  Poisson request stream; each request an R-MAT graph with n
      log-uniform in [1e3, 2e5], edge factor U{4..24}, k in U{3,4,5,6}
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SEED_BASE = 0x4B434453


class _G(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("m", ctypes.c_uint64),
        ("offset", ctypes.POINTER(ctypes.c_uint32)),
        ("nbr", ctypes.POINTER(ctypes.c_uint32)),
        ("nplanted", ctypes.c_uint32),
        ("planted", ctypes.POINTER(ctypes.c_uint32)),
        ("weight_scale", ctypes.c_double),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libkcdsgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_G)
        L.kcdsgen_er.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, P]
        L.kcdsgen_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.c_uint32, ctypes.c_uint64, P]
        L.kcdsgen_chunglu.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, P]
        L.kcdsgen_free.argtypes = [P]
        L.kcdsgen_write_edgelist.argtypes = [ctypes.c_char_p, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        L.kcdsgen_write_adjlist.argtypes = L.kcdsgen_write_edgelist.argtypes
        for f in ("kcdsgen_er", "kcdsgen_rmat", "kcdsgen_chunglu", "kcdsgen_write_edgelist",
                  "kcdsgen_write_adjlist"):
            getattr(L, f).restype = ctypes.c_int
        _LIB = L
    return _LIB


class Graph:
    """Undirected CSR graph: offset[n+1] uint32, nbr[2m] uint32 (sorted lists)."""

    def __init__(self, offset, nbr, planted=None, weight_scale=1.0, name=""):
        self.offset = np.ascontiguousarray(offset, dtype=np.uint32)
        self.nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
        self.n = len(self.offset) - 1
        self.m = int(self.offset[-1]) // 2
        self.planted = np.asarray(planted if planted is not None else [], dtype=np.uint32)
        self.weight_scale = weight_scale
        self.name = name

    def write_edgelist(self, path):
        rc = lib().kcdsgen_write_edgelist(path.encode(), self.n, _p32(self.offset), _p32(self.nbr))
        if rc != 0:
            raise OSError(f"cannot write {path}")

    def write_adjlist(self, path):
        rc = lib().kcdsgen_write_adjlist(path.encode(), self.n, _p32(self.offset), _p32(self.nbr))
        if rc != 0:
            raise OSError(f"cannot write {path}")


def _p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _take(g, name):
    n = g.n
    off = np.ctypeslib.as_array(g.offset, shape=(n + 1,)).copy()
    nnz = int(off[-1])
    nbr = np.ctypeslib.as_array(g.nbr, shape=(max(nnz, 1),))[:nnz].copy()
    pl = np.ctypeslib.as_array(g.planted, shape=(max(g.nplanted, 1),))[: g.nplanted].copy()
    ws = g.weight_scale
    lib().kcdsgen_free(ctypes.byref(g))
    return Graph(off, nbr, pl, ws, name)


def er(n, p, clique, seed, name="er"):
    g = _G()
    rc = lib().kcdsgen_er(n, p, clique, seed, ctypes.byref(g))
    if rc != 0:
        raise RuntimeError(f"kcdsgen_er failed ({rc})")
    return _take(g, name)


def rmat(scale, edge_factor, cliques=(), seed=0, n=None, abc=(0.57, 0.19, 0.19), name="rmat"):
    n = (1 << scale) if n is None else n
    arr = (ctypes.c_uint32 * max(1, len(cliques)))(*cliques)
    g = _G()
    rc = lib().kcdsgen_rmat(n, scale, edge_factor, abc[0], abc[1], abc[2], arr, len(cliques), seed,
                            ctypes.byref(g))
    if rc != 0:
        raise RuntimeError(f"kcdsgen_rmat failed ({rc})")
    return _take(g, name)


def chunglu(n, exponent, wmin, wmax, target_edges, nc, nc_density, seed, name="chunglu"):
    g = _G()
    rc = lib().kcdsgen_chunglu(n, exponent, wmin, wmax, target_edges, nc, nc_density, seed, ctypes.byref(g))
    if rc != 0:
        raise RuntimeError(f"kcdsgen_chunglu failed ({rc})")
    return _take(g, name)


def config(idx):
    """Returns (graph, k) for single-graph configs 1..3."""
    seed = SEED_BASE + idx
    if idx == 1:
        return er(2000, 0.005, 40, seed, name="C1 ER(2000,0.005)+K40"), 3
    if idx == 2:
        return rmat(18, 16, (60, 30, 30, 30, 30, 30), seed, name="C2 RMAT s18 ef16 +K60+5xK30"), 4
    if idx == 3:
        return chunglu(1_000_000, 2.1, 2.0, 5000.0, 10_000_000, 80, 0.95, seed,
                       name="C3 ChungLu 1e6 Pareto2.1 +near-K80"), 5
    raise ValueError(idx)


def batch_specs(count=4096, seed=SEED_BASE + 4):
    """Per-graph parameters of the batched workload (C4)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(50), math.log(5000)))))
        d = math.exp(rng.uniform(math.log(4), math.log(40)))
        s = int(rng.integers(8, 31))
        specs.append((n, min(1.0, d / max(1, n - 1)), min(s, n), (seed ^ i) & 0xFFFFFFFFFFFFFFFF))
    return specs


def batch_graph(spec, name="er"):
    n, p, s, sd = spec
    return er(n, p, s, sd, name=name)


def request_specs(rate=200.0, duration=60.0, seed=SEED_BASE + 5):
    """Poisson arrivals of R-MAT requests (C5): list of (t_arrival, n, ef, k, seed)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    i = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(math.floor(math.exp(rng.uniform(math.log(1e3), math.log(2e5)))))
        ef = int(rng.integers(4, 25))
        k = int(rng.integers(3, 7))
        out.append((t, n, ef, k, (seed ^ i) & 0xFFFFFFFFFFFFFFFF))
        i += 1
    return out


def request_graph(spec):
    _, n, ef, k, sd = spec
    scale = max(1, int(math.ceil(math.log2(n))))
    return rmat(scale, ef * n / float(1 << scale), (), sd, n=n, name=f"req n={n} ef={ef}"), k
