"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product library never uses it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U32P = ctypes.POINTER(ctypes.c_uint32)
I64P = ctypes.POINTER(ctypes.c_int64)
I32P = ctypes.POINTER(ctypes.c_int32)
U8P = ctypes.POINTER(ctypes.c_uint8)


class PeelStats(ctypes.Structure):
    _fields_ = [("max_num", ctypes.c_uint64), ("max_den", ctypes.c_uint64), ("k_prime", ctypes.c_uint32),
                ("core_size", ctypes.c_uint32), ("core_cliques", ctypes.c_uint64),
                ("max_core", ctypes.c_int32), ("levels", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("T", ctypes.c_uint64), ("rho_num", ctypes.c_uint64), ("rho_den", ctypes.c_uint64),
                ("k_prime", ctypes.c_uint32), ("core_vertices", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("flow_calls", ctypes.c_uint32), ("peel", PeelStats)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.oracle_load_adjlist.argtypes = [ctypes.c_char_p, U32P, U32P, ctypes.POINTER(U32P), ctypes.POINTER(U32P)]
        L.oracle_load_edgelist.argtypes = [ctypes.c_char_p, ctypes.c_uint32, U32P, U32P, ctypes.POINTER(U32P),
                                           ctypes.POINTER(U32P)]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_core_decompose.argtypes = [ctypes.c_uint32, U32P, U32P, U32P, U32P, U32P, U32P]
        L.oracle_count.argtypes = [ctypes.c_uint32, U32P, U32P, ctypes.c_uint32, I64P, ctypes.c_int]
        L.oracle_count.restype = ctypes.c_uint64
        L.oracle_count_roots.argtypes = [ctypes.c_uint32, U32P, U32P, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_int]
        L.oracle_count_roots.restype = ctypes.c_uint64
        L.oracle_list.argtypes = [ctypes.c_uint32, U32P, U32P, ctypes.c_uint32, U32P, ctypes.c_uint64]
        L.oracle_list.restype = ctypes.c_uint64
        L.oracle_peel.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, U32P, I32P, I32P,
                                  ctypes.POINTER(PeelStats)]
        L.oracle_clique_maxflow.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, U32P,
                                            ctypes.c_int64, ctypes.c_int64, U8P]
        L.oracle_clique_maxflow.restype = ctypes.c_int64
        L.oracle_solve.argtypes = [ctypes.c_uint32, U32P, U32P, ctypes.c_uint32, ctypes.POINTER(Result), I64P,
                                   I32P, U32P, U8P]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def _csr(offset, nbr):
    return (np.ascontiguousarray(offset, dtype=np.uint32), np.ascontiguousarray(nbr, dtype=np.uint32))


def _take_csr(n, off, nbr):
    o = np.ctypeslib.as_array(off, shape=(n.value + 1,)).copy()
    nn = int(o[-1])
    b = np.ctypeslib.as_array(nbr, shape=(max(nn, 1),))[:nn].copy()
    lib().oracle_free(ctypes.cast(off, ctypes.c_void_p))
    lib().oracle_free(ctypes.cast(nbr, ctypes.c_void_p))
    return o, b


def load_adjlist(path):
    n, m, off, nbr = ctypes.c_uint32(), ctypes.c_uint32(), U32P(), U32P()
    rc = lib().oracle_load_adjlist(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(off),
                                   ctypes.byref(nbr))
    if rc != 0:
        raise ValueError(f"oracle_load_adjlist({path}) failed: {rc}")
    o, b = _take_csr(n, off, nbr)
    return o, b, m.value


def load_edgelist(path, n_hint=0):
    n, m, off, nbr = ctypes.c_uint32(), ctypes.c_uint32(), U32P(), U32P()
    rc = lib().oracle_load_edgelist(path.encode(), n_hint, ctypes.byref(n), ctypes.byref(m), ctypes.byref(off),
                                    ctypes.byref(nbr))
    if rc != 0:
        raise ValueError(f"oracle_load_edgelist({path}) failed: {rc}")
    o, b = _take_csr(n, off, nbr)
    return o, b, m.value


def core_decompose(offset, nbr):
    offset, nbr = _csr(offset, nbr)
    n = len(offset) - 1
    core = np.zeros(n, np.uint32)
    seq = np.zeros(n, np.uint32)
    lo = np.zeros(n, np.uint32)
    kmax = ctypes.c_uint32()
    lib().oracle_core_decompose(n, _p(offset, U32P), _p(nbr, U32P), _p(core, U32P), _p(seq, U32P),
                                _p(lo, U32P), ctypes.byref(kmax))
    return core, seq, lo, kmax.value


def count(offset, nbr, k, per_vertex=True, threads=0):
    offset, nbr = _csr(offset, nbr)
    n = len(offset) - 1
    pv = np.zeros(n, np.int64) if per_vertex else None
    T = lib().oracle_count(n, _p(offset, U32P), _p(nbr, U32P), k, _p(pv, I64P) if pv is not None else None,
                           threads)
    return int(T), pv


def count_roots(offset, nbr, k, begin, end, threads=0):
    offset, nbr = _csr(offset, nbr)
    n = len(offset) - 1
    return int(lib().oracle_count_roots(n, _p(offset, U32P), _p(nbr, U32P), k, begin, end, threads))


def list_cliques(offset, nbr, k):
    offset, nbr = _csr(offset, nbr)
    n = len(offset) - 1
    T, _ = count(offset, nbr, k, per_vertex=False)
    out = np.zeros((max(T, 1), k), np.uint32)
    T2 = lib().oracle_list(n, _p(offset, U32P), _p(nbr, U32P), k, _p(out, U32P), T)
    assert T2 == T
    return out[:T]


def peel(n, k, cliques):
    cl = np.ascontiguousarray(cliques, dtype=np.uint32).reshape(-1, k) if len(cliques) else np.zeros((0, k), np.uint32)
    T = cl.shape[0]
    core = np.zeros(n, np.int32)
    status = np.zeros(max(T, 1), np.int32)
    st = PeelStats()
    lib().oracle_peel(n, k, T, _p(cl, U32P), _p(core, I32P), _p(status, I32P), ctypes.byref(st))
    return core, status[:T], st


def clique_maxflow(nc, k, cliques, p, q, want_side=False):
    cl = np.ascontiguousarray(cliques, dtype=np.uint32).reshape(-1, k)
    side = np.zeros(max(nc, 1), np.uint8) if want_side else None
    f = lib().oracle_clique_maxflow(nc, k, cl.shape[0], _p(cl, U32P), p, q,
                                    _p(side, U8P) if side is not None else None)
    return (int(f), side[:nc]) if want_side else int(f)


def solve(offset, nbr, k):
    """Returns (Result, per_vertex counts, clique cores, component ids, in_set)."""
    offset, nbr = _csr(offset, nbr)
    n = len(offset) - 1
    res = Result()
    pv = np.zeros(n, np.int64)
    core = np.zeros(n, np.int32)
    comp = np.zeros(n, np.uint32)
    ins = np.zeros(max(n, 1), np.uint8)
    rc = lib().oracle_solve(n, _p(offset, U32P), _p(nbr, U32P), k, ctypes.byref(res), _p(pv, I64P),
                            _p(core, I32P), _p(comp, U32P), _p(ins, U8P))
    if rc != 0:
        raise RuntimeError(f"oracle_solve failed: {rc}")
    return res, pv, core, comp, ins[:n].astype(bool)
