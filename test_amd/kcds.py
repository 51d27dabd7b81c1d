"""Python front end of libkcds.so (the C ABI in include/kcds.h).

Mirrors the reference pipeline's stage functions (main.cu): upload the graph
(memoryAllocationGraph), orient (generateDAG), count / list cliques
(listAllCliques), peel (cliqueCoreDecompose), components (generateDensestCore
+ prune + componentDecompose), max-flow and the whole solve (main).  Every
call goes to the HIP library; there is no CPU fallback: if libkcds.so is
missing or no GPU is visible the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkcds.so")
_LIB = None

U32P = ctypes.POINTER(ctypes.c_uint32)
I64P = ctypes.POINTER(ctypes.c_int64)
I32P = ctypes.POINTER(ctypes.c_int32)
U8P = ctypes.POINTER(ctypes.c_uint8)
U64P = ctypes.POINTER(ctypes.c_uint64)

# every entry point include/kcds.h declares
EXPORTS = ("kcds_create", "kcds_destroy", "kcds_last_error", "kcds_upload_csr", "kcds_upload_edges",
           "kcds_download_csr", "kcds_orient", "kcds_count",
           "kcds_list", "kcds_download_cliques", "kcds_peel", "kcds_components", "kcds_maxflow", "kcds_solve",
           "kcds_sync", "kcds_set_tuning", "kcds_get_list_info", "kcds_solve_batch", "kcds_set_list_shape",
           "kcds_autotune_list", "kcds_set_graph_budget", "kcds_set_dense_budget", "kcds_server_create", "kcds_submit", "kcds_poll",
           "kcds_server_clock", "kcds_server_cancel_pending", "kcds_server_destroy", "kcds_count_dag_noalloc")

GRAPH_SOLVED, GRAPH_OVER_BUDGET = 0, 1
TUNE_GRID_PEEL, TUNE_GRID_FLOW, TUNE_EXACT_ORDER, TUNE_NO_GRAPHS, TUNE_BLOCK_FLOW = 1, 2, 4, 8, 16
TUNE_ENUM_PEEL, TUNE_DENSE, TUNE_DENSE_PEEL, TUNE_FLOW_ALL, TUNE_DEGENERACY_ORDER = 32, 64, 128, 256, 512
TUNE_BUILD_FRAMES, TUNE_BLOCK_PEEL = 1024, 2048
E_INVALID, E_HIP, E_NOMEM, E_STATE, E_LIMIT, E_PENDING = -1, -2, -3, -4, -5, -6


class PeelStats(ctypes.Structure):
    _fields_ = [("max_num", ctypes.c_uint64), ("max_den", ctypes.c_uint64), ("k_prime", ctypes.c_uint64),
                ("core_cliques", ctypes.c_uint64), ("max_core", ctypes.c_int64), ("core_size", ctypes.c_uint32),
                ("levels", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("T", ctypes.c_uint64), ("rho_num", ctypes.c_uint64), ("rho_den", ctypes.c_uint64),
                ("k_prime", ctypes.c_uint64), ("core_vertices", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("flow_calls", ctypes.c_uint32), ("peel", PeelStats),
                ("ms_orient", ctypes.c_float), ("ms_list", ctypes.c_float), ("ms_peel", ctypes.c_float),
                ("ms_components", ctypes.c_float), ("ms_flow", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("dense_passes", ctypes.c_uint32), ("dense_critical", ctypes.c_uint32)]


class GraphResult(ctypes.Structure):
    _fields_ = [("T", ctypes.c_uint64), ("rho_num", ctypes.c_uint64), ("rho_den", ctypes.c_uint64),
                ("k_prime", ctypes.c_uint64), ("core_vertices", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("flow_calls", ctypes.c_uint32), ("status", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class ListInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("arcs", ctypes.c_uint64), ("w2", ctypes.c_uint64),
                ("max_outdeg", ctypes.c_uint32), ("kernel_ms", ctypes.c_float),
                ("kernel_launches", ctypes.c_uint32), ("class_roots", ctypes.c_uint32 * 4),
                ("waves_per_root", ctypes.c_uint32 * 4), ("order_replaced", ctypes.c_uint32),
                ("order_kind", ctypes.c_uint32)]


class KcdsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kcds error {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        V = ctypes.c_void_p
        L.kcds_create.argtypes = [ctypes.POINTER(V), ctypes.c_int, ctypes.c_uint64]
        L.kcds_destroy.argtypes = [V]
        L.kcds_destroy.restype = None
        L.kcds_last_error.argtypes = [V]
        L.kcds_last_error.restype = ctypes.c_char_p
        L.kcds_upload_csr.argtypes = [V, ctypes.c_uint32, U32P, U32P]
        L.kcds_upload_edges.argtypes = [V, ctypes.c_uint32, ctypes.c_uint64, U32P, U64P]
        L.kcds_download_csr.argtypes = [V, U32P, U32P, ctypes.c_uint64, U64P]
        L.kcds_orient.argtypes = [V, U32P]
        L.kcds_count.argtypes = [V, ctypes.c_uint32, U64P, I64P]
        L.kcds_list.argtypes = [V, ctypes.c_uint32, U64P]
        L.kcds_download_cliques.argtypes = [V, U32P, ctypes.c_uint64, U64P]
        L.kcds_peel.argtypes = [V, I64P, ctypes.POINTER(PeelStats)]
        L.kcds_components.argtypes = [V, U32P, U32P]
        L.kcds_maxflow.argtypes = [V, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, I64P, U8P]
        L.kcds_solve.argtypes = [V, ctypes.c_uint32, ctypes.POINTER(Result), U8P]
        L.kcds_solve_batch.argtypes = [V, ctypes.c_uint32, ctypes.c_uint32, U32P, ctypes.POINTER(GraphResult), U8P]
        L.kcds_sync.argtypes = [V]
        L.kcds_count_dag_noalloc.argtypes = [ctypes.c_int, ctypes.c_uint32, V, V, ctypes.c_uint32, V,
                                             ctypes.POINTER(ctypes.c_uint64), V]
        L.kcds_set_tuning.argtypes = [V, ctypes.c_uint32, ctypes.c_uint32]
        L.kcds_get_list_info.argtypes = [V, ctypes.POINTER(ListInfo)]
        L.kcds_set_list_shape.argtypes = [V, U32P]
        L.kcds_autotune_list.argtypes = [V, ctypes.c_uint32, U32P, ctypes.POINTER(ctypes.c_float)]
        L.kcds_set_graph_budget.argtypes = [V, ctypes.c_uint64]
        L.kcds_set_dense_budget.argtypes = [V, ctypes.c_uint32]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Solver:
    """One device context (one HIP stream, its own buffers)."""

    def __init__(self, device=0, hbm_budget=0):
        self._h = ctypes.c_void_p()
        rc = lib().kcds_create(ctypes.byref(self._h), device, hbm_budget)
        if rc != 0:
            raise KcdsError(rc, "kcds_create failed (no GPU visible?)")
        self.n = 0

    def close(self):
        if self._h:
            lib().kcds_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise KcdsError(rc, lib().kcds_last_error(self._h).decode())

    def upload(self, offset, nbr):
        self._off = np.ascontiguousarray(offset, dtype=np.uint32)
        self._nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
        self.n = len(self._off) - 1
        self._chk(lib().kcds_upload_csr(self._h, self.n, _p(self._off, U32P), _p(self._nbr, U32P)))

    def upload_edges(self, n, uv):
        """Edge list -> CSR on the GPU (kcds_upload_edges); uv: m x 2 ids < n.
        Returns the number of distinct undirected edges."""
        self._uv = np.ascontiguousarray(uv, dtype=np.uint32).reshape(-1, 2)
        self._off = self._nbr = None
        self.n = int(n)
        m = ctypes.c_uint64()
        self._chk(lib().kcds_upload_edges(self._h, self.n, len(self._uv), _p(self._uv, U32P), ctypes.byref(m)))
        return m.value

    def download_csr(self):
        """The device CSR (offset[n+1], nbr[nnz]) as numpy arrays."""
        nnz = ctypes.c_uint64()
        self._chk(lib().kcds_download_csr(self._h, None, None, 0, ctypes.byref(nnz)))
        off = np.zeros(self.n + 1, np.uint32)
        nbr = np.zeros(max(1, nnz.value), np.uint32)
        self._chk(lib().kcds_download_csr(self._h, _p(off, U32P), _p(nbr, U32P), nnz.value, ctypes.byref(nnz)))
        return off, nbr[:nnz.value]

    def orient(self, listing_order=None):
        lo = None if listing_order is None else np.ascontiguousarray(listing_order, dtype=np.uint32)
        self._chk(lib().kcds_orient(self._h, _p(lo, U32P)))

    def count(self, k, per_vertex=True):
        T = ctypes.c_uint64()
        pv = np.zeros(self.n, np.int64) if per_vertex else None
        self._chk(lib().kcds_count(self._h, k, ctypes.byref(T), _p(pv, I64P)))
        return T.value, pv

    def list(self, k):
        T = ctypes.c_uint64()
        self._chk(lib().kcds_list(self._h, k, ctypes.byref(T)))
        return T.value

    def download_cliques(self, k):
        T = ctypes.c_uint64()
        self._chk(lib().kcds_download_cliques(self._h, None, 0, ctypes.byref(T)))
        out = np.zeros((max(T.value, 1), k), np.uint32)
        self._chk(lib().kcds_download_cliques(self._h, _p(out, U32P), T.value, ctypes.byref(T)))
        return out[: T.value]

    def peel(self):
        core = np.zeros(self.n, np.int64)
        st = PeelStats()
        self._chk(lib().kcds_peel(self._h, _p(core, I64P), ctypes.byref(st)))
        return core, st

    def components(self):
        comp = np.zeros(self.n, np.uint32)
        nc = ctypes.c_uint32()
        self._chk(lib().kcds_components(self._h, _p(comp, U32P), ctypes.byref(nc)))
        return comp, nc.value

    def maxflow(self, comp_id, p, q, want_side=False):
        f = ctypes.c_int64()
        side = np.zeros(self.n, np.uint8) if want_side else None
        self._chk(lib().kcds_maxflow(self._h, comp_id, p, q, ctypes.byref(f), _p(side, U8P)))
        return (f.value, side.astype(bool)) if want_side else f.value

    def solve(self, k, want_set=True):
        res = Result()
        ins = np.zeros(max(self.n, 1), np.uint8) if want_set else None
        self._chk(lib().kcds_solve(self._h, k, ctypes.byref(res), _p(ins, U8P)))
        return res, (ins[: self.n].astype(bool) if want_set else None)

    def solve_batch(self, graphs, k, want_set=True):
        """Solves many disjoint graphs [(offset, nbr), ...] in one batched pass.

        Returns (list of GraphResult, list of per-graph boolean in_set arrays)."""
        off, nbr, voff = concat_graphs(graphs)
        self.upload(off, nbr)
        G = len(graphs)
        res = (GraphResult * G)()
        ins = np.zeros(max(self.n, 1), np.uint8) if want_set else None
        self._chk(lib().kcds_solve_batch(self._h, k, G, _p(voff, U32P), res, _p(ins, U8P)))
        sets = [ins[voff[i]:voff[i + 1]].astype(bool) for i in range(G)] if want_set else None
        return list(res), sets

    def solve_resident_batch(self, voff, k, want_set=False):
        """Solves the disjoint graphs of the already uploaded union, split at voff (G+1 offsets)."""
        voff = np.ascontiguousarray(voff, dtype=np.uint32)
        G = len(voff) - 1
        res = (GraphResult * G)()
        ins = np.zeros(max(self.n, 1), np.uint8) if want_set else None
        self._chk(lib().kcds_solve_batch(self._h, k, G, _p(voff, U32P), res, _p(ins, U8P)))
        sets = [ins[voff[i]:voff[i + 1]].astype(bool) for i in range(G)] if want_set else None
        return list(res), sets

    def sync(self):
        self._chk(lib().kcds_sync(self._h))

    def set_tuning(self, waves_per_cu=0, flags=0):
        self._chk(lib().kcds_set_tuning(self._h, waves_per_cu, flags))

    def set_graph_budget(self, nbytes):
        """Per-graph cap on T*(8k+4) bytes of materialised cliques (0 = none)."""
        self._chk(lib().kcds_set_graph_budget(self._h, int(nbytes)))

    def set_dense_budget(self, major_cycles):
        """Dense search cycles before a listable k'-core takes the clique max-flow (0 = default)."""
        self._chk(lib().kcds_set_dense_budget(self._h, int(major_cycles)))

    def set_list_shape(self, waves_per_root=None):
        """Waves per root of the four width classes (None = library defaults)."""
        w = None if waves_per_root is None else np.ascontiguousarray(waves_per_root, dtype=np.uint32)
        if w is not None and len(w) != 4:
            raise ValueError("waves_per_root needs 4 entries")
        self._chk(lib().kcds_set_list_shape(self._h, _p(w, U32P)))

    def autotune_list(self, k):
        """Times every waves-per-root choice per width class; keeps and returns the fastest."""
        w = np.zeros(4, np.uint32)
        ms = (ctypes.c_float * 4)()
        self._chk(lib().kcds_autotune_list(self._h, k, _p(w, U32P), ms))
        return w.tolist(), list(ms)

    def list_info(self):
        info = ListInfo()
        self._chk(lib().kcds_get_list_info(self._h, ctypes.byref(info)))
        return info


def concat_graphs(graphs):
    """Disjoint union of CSR graphs: (offset, nbr, graph_voff)."""
    ns = [len(o) - 1 for o, _ in graphs]
    voff = np.zeros(len(graphs) + 1, np.uint32)
    voff[1:] = np.cumsum(ns, dtype=np.uint64).astype(np.uint32)
    offs, nbrs = [np.zeros(1, np.uint64)], []
    base_e = 0
    for (o, b), v0 in zip(graphs, voff[:-1]):
        o = np.asarray(o, np.uint64)
        offs.append(o[1:] + base_e)
        nbrs.append(np.asarray(b, np.uint32) + np.uint32(v0))
        base_e += int(o[-1])
    off = np.concatenate(offs).astype(np.uint32)
    nbr = np.concatenate(nbrs) if nbrs else np.zeros(0, np.uint32)
    return off, nbr.astype(np.uint32), voff


def count_dag_noalloc(n, doff_ptr, dnbr_ptr, k, per_vertex_ptr, stream_ptr=0, device=0):
    """kcds_count_dag_noalloc over device pointers (a DAG in listing form:
    arcs to higher ids, out-lists ascending); per_vertex_ptr receives n uint64
    counts.  Returns T; raises KcdsError on a nonzero status."""
    T = ctypes.c_uint64()
    rc = lib().kcds_count_dag_noalloc(device, n, ctypes.c_void_p(doff_ptr), ctypes.c_void_p(dnbr_ptr), k,
                                      ctypes.c_void_p(per_vertex_ptr), ctypes.byref(T),
                                      ctypes.c_void_p(stream_ptr) if stream_ptr else None)
    if rc != 0:
        raise KcdsError(rc, f"kcds_count_dag_noalloc failed ({rc})")
    return T.value
