"""GPU: kcds_count_dag_noalloc, the allocation-free count over a DAG the caller
holds in device memory, against kcds_count on the same graph: T and every
vertex's count bit-exact, over calls in a row (each later
one smaller, so contents the shared scratch array kept from an earlier call
would show), and its error returns for a DAG not in listing form and for n
above its limit, after which it still counts correctly.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def listing_dag(off, nbr, lo):
    """The DAG in listing form: vertex v renamed rank[v] = lo[v] - 1, arcs to
    the higher rank, every out-list ascending."""
    off = np.asarray(off, np.int64)
    nbr = np.asarray(nbr, np.int64)
    n = len(off) - 1
    rank = np.asarray(lo, np.int64) - 1
    src = np.repeat(np.arange(n), np.diff(off))
    a, b = rank[src], rank[nbr]
    keep = a < b
    a, b = a[keep], b[keep]
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    doff = np.concatenate([[0], np.cumsum(np.bincount(a, minlength=n))])
    return doff.astype(np.uint32), b.astype(np.uint32), rank


def on_device(a):
    a = np.asarray(a, np.uint32)
    return torch.from_numpy((a if len(a) else np.zeros(1, np.uint32)).view(np.int32).copy()).cuda()


def run_noalloc(doff, dnbr, k):
    n = len(doff) - 1
    d_off, d_nbr = on_device(doff), on_device(dnbr)
    pv = torch.full((max(n, 1),), -1, dtype=torch.int64, device="cuda")  # the call must overwrite these
    torch.cuda.synchronize()
    T = kcds.count_dag_noalloc(n, d_off.data_ptr(), d_nbr.data_ptr(), k, pv.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)
    return T, pv.cpu().numpy()[:n]


def cases():
    g = gen.rmat(12, 16, (30, 20), seed=5)
    yield "rmat12_hubs", g.offset, g.nbr, 4
    g = gen.chunglu(20000, 2.1, 2.0, 900.0, 120000, 25, 0.95, seed=17)
    yield "chunglu20k", g.offset, g.nbr, 3
    g = gen.er(400, 0.04, 18, seed=21)
    yield "er400", g.offset, g.nbr, 5
    g = gen.er(300, 0.05, 10, seed=400)
    yield "er300", g.offset, g.nbr, 2


def test_noalloc_count_matches_kcds_count(solver):
    """Calls in a row, each smaller than the one before (n 4096 -> 20000 is
    the exception: larger, then 400, 300): every one equals kcds_count."""
    for name, off, nbr, k in cases():
        _, _, lo, _ = O.core_decompose(off, nbr)
        # the reverse of the reference's listing order: out-degrees at most the degeneracy
        lo = (len(off) - lo.astype(np.int64)).astype(np.uint32)
        doff, dnbr, rank = listing_dag(off, nbr, lo)
        print(f"  {name}: n={len(doff) - 1} arcs={len(dnbr)} k={k} max_outdeg={np.diff(doff.astype(np.int64)).max()}",
              flush=True)
        T, pv = run_noalloc(doff, dnbr, k)
        solver.upload(off, nbr)
        solver.orient(None)  # any acyclic orientation gives the same per-vertex counts
        T_ref, pv_ref = solver.count(k, per_vertex=True)
        assert T == T_ref, name
        np.testing.assert_array_equal(pv[rank].astype(np.int64), pv_ref)


def test_noalloc_count_errors_then_counts():
    # an arc to a lower id: not in listing form
    doff = np.array([0, 1, 2, 2], np.uint32)
    dnbr = np.array([1, 0], np.uint32)
    with pytest.raises(kcds.KcdsError) as e:
        run_noalloc(doff, dnbr, 3)
    assert e.value.code == kcds.E_INVALID
    # n above the limit is refused before any device work
    d = on_device(np.zeros(4, np.uint32))
    T = ctypes.c_uint64()
    rc = kcds.lib().kcds_count_dag_noalloc(0, (1 << 20) + 1, ctypes.c_void_p(d.data_ptr()),
                                           ctypes.c_void_p(d.data_ptr()), 3, ctypes.c_void_p(d.data_ptr()),
                                           ctypes.byref(T), None)
    assert rc == kcds.E_LIMIT
    # then a valid count: K5 has C(5,3) = 10 triangles, each vertex in C(4,2) = 6
    doff = np.array([0, 4, 7, 9, 10, 10], np.uint32)
    dnbr = np.array([1, 2, 3, 4, 2, 3, 4, 3, 4, 4], np.uint32)
    T, pv = run_noalloc(doff, dnbr, 3)
    assert T == 10
    np.testing.assert_array_equal(pv, [6] * 5)
