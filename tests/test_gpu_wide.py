"""Roots wider than the 512-row LDS frames (graphs of degeneracy above 512):
their bitmask rows live in HBM (k_list_count_wide / k_list_emit_wide).  The
reference lists any out-degree (its masks are sized from the widest out-list,
src/gpuMemoryAllocation.cu:58).

Checked against the oracle and closed forms:
  - a complete graph K_s (s = 514..600, so the DAG's widest out-list is
    s - 1 > 512) beside a sparse ER graph with no edge between them: T and
    every vertex's count are C(s, k) / C(s-1, k-1) on the clique and the
    oracle's on the ER part;
  - a dense random graph G(600, 0.9) (degeneracy about 520): the oracle's
    per-vertex triangle counts;
  - the materialised clique list of K514 (k = 3): every 3-subset once;
  - the solve of K530 + ER at k = 3: rho* = C(530,3)/530 on the clique, whose
    clique-core numbers are C(529, 2), through the materialising peel.
"""
import math

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def union_with_clique(er, s):
    """CSR of the ER graph on [0, n) and K_s on [n, n + s), no edge between."""
    off = np.asarray(er.offset, np.int64)
    nbr = np.asarray(er.nbr, np.int64)
    n = len(off) - 1
    deg = np.concatenate([np.diff(off), np.full(s, s - 1, np.int64)])
    o = np.zeros(n + s + 1, np.int64)
    o[1:] = np.cumsum(deg)
    cl = np.arange(n, n + s, dtype=np.int64)
    rows = [nbr] + [np.delete(cl, i) for i in range(s)]
    return o.astype(np.uint32), np.concatenate(rows).astype(np.uint32), n


@pytest.mark.parametrize("s,ks", [(600, (3, 4)), (514, (3, 4))])
def test_wide_clique_counts(solver, s, ks):
    er = gen.er(2000, 0.004, 0, seed=s)
    off, nbr, n0 = union_with_clique(er, s)
    for k in ks:
        solver.set_tuning(0, 0)
        solver.upload(off, nbr)
        solver.orient(None)
        T, pv = solver.count(k)
        info = solver.list_info()
        Te, pve = O.count(er.offset, er.nbr, k)
        print(f"  K{s}+ER k={k}: T={T} widest out-list {info.max_outdeg} order {info.order_kind}", flush=True)
        assert info.max_outdeg == s - 1
        assert T == Te + math.comb(s, k)
        np.testing.assert_array_equal(pv[:n0], pve)
        assert np.all(pv[n0:] == math.comb(s - 1, k - 1))


def test_wide_dense_random_triangles(solver):
    g = gen.er(600, 0.9, 0, seed=5)
    solver.upload(g.offset, g.nbr)
    solver.orient(None)
    T, pv = solver.count(3)
    info = solver.list_info()
    Tr, pvr = O.count(g.offset, g.nbr, 3)
    print(f"  G(600, 0.9): T={T} widest out-list {info.max_outdeg}", flush=True)
    assert info.max_outdeg > 512
    assert T == Tr
    np.testing.assert_array_equal(pv, pvr)


def test_wide_materialised_cliques(solver):
    s = 514
    er = gen.er(300, 0.01, 0, seed=3)
    off, nbr, n0 = union_with_clique(er, s)
    solver.upload(off, nbr)
    solver.orient(None)
    T = solver.list(3)
    Te, _ = O.count(er.offset, er.nbr, 3)
    assert T == Te + math.comb(s, 3)
    cl = solver.download_cliques(3).astype(np.int64)
    cl.sort(axis=1)
    inK = cl[:, 0] >= n0
    kc = cl[inK]
    assert len(kc) == math.comb(s, 3)
    assert np.all(kc[:, 0] < kc[:, 1]) and np.all(kc[:, 1] < kc[:, 2])
    key = (kc[:, 0] - n0) * (s * s) + (kc[:, 1] - n0) * s + (kc[:, 2] - n0)
    assert len(np.unique(key)) == len(key)  # every 3-subset of the clique exactly once
    er_rows = {tuple(r) for r in cl[~inK].tolist()}
    assert er_rows == {tuple(sorted(r)) for r in O.list_cliques(er.offset, er.nbr, 3).tolist()}


def test_wide_solve(solver):
    s = 530
    er = gen.er(1500, 0.004, 0, seed=11)
    off, nbr, n0 = union_with_clique(er, s)
    solver.set_tuning(0, 0)
    solver.upload(off, nbr)
    res, ins = solver.solve(3)
    print(f"  K{s}+ER solve: rho={res.rho_num}/{res.rho_den} k'={res.k_prime} flows={res.flow_calls}", flush=True)
    assert (res.rho_num, res.rho_den) == (math.comb(s, 3), s)
    assert np.flatnonzero(ins).tolist() == list(range(n0, n0 + s))
    solver.orient(None)
    solver.list(3)
    core, st = solver.peel()
    assert np.all(core[n0:] == math.comb(s - 1, 2))
