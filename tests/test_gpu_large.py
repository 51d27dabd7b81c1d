"""Whole exact solves of the two large single-graph workloads on the GPU:
R-MAT scale 18 (n = 262,144, edge factor 16, + K60 + 5 x K30), k = 4, and
Chung-Lu n = 10^6 (Pareto 2.1 weights, + near-K80), k = 5.

Their k'-cores hold 1.1e9 and 4.6e10 cliques on 988 and 574 vertices, so the
solve runs the dense exact search, whose answer carries its own integer
certificate.  What the oracle can check at this size is checked here:
  - the returned set S is recounted by the oracle on the induced subgraph:
    c_k(S) and |S| must equal the reported rational;
  - the planted K60 gives a closed-form floor C(60,4)/60 for the R-MAT graph;
  - S lies inside the k'-core and the peel statistics are those of the
    materialising peel (R-MAT) / of the enumerating peel (Chung-Lu, whose
    7.3e10 cliques cannot be listed);
  - the two peels give the same clique-core numbers on the R-MAT graph.
"""
import math

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def induced_count(g, S, k):
    mask = np.zeros(g.n, bool)
    mask[S] = True
    idx = -np.ones(g.n, np.int64)
    idx[S] = np.arange(len(S))
    deg = np.diff(g.offset.astype(np.int64))
    src = np.repeat(np.arange(g.n), deg)
    keep = mask[src] & mask[g.nbr]
    a, b = idx[src[keep]], idx[g.nbr[keep]]
    o = np.lexsort((b, a))
    a, b = a[o], b[o]
    off = np.zeros(len(S) + 1, np.int64)
    np.add.at(off, a + 1, 1)
    T, _ = O.count(np.cumsum(off).astype(np.uint32), b.astype(np.uint32), k, per_vertex=False)
    return T


def test_config2_solve(solver):
    g, k = gen.config(2)
    solver.set_tuning(0, 0)
    solver.upload(g.offset, g.nbr)
    res, ins = solver.solve(k)
    S = np.flatnonzero(ins)
    print(f"  C2 rho={res.rho_num}/{res.rho_den} k'={res.k_prime} core={res.core_vertices} "
          f"core_cliques={res.peel.core_cliques} passes={res.dense_passes} total={res.ms_total:.0f} ms", flush=True)
    assert res.T == 2281884915
    assert res.dense_passes > 0
    assert (len(S), induced_count(g, S, k)) == (res.rho_den, res.rho_num)
    assert res.rho_num * 60 >= math.comb(60, 4) * res.rho_den  # the planted K60
    # the peel's best core is a lower bound, and S lies in the k'-core
    assert res.rho_num * res.peel.max_den >= res.peel.max_num * res.rho_den
    solver.orient(None)
    solver.list(k)
    core, st = solver.peel()
    assert np.all(core[S] >= st.k_prime)
    assert (st.k_prime, st.core_size, st.core_cliques) == (res.k_prime, res.core_vertices, res.peel.core_cliques)
    # the enumerating peel gives the same clique cores
    solver.orient(None)
    solver.count(k, per_vertex=False)
    core_e, st_e = solver.peel()
    np.testing.assert_array_equal(core_e, core)
    assert (st_e.max_num, st_e.max_den, st_e.levels) == (st.max_num, st.max_den, st.levels)


def test_config3_solve(solver):
    g, k = gen.config(3)
    solver.set_tuning(0, 0)
    solver.upload(g.offset, g.nbr)
    res, ins = solver.solve(k)
    S = np.flatnonzero(ins)
    print(f"  C3 rho={res.rho_num}/{res.rho_den} k'={res.k_prime} core={res.core_vertices} "
          f"core_cliques={res.peel.core_cliques} passes={res.dense_passes} total={res.ms_total:.0f} ms", flush=True)
    assert res.T == 72726778604
    assert res.dense_passes > 0
    assert (len(S), induced_count(g, S, k)) == (res.rho_den, res.rho_num)
    assert res.rho_num * res.peel.max_den >= res.peel.max_num * res.rho_den
    assert len(S) <= res.core_vertices
