"""GPU: per-graph HBM admission on the count pass.

A graph is admitted iff its materialised cliques, incidence and status fit
the per-graph budget: T * (8k + 4) <= budget.  The count pass stops counting
a graph once its running count passes that cap, so an over-budget graph
costs little more than the cap to reject.  Checked here: the boundary is
exact (T at the cap is admitted, one byte less rejects it), an over-budget
graph in a batch reports a lower bound above the cap and at most its true
count while the other graphs' results are unchanged, and a context that
rejected a graph counts it exactly afterwards.
"""
from math import comb

import numpy as np
import pytest

from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def clique(n):
    off = np.arange(n + 1, dtype=np.uint32) * (n - 1)
    nbr = np.array([u for v in range(n) for u in range(n) if u != v], np.uint32)
    return off, nbr


def test_admission_boundary_exact():
    off, nbr = clique(60)
    k = 4
    T = comb(60, 4)
    with kcds.Solver(0) as s:
        s.upload(off, nbr)
        s.set_graph_budget(T * (8 * k + 4))
        res, _ = s.solve(k, want_set=False)
        assert (res.T, res.rho_num, res.rho_den) == (T, T, 60)
        s.set_graph_budget(T * (8 * k + 4) - 1)
        with pytest.raises(kcds.KcdsError) as ei:
            s.solve(k, want_set=False)
        assert ei.value.code == kcds.E_NOMEM
        # the stopped count is not reused: a plain count is exact
        s.orient(None)
        Tc, _ = s.count(k, per_vertex=False)
        assert Tc == T


def test_admission_stops_big_graph_in_batch():
    k = 6
    big = clique(200)  # C(200, 6) = 82,408,626,300 6-cliques
    small = []
    for i in range(6):
        g = gen.er(300 + 40 * i, 0.03, 12, seed=900 + i)
        small.append((np.asarray(g.offset, np.uint32), np.asarray(g.nbr, np.uint32)))
    graphs = small[:3] + [big] + small[3:]
    budget = 8 << 20
    cap = budget // (8 * k + 4)
    with kcds.Solver(0) as s:
        s.set_graph_budget(budget)
        solo = []
        for off, nbr in small:
            s.upload(off, nbr)
            r, _ = s.solve(k, want_set=False)
            solo.append((r.T, r.rho_num, r.rho_den))
        with pytest.raises(kcds.KcdsError) as ei:
            s.upload(*big)
            s.solve(k, want_set=False)
        assert ei.value.code == kcds.E_NOMEM
        res, _ = s.solve_batch(graphs, k, want_set=False)
    got = [(int(g.T), int(g.rho_num), int(g.rho_den)) for g in res]
    assert res[3].status == kcds.GRAPH_OVER_BUDGET
    assert cap < res[3].T <= comb(200, 6)
    print(f"  K200 stopped at {res[3].T} of {comb(200, 6)} 6-cliques (cap {cap})", flush=True)
    assert got[:3] + got[4:] == solo
    assert all(res[i].status == kcds.GRAPH_SOLVED for i in range(7) if i != 3)
