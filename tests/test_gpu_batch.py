"""GPU parity of the batched solve: many disjoint graphs solved in one pass
(one orientation, one listing, one peel with per-graph bookkeeping, one
launch for every component's exact search) must give each graph exactly the
answer it gets alone -- and the oracle's.
"""
import math

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def mixed_batch():
    gs = []
    for i in range(14):
        g = gen.er(60 + 17 * i, 0.08, 6 + i % 9, seed=300 + i)
        gs.append((g.offset, g.nbr))
    g = gen.rmat(9, 8, (10,), seed=41)
    gs.append((g.offset, g.nbr))
    gs.append((np.zeros(6, np.uint32), np.zeros(0, np.uint32)))           # 5 isolated vertices
    gs.append((np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32)))  # one edge, no clique
    # two disjoint K6 (a tie) inside one graph
    n = 12
    adj = [[u for u in range(6 * (v // 6), 6 * (v // 6) + 6) if u != v] for v in range(n)]
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum([len(a) for a in adj])
    gs.append((off, np.array([u for a in adj for u in a], np.uint32)))
    return gs


@pytest.mark.parametrize("grid_peel", [False, True], ids=["block_peel", "grid_peel"])
@pytest.mark.parametrize("k", [3, 4])
def test_batch_matches_oracle_and_solo(solver, k, grid_peel):
    solver.set_tuning(0, 1 if grid_peel else 0)
    try:
        _batch_matches(solver, k)
    finally:
        solver.set_tuning(0, 0)


def _batch_matches(solver, k):
    gs = mixed_batch()
    res, sets = solver.solve_batch(gs, k)
    for i, (off, nbr) in enumerate(gs):
        r, pv, core, comp, ins_r = O.solve(off, nbr, k)
        got = res[i]
        print(f"  graph {i}: n={len(off)-1} T={got.T} rho={got.rho_num}/{got.rho_den}", flush=True)
        assert got.T == r.T
        if r.T == 0:
            assert got.rho_den == 0 and not sets[i].any()
            continue
        assert (got.rho_num, got.rho_den) == (r.rho_num, r.rho_den)
        assert got.k_prime == r.k_prime
        np.testing.assert_array_equal(sets[i], ins_r)
    # solo solves of a few graphs give the same answers
    for i in (0, 5, 14, len(gs) - 1):
        solver.upload(*gs[i])
        solo, ins = solver.solve(k)
        assert (solo.T, solo.rho_num, solo.rho_den) == (res[i].T, res[i].rho_num, res[i].rho_den)
        np.testing.assert_array_equal(ins, sets[i])


def test_batch_config4_subset(solver):
    specs = gen.batch_specs()[:96]
    gs = [gen.batch_graph(s) for s in specs]
    res, sets = solver.solve_batch([(g.offset, g.nbr) for g in gs], 4)
    for i, g in enumerate(gs):
        r, pv, core, comp, ins_r = O.solve(g.offset, g.nbr, 4)
        assert res[i].T == r.T
        assert (res[i].rho_num, res[i].rho_den) == (r.rho_num, r.rho_den)
        np.testing.assert_array_equal(sets[i], ins_r)
        s = len(g.planted)
        assert res[i].rho_num * s >= math.comb(s, 4) * res[i].rho_den  # at least the planted clique


def test_batch_rejects_cross_edges(solver):
    from test_amd import kcds
    g = gen.er(50, 0.2, 5, seed=1)
    solver.upload(g.offset, g.nbr)
    import ctypes
    voff = np.array([0, 25, 50], np.uint32)
    res = (kcds.GraphResult * 2)()
    rc = kcds.lib().kcds_solve_batch(solver._h, 3, 2, voff.ctypes.data_as(kcds.U32P), res, None)
    assert rc == -1  # KCDS_E_INVALID: edges between the two declared graphs


def test_batch_mixed_peels(solver):
    """A batch whose graphs go to different peels: one with more cliques than a
    workgroup peel takes (K40 plus noise at k = 5: C(40,5) = 658,008 cliques)
    is peeled by the grid kernels while the small ones are peeled a workgroup
    each; every graph's answer equals the oracle's."""
    gs = mixed_batch()[:6]
    g = gen.er(120, 0.05, 40, seed=77)
    gs.append((g.offset, g.nbr))
    res, sets = solver.solve_batch(gs, 5)
    for i, (off, nbr) in enumerate(gs):
        r, pv, core, comp, ins_r = O.solve(off, nbr, 5)
        assert res[i].T == r.T
        if r.T == 0:
            continue
        assert (res[i].rho_num, res[i].rho_den) == (r.rho_num, r.rho_den)
        np.testing.assert_array_equal(sets[i], ins_r)
    assert res[-1].T > (1 << 18)


@pytest.mark.parametrize("count", [40, 80])
def test_batch_grid_and_workgroup_flows_together(solver, count):
    """A batch whose big components go to the grid max-flow while the rest
    run in the workgroup launch on a stream of its own.  Graph 546 of the
    4,096-graph batch has 100,529 4-cliques; with 41 components every
    component of 4,096+ core cliques takes the grid solver, with 81 the
    49,152 threshold applies.  Every graph's result equals its own solve
    (which the oracle tests pin)."""
    specs = gen.batch_specs()
    gs = [gen.batch_graph(s) for s in specs[:count]] + [gen.batch_graph(specs[546])]
    pairs = [(g.offset, g.nbr) for g in gs]
    # every component through a max-flow (the planted cliques are complete
    # components, which the default answers without one)
    solver.set_tuning(0, kcds.TUNE_FLOW_ALL)
    try:
        res, sets = solver.solve_batch(pairs, 4)
        assert sum(r.flow_calls for r in res) >= len(gs)
    finally:
        solver.set_tuning(0, 0)
    for i, g in enumerate(gs):
        solver.upload(g.offset, g.nbr)
        solo, ins = solver.solve(4)
        assert (solo.T, solo.rho_num, solo.rho_den) == (res[i].T, res[i].rho_num, res[i].rho_den), i
        np.testing.assert_array_equal(ins, sets[i])
    assert res[-1].T == 100529
    # the default answers the same
    res_d, sets_d = solver.solve_batch(pairs, 4)
    for i in range(len(gs)):
        assert (res_d[i].T, res_d[i].rho_num, res_d[i].rho_den) == (res[i].T, res[i].rho_num, res[i].rho_den), i
        np.testing.assert_array_equal(sets_d[i], sets[i])


def test_complete_components_answered_without_flow(solver):
    """Components that are complete graphs (clique count C(n_c, k)) take no
    max-flow, and their answers equal the oracle's: three disjoint cliques
    K9, K9, K7 (the two K9 tie, so both are returned), and planted K20 / K26 in
    ER graphs, whose k'-cores are the cliques alone."""
    sizes = (9, 9, 7)
    adj, base = [], 0
    for s_ in sizes:
        adj += [[base + b for b in range(s_) if b != a] for a in range(s_)]
        base += s_
    off = np.zeros(base + 1, np.uint32)
    off[1:] = np.cumsum([len(x) for x in adj])
    cases = [("K9+K9+K7", off, np.array([v for x in adj for v in x], np.uint32))]
    for n_, s_ in ((300, 20), (800, 26)):
        g = gen.er(n_, 0.02, s_, seed=n_)
        cases.append((f"er{n_}+K{s_}", g.offset, g.nbr))
    solver.set_tuning(0, 0)
    for k in (3, 4):
        for name, off, nbr in cases:
            solver.upload(off, nbr)
            res, ins = solver.solve(k)
            r, _, _, _, ins_r = O.solve(off, nbr, k)
            assert (res.T, res.rho_num, res.rho_den) == (r.T, r.rho_num, r.rho_den), (name, k)
            np.testing.assert_array_equal(ins, ins_r)
            assert res.flow_calls == 0, (name, k, res.flow_calls)
    # the tie: both K9 are in the answer
    solver.upload(cases[0][1], cases[0][2])
    res, ins = solver.solve(3)
    assert int(ins.sum()) == 18 and (res.rho_num, res.rho_den) == (2 * math.comb(9, 3), 18)
