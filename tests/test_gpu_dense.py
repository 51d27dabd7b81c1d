"""GPU parity of the dense exact search (k_dense.hip): the k'-core held as a
bitmask matrix, greedy clique splits combined by Wolfe's minimum-norm-point
method, and the answer certified by an exact integer max-flow over vertex
pairs (with the critical-set rule for vertices the flow leaves unfilled).  Core sizes cover every matrix width (1, 2, 4, 8, 10, 12 and 16 words
per row).

The search is forced (KCDS_TUNE_DENSE) on graphs whose k'-core has at most
1,024 vertices, after either peel, and must return the oracle's density
rational and maximal densest set, including ties that span components.
"""
import itertools
import math
import os

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def disjoint_cliques(sizes, extra_edges=()):
    n = sum(sizes)
    adj = [set() for _ in range(n)]
    at = 0
    for s in sizes:
        for i in range(at, at + s):
            adj[i].update(j for j in range(at, at + s) if j != i)
        at += s
    for u, v in extra_edges:
        adj[u].add(v)
        adj[v].add(u)
    off = np.zeros(n + 1, np.uint32)
    for v in range(n):
        off[v + 1] = off[v] + len(adj[v])
    nbr = np.array([u for v in range(n) for u in sorted(adj[v])], np.uint32)
    return off, nbr


def graphs():
    off, nbr, _ = O.load_adjlist(os.path.join(GOLD, "ref_data_test.txt"))
    yield "ref_data_test", off, nbr, (3, 4)
    g = gen.er(400, 0.04, 18, seed=21)
    yield "er400", g.offset, g.nbr, (3, 4, 5, 6)
    g = gen.rmat(10, 8, (12,), seed=9)
    yield "rmat10", g.offset, g.nbr, (3, 4)
    off, nbr = disjoint_cliques([6, 6, 5], extra_edges=[(0, 6), (11, 12)])
    yield "tied_cliques", off, nbr, (3, 4)
    g = gen.rmat(12, 16, (30, 20), seed=5)
    yield "rmat12_hubs", g.offset, g.nbr, (3,)
    g = gen.er(1500, 0.02, 30, seed=3)
    yield "er1500", g.offset, g.nbr, (3, 4)
    g = gen.chunglu(20000, 2.1, 2.0, 900.0, 120000, 25, 0.95, seed=17)
    yield "chunglu20k", g.offset, g.nbr, (4,)
    g = gen.er(600, 0.07, 0, seed=11)  # k'-core of about 600 vertices: 10-word rows
    yield "er600", g.offset, g.nbr, (3, 4)
    g = gen.er(700, 0.06, 0, seed=11)  # 696 vertices: 12-word rows
    yield "er700", g.offset, g.nbr, (3,)
    off, nbr = disjoint_cliques([6] * 100)  # 600 vertices in 100 tied components
    yield "k6_x100", off, nbr, (3,)


GRAPHS = list(graphs())
PEELS = {"listed_peel": kcds.TUNE_DENSE, "enum_peel": kcds.TUNE_DENSE | kcds.TUNE_ENUM_PEEL}


@pytest.mark.parametrize("peel", list(PEELS))
@pytest.mark.parametrize("name,off,nbr,ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_dense_solve_matches_oracle(solver, name, off, nbr, ks, peel):
    solver.set_tuning(0, PEELS[peel])
    try:
        for k in ks:
            solver.upload(off, nbr)
            res, ins = solver.solve(k)
            r, _, _, comp_r, ins_r = O.solve(off, nbr, k)
            print(f"  {name} k={k} rho={res.rho_num}/{res.rho_den} core={res.core_vertices} "
                  f"passes={res.dense_passes} ncomp={res.ncomp}", flush=True)
            assert res.dense_passes > 0  # the dense search ran
            assert res.T == r.T
            assert (res.rho_num, res.rho_den) == (r.rho_num, r.rho_den)
            assert (res.k_prime, res.core_vertices) == (r.k_prime, r.core_vertices)
            assert res.ncomp == len(set(int(x) for x in comp_r if x != 0xFFFFFFFF))
            np.testing.assert_array_equal(ins, ins_r)
    finally:
        solver.set_tuning(0, 0)


def test_dense_tied_components_union(solver):
    """Two disjoint K6 and a K5: the maximal densest set is both K6."""
    off, nbr = disjoint_cliques([6, 6, 5])
    solver.set_tuning(0, kcds.TUNE_DENSE)
    try:
        solver.upload(off, nbr)
        res, ins = solver.solve(3)
    finally:
        solver.set_tuning(0, 0)
    assert res.dense_passes > 0
    assert (res.rho_num, res.rho_den) == (40, 12)
    assert set(np.flatnonzero(ins).tolist()) == set(range(12))


def brute_densest(n, cl):
    best, union = (0, 1), 0
    sets = []
    for m in range(1, 1 << n):
        inside = sum(1 for c in cl if all((m >> v) & 1 for v in c))
        size = bin(m).count("1")
        sets.append((m, inside, size))
        if inside * best[1] > best[0] * size:
            best = (inside, size)
    for m, inside, size in sets:
        if inside and inside * best[1] == best[0] * size:
            union |= m
    return best, union


def test_dense_brute_force_small(solver):
    rng = np.random.default_rng(11)
    solver.set_tuning(0, kcds.TUNE_DENSE)
    try:
        for trial in range(8):
            n = int(rng.integers(7, 12))
            g = gen.er(n, float(rng.uniform(0.45, 0.85)), 0, seed=700 + trial)
            adj = [set(g.nbr[g.offset[v]:g.offset[v + 1]].tolist()) for v in range(n)]
            for k in (3, 4):
                cl = [c for c in itertools.combinations(range(n), k)
                      if all(b in adj[a] for a, b in itertools.combinations(c, 2))]
                if not cl:
                    continue
                solver.upload(g.offset, g.nbr)
                res, ins = solver.solve(k)
                best, union = brute_densest(n, cl)
                assert res.rho_num * best[1] == best[0] * res.rho_den
                assert int(sum(1 << int(v) for v in np.flatnonzero(ins))) == union
    finally:
        solver.set_tuning(0, 0)


# Graphs whose certificate needs the critical set: a member of the densest set
# whose clique-degree inside it barely exceeds rho*, or outsiders that must
# take the whole share of their cliques with it (found by a sweep of R-MAT and
# Chung-Lu seeds; each answer checked against the oracle).
CRITICAL_GRAPHS = [
    ("rmat11_s4", lambda: gen.rmat(11, 12, (), seed=4), (4, 5)),
    ("rmat12_s2", lambda: gen.rmat(12, 10, (), seed=2), (4,)),
    ("chunglu8k_s5", lambda: gen.chunglu(8000, 2.1, 2.0, 400.0, 40000, 0, 0.95, seed=5), (4, 5)),
    ("chunglu8k_s6", lambda: gen.chunglu(8000, 2.1, 2.0, 400.0, 40000, 0, 0.95, seed=6), (4,)),
]


def test_dense_critical_certificate(solver):
    used = 0
    solver.set_tuning(0, kcds.TUNE_DENSE)
    try:
        for name, make, ks in CRITICAL_GRAPHS:
            g = make()
            for k in ks:
                solver.upload(g.offset, g.nbr)
                res, ins = solver.solve(k)
                r, _, _, _, ins_r = O.solve(g.offset, g.nbr, k)
                print(f"  {name} k={k} rho={res.rho_num}/{res.rho_den} core={res.core_vertices} "
                      f"passes={res.dense_passes} critical={res.dense_critical}", flush=True)
                assert res.dense_passes > 0
                assert (res.rho_num, res.rho_den) == (r.rho_num, r.rho_den)
                np.testing.assert_array_equal(ins, ins_r)
                used += res.dense_critical > 0
    finally:
        solver.set_tuning(0, 0)
    assert used >= 3  # the critical rule took part in most of these certificates


def test_dense_budget_falls_back_to_flow(solver):
    """A k'-core routed to the dense search by default (a planted K80 less one
    edge at k = 5: about 2.3e7 cliques, from 2^24 up) whose clique list fits:
    with a budget of one major cycle the search gives up and the solve lists
    the core and runs the clique max-flow.  Both exact paths must agree, and
    the answer lies inside the planted clique."""
    g = gen.er(3000, 0.002, 80, seed=41)
    pl = [int(v) for v in g.planted]
    off = np.asarray(g.offset, np.int64)
    adj = [set(g.nbr[off[v]:off[v + 1]].tolist()) for v in range(len(off) - 1)]
    a, b = pl[0], pl[1]
    adj[a].discard(b)
    adj[b].discard(a)
    o = np.zeros(len(adj) + 1, np.int64)
    o[1:] = np.cumsum([len(s) for s in adj])
    nbr = np.array([u for v in range(len(adj)) for u in sorted(adj[v])], np.uint32)
    o = o.astype(np.uint32)
    solver.set_tuning(0, 0)
    solver.upload(o, nbr)
    res_d, ins_d = solver.solve(5)
    solver.set_dense_budget(1)
    try:
        solver.upload(o, nbr)
        res_f, ins_f = solver.solve(5)
    finally:
        solver.set_dense_budget(0)
    print(f"  K80-e k=5: dense rho={res_d.rho_num}/{res_d.rho_den} passes={res_d.dense_passes}; "
          f"budget 1: rho={res_f.rho_num}/{res_f.rho_den} flows={res_f.flow_calls}", flush=True)
    assert res_d.dense_passes > 0 and res_d.flow_calls == 0
    assert res_f.dense_passes == 0 and res_f.flow_calls > 0
    assert (res_f.rho_num, res_f.rho_den) == (res_d.rho_num, res_d.rho_den)
    np.testing.assert_array_equal(ins_f, ins_d)
    assert set(np.flatnonzero(ins_d).tolist()) <= set(pl)
    # K80 less an edge keeps both ends: (C(80, 5) - C(78, 3)) / 80 = 299,549.25
    # beats C(79, 5) / 79 = 285,285 with one end dropped
    assert res_d.rho_num * 80 == (math.comb(80, 5) - math.comb(78, 3)) * res_d.rho_den
