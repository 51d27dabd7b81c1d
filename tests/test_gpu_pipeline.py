"""GPU parity of the stages after listing: peeling, densest core + components,
max-flow, and the whole exact solve, against the CPU oracle.

Bit-exact: clique-core numbers, the peel statistics (maxDensity as a
rational, k', core size), canonical component labels, max-flow values at
rational alpha, maximal min-cut sides, and the densest density as an exact
rational with the unique maximal densest vertex set.
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
            for j in range(at, at + s):
                if i != j:
                    adj[i].add(j)
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
    yield "ref_data_test", off, nbr, (3, 4, 2)
    g = gen.er(400, 0.04, 18, seed=21)
    yield "er400", g.offset, g.nbr, (3, 4, 5, 2)
    g = gen.rmat(10, 8, (12,), seed=9)
    yield "rmat10", g.offset, g.nbr, (3, 4, 2)
    off, nbr = disjoint_cliques([6, 6, 5], extra_edges=[(0, 6), (11, 12)])
    yield "tied_cliques", off, nbr, (3, 4, 2)
    g = gen.rmat(9, 6, (), seed=13)  # edge density only (k = 2 first: the max-flow test uses ks[0])
    yield "rmat9_edges", g.offset, g.nbr, (2, 3)


GRAPHS = list(graphs())


@pytest.mark.parametrize("grid_peel", [False, True], ids=["block_peel", "grid_peel"])
@pytest.mark.parametrize("name,off,nbr,ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_peel_and_components(solver, name, off, nbr, ks, grid_peel):
    """Both peels (workgroup per graph, and grid-wide step kernels) against the oracle."""
    solver.set_tuning(0, kcds.TUNE_GRID_PEEL if grid_peel else kcds.TUNE_BLOCK_PEEL)
    try:
        _peel_and_components(solver, name, off, nbr, ks)
    finally:
        solver.set_tuning(0, 0)


def _peel_and_components(solver, name, off, nbr, ks):
    n = len(off) - 1
    for k in ks:
        solver.upload(off, nbr)
        solver.orient(None)
        T = solver.list(k)
        cl = O.list_cliques(off, nbr, k)
        assert T == len(cl)
        if T == 0:
            continue
        core, st = solver.peel()
        core_r, status_r, st_r = O.peel(n, k, cl)
        print(f"  {name} k={k} T={T} maxDensity={st.max_num}/{st.max_den} k'={st.k_prime}", flush=True)
        np.testing.assert_array_equal(core, core_r)
        assert (st.max_num, st.max_den, st.k_prime, st.core_size, st.core_cliques, st.max_core) == \
            (st_r.max_num, st_r.max_den, st_r.k_prime, st_r.core_size, st_r.core_cliques, st_r.max_core)
        comp, ncomp = solver.components()
        _, _, _, comp_r, _ = O.solve(off, nbr, k)
        np.testing.assert_array_equal(comp, comp_r)
        assert ncomp == len(set(int(x) for x in comp_r if x != 0xFFFFFFFF))


# block_flow: the workgroup max-flow for every component; default: each
# component to the solver the library picks (complete components answered
# without a max-flow); grid_flow: the grid-wide one, its relabel and sweep
# sequences replayed as captured HIP graphs (many replays per flow, with host
# reads and syncs between them); grid_flow_eager: the same launches one at a
# time.  The forced forms run a max-flow on every component (256 =
# KCDS_TUNE_FLOW_ALL), complete ones included.
FLOWS = {"block_flow": 16 | 256, "default_flow": 0, "grid_flow": 2 | 256, "grid_flow_eager": 2 | 8 | 256}


@pytest.mark.parametrize("flow", list(FLOWS))
@pytest.mark.parametrize("name,off,nbr,ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_maxflow_values(solver, name, off, nbr, ks, flow):
    """Max-flow values and maximal min-cut sides of the workgroup and the grid solver."""
    solver.set_tuning(0, FLOWS[flow])
    try:
        _maxflow_values(solver, name, off, nbr, ks)
    finally:
        solver.set_tuning(0, 0)


def _maxflow_values(solver, name, off, nbr, ks):
    n = len(off) - 1
    k = ks[0]
    solver.upload(off, nbr)
    solver.orient(None)
    if solver.list(k) == 0:
        return
    solver.peel()
    comp, ncomp = solver.components()
    cl = O.list_cliques(off, nbr, k)
    core_r, status_r, st_r = O.peel(n, k, cl)
    for cid in sorted(set(int(x) for x in comp if x != 0xFFFFFFFF)):
        verts = np.flatnonzero(comp == cid)
        local = {int(v): i for i, v in enumerate(verts)}
        ccl = np.array([[local[int(u)] for u in row] for row, s in zip(cl, status_r)
                        if s >= st_r.k_prime and int(row[0]) in local], np.uint32).reshape(-1, k)
        Tc, nc = len(ccl), len(verts)
        for (p, q) in [(0, 1), (Tc, nc), (Tc + 1, nc), (2 * Tc, 3 * nc + 1), (st_r.max_core, 1), (1, 7)]:
            f, side = solver.maxflow(cid, p, q, want_side=True)
            fr, side_r = O.clique_maxflow(nc, k, ccl, p, q, want_side=True)
            assert f == fr, (cid, p, q)
            got = set(np.flatnonzero(side).tolist())
            exp = set(int(verts[i]) for i in np.flatnonzero(side_r))
            assert got == exp, (cid, p, q)


@pytest.mark.parametrize("flow", list(FLOWS))
@pytest.mark.parametrize("name,off,nbr,ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_solve_matches_oracle(solver, name, off, nbr, ks, flow):
    solver.set_tuning(0, FLOWS[flow])
    try:
        _solve_matches(solver, name, off, nbr, ks)
    finally:
        solver.set_tuning(0, 0)


def _solve_matches(solver, name, off, nbr, ks):
    for k in ks:
        solver.upload(off, nbr)
        res, ins = solver.solve(k)
        r, pv, core, comp, ins_r = O.solve(off, nbr, k)
        print(f"  {name} k={k} rho={res.rho_num}/{res.rho_den} flows={res.flow_calls}", flush=True)
        assert res.T == r.T
        assert (res.rho_num, res.rho_den) == (r.rho_num, r.rho_den)
        np.testing.assert_array_equal(ins, ins_r)


def test_tied_cliques_union(solver):
    # two disjoint K6 and a K5: both K6 reach the optimum C(6,3)/6 for k=3,
    # so the maximal densest set is their union
    off, nbr = disjoint_cliques([6, 6, 5])
    solver.upload(off, nbr)
    res, ins = solver.solve(3)
    assert (res.rho_num, res.rho_den) == (2 * math.comb(6, 3), 12)
    assert set(np.flatnonzero(ins).tolist()) == set(range(12))


def brute_densest(n, cl):
    """(cliques, vertices) of the optimum and the union of all optimal sets."""
    masks = np.arange(1, 1 << n, dtype=np.int64)
    sizes = np.array([bin(int(m)).count("1") for m in masks])
    cm = np.array([sum(1 << v for v in c) for c in cl], dtype=np.int64)
    tin = ((masks[:, None] & cm[None, :]) == cm[None, :]).sum(1)
    best = (0, 1)
    for i in range(len(masks)):
        if tin[i] * best[1] > best[0] * sizes[i]:
            best = (int(tin[i]), int(sizes[i]))
    union = 0
    for i in np.flatnonzero((tin > 0) & (tin * best[1] == best[0] * sizes)):
        union |= int(masks[i])
    return best, union


def test_brute_force_small(solver):
    rng = np.random.default_rng(7)
    for trial in range(10):
        n = int(rng.integers(6, 12))
        g = gen.er(n, float(rng.uniform(0.35, 0.8)), 0, seed=100 + trial)
        adj = [set(g.nbr[g.offset[v]:g.offset[v + 1]].tolist()) for v in range(n)]
        for k in (3, 4):
            cl = [c for c in itertools.combinations(range(n), k)
                  if all(b in adj[a] for a, b in itertools.combinations(c, 2))]
            solver.upload(g.offset, g.nbr)
            res, ins = solver.solve(k)
            assert res.T == len(cl)
            if not cl:
                continue
            best, union = brute_densest(n, cl)
            assert res.rho_num * best[1] == best[0] * res.rho_den
            assert int(sum(1 << int(v) for v in np.flatnonzero(ins))) == union


def test_config1_closed_form(solver):
    g, k = gen.config(1)
    solver.upload(g.offset, g.nbr)
    res, ins = solver.solve(k)
    assert (res.rho_num, res.rho_den) == (math.comb(40, 3), 40)
    assert set(np.flatnonzero(ins).tolist()) == set(g.planted.tolist())


def test_cli_matches_reference_output_lines():
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "test_amd", "kcds")
    out = subprocess.run([exe, os.path.join(GOLD, "ref_data_test.txt"), "3", "1000", "1000", "100", "10"],
                         stdout=subprocess.PIPE, text=True, check=True).stdout.splitlines()
    assert out[-2] == "Max Density 1 Size 4"
    assert out[-1].startswith("Time taken: ") and out[-1].endswith(" ms")
    assert "Exact Density 4/4" in out
    out = subprocess.run([exe, os.path.join(GOLD, "ref_data_test.txt"), "4"],
                         stdout=subprocess.PIPE, text=True, check=True).stdout.splitlines()
    assert out[-2] == "Max Density 0.25 Size 4"


def test_cli_on_c1_edge_list(tmp_path):
    """C1 (ER(2000, 0.005) + K40, k = 3) written as a 'u v' edge list and run
    through the drop-in CLI: the oracle's exact density 9880/40 and the K40."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "test_amd", "kcds")
    g, k = gen.config(1)
    off, nbr = np.asarray(g.offset, np.int64), np.asarray(g.nbr, np.int64)
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    keep = src < nbr
    path = tmp_path / "c1_edges.txt"
    np.savetxt(path, np.stack([src[keep], nbr[keep]], 1), fmt="%d")
    r, _, _, _, _ = O.solve(off, nbr, k)
    assert (r.rho_num, r.rho_den) == (9880, 40)
    out = subprocess.run([exe, str(path), str(k)], stdout=subprocess.PIPE, text=True, check=True).stdout.splitlines()
    assert "Exact Density 9880/40" in out
    assert out[-2] == "Max Density 247 Size 40"
