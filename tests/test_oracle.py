"""Pins the CPU oracle (runs without a GPU).

Each oracle function is checked against something independent of it:
  * loader + BZ ordering: the reference's own src/graph.cpp compiled into
    oracle/_ref (skipped where the reference sources are absent) and the
    hand-derived values for the reference's data/test.txt;
  * counting / listing: brute force over all k-subsets, networkx clique
    enumeration, closed forms C(s,k) and C(s-1,k-1), sum(counts) == k*T;
  * peeling: the clique-core definition (L-core = fixpoint of deleting
    vertices in fewer than L cliques), computed independently in Python;
  * max-flow: networkx maximum_flow / minimum_cut on the same network;
  * solve: brute force over every vertex subset, the C1 closed form 247.
A planted mistake in the oracle (wrong orientation test, dropped clique,
wrong capacity) breaks at least one of these.
"""
import itertools
import json
import math
import os
import subprocess

import networkx as nx
import numpy as np
import pytest

import oracle as O
from test_amd import gen

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_graph")


def run_ref(path):
    out = subprocess.run([REF_BIN, path], stdout=subprocess.PIPE, check=True, text=True).stdout
    lines = out.strip("\n").split("\n")[-7:]  # the reference prints two log lines first
    head = list(map(int, lines[0].split()))
    arrs = [np.array(list(map(int, ln.split())), dtype=np.int64) for ln in lines[1:]]
    return head, arrs


def brute_cliques(n, adj, k):
    return [c for c in itertools.combinations(range(n), k)
            if all(b in adj[a] for a, b in itertools.combinations(c, 2))]


def adjsets(off, nbr):
    return [set(nbr[off[v]:off[v + 1]].tolist()) for v in range(len(off) - 1)]


# ---------------------------------------------------------------- loader / BZ
def test_reference_fixture_hand_values():
    off, nbr, m = O.load_adjlist(os.path.join(GOLD, "ref_data_test.txt"))
    assert len(off) == 11 and m == 14
    assert np.diff(off).tolist() == [1, 3, 2, 2, 3, 3, 4, 4, 3, 3]
    core, seq, lo, kmax = O.core_decompose(off, nbr)
    assert core.tolist() == [1, 2, 2, 2, 2, 2, 3, 3, 3, 3]
    assert seq.tolist() == [6, 7, 9, 8, 4, 5, 1, 2, 3, 0]
    assert lo.tolist() == [10, 7, 8, 9, 5, 6, 1, 2, 4, 3]
    assert kmax == 3


def test_golden_file_matches_oracle():
    with open(os.path.join(GOLD, "ref_graph_outputs.json")) as f:
        gold = json.load(f)
    for case in gold["cases"]:
        path = os.path.join(GOLD, case["file"])
        off, nbr, m = O.load_adjlist(path)
        core, seq, lo, kmax = O.core_decompose(off, nbr)
        assert off.tolist() == case["offset"]
        assert nbr.tolist() == case["neighbors"]
        assert core.tolist() == case["core"]
        assert seq.tolist() == case["corePeelSequence"]
        assert lo.tolist() == case["listingOrder"]
        assert kmax == case["kmax"]


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_loader_and_bz_vs_compiled_reference(tmp_path, seed):
    g = gen.rmat(10, 6, (9,), seed=seed)
    path = str(tmp_path / "g.txt")
    g.write_adjlist(path)
    (n, m, kmax), (roff, rnbr, rdeg, rcore, rseq, rlo) = run_ref(path)
    off, nbr, m2 = O.load_adjlist(path)
    core, seq, lo, km = O.core_decompose(off, nbr)
    assert (n, m, kmax) == (len(off) - 1, m2, km)
    np.testing.assert_array_equal(off, roff)
    np.testing.assert_array_equal(nbr, rnbr)
    np.testing.assert_array_equal(core, rcore)
    np.testing.assert_array_equal(seq, rseq)
    np.testing.assert_array_equal(lo, rlo)
    np.testing.assert_array_equal(rdeg, rcore)  # graph.cpp leaves degree == core


def test_edgelist_and_adjlist_agree(tmp_path):
    g = gen.er(300, 0.02, 10, seed=4)
    g.write_edgelist(str(tmp_path / "e.txt"))
    g.write_adjlist(str(tmp_path / "a.txt"))
    o1, n1, m1 = O.load_edgelist(str(tmp_path / "e.txt"), n_hint=g.n)
    o2, n2, m2 = O.load_adjlist(str(tmp_path / "a.txt"))
    np.testing.assert_array_equal(o1, g.offset)
    np.testing.assert_array_equal(n1, g.nbr)
    np.testing.assert_array_equal(o2, g.offset)
    np.testing.assert_array_equal(n2, g.nbr)
    assert m1 == m2 == g.m


def test_adjlist_rejects_out_of_order_lines(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("3 1\n1 2\n0\n2 1\n")
    with pytest.raises(ValueError):
        O.load_adjlist(str(p))


def test_generator_properties():
    for g in (gen.er(500, 0.02, 12, seed=1), gen.rmat(9, 8, (10, 7), seed=2)):
        adj = adjsets(g.offset, g.nbr)
        for v in range(g.n):
            lst = g.nbr[g.offset[v]:g.offset[v + 1]]
            assert np.all(np.diff(lst.astype(np.int64)) > 0)
            assert v not in adj[v]
            assert all(v in adj[u] for u in lst)
        pl = g.planted.tolist()
        s = 12 if g.n == 500 else 10
        assert all(b in adj[a] for a, b in itertools.combinations(pl[:s], 2))
    a, b = gen.er(400, 0.03, 5, seed=9), gen.er(400, 0.03, 5, seed=9)
    np.testing.assert_array_equal(a.nbr, b.nbr)


# ---------------------------------------------------------------- counting
@pytest.mark.parametrize("seed", range(6))
def test_count_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(8, 20))
    g = gen.er(n, float(rng.uniform(0.3, 0.8)), 0, seed=seed)
    adj = adjsets(g.offset, g.nbr)
    for k in (2, 3, 4, 5):
        cl = brute_cliques(n, adj, k)
        T, pv = O.count(g.offset, g.nbr, k)
        assert T == len(cl)
        exp = np.zeros(n, np.int64)
        for c in cl:
            exp[list(c)] += 1
        np.testing.assert_array_equal(pv, exp)
        lst = O.list_cliques(g.offset, g.nbr, k)
        assert [tuple(r) for r in lst.tolist()] == cl


def test_count_vs_networkx():
    g = gen.rmat(10, 8, (14,), seed=3)
    G = nx.Graph()
    G.add_nodes_from(range(g.n))
    for v in range(g.n):
        for u in g.nbr[g.offset[v]:g.offset[v + 1]]:
            G.add_edge(v, int(u))
    by_size = {}
    for c in nx.enumerate_all_cliques(G):
        if len(c) > 5:
            break
        by_size.setdefault(len(c), []).append(c)
    for k in (3, 4, 5):
        T, pv = O.count(g.offset, g.nbr, k)
        assert T == len(by_size.get(k, []))
        exp = np.zeros(g.n, np.int64)
        for c in by_size.get(k, []):
            exp[c] += 1
        np.testing.assert_array_equal(pv, exp)
        assert pv.sum() == k * T


@pytest.mark.parametrize("s,k", [(10, 3), (20, 4), (30, 5), (12, 6)])
def test_count_closed_form(s, k):
    off = np.arange(s + 1, dtype=np.uint32) * (s - 1)
    nbr = np.array([u for v in range(s) for u in range(s) if u != v], np.uint32)
    T, pv = O.count(off, nbr, k)
    assert T == math.comb(s, k)
    assert np.all(pv == math.comb(s - 1, k - 1))


# ---------------------------------------------------------------- peeling
def clique_cores_by_definition(n, k, cl):
    """core(v) = max L with v in the L-core (fixpoint of deleting vertices that
    lie in fewer than L cliques of the remaining graph)."""
    cores = np.zeros(n, np.int64)
    dens = []
    L = 0
    while True:
        alive = np.ones(n, bool)
        while True:
            cnt = np.zeros(n, np.int64)
            live = [c for c in cl if alive[list(c)].all()]
            for c in live:
                cnt[list(c)] += 1
            drop = alive & (cnt < L)
            if not drop.any():
                break
            alive &= ~drop
        if not alive.any():
            break
        cores[alive] = L
        dens.append((len(live), int(alive.sum())))
        L += 1
    return cores, dens


@pytest.mark.parametrize("seed", range(5))
def test_peel_vs_definition(seed):
    g = gen.er(40, 0.3, 9, seed=50 + seed)
    for k in (3, 4):
        cl = [tuple(r) for r in O.list_cliques(g.offset, g.nbr, k).tolist()]
        if not cl:
            continue
        core, status, st = O.peel(g.n, k, np.array(cl, np.uint32))
        exp, dens = clique_cores_by_definition(g.n, k, cl)
        np.testing.assert_array_equal(core, exp)
        for i, c in enumerate(cl):
            assert status[i] == min(core[list(c)])
        best = max(dens, key=lambda x: x[0] / x[1])
        assert st.max_num * best[1] == best[0] * st.max_den
        kp = math.ceil(best[0] / best[1])
        assert st.k_prime == kp
        assert st.core_size == int((exp >= kp).sum())


def test_peel_disjoint_cliques_closed_form():
    s, k = 9, 4
    sizes = [s, s - 2]
    cl = []
    at = 0
    for sz in sizes:
        cl += [tuple(at + x for x in c) for c in itertools.combinations(range(sz), k)]
        at += sz
    core, status, st = O.peel(at, k, np.array(cl, np.uint32))
    assert core[:s].tolist() == [math.comb(s - 1, k - 1)] * s
    assert core[s:].tolist() == [math.comb(s - 3, k - 1)] * (s - 2)


# ---------------------------------------------------------------- max-flow
def nx_clique_flow(nc, k, cl, p, q):
    G = nx.DiGraph()
    s, t = "s", "t"
    deg = np.zeros(nc, np.int64)
    for c in cl:
        deg[list(c)] += 1
    for v in range(nc):
        G.add_edge(s, v, capacity=int(q * deg[v]))
        G.add_edge(v, t, capacity=int(k * p))
    for i, c in enumerate(cl):
        for v in c:
            G.add_edge(v, ("c", i), capacity=int(q))
            G.add_edge(("c", i), v, capacity=int(q * (k - 1)))
    val, (S, _) = nx.minimum_cut(G, s, t)
    return val


@pytest.mark.parametrize("seed", range(4))
def test_maxflow_vs_networkx(seed):
    g = gen.er(24, 0.45, 7, seed=80 + seed)
    k = 3
    cl = O.list_cliques(g.offset, g.nbr, k)
    T = len(cl)
    for p, q in [(0, 1), (T, g.n), (T, 2 * g.n), (3, 1), (T * 2, g.n)]:
        f = O.clique_maxflow(g.n, k, cl, p, q)
        assert f == nx_clique_flow(g.n, k, [tuple(r) for r in cl.tolist()], p, q)


def test_maxflow_side_is_maximal_min_cut():
    # K5 plus a pendant triangle: at alpha = rho* = 2 (10 triangles / 5) the
    # maximal min cut source side is exactly the K5
    adj = {v: set() for v in range(7)}
    for a, b in itertools.combinations(range(5), 2):
        adj[a].add(b); adj[b].add(a)
    for a, b in [(4, 5), (5, 6), (4, 6)]:
        adj[a].add(b); adj[b].add(a)
    off = np.zeros(8, np.uint32)
    for v in range(7):
        off[v + 1] = off[v] + len(adj[v])
    nbr = np.array([u for v in range(7) for u in sorted(adj[v])], np.uint32)
    cl = O.list_cliques(off, nbr, 3)
    f, side = O.clique_maxflow(7, 3, cl, 2, 1, want_side=True)
    assert f == 3 * len(cl)
    assert np.flatnonzero(side).tolist() == [0, 1, 2, 3, 4]


# ---------------------------------------------------------------- solve
def brute_densest(n, cl):
    best = (0, 1)
    union = 0
    masks = np.arange(1, 1 << n, dtype=np.int64)
    sizes = np.array([bin(int(m)).count("1") for m in masks])
    cm = np.array([sum(1 << v for v in c) for c in cl], dtype=np.int64)
    tin = ((masks[:, None] & cm[None, :]) == cm[None, :]).sum(1)
    for i in range(len(masks)):
        if tin[i] * best[1] > best[0] * sizes[i]:
            best = (int(tin[i]), int(sizes[i]))
    for i in range(len(masks)):
        if tin[i] > 0 and tin[i] * best[1] == best[0] * sizes[i]:
            union |= int(masks[i])
    return best, union


@pytest.mark.parametrize("seed", range(8))
def test_solve_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(7, 13))
    g = gen.er(n, float(rng.uniform(0.35, 0.85)), 0, seed=200 + seed)
    adj = adjsets(g.offset, g.nbr)
    for k in (2, 3, 4):  # k = 2: edge density (main.cu:1365-1423)
        cl = brute_cliques(n, adj, k)
        r, pv, core, comp, ins = O.solve(g.offset, g.nbr, k)
        assert r.T == len(cl)
        if not cl:
            continue
        (bt, bs), union = brute_densest(n, cl)
        assert r.rho_num * bs == bt * r.rho_den
        assert sum(1 << int(v) for v in np.flatnonzero(ins)) == union


def test_solve_config1_closed_form():
    g, k = gen.config(1)
    r, pv, core, comp, ins = O.solve(g.offset, g.nbr, k)
    assert (r.rho_num, r.rho_den) == (math.comb(40, 3), 40)
    assert set(np.flatnonzero(ins).tolist()) == set(g.planted.tolist())


def test_reference_fixture_solve():
    off, nbr, _ = O.load_adjlist(os.path.join(GOLD, "ref_data_test.txt"))
    r, pv, core, comp, ins = O.solve(off, nbr, 3)
    assert (r.T, r.rho_num, r.rho_den) == (4, 4, 4)
    assert np.flatnonzero(ins).tolist() == [6, 7, 8, 9]
    assert core.tolist() == [0, 0, 0, 0, 0, 0, 3, 3, 3, 3]
    r, pv, core, comp, ins = O.solve(off, nbr, 4)
    assert (r.T, r.rho_num, r.rho_den) == (1, 1, 4)
    assert np.flatnonzero(ins).tolist() == [6, 7, 8, 9]


def _core_clique_components(off, nbr, k):
    """The prune rule (reference src/helpers.cu:874-906: an edge of the densest
    core survives iff it lies in a clique whose status >= k') and networkx
    connected components of the surviving edges, from the oracle's listing and
    peel (pinned above) — not from its component code."""
    n = len(off) - 1
    cl = O.list_cliques(off, nbr, k)
    core, status, st = O.peel(n, k, cl)
    kp = int(st.k_prime)
    G = nx.Graph()
    G.add_nodes_from(int(v) for v in np.flatnonzero(core >= kp))
    for c in cl[status >= kp]:
        G.add_edges_from(itertools.combinations((int(x) for x in c), 2))
    return core, kp, {frozenset(s) for s in nx.connected_components(G)}


def _pruned_test_graphs():
    yield "two_k6_and_k5_bridged", disjoint_with_bridges([6, 6, 5], [(0, 6), (5, 12)]), 3
    yield "three_k7", disjoint_with_bridges([7, 7, 7], [(0, 7), (8, 14)]), 4
    g = gen.er(300, 0.05, 12, seed=3)
    yield "er300", (g.offset, g.nbr), 3
    g = gen.rmat(9, 6, (10, 9), seed=4)
    yield "rmat9", (g.offset, g.nbr), 4


def disjoint_with_bridges(sizes, bridges):
    n = sum(sizes)
    adj = [set() for _ in range(n)]
    at = 0
    for s in sizes:
        for i in range(at, at + s):
            adj[i].update(j for j in range(at, at + s) if j != i)
        at += s
    for u, v in bridges:
        adj[u].add(v)
        adj[v].add(u)
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum([len(a) for a in adj])
    nbr = np.array([w for a in adj for w in sorted(a)], np.uint32)
    return off, nbr


@pytest.mark.parametrize("name,g,k", list(_pruned_test_graphs()), ids=[x[0] for x in _pruned_test_graphs()])
def test_components_vs_networkx(name, g, k):
    off, nbr = g
    res, _, core_o, comp, _ = O.solve(off, nbr, k)
    core, kp, parts = _core_clique_components(off, nbr, k)
    assert res.k_prime == kp
    np.testing.assert_array_equal(core_o, core)
    assert res.ncomp == len(parts)
    # the oracle labels each densest-core vertex with its component, others with 2^32 - 1
    in_core = core >= kp
    assert np.all(comp[~in_core] == 0xFFFFFFFF)
    mine = {}
    for v in np.flatnonzero(in_core):
        mine.setdefault(int(comp[v]), set()).add(int(v))
    assert {frozenset(s) for s in mine.values()} == parts
    # every label is its component's smallest vertex (the canonical label)
    assert all(lbl == min(s) for lbl, s in mine.items())
