"""Edge list -> CSR on the GPU (kcds_upload_edges, test_amd/csrc/k_edges.hip).

The reference builds its graph on the host (src/graph.cpp:7-40).  Here the
pairs go to the device as given and the CSR is built there: both directions,
self-loops dropped, each neighbour list ascending without duplicates.

Checked against:
  - the generators' own CSR (built on the CPU by the C++ generator): the
    graph's edges, written as a shuffled list with every edge in both
    directions, extra repeats and self-loops, must give back that CSR exactly
    (an ER graph, a hub-heavy R-MAT graph, and the C2 R-MAT graph at its full
    size, 2^18 vertices with hubs of degree about 25,000);
  - numpy's np.unique over the symmetrised pairs, on random lists with many
    repeats;
  - the edge cases: no pairs, only self-loops, one vertex, a star whose
    centre has 100,000 arcs given five times over, and an id >= n (an error
    that leaves no graph behind; the next upload works);
  - the solve: kcds_solve after kcds_upload_edges equals kcds_solve after
    kcds_upload_csr of the same graph, and the oracle.
"""
import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu


def csr_pairs(off, nbr):
    """The undirected edges u < v of a CSR, m x 2."""
    off = np.asarray(off, np.int64)
    nbr = np.asarray(nbr, np.int64)
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    keep = src < nbr
    return np.stack([src[keep], nbr[keep]], axis=1)


def noisy_list(pairs, n, seed, repeat=0.2, loops=0.01):
    """pairs in both directions, with repeats and self-loops, shuffled."""
    rng = np.random.default_rng(seed)
    both = np.concatenate([pairs, pairs[:, ::-1]])
    rep = both[rng.integers(0, max(1, len(both)), int(repeat * len(both)))] if len(both) else both[:0]
    lp = rng.integers(0, max(1, n), int(loops * len(both)) + 1)
    lp = np.stack([lp, lp], axis=1) if n else np.zeros((0, 2), np.int64)
    out = np.concatenate([both, rep, lp]).astype(np.uint32)
    return out[rng.permutation(len(out))]


def numpy_csr(n, uv):
    uv = np.asarray(uv, np.int64).reshape(-1, 2)
    arcs = np.concatenate([uv, uv[:, ::-1]])
    arcs = arcs[arcs[:, 0] != arcs[:, 1]]
    key = np.unique(arcs[:, 0] * n + arcs[:, 1])
    src, dst = key // n, key % n
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(src, minlength=n))
    return off.astype(np.uint32), dst.astype(np.uint32)


def check_roundtrip(solver, g, seed):
    pairs = csr_pairs(g.offset, g.nbr)
    uv = noisy_list(pairs, g.n, seed)
    m = solver.upload_edges(g.n, uv)
    off, nbr = solver.download_csr()
    print(f"  {g.name}: {len(uv)} pairs -> {m} edges", flush=True)
    assert m == len(pairs)
    np.testing.assert_array_equal(off, np.asarray(g.offset, np.uint32))
    np.testing.assert_array_equal(nbr, np.asarray(g.nbr, np.uint32))


@pytest.mark.parametrize("which", ["er", "rmat"])
def test_edges_match_generator_csr(solver, which):
    g = gen.er(3000, 0.01, 12, seed=4) if which == "er" else gen.rmat(14, 12, (20,), seed=9)
    check_roundtrip(solver, g, seed=1)
    solver.orient(None)
    T, pv = solver.count(4)
    Te, pve = O.count(g.offset, g.nbr, 4)
    assert T == Te
    np.testing.assert_array_equal(pv, pve)


def test_edges_c2_full_size(solver):
    g, k = gen.config(2)
    check_roundtrip(solver, g, seed=2)
    solver.orient(None)
    assert solver.count(k, per_vertex=False)[0] == 2281884915  # C2's k = 4 count, as in test_gpu_list


@pytest.mark.parametrize("n,m,seed", [(50, 2000, 0), (1000, 20000, 1), (70000, 300000, 2)])
def test_edges_match_numpy(solver, n, m, seed):
    rng = np.random.default_rng(seed)
    uv = rng.integers(0, n, (m, 2)).astype(np.uint32)
    uv = np.concatenate([uv, uv[rng.integers(0, m, m // 2)]])  # many repeats
    got = solver.upload_edges(n, uv)
    off, nbr = solver.download_csr()
    eo, en = numpy_csr(n, uv)
    assert got == len(en) // 2
    np.testing.assert_array_equal(off, eo)
    np.testing.assert_array_equal(nbr, en)


def test_edges_edge_cases(solver):
    assert solver.upload_edges(10, np.zeros((0, 2), np.uint32)) == 0
    off, nbr = solver.download_csr()
    assert off.tolist() == [0] * 11 and len(nbr) == 0
    assert solver.upload_edges(5, np.array([[3, 3], [0, 0], [4, 4]], np.uint32)) == 0
    off, _ = solver.download_csr()
    assert off.tolist() == [0] * 6
    assert solver.upload_edges(1, np.array([[0, 0]], np.uint32)) == 0
    # a star: centre 0 with 100,000 leaves, every edge five times, both ways
    leaves = np.arange(1, 100_001, dtype=np.uint32)
    star = np.stack([np.zeros_like(leaves), leaves], axis=1)
    uv = np.concatenate([star, star[:, ::-1], star, star[:, ::-1], star])
    uv = uv[np.random.default_rng(3).permutation(len(uv))]
    assert solver.upload_edges(100_001, uv) == 100_000
    off, nbr = solver.download_csr()
    assert off[1] == 100_000 and np.array_equal(nbr[:100_000], leaves)
    assert np.all(np.diff(off[1:].astype(np.int64)) == 1) and np.all(nbr[100_000:] == 0)


def test_edges_bad_id_leaves_no_graph(solver):
    with pytest.raises(kcds.KcdsError) as e:
        solver.upload_edges(10, np.array([[1, 2], [3, 10]], np.uint32))
    assert e.value.code == kcds.E_INVALID
    with pytest.raises(kcds.KcdsError) as e:
        solver.orient(None)
    assert e.value.code == kcds.E_STATE
    assert solver.upload_edges(10, np.array([[1, 2], [2, 3], [3, 1]], np.uint32)) == 3
    solver.orient(None)
    assert solver.count(3)[0] == 1


@pytest.mark.parametrize("k", [3, 4])
def test_edges_solve_equals_csr_solve(solver, k):
    g = gen.er(400, 0.05, 14, seed=21)
    solver.upload(g.offset, g.nbr)
    r0, s0 = solver.solve(k)
    solver.upload_edges(g.n, noisy_list(csr_pairs(g.offset, g.nbr), g.n, seed=5))
    r1, s1 = solver.solve(k)
    assert (r1.rho_num, r1.rho_den) == (r0.rho_num, r0.rho_den)
    np.testing.assert_array_equal(s1, s0)
    r, _, _, _, ins = O.solve(g.offset, g.nbr, k)
    assert (r1.rho_num, r1.rho_den) == (r.rho_num, r.rho_den)
    np.testing.assert_array_equal(s1, ins)
