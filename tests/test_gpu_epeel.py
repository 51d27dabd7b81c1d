"""GPU parity of the enumerating peel: the clique-core peel that keeps no clique
list and enumerates each removed vertex's live cliques from its
neighbourhood (narrow frames for degree <= 512, stored before the first
level or built at the step; candidate-split frames for wider vertices), then
lists only the k'-core's cliques.

Bit-exact against the oracle's peel over the materialised clique list (core
numbers, maxDensity as a rational, k', core size, core cliques, largest
core) and against the library's own materialising peel; whole solves along
this path must equal the oracle's (density rational and maximal set).
"""
import os

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def graphs():
    """(name, offset, nbr, k values peeled, k values solved)."""
    off, nbr, _ = O.load_adjlist(os.path.join(GOLD, "ref_data_test.txt"))
    yield "ref_data_test", off, nbr, (2, 3, 4), (2, 3, 4)
    g = gen.er(400, 0.04, 18, seed=21)
    yield "er400", g.offset, g.nbr, (2, 3, 4, 5, 6), (2, 3, 4, 5, 6)
    g = gen.rmat(10, 8, (12,), seed=9)
    yield "rmat10", g.offset, g.nbr, (3, 4), (3, 4)
    # 13 hubs of degree above 512 (up to 1,314) take the candidate-split form
    g = gen.rmat(12, 16, (30, 20), seed=5)
    yield "rmat12_hubs", g.offset, g.nbr, (3, 4), (3,)
    g = gen.er(1500, 0.02, 30, seed=3)
    yield "er1500", g.offset, g.nbr, (3, 4, 7), (3, 4)
    # 25 vertices of degree above 512 (up to 875)
    g = gen.chunglu(20000, 2.1, 2.0, 900.0, 120000, 25, 0.95, seed=17)
    yield "chunglu20k", g.offset, g.nbr, (4,), (4,)


GRAPHS = list(graphs())


MODES = {"dense_mode": kcds.TUNE_DENSE_PEEL, "steps_only": 0, "built_frames": kcds.TUNE_BUILD_FRAMES}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name,off,nbr,ks,_solve_ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_enum_peel_matches_oracle(solver, name, off, nbr, ks, _solve_ks, mode):
    """dense_mode: a single graph's peel goes over to the bitmask-matrix mode
    once at most 1,024 vertices are live (from the start for graphs that
    small); steps_only: the default, the step kernels to the end, narrow
    frames loaded from the frames stored before the first level;
    built_frames: every frame built from the out-lists at its step."""
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    tune = MODES[mode]
    for k in ks:
        solver.set_tuning(0, tune)
        solver.upload(off, nbr)
        solver.orient(None)
        T, _ = solver.count(k, per_vertex=False)
        cl = O.list_cliques(off, nbr, k)
        assert T == len(cl)
        if T == 0:
            continue
        core, st = solver.peel()  # counted, not listed: the enumerating peel
        core_r, _, st_r = O.peel(n, k, cl)
        print(f"  {name} k={k} T={T} maxdeg={deg.max()} maxDensity={st.max_num}/{st.max_den} k'={st.k_prime} "
              f"levels={st.levels}", flush=True)
        np.testing.assert_array_equal(core, core_r)
        assert (st.max_num, st.max_den, st.k_prime, st.core_size, st.core_cliques, st.max_core) == \
            (st_r.max_num, st_r.max_den, st_r.k_prime, st_r.core_size, st_r.core_cliques, st_r.max_core)
        # the materialising peel on the same graph gives the same cores
        solver.upload(off, nbr)
        solver.orient(None)
        solver.list(k)
        core_m, st_m = solver.peel()
        np.testing.assert_array_equal(core, core_m)
        assert (st.max_num, st.max_den, st.levels) == (st_m.max_num, st_m.max_den, st_m.levels)
    solver.set_tuning(0, 0)


def test_enum_peel_reference_order(solver):
    """The reference's own listing order (high core -> low core arcs) as the DAG."""
    g = gen.er(400, 0.04, 18, seed=21)
    n = g.n
    _, _, lo, _ = O.core_decompose(g.offset, g.nbr)
    for k in (3, 4):
        solver.upload(g.offset, g.nbr)
        solver.orient(lo)
        solver.count(k, per_vertex=False)
        core, st = solver.peel()
        core_r, _, st_r = O.peel(n, k, O.list_cliques(g.offset, g.nbr, k))
        np.testing.assert_array_equal(core, core_r)
        assert (st.max_num, st.max_den, st.k_prime) == (st_r.max_num, st_r.max_den, st_r.k_prime)


@pytest.mark.parametrize("name,off,nbr,_peel_ks,ks", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_enum_solve_matches_oracle(solver, name, off, nbr, _peel_ks, ks):
    """Whole solves along the enumerating path: peel, k'-core listing, components, flow."""
    solver.set_tuning(0, kcds.TUNE_ENUM_PEEL)
    try:
        for k in ks:
            solver.upload(off, nbr)
            res, ins = solver.solve(k)
            r, _, _, _, ins_r = O.solve(off, nbr, k)
            print(f"  {name} k={k} rho={res.rho_num}/{res.rho_den} core_cliques={res.peel.core_cliques}", flush=True)
            assert res.T == r.T
            assert (res.rho_num, res.rho_den) == (r.rho_num, r.rho_den)
            assert (res.k_prime, res.core_vertices) == (r.k_prime, r.core_vertices)
            assert res.peel.core_cliques == r.peel.core_cliques
            np.testing.assert_array_equal(ins, ins_r)
    finally:
        solver.set_tuning(0, 0)


def test_enum_batch_equals_solo(solver):
    """A batch along the enumerating path equals each graph solved alone."""
    graphs = [gen.er(300 + 40 * i, 0.05, 10 + i, seed=400 + i) for i in range(6)]
    graphs.append(gen.rmat(11, 12, (16,), seed=77))
    pairs = [(g.offset, g.nbr) for g in graphs]
    k = 4
    solver.set_tuning(0, kcds.TUNE_ENUM_PEEL)
    try:
        res, sets = solver.solve_batch(pairs, k)
        for (o, b), rb, sb in zip(pairs, res, sets):
            solver.upload(o, b)
            r1, s1 = solver.solve(k)
            assert (rb.T, rb.rho_num, rb.rho_den, rb.k_prime) == (r1.T, r1.rho_num, r1.rho_den, r1.k_prime)
            np.testing.assert_array_equal(sb, s1)
    finally:
        solver.set_tuning(0, 0)
