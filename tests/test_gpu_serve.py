"""GPU: the serving loop (native server behind kcds_submit / kcds_poll)
answers every request exactly as a solo solve does.

Requests of different sizes and k arrive close together, are batched
continuously across several device slots, and each result (T, the exact
density as a rational, the maximal densest vertex set) must equal the
request solved alone.  A request over the per-graph clique budget moves to
the large lane (solved alone without a clique list) and still equals its
solo solve; with no large lane it is answered over budget.  A seeded slice
of the C5 request stream (R-MAT, n log-uniform in 1e3..2e5, edge factor
4..24, k in 3..6) is served against its solo solves and, where it is small
enough, against the oracle.
"""
import math

import numpy as np
import pytest

import oracle as O
from test_amd import gen, kcds
from test_amd.serve import Request, Server, c5_requests

pytestmark = pytest.mark.gpu


def requests():
    out = []
    rng = np.random.default_rng(11)
    for i in range(30):
        n = int(rng.integers(60, 900))
        k = int(rng.integers(3, 7))
        g = gen.rmat(int(np.ceil(np.log2(n))), float(rng.uniform(3, 8)), (int(rng.integers(6, 14)),),
                     seed=500 + i, n=n)
        out.append(Request(rid=i, t_arrival=0.002 * i, offset=np.asarray(g.offset, np.uint32),
                           nbr=np.asarray(g.nbr, np.uint32), k=k))
    # one request far over the per-graph budget below: K60 at k = 4 has C(60,4) = 487,635 cliques
    off = np.arange(61, dtype=np.uint32) * 59
    nbr = np.array([u for v in range(60) for u in range(60) if u != v], np.uint32)
    out.append(Request(rid=30, t_arrival=0.03, offset=off, nbr=nbr, k=4))
    return out


BUDGET = 4 << 20  # bytes of clique list per graph in a batch


def solo_results(reqs, max_cliques=0):
    solo = {}
    with kcds.Solver(0) as s:
        for r in reqs:
            s.set_graph_budget(max_cliques * (8 * r.k + 4) if max_cliques else 0)
            s.upload(r.offset, r.nbr)
            try:
                res, ins = s.solve(r.k)
                solo[r.rid] = (kcds.GRAPH_SOLVED, res.T, res.rho_num, res.rho_den, ins)
            except kcds.KcdsError as e:
                assert e.code == kcds.E_NOMEM
                solo[r.rid] = (kcds.GRAPH_OVER_BUDGET, None, None, None, None)
    return solo


def check_against(out, solo):
    for r in out:
        exp = solo[r.rid]
        assert r.status == exp[0], (r.rid, r.status, r.code, r.error)
        if r.status == kcds.GRAPH_SOLVED:
            assert (r.T, r.rho_num, r.rho_den) == exp[1:4], r.rid
            np.testing.assert_array_equal(r.in_set, exp[4])


def test_serve_matches_solo():
    reqs = requests()
    solo = solo_results(reqs)
    assert solo[30][:3] == (kcds.GRAPH_SOLVED, math.comb(60, 4), math.comb(60, 4))
    out, st = Server(slots=3, large_slots=1, batch_edges=200_000, graph_budget=BUDGET, want_sets=True).run(reqs)
    print(f"  served {st.offered}: solved {st.solved}, large lane {st.large_lane}, batches {st.batches}, "
          f"mean batch {st.mean_batch:.1f}, p50 {st.p50_ms:.1f} ms, p99 {st.p99_ms:.1f} ms", flush=True)
    assert st.failed == 0 and st.solved == len(reqs)
    assert st.batches < len(reqs)  # requests were batched
    assert [r for r in out if r.rid == 30][0].lane == 1  # the K60 took the large lane
    check_against(out, solo)


def test_serve_without_large_lane_answers_over_budget():
    reqs = requests()
    out, st = Server(slots=2, large_slots=0, batch_edges=200_000, graph_budget=BUDGET).run(reqs)
    k60 = [r for r in out if r.rid == 30][0]
    assert k60.status == kcds.GRAPH_OVER_BUDGET and k60.code == kcds.E_NOMEM
    assert st.solved == len(reqs) - 1


def test_serve_c5_slice():
    """The first requests of the C5 stream up to 20,000 vertices (every k), served
    with a small batch budget so several take the large lane."""
    reqs = c5_requests(count=36, max_n=20_000, time_scale=0.05)
    cap = 50_000_000
    solo = solo_results(reqs, max_cliques=cap)
    out, st = Server(slots=3, large_slots=2, graph_budget=16 << 20, max_cliques=cap, want_sets=True).run(reqs)
    print(f"  C5 slice {st.offered}: solved {st.solved}, over budget {st.over_budget}, large lane {st.large_lane}, "
          f"batches {st.batches}, max queue {st.max_pending}, p50 {st.p50_ms:.1f} ms, p99 {st.p99_ms:.1f} ms",
          flush=True)
    assert st.failed == 0 and st.cancelled == 0
    assert st.large_lane > 0
    check_against(out, solo)
    # the oracle where it is cheap
    checked = 0
    for r in out:
        if r.status == kcds.GRAPH_SOLVED and r.T <= 200_000 and r.n <= 6000:
            ro, _, _, _, ins_r = O.solve(r.offset, r.nbr, r.k)
            assert (r.T, r.rho_num, r.rho_den) == (ro.T, ro.rho_num, ro.rho_den), r.rid
            np.testing.assert_array_equal(r.in_set, ins_r)
            checked += 1
    print(f"  {checked} checked against the oracle", flush=True)
    assert checked > 0


def test_serve_deadline_cancels_unstarted():
    """Past the deadline, requests no slot has started are answered cancelled;
    every request is answered and the served ones equal their solo solves."""
    base = requests()
    reqs = [Request(rid=i, t_arrival=0.0, offset=r.offset, nbr=r.nbr, k=r.k)
            for i, r in enumerate(base * 6)]
    out, st = Server(slots=1, large_slots=0, max_batch=1, graph_budget=BUDGET).run(reqs, deadline_s=0.0)
    assert st.cancelled > 0
    assert st.solved + st.over_budget + st.cancelled == len(reqs)
    assert all(r.status in (kcds.GRAPH_SOLVED, kcds.GRAPH_OVER_BUDGET) for r in out)
