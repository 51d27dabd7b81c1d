"""Benchmark: k-cliques listed per second on one MI355X.

Workload (one step = one pass of the hot path over the graph already resident
in HBM): config C2 -- R-MAT scale 18 (n = 262,144), edge factor 16,
(a,b,c,d) = (.57,.19,.19,.05), symmetrised and deduplicated, plus a planted
60-clique and five 30-cliques; k = 4.  A step is the GPU listing order (the
vertices sorted by (degree, id)) + DAG build (generateDAG) followed by the
wave64 listing pass that counts every 4-clique and every per-vertex 4-clique
count (listAllCliques' T and cliqueDegree).  value = T / seconds per step.

Also reported on the same JSON line:
  roofline      the listing kernels' algorithmic HBM bytes per pass
                B = 4(n+1) + 4*arcs + 4*W2 + 8n (DAG CSR read once, every
                root's out-neighbour rows staged, int64 counts written) over
                their hipEvent-timed duration, against 8 TB/s;
  cpu_baseline  the CPU oracle's count of the same graph (all host threads);
  extras        exact k-CDS solves/s on config C1 (ER 2000 + K40, k = 3);
                the exact solves of C2 and C3 (dense exact search on their
                k'-cores); the same counting step on C3 (Chung-Lu 1M, k = 5)
                with its roofline figures; C4 batched vs plain solves/s (4,096
                graphs, k = 4); C5 served vs plain over the first second of the
                Poisson 200/s stream (solves/s, p50/p99, rejections, queue
                depth); the CPU oracle on one thread and the C1 oracle solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def bench_c4(np, gen, kcds, plain_sample):
    """C4: 4,096 disjoint graphs (ER + planted clique, k = 4) solved exactly.

    batched: one kcds_solve_batch over the union already resident in HBM (one
    orientation, one listing, one workgroup-per-graph peel, one launch for all
    components' exact searches); plain: one kcds_solve per graph on a sample,
    each graph uploaded first (upload time excluded)."""
    k = 4
    specs = gen.batch_specs()
    graphs = [gen.batch_graph(sp) for sp in specs]
    pairs = [(g.offset, g.nbr) for g in graphs]
    off, nbr, voff = kcds.concat_graphs(pairs)
    s = kcds.Solver(0)
    t0 = time.perf_counter()
    s.upload(off, nbr)
    s.sync()
    t_up = time.perf_counter() - t0
    s.orient(None)
    tuned = s.autotune_list(k)  # setup: launch shape for this union
    s.solve_resident_batch(voff, k)  # warm
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        res, _ = s.solve_resident_batch(voff, k)
    t_b = (time.perf_counter() - t0) / reps
    nb = len(pairs)
    # plain form: one graph per call
    m = min(plain_sample, nb)
    t_solve = 0.0
    for i in range(m):
        s.upload(*pairs[i])
        s.sync()
        t1 = time.perf_counter()
        r1, _ = s.solve(k, want_set=False)
        t_solve += time.perf_counter() - t1
        assert (r1.T, r1.rho_num, r1.rho_den) == (res[i].T, res[i].rho_num, res[i].rho_den)
    s.close()
    return {"C4_graphs": nb, "C4_batched_solves_per_s": nb / t_b, "C4_batched_ms": t_b * 1e3,
            "C4_plain_solves_per_s": m / t_solve, "C4_plain_sample": m,
            "C4_batched_over_plain": (nb / t_b) / (m / t_solve), "C4_union_upload_ms": t_up * 1e3,
            "C4_T_total": int(sum(r.T for r in res)), "C4_waves_per_root": tuned[0]}


def bench_c5(window_s, graph_budget, max_cliques, slots, large_slots, deadline_s):
    """C5: the Poisson 200/s R-MAT stream (k in 3..6) over its first window_s seconds.

    served: the native continuous-batching server (kcds_submit / kcds_poll)
    with `slots` batch slots and `large_slots` large-lane slots; a graph whose
    clique list exceeds graph_budget bytes is solved alone in the large lane
    without a clique list; one with more than max_cliques cliques is answered
    over budget.  plain: the same stream through one slot, one graph per call,
    the same clique cap.  Requests not started deadline_s after the last
    arrival are cancelled.  Latency runs from the scheduled arrival to the
    answer; p50/p99 are given over the solved requests and over every offered
    request (a cancelled one counts as waiting to the end of the run)."""
    import copy
    from test_amd import kcds, serve
    t0 = time.perf_counter()
    reqs = serve.c5_requests(duration=window_s)
    t_gen = time.perf_counter() - t0
    out = {"C5_window_s": window_s, "C5_requests": len(reqs), "C5_graph_budget_bytes": graph_budget,
           "C5_max_cliques": max_cliques, "C5_slots": slots, "C5_large_slots": large_slots,
           "C5_deadline_s": deadline_s, "C5_generate_s": t_gen}
    warm = [copy.copy(r) for r in reqs[: min(8, len(reqs))]]
    for r in warm:
        r.t_arrival = 0.0
    serve.Server(slots=slots, large_slots=large_slots, graph_budget=graph_budget,
                 max_cliques=max_cliques).run(warm, deadline_s=deadline_s)  # warm every path's kernels
    forms = (("served", serve.Server(slots=slots, large_slots=large_slots, graph_budget=graph_budget,
                                     max_cliques=max_cliques), False),
             ("plain", serve.Server(slots=1, large_slots=0, graph_budget=0, max_cliques=max_cliques), True))
    results = {}
    for form, srv, plain in forms:
        done, st = srv.run([copy.copy(r) for r in reqs], plain=plain, deadline_s=deadline_s)
        results[form] = {r.rid: (r.status, r.T, r.rho_num, r.rho_den) for r in done}
        for key in ("solves_per_s", "p50_ms", "p99_ms", "p50_all_ms", "p99_all_ms", "wall_s", "solved",
                    "over_budget", "cancelled", "failed", "large_lane", "max_pending", "mean_batch"):
            out[f"C5_{form}_{key}"] = getattr(st, key)
        out[f"C5_{form}_rejected_frac"] = (st.over_budget + st.cancelled + st.failed) / max(1, st.offered)
    both = [rid for rid, v in results["served"].items()
            if v[0] == kcds.GRAPH_SOLVED and results["plain"].get(rid, (None,))[0] == kcds.GRAPH_SOLVED]
    out["C5_served_equals_plain"] = all(results["served"][rid] == results["plain"][rid] for rid in both)
    out["C5_compared"] = len(both)
    return out


def bench_c2_solve(g, k, kcds):
    """The exact solve of the headline graph (peel, dense exact search), one run."""
    s = kcds.Solver(0)
    s.upload(g.offset, g.nbr)
    t0 = time.perf_counter()
    res, _ = s.solve(k, want_set=False)
    dt = time.perf_counter() - t0
    s.close()
    return {"C2_solve_ms": dt * 1e3, "C2_rho": f"{res.rho_num}/{res.rho_den}", "C2_k_prime": res.k_prime,
            "C2_core_vertices": res.core_vertices, "C2_core_cliques": res.peel.core_cliques,
            "C2_dense_passes": res.dense_passes,
            "C2_stage_ms": {"orient": res.ms_orient, "list": res.ms_list, "peel": res.ms_peel,
                            "dense_or_components": res.ms_components, "flow": res.ms_flow}}


def bench_c2_edges(np, g, kcds, reps=3):
    """The headline graph handed over as an edge list: every edge once in a random direction plus 20 % repeats,
    shuffled; kcds_upload_edges builds the CSR on the GPU (host pairs, so the time includes their PCIe copy)."""
    rng = np.random.default_rng(7)
    off = np.asarray(g.offset, np.int64)
    src = np.repeat(np.arange(g.n), np.diff(off))
    dst = np.asarray(g.nbr, np.int64)
    pairs = np.stack([src[src < dst], dst[src < dst]], axis=1)
    flip = rng.random(len(pairs)) < 0.5
    pairs[flip] = pairs[flip][:, ::-1]
    uv = np.concatenate([pairs, pairs[rng.integers(0, len(pairs), len(pairs) // 5)]]).astype(np.uint32)
    uv = np.ascontiguousarray(uv[rng.permutation(len(uv))])
    s = kcds.Solver(0)
    s.upload_edges(g.n, uv)  # warm-up: buffers reserved
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        m = s.upload_edges(g.n, uv)
        times.append(time.perf_counter() - t0)
    o2, n2 = s.download_csr()
    same = bool(np.array_equal(o2, g.offset) and np.array_equal(n2, g.nbr))
    s.close()
    t = min(times)
    return {"C2_edges_pairs": len(uv), "C2_edges_distinct": m, "C2_edges_to_csr_ms": t * 1e3,
            "C2_edges_pairs_per_s": len(uv) / t, "C2_edges_csr_equal": same}


def bench_c3(gen, kcds, reps=3, solve=True, shape=None):
    """C3 (Chung-Lu n = 1M, ~10M edges, near-K80, k = 5): the same step as the
    headline (GPU (degree, id) order + DAG + per-vertex counting pass) with its own roofline
    figures, and one exact solve (enumerating peel, dense exact search: T =
    7.3e10 cliques are never listed)."""
    g, k = gen.config(3)
    s = kcds.Solver(0)
    s.upload(g.offset, g.nbr)
    s.orient(None)
    if shape:
        s.set_list_shape(shape)
    else:
        s.autotune_list(k)
    s.count(k, per_vertex=False)
    s.sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(reps):
        s.orient(None)
        T, _ = s.count(k, per_vertex=False)
        kms.append(s.list_info().kernel_ms)
    s.sync()
    dt = (time.perf_counter() - t0) / reps
    info = s.list_info()
    km = sum(kms) / len(kms)
    b = 4 * (info.n + 1) + 4 * info.arcs + 4 * info.w2 + 8 * info.n
    traffic, traffic_src = committed_traffic("list_count_c3_traffic.json")
    out = {"C3_T": int(T), "C3_step_ms": dt * 1e3, "C3_cliques_per_s": T / dt, "C3_list_pass_ms": km,
           "C3_max_outdeg": info.max_outdeg, "C3_waves_per_root": list(info.waves_per_root),
           "C3_roofline": {"bound": "hbm", "achieved": b / (km / 1e3) / 1e9, "peak": HBM_PEAK / 1e9,
                           "unit": "GB/s", "frac": b / (km / 1e3) / HBM_PEAK, "traffic": traffic,
                           "traffic_source": traffic_src, "kernel": "k_list_count<W,5>", "kernel_ms": km,
                           "w2": int(info.w2), "bytes_per_pass": b}}
    if solve:
        t1 = time.perf_counter()
        res, _ = s.solve(k, want_set=False)
        out.update({"C3_solve_ms": (time.perf_counter() - t1) * 1e3, "C3_rho": f"{res.rho_num}/{res.rho_den}",
                    "C3_k_prime": res.k_prime, "C3_core_vertices": res.core_vertices,
                    "C3_core_cliques": res.peel.core_cliques, "C3_dense_passes": res.dense_passes,
                    "C3_stage_ms": {"orient": res.ms_orient, "list": res.ms_list, "peel": res.ms_peel,
                                    "dense_or_components": res.ms_components, "flow": res.ms_flow}})
    s.close()
    return out


def committed_traffic(name):
    """HBM bytes per pass from a committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, each in its own
    run, scaled by a plain-copy calibration); counters cannot be read from inside this process."""
    tpath = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(tpath):
        return None, None
    with open(tpath) as fh:
        tj = json.load(fh)
    return tj.get("traffic_bytes"), tj.get("source")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default waves per root (skip the autotuner)")
    ap.add_argument("--shape", default=None, help="waves per root of the 4 width classes, e.g. 2,2,4,8 (skips the autotuner)")
    ap.add_argument("--c4-plain-sample", type=int, default=256,
                    help="graphs of C4 solved one per call for the plain-form rate")
    ap.add_argument("--c5-window", type=float, default=1.0,
                    help="seconds of the C5 stream served (0 skips the serving leg)")
    ap.add_argument("--c5-graph-budget", type=int, default=2 << 30,
                    help="clique-list bytes per graph in a serving batch (larger graphs take the large lane)")
    ap.add_argument("--c5-max-cliques", type=int, default=1 << 28,
                    help="graphs with more cliques are answered over budget")
    ap.add_argument("--c5-slots", type=int, default=4)
    ap.add_argument("--c5-large-slots", type=int, default=4)
    ap.add_argument("--c5-deadline", type=float, default=60.0,
                    help="seconds after the last arrival before unstarted requests are cancelled")
    ap.add_argument("--no-solves", action="store_true", help="skip the exact C2 and C3 solves")
    ap.add_argument("--only", choices=["c2solve", "c3", "c4", "c5"], default=None,
                    help="run one extras leg alone and print its figures (for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the listing shards by root ranges with one all-reduce of int64 counts; not built here)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from test_amd import gen, kcds

    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.int32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "copies": 5}))
        return

    if args.only:
        if args.only == "c2solve":
            g, k = gen.config(2)
            out = bench_c2_solve(g, k, kcds)
            out.update(bench_c2_edges(np, g, kcds))
        elif args.only == "c3":
            out = bench_c3(gen, kcds, solve=not args.no_solves,
                           shape=[int(x) for x in args.shape.split(",")] if args.shape else None)
        elif args.only == "c4":
            out = bench_c4(np, gen, kcds, args.c4_plain_sample)
        else:
            out = bench_c5(args.c5_window, args.c5_graph_budget, args.c5_max_cliques, args.c5_slots,
                           args.c5_large_slots, args.c5_deadline)
        print(json.dumps({"only": args.only, **out}))
        return

    g, k = gen.config(2)
    s = kcds.Solver(0)
    s.upload(g.offset, g.nbr)

    def step():
        s.orient(None)
        T, _ = s.count(k, per_vertex=False)
        return T

    tuned = None
    if args.shape:
        s.set_list_shape([int(x) for x in args.shape.split(",")])
    elif not args.no_tune:  # setup, outside the timed region: pick waves per root for this graph
        s.orient(None)
        tuned = s.autotune_list(k)

    for _ in range(args.warmup):
        step()
    s.sync()
    kern_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        T = step()
        kern_ms.append(s.list_info().kernel_ms)
    s.sync()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    info = s.list_info()
    value = T / (ms / 1e3)

    n, arcs, w2 = info.n, info.arcs, info.w2
    bytes_list = 4 * (n + 1) + 4 * arcs + 4 * w2 + 8 * n
    kms = float(np.mean(kern_ms))
    achieved = bytes_list / (kms / 1e3)
    # HBM bytes per pass from the committed rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE scaled by a
    # plain-copy calibration); counters cannot be read from inside this process
    traffic, traffic_src = committed_traffic("list_count_c2_traffic.json")
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "k_list_count<W,4> (one launch per width class W=1,2,4,8, concurrently on four streams)",
                "kernel_ms": kms, "bytes_per_pass": bytes_list}

    cpu = None
    cpu_extra = {}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = os.cpu_count() or 1
        t1 = time.perf_counter()
        Tc, _ = O.count(g.offset, g.nbr, k, per_vertex=True, threads=threads)
        tc = time.perf_counter() - t1
        assert Tc == T, (Tc, T)
        cpu = {"value": Tc / tc, "unit": "cliques/s", "cores": threads, "kind": "port",
               "sample": f"full C2 graph, k={k}, per-vertex counts (oracle kClist-style listing, OpenMP)",
               "seconds": tc, "cpu_model": cpu_model()}
        # one thread on a bounded sample: the first roots of the oracle's own
        # degeneracy DAG, grown until the sample takes about 10 s
        lo, hi, t_one, T_one = 0, 4096, 0.0, 0
        while hi <= g.n:
            t1 = time.perf_counter()
            T_one = O.count_roots(g.offset, g.nbr, k, lo, hi, threads=1)
            t_one = time.perf_counter() - t1
            if t_one > 5.0 or hi == g.n:
                break
            hi = min(g.n, hi * 2 if t_one < 2.5 else int(hi * 10.0 / max(t_one, 1e-3)))
        cpu_extra = {"cpu_single_thread_cliques_per_s": T_one / t_one if t_one > 0 else None,
                     "cpu_single_thread_sample": f"C2 k={k}, roots [0, {hi}) of the oracle's degeneracy DAG: "
                                                 f"{T_one} cliques in {t_one:.2f} s",
                     "cpu_model": cpu_model(), "cpu_threads_all_cores": threads}
        # the reference-compatible solve window on C1 (CSR in memory -> final density;
        # the oracle's solve includes its BZ ordering)
        g1, k1 = gen.config(1)
        t1 = time.perf_counter()
        r1 = O.solve(g1.offset, g1.nbr, k1)[0]
        cpu_extra["cpu_C1_oracle_solve_ms"] = (time.perf_counter() - t1) * 1e3
        cpu_extra["cpu_C1_oracle_rho"] = f"{r1.rho_num}/{r1.rho_den}"

    extras = {}
    if not args.no_extras:
        g1, k1 = gen.config(1)
        s1 = kcds.Solver(0)
        s1.upload(g1.offset, g1.nbr)
        s1.solve(k1, want_set=False)
        reps = 5
        t2 = time.perf_counter()
        for _ in range(reps):
            res, _ = s1.solve(k1, want_set=False)
        t_solve = (time.perf_counter() - t2) / reps
        extras["C1_exact_solves_per_s"] = 1.0 / t_solve
        extras["C1_solve_ms"] = t_solve * 1e3
        extras["C1_rho"] = f"{res.rho_num}/{res.rho_den}"
        s1.close()
        if not args.no_solves:
            extras.update(bench_c2_solve(g, k, kcds))
        extras.update(bench_c2_edges(np, g, kcds))
        extras.update(bench_c3(gen, kcds, solve=not args.no_solves))
        extras.update(bench_c4(np, gen, kcds, args.c4_plain_sample))
        if args.c5_window > 0:
            extras.update(bench_c5(args.c5_window, args.c5_graph_budget, args.c5_max_cliques, args.c5_slots,
                                   args.c5_large_slots, args.c5_deadline))
    extras.update(cpu_extra)

    line = {
        "metric": "k-cliques listed/s", "value": value, "unit": "cliques/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "C2 R-MAT scale 18 ef16 +K60+5xK30, k=4: GPU (degree, id) order + DAG + "
                               "per-vertex 4-clique counts", "n": g.n, "m": g.m, "k": k, "T": T,
                   "order": ["degree", "slack degeneracy", "exact degeneracy", "caller"][info.order_kind],
                   "dag_arcs": arcs, "w2": w2, "max_outdeg": info.max_outdeg,
                   "class_roots": list(info.class_roots), "waves_per_root": list(info.waves_per_root),
                   "autotune_class_ms": tuned[1] if tuned else None},
        "roofline": roofline, "cpu_baseline": cpu, "extras": extras,
    }
    print(json.dumps(line))
    s.close()


if __name__ == "__main__":
    main()
