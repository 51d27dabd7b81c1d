"""Writes tests/golden/c3_counts.json: the oracle's 5-clique counts of the
C3 graph (Chung-Lu n = 1,000,000, Pareto exponent 2.1, degrees 2..5000,
~10M edges, planted 80-vertex near-clique of density 0.95; k = 5).

The full per-vertex vector is 8 MB, so the fixture keeps T, the sum and two
position-weighted checksums of the per-vertex counts (mod 2^64), their
maximum and 64 sampled entries.  The graph is regenerated from its seed by
test_amd.gen.config(3).  The oracle takes about a minute on 8 host threads:
  python tests/golden/make_c3_counts.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from test_amd import gen  # noqa: E402
import oracle as O  # noqa: E402


def checksums(pv):
    pv = np.asarray(pv, dtype=np.uint64)
    idx = np.arange(1, len(pv) + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        c1 = int(np.sum(pv * idx, dtype=np.uint64))
        c2 = int(np.sum(pv * (idx * idx + np.uint64(7)), dtype=np.uint64))
    return c1, c2


def main():
    g, k = gen.config(3)
    T, pv = O.count(g.offset, g.nbr, k, per_vertex=True, threads=os.cpu_count() or 1)
    pv = np.asarray(pv, dtype=np.int64)
    rng = np.random.default_rng(2024)
    sample = sorted(set(rng.integers(0, g.n, 48).tolist()) | set(np.argsort(-pv)[:16].tolist()))
    c1, c2 = checksums(pv)
    out = dict(config="C3", n=int(g.n), m=int(g.m), k=int(k), T=int(T), sum=int(pv.sum()), check1=c1, check2=c2,
               max=int(pv.max()), sample={str(v): int(pv[v]) for v in sample})
    with open(os.path.join(HERE, "c3_counts.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("T", T, "sum", int(pv.sum()))


if __name__ == "__main__":
    main()
