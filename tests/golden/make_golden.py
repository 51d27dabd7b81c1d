"""Writes tests/golden/ref_graph_outputs.json from the reference's own
src/graph.cpp (compiled into oracle/_ref/ref_graph by oracle/Makefile).

The fixtures are data only: the adjacency-list inputs below and what the
reference loader and Batagelj-Zaversnik peel produce for them (CSR, core
numbers, corePeelSequence, listing order).  Run in a container where the
reference sources are mounted:  python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from test_amd import gen  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_graph")


def run(path):
    out = subprocess.run([REF_BIN, path], stdout=subprocess.PIPE, check=True, text=True).stdout
    lines = out.strip("\n").split("\n")[-7:]
    n, m, kmax = map(int, lines[0].split())
    arr = [list(map(int, ln.split())) for ln in lines[1:]]
    return dict(n=n, m=m, kmax=kmax, offset=arr[0], neighbors=arr[1], degree_after=arr[2], core=arr[3],
                corePeelSequence=arr[4], listingOrder=arr[5])


def main():
    cases = []
    files = ["ref_data_test.txt"]
    for name, g in [("golden_er60.txt", gen.er(60, 0.12, 8, seed=31)),
                    ("golden_rmat7.txt", gen.rmat(7, 5, (6,), seed=32))]:
        g.write_adjlist(os.path.join(HERE, name))
        files.append(name)
    for f in files:
        d = run(os.path.join(HERE, f))
        d["file"] = f
        cases.append(d)
    with open(os.path.join(HERE, "ref_graph_outputs.json"), "w") as fh:
        json.dump({"source": "wajidmanzoor/test src/graph.cpp via oracle/_ref/ref_graph", "cases": cases}, fh)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
