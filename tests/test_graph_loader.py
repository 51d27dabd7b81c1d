"""The product's host Graph surface (test_amd/csrc/graph.cpp, the drop-in for
the reference's src/graph.cpp:7-113), compiled with g++ into a small driver
(tests/graph_dump.cpp): its CSR, core numbers, peel sequence and listing
order on the reference's golden inputs equal the vectors the compiled
reference wrote (tests/golden/ref_graph_outputs.json); the same graphs
written as 'u v' edge lists load to the same CSR and order; and a file whose
first lines fit the adjacency pattern but whose lists are not symmetric is
read as an edge list.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("graph_dump") / "graph_dump")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(ROOT, "tests", "graph_dump.cpp"), os.path.join(ROOT, "test_amd", "csrc", "graph.cpp")],
                   check=True)

    def run(path):
        r = subprocess.run([exe, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        assert r.returncode == 0, r.stderr
        return json.loads(r.stdout)
    return run


def golden_cases():
    with open(os.path.join(GOLD, "ref_graph_outputs.json")) as f:
        return json.load(f)["cases"]


CASES = golden_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["file"] for c in CASES])
def test_graph_matches_compiled_reference(dump, case):
    g = dump(os.path.join(GOLD, case["file"]))
    assert (g["n"], g["m"], g["kmax"]) == (case["n"], case["m"], case["kmax"])
    assert g["offset"] == case["offset"]
    assert g["neighbors"] == case["neighbors"]
    assert g["core"] == case["core"]
    assert g["corePeelSequence"] == case["corePeelSequence"]
    assert g["listingOrder"] == case["listingOrder"]
    # the reference overwrites degree[] with core numbers during the peel; ours keeps it
    assert g["degree"] == np.diff(case["offset"]).tolist()


@pytest.mark.parametrize("case", CASES, ids=[c["file"] for c in CASES])
def test_edge_list_loads_the_same_graph(dump, case, tmp_path):
    off = np.asarray(case["offset"], np.int64)
    nbr = np.asarray(case["neighbors"], np.int64)
    src = np.repeat(np.arange(case["n"]), np.diff(off))
    keep = src < nbr
    pairs = np.stack([src[keep], nbr[keep]], 1)
    # shuffled, both directions for some edges, a duplicate, a self-loop and comments
    rng = np.random.default_rng(7)
    pairs = pairs[rng.permutation(len(pairs))]
    flip = rng.random(len(pairs)) < 0.5
    pairs[flip] = pairs[flip][:, ::-1]
    lines = ["# u v edge list", "% comment"] + [f"{u} {v}" for u, v in pairs] + [f"{pairs[0][1]} {pairs[0][0]}", "0 0"]
    path = tmp_path / "edges.txt"
    path.write_text("\r\n".join(lines) + "\r\n")
    g = dump(path)
    hi = int(pairs.max()) + 1  # an edge list's n is its largest id + 1
    assert g["n"] == hi and g["m"] == case["m"]
    assert g["offset"] == case["offset"][: hi + 1]
    assert g["neighbors"] == case["neighbors"]
    if hi == case["n"]:
        assert g["corePeelSequence"] == case["corePeelSequence"]
        assert g["listingOrder"] == case["listingOrder"]


def test_ambiguous_header_reads_as_edge_list(dump, tmp_path):
    # "3 2" then lines starting 0, 1, 2: the adjacency pattern, but 0->1, 1->2, 2->0 is not symmetric
    path = tmp_path / "amb.txt"
    path.write_text("3 2\n0 1\n1 2\n2 0\n")
    g = dump(path)
    # as edges: {3,2}, {0,1}, {1,2}, {2,0} on 4 vertices
    assert g["n"] == 4 and g["m"] == 4
    assert g["offset"] == [0, 2, 4, 7, 8]
    assert g["neighbors"] == [1, 2, 0, 2, 0, 1, 3, 2]


def test_symmetric_adjacency_with_header_reads_as_adjacency(dump, tmp_path):
    # the same pattern with symmetric lists and 2m entries is the reference format
    path = tmp_path / "adj.txt"
    path.write_text("3 3\n0 1 2\n1 0 2\n2 0 1\n")
    g = dump(path)
    assert (g["n"], g["m"]) == (3, 3)
    assert g["offset"] == [0, 2, 4, 6] and g["neighbors"] == [1, 2, 0, 2, 0, 1]


def test_adjacency_with_too_many_entries_reads_as_edge_list(dump, tmp_path):
    # symmetric lists but more than 2m entries: not the reference's format (it sizes neighbors 2m)
    path = tmp_path / "big.txt"
    path.write_text("2 0\n0 1\n1 0\n")
    g = dump(path)
    assert (g["n"], g["m"]) == (3, 2)  # edges {2,0} and {0,1}
