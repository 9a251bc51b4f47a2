"""The merge model of the GPU path (unitig walks + Cartesian-tree multiplicities
keyed by the reference's visit order) against the oracle's sequential merge."""
import gzip
import json
import os

import numpy as np
import pytest

from oracle_binding import Oracle
import merge_model

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def parse_contig_graph(text):
    out = {}
    for line in text.decode().splitlines():
        if not line.strip():
            continue
        parts = line.split("#")
        head = parts[0].split(" ", 1)
        node = int(head[0])
        arcs = [head[1].strip()] + [p.strip() for p in parts[1:8]]
        mults = [int(v) for v in parts[8].split()]
        out[node] = (arcs, mults)
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_cartesian_model_matches_sequential_merge(name):
    case = CASES[name]
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        fa = f.read()
    k, c = case["k"], case["c"]
    o = Oracle(fa, k, c)
    keys, mult = o.survivors()
    visit = o.visit_order()
    pos = {int(x): i for i, x in enumerate(visit)}
    prio = np.array([pos[int(x)] for x in keys], dtype=np.int64)
    model = merge_model.compacted(keys, mult, prio, k)
    graph = parse_contig_graph(o.file(1))
    checked = 0
    for (X, s), (string, m) in model.items():
        arcs, mults = graph[X]
        assert arcs[s] == string, (X, s)
        assert mults[s] == m, (X, s, mults[s], m)
        checked += 1
    assert checked > 0
