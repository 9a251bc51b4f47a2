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


@pytest.mark.parametrize("k", [9, 13, 19, 31])
def test_chain_bookkeeping_is_consistent(k):
    """The device merge's bookkeeping, restated on the CPU, on a random genome:
    every chain edge agrees with its other end and every chain array position
    is written exactly once.  k = 9 is the case whose GPU run once stalled (in
    an earlier version of the kernels): short k gives many branching nodes and
    cycle / fold-back components."""
    import swapamd
    fa = swapamd.generate_params(15000, 100, 25.0, 0.01, 300 + k)
    o = Oracle(fa, k, 2)
    keys, mult = o.survivors()
    problems, stats = merge_model.chain_bookkeeping(keys, mult, k)
    assert problems == []
    assert stats["chains"] > 0
    if k == 9:
        assert stats["special_edges"] > 0


def test_chain_bookkeeping_sees_a_missing_node():
    import swapamd
    fa = swapamd.generate_params(5000, 100, 25.0, 0.0, 77)
    o = Oracle(fa, 21, 2)
    keys, mult = o.survivors()
    one = [i for i in range(len(keys)) if (mult[i][:4] > 0).sum() == 1 and (mult[i][4:] > 0).sum() == 1]
    drop = np.ones(len(keys), dtype=bool)
    drop[one[len(one) // 2] + 1 if one[len(one) // 2] + 1 < len(keys) else 0] = False
    problems, _ = merge_model.chain_bookkeeping(keys[drop], mult[drop], 21)
    assert problems
