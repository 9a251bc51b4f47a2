"""Contig extension (host code of the product library) against the reference's
own rank-0 code: for each fixture graph (the contigGraph.txt of a single-rank
merge), the noCE contigs, CEContig.fasta and the -j files written by the
reference (oracle/_ref/ref_ce, tests/golden/make_golden_ce.py) byte for byte.
No device is needed: swap_contig_extension reads the graph from its text."""
import gzip
import json
import os

import pytest

import swapamd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(REPO, "tests", "golden")
CASES = json.load(open(os.path.join(GOLD, "ce_cases.json")))

FILES = [("noce", swapamd.SWAP_FILE_NOCE), ("cecontig", swapamd.SWAP_FILE_CE_CONTIGS),
         ("jungarc", swapamd.SWAP_FILE_JUNG_ARC), ("jungmul", swapamd.SWAP_FILE_JUNG_MUL)]


def gold(name, what):
    with gzip.open(os.path.join(GOLD, "ce_%s.%s.gz" % (name, what)), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("what,which", FILES, ids=[f[0] for f in FILES])
def test_extension_matches_reference(name, what, which):
    case = CASES[name]
    got = swapamd.contig_extension(case["k"], case["survivors"], gold(name, "contigGraph"), which)
    want = gold(name, what)
    if got != want:
        g, w = got.decode().splitlines(), want.decode().splitlines()
        first = next((i for i, (a, b) in enumerate(zip(g, w)) if a != b), min(len(g), len(w)))
        print("first difference at line %d of %d/%d" % (first, len(g), len(w)))
    assert got == want


def test_fixtures_exercise_the_passes():
    # the fixture set holds removed bubble paths and many-node graphs
    assert any(c["removed"] > 0 for c in CASES.values())
    assert max(c["kept"] for c in CASES.values()) > 1000


def test_graph_without_edges_is_reported():
    with pytest.raises(swapamd.SwapError) as e:
        swapamd.contig_extension(21, 10, b"", swapamd.SWAP_FILE_CE_CONTIGS)
    assert e.value.code == -2


# The reference run end to end as one rank (tests/golden/make_golden_np1.py):
# its contigGraph.txt (equal to the oracle's, tests/test_oracle.py) into the
# library's contig extension, against the files that run wrote.
from test_oracle import NP1, NP1_SMALL, np1_input, same_as_np1  # noqa: E402
from oracle_binding import Oracle  # noqa: E402

NP1_FILES = [("noce", swapamd.SWAP_FILE_NOCE), ("cecontig", swapamd.SWAP_FILE_CE_CONTIGS),
             ("jungarc", swapamd.SWAP_FILE_JUNG_ARC), ("jungmul", swapamd.SWAP_FILE_JUNG_MUL)]


@pytest.mark.parametrize("name", NP1_SMALL)
def test_extension_matches_single_rank_reference(name):
    case = NP1[name]
    o = Oracle(np1_input(name), case["k"], case["c"])
    graph = o.file(1)
    assert same_as_np1(name, "contigGraph", graph)
    for what, which in NP1_FILES:
        got = swapamd.contig_extension(case["k"], o.stats()["survivors"], graph, which)
        assert same_as_np1(name, what, got), what
