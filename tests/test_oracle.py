"""The CPU oracle against the reference's own outputs (golden fixtures made by
tests/golden/make_golden.py from the unmodified reference) and against the
textbook definition of the k-mer graph."""
import gzip
import json
import os
import random
import re

import numpy as np
import pytest

from oracle_binding import Oracle
import kmer_textbook

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def gold(name, what):
    with gzip.open(os.path.join(GOLD, "%s.%s.gz" % (name, what)), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(CASES))
def test_kmer_graph_matches_reference(name):
    case = CASES[name]
    fa = gold(name, "fa")
    o = Oracle(fa, case["k"], case["c"], kmer_graph_only=True)
    got = sorted(o.file(0).decode().splitlines())
    want = gold(name, "kmerGraph").decode().splitlines()
    assert len(got) == case["kmer_graph_lines"]
    assert got == want


@pytest.mark.parametrize("name", sorted(CASES))
def test_noce_contigs_match_reference(name):
    case = CASES[name]
    fa = gold(name, "fa")
    o = Oracle(fa, case["k"], case["c"])
    got = sorted(o.file(2).decode().splitlines())
    want = gold(name, "noce").decode().splitlines()
    assert got == want


def _random_fasta(seed, G, L, n, err, extra=False):
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(G))
    out = []
    for i in range(n):
        st = rng.randrange(G - L + 1)
        s = g[st:st + L]
        if rng.random() < 0.5:
            s = kmer_textbook.rc(s)
        s = "".join(ch if rng.random() >= err else rng.choice("ACGT") for ch in s)
        if extra and rng.random() < 0.1:
            s = s.lower()
        if extra and rng.random() < 0.05:
            s = s[: rng.randrange(len(s))]
        out.append(">r%d\n%s\n" % (i, s))
    return "".join(out).encode()


@pytest.mark.parametrize("seed,k,c", [(1, 21, 0), (2, 19, 2), (3, 31, 1), (4, 25, 3)])
def test_survivors_match_textbook(seed, k, c):
    fa = _random_fasta(seed, 600, 60, 400, 0.01, extra=True)
    o = Oracle(fa, k, c)
    keys, mult = o.survivors()
    want = kmer_textbook.graph(fa, k, c)
    assert len(keys) == len(want)
    for X, m in zip(keys.tolist(), mult.tolist()):
        assert want[X] == m


def test_saturation_at_255():
    # one 40-base read repeated 300 times: every interior edge is seen 300 times
    rng = random.Random(7)
    s = "".join(rng.choice("ACGT") for _ in range(40))
    fa = "".join(">r%d\n%s\n" % (i, s) for i in range(300)).encode()
    o = Oracle(fa, 21, 1)
    keys, mult = o.survivors()
    assert mult.max() == 255
    want = kmer_textbook.graph(fa, 21, 1)
    for X, m in zip(keys.tolist(), mult.tolist()):
        assert want[X] == m


def test_palindromic_edge_counts_twice():
    # k=19: a 20-mer that is its own reverse complement
    half = "ACGTTGCAAGCTTAGC"[:10]
    w = half + kmer_textbook.rc(half)
    assert w == kmer_textbook.rc(w)
    fa = (">r\n%s\n" % w).encode()
    o = Oracle(fa, 19, 0)
    keys, mult = o.survivors()
    # one node (prefix and suffix of w are reverse complements), one slot counted twice
    assert len(keys) == 1
    assert sorted(mult[0].tolist()) == [0] * 7 + [2]


def test_edge_symmetry_invariant():
    # every edge X -b-> Y is seen from Y with the same multiplicity
    fa = _random_fasta(11, 2000, 80, 900, 0.01)
    k = 21
    o = Oracle(fa, k, 2)
    keys, mult = o.survivors()
    table = {int(x): m for x, m in zip(keys.tolist(), mult.tolist())}
    mask = (1 << (2 * k)) - 1

    def rcv(x):
        r = 0
        for _ in range(k):
            r = (r << 2) | (3 - (x & 3))
            x >>= 2
        return r

    for X, m in table.items():
        for s in range(8):
            if not m[s]:
                continue
            o_ = X if s < 4 else rcv(X)
            t = ((o_ << 2) | (s & 3)) & mask
            Y = max(t, rcv(t))
            assert Y in table
            # back edge: from rc(t) followed by complement of o_'s first base
            first = (o_ >> (2 * (k - 1))) & 3
            back = rcv(t)
            bslot = (3 - first) if back == Y else 4 + (3 - first)
            assert table[Y][bslot] == m[s]


def test_unterminated_last_record_is_ignored():
    fa = b">a\nACGTACGTACGTACGTACGTACGTAC\n>b\nTTTTGGGGCCCCAAAATTTTGGGG"
    o = Oracle(fa, 21, 0)
    assert o.stats()["reads"] == 1


def test_records_before_first_gt_are_skipped():
    fa = b"junk\nmore junk\n>a\nACGTACGTACGTACGTACGTACGTAC\n"
    o = Oracle(fa, 21, 0)
    assert o.stats()["reads"] == 1 and o.stats()["kmers"] == 6


def decode_kmer(x, k):
    """2-bit MSB-first value -> bases (A=0 C=1 G=2 T=3)."""
    return "".join("ACGT"[(x >> (2 * (k - 1 - i))) & 3] for i in range(k))


@pytest.mark.parametrize("name", sorted(CASES))
def test_perf_contigs_are_the_noce_selection(name):
    """The -p writer (stringContigs) selects exactly printContigs' contigs: on a
    single rank nothing is left for the finisher on these inputs, so the
    sequences equal the noCE set, and each header's fields describe its line."""
    case = CASES[name]
    o = Oracle(gold(name, "fa"), case["k"], case["c"])
    perf = o.file(3).decode().splitlines()
    noce = sorted(l[len(">proc0: "):] for l in o.file(2).decode().splitlines())
    assert len(perf) % 2 == 0
    seqs = []
    for i in range(0, len(perf), 2):
        m = re.match(r">contig_NodeID(\d+)_Index(\d)_Len(\d+)_Mul(\d+)_(\d+)$", perf[i])
        assert m, perf[i]
        node, slot, ln, mul, cnt = (int(g) for g in m.groups())
        seq = perf[i + 1]
        assert ln == len(seq) and cnt == i // 2 + 1 and 0 < mul <= 255
        first = seq[:case["k"]]
        assert (first if slot < 4 else kmer_textbook.rc(first)) == decode_kmer(node, case["k"])
        seqs.append(seq)
    assert sorted(seqs) == noce


# ---------------------------------------------------------------------------
# The unmodified reference run as one rank (tests/golden/make_golden_np1.py):
# exact bytes of every file, line order and merged multiplicities included.

NP1 = json.load(open(os.path.join(GOLD, "np1_cases.json")))
NP1_SMALL = sorted(n for n, c in NP1.items() if not c["input"].startswith("workload:"))
NP1_SPECIAL = sorted(n for n in NP1_SMALL if NP1[n]["input"].startswith("np1_"))


def np1_input(name):
    with gzip.open(os.path.join(GOLD, NP1[name]["input"]), "rb") as f:
        return f.read()


def np1_file(name, suffix):
    with gzip.open(os.path.join(GOLD, "np1_%s.%s.gz" % (name, suffix)), "rb") as f:
        return f.read()


def sha(b):
    import hashlib
    return hashlib.sha256(b).hexdigest()


def same_as_np1(name, suffix, data):
    want = NP1[name]["files"][suffix]
    if want is None:
        return False
    if sha(data) == want["sha256"]:
        return True
    if want.get("stored"):
        g, w = data.decode().splitlines(), np1_file(name, suffix).decode().splitlines()
        first = next((i for i, (a, b) in enumerate(zip(g, w)) if a != b), min(len(g), len(w)))
        print("%s %s: first difference at line %d of %d/%d" % (name, suffix, first, len(g), len(w)))
    return False


@pytest.mark.parametrize("name", NP1_SMALL)
def test_oracle_equals_single_rank_reference(name):
    case = NP1[name]
    fa = np1_input(name)
    o = Oracle(fa, case["k"], case["c"])
    assert same_as_np1(name, "contigGraph", o.file(1))
    assert same_as_np1(name, "noce", o.file(2))
    assert same_as_np1(name, "perf", o.file(3))
    s = Oracle(fa, case["k"], case["c"], kmer_graph_only=True)
    assert same_as_np1(name, "kmerGraph", s.file(0))


def test_oracle_equals_single_rank_reference_1mbp_workload():
    import swapamd
    fa, k, c = swapamd.generate("small")
    name = "workload_small"
    assert (k, c) == (NP1[name]["k"], NP1[name]["c"])
    o = Oracle(fa, k, c)
    assert same_as_np1(name, "contigGraph", o.file(1))
    assert same_as_np1(name, "noce", o.file(2))
    assert same_as_np1(name, "perf", o.file(3))


def _canon_neighbour(X, slot, edge, k):
    s = decode_kmer(X, k)
    if slot >= 4:
        s = kmer_textbook.rc(s)
    t = (s + edge)[-k:]
    t = max(t, kmer_textbook.rc(t))
    return int("".join("%d" % "ACGT".index(ch) for ch in t), 4)


def _one_one_after_merge(name):
    """(1,1) nodes of the reference's contigGraph.txt (the table right after its
    merge) and, per node, whether a neighbour is the node itself."""
    k = NP1[name]["k"]
    out = []
    for line in np1_file(name, "contigGraph").decode().splitlines():
        if not line.strip():
            continue
        parts = line.split("#")
        head = parts[0].split(" ")
        X = int(head[0])
        edges = [head[1]] + [p.strip() for p in parts[1:8]]
        right = [i for i in range(4) if edges[i]]
        left = [i for i in range(4, 8) if edges[i]]
        if len(right) == 1 and len(left) == 1:
            nb = {_canon_neighbour(X, i, edges[i], k) for i in right + left}
            out.append(X in nb)
    return out


def test_single_rank_fixtures_reach_the_special_merges():
    # the reference left (1,1) nodes behind its merge: ones that are their own
    # neighbour (cycles, fold-backs) and ones that are not (skipped while still
    # locked by the previous visit)
    self_nb = sum(sum(_one_one_after_merge(n)) for n in NP1_SPECIAL)
    other = sum(len(v) - sum(v) for v in map(_one_one_after_merge, NP1_SPECIAL))
    assert self_nb >= 10 and other >= 2


@pytest.mark.parametrize("plant", [Oracle.PLANT_EAGER_UNLOCK, Oracle.PLANT_DOUBLE_LOCK])
def test_single_rank_fixtures_catch_planted_merge_mistakes(plant):
    broken = []
    for name in NP1_SPECIAL:
        case = NP1[name]
        o = Oracle(np1_input(name), case["k"], case["c"], plant=plant)
        if not same_as_np1(name, "contigGraph", o.file(1)):
            broken.append(name)
    assert broken, "no fixture detects the planted mistake"
