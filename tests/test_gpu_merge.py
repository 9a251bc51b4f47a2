"""GPU k-molecule merge (through the C ABI) against the oracle: contigGraph.txt
and noCEcontigs.fasta byte for byte in the reference order, and the same
compacted topology / contig set in the GPU-native hash order."""
import gzip
import json
import os
import random

import numpy as np
import pytest

import swapamd
from oracle_binding import Oracle
import kmer_textbook

from test_oracle import NP1, NP1_SMALL, np1_input, same_as_np1

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))

pytestmark = pytest.mark.gpu


def gold_fa(name):
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        return f.read()


def strip_mult(text):
    """contigGraph lines without the 8 multiplicities and padding."""
    out = []
    for line in text.decode().splitlines():
        if not line.strip():
            continue
        parts = line.split("#")
        out.append("#".join(parts[:8]))
    return sorted(out)


def check_merge(fa, k, c, hash_order=False, table_spill=False, wide=False, single_fill=False):
    print("merge case k=%d c=%d bytes=%d hash_order=%s table_spill=%s wide=%s single_fill=%s" % (
        k, c, len(fa), hash_order, table_spill, wide, single_fill), flush=True)
    o = Oracle(fa, k, c)
    a = swapamd.Assembler(k, c, hash_order=hash_order, table_spill=table_spill, wide_partitions=wide,
                          single_fill=single_fill)
    a.build(fa)
    a.merge()
    cg = a.file_bytes(swapamd.SWAP_FILE_CONTIG_GRAPH)
    noce = a.file_bytes(swapamd.SWAP_FILE_NOCE)
    ocg, onoce = o.file(1), o.file(2)
    st = a.stats()
    assert st["kept"] == o.stats()["kept"]
    if hash_order:
        assert strip_mult(cg) == strip_mult(ocg)
        assert sorted(noce.splitlines()) == sorted(onoce.splitlines())
    else:
        if cg != ocg:
            got, want = cg.decode().splitlines(), ocg.decode().splitlines()
            diff = [(i, g, w) for i, (g, w) in enumerate(zip(got, want)) if g != w]
            print("contigGraph differs: %d vs %d lines, %d differing, first: %s" % (
                len(got), len(want), len(diff), diff[:3]), flush=True)
            print("as sets equal:", sorted(got) == sorted(want), " without mult:", strip_mult(cg) == strip_mult(ocg),
                  flush=True)
        assert cg == ocg
        assert noce == onoce
        # the -p contigs file: same merged table, written in the merge's own order
        assert a.file_bytes(swapamd.SWAP_FILE_PERF_CONTIGS) == o.file(3)
    a.close()
    return st


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_inputs(name):
    case = CASES[name]
    check_merge(gold_fa(name), case["k"], case["c"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_inputs_hash_order(name):
    case = CASES[name]
    check_merge(gold_fa(name), case["k"], case["c"], hash_order=True)


@pytest.mark.parametrize("name", ["repeats_k25_c3", "uniform_k31_c3"])
def test_device_memory_table(name):
    # survivors counted by the device-memory table come in several runs per
    # partition: the survivor index builds those regions with device atomics
    case = CASES[name]
    check_merge(gold_fa(name), case["k"], case["c"], table_spill=True)


@pytest.mark.parametrize("name", ["repeats_k25_c3", "uniform_k31_c3", "dirty_k21_c2"])
def test_wide_partitions(name):
    # 2^18 count partitions: the survivor index has as many regions
    case = CASES[name]
    check_merge(gold_fa(name), case["k"], case["c"], wide=True)


def test_single_fill():
    case = CASES["uniform_k21_c2"]
    check_merge(gold_fa("uniform_k21_c2"), case["k"], case["c"], single_fill=True)


def test_wide_partitions_hash_order():
    case = CASES["uniform_k21_c2"]
    check_merge(gold_fa("uniform_k21_c2"), case["k"], case["c"], hash_order=True, wide=True)


def test_partition_past_16bit_counters():
    from test_gpu_build import poly_a_mix
    check_merge(poly_a_mix(7), 31, 2)


def circular_reads(seed, G, L, n, err):
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(G))
    out = []
    for i in range(n):
        st = rng.randrange(G)
        s = (g + g[:L])[st:st + L]
        if rng.random() < 0.5:
            s = kmer_textbook.rc(s)
        s = "".join(ch if rng.random() >= err else rng.choice("ACGT") for ch in s)
        out.append(">r%d\n%s\n" % (i, s))
    return "".join(out).encode()


def test_isolated_cycle():
    # a circular genome without errors: one isolated cycle of (1,1) nodes; it is
    # merged on the device as the chain from its highest-priority node around to
    # itself (that node kept), not by the host replay
    fa = circular_reads(21, 3000, 100, 900, 0.0)
    st = check_merge(fa, 21, 1)
    assert st["special"] == 0 and st["kept"] == 1


def test_fold_back_palindrome():
    rng = random.Random(5)
    left = "".join(rng.choice("ACGT") for _ in range(400))
    arm = "".join(rng.choice("ACGT") for _ in range(120))
    g = left + arm + kmer_textbook.rc(arm)
    reads = []
    for i in range(1500):
        st = rng.randrange(len(g) - 100 + 1)
        s = g[st:st + 100]
        if rng.random() < 0.5:
            s = kmer_textbook.rc(s)
        reads.append(">r%d\n%s\n" % (i, s))
    fa = "".join(reads).encode()
    check_merge(fa, 21, 2)
    check_merge(fa, 21, 2, hash_order=True)


@pytest.mark.parametrize("k", [
    pytest.param(9, marks=pytest.mark.skip(reason="k=9 on a 40 kbp genome stalled the GPU run (killed after 180 s "
                                                  "without output); cause not yet identified, case kept off the GPU")),
    19, 25, 31])
def test_random_genomes(k):
    fa = swapamd.generate_params(40000, 100, 25.0, 0.01, 200 + k)
    check_merge(fa, k, 2)


def test_small_workload_full_size():
    fa, k, c = swapamd.generate("small")
    check_merge(fa, k, c)


CE_CASES = json.load(open(os.path.join(GOLD, "ce_cases.json")))


def ce_gold(name, what):
    with gzip.open(os.path.join(GOLD, "ce_%s.%s.gz" % (name, what)), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(CE_CASES))
def test_contig_extension_from_reads(name):
    """Reads -> GPU build and merge -> host contig extension: contigGraph.txt and
    the extension's files equal what the reference's own rank-0 code wrote for
    that graph."""
    case = CE_CASES[name]
    fa = ce_gold(name, "fa")
    print("contig extension case %s k=%d c=%d" % (name, case["k"], case["c"]), flush=True)
    a = swapamd.Assembler(case["k"], case["c"])
    a.build(fa)
    a.merge()
    assert a.stats()["survivors"] == case["survivors"]
    assert a.file_bytes(swapamd.SWAP_FILE_CONTIG_GRAPH) == ce_gold(name, "contigGraph")
    assert a.file_bytes(swapamd.SWAP_FILE_NOCE) == ce_gold(name, "noce")
    assert a.file_bytes(swapamd.SWAP_FILE_CE_CONTIGS) == ce_gold(name, "cecontig")
    assert a.file_bytes(swapamd.SWAP_FILE_JUNG_ARC) == ce_gold(name, "jungarc")
    assert a.file_bytes(swapamd.SWAP_FILE_JUNG_MUL) == ce_gold(name, "jungmul")
    a.close()


NP1_FILES = [("contigGraph", swapamd.SWAP_FILE_CONTIG_GRAPH), ("noce", swapamd.SWAP_FILE_NOCE),
             ("perf", swapamd.SWAP_FILE_PERF_CONTIGS), ("cecontig", swapamd.SWAP_FILE_CE_CONTIGS),
             ("jungarc", swapamd.SWAP_FILE_JUNG_ARC), ("jungmul", swapamd.SWAP_FILE_JUNG_MUL)]


def check_single_rank_reference(name, fa, **flags):
    """Every file of the GPU path equals what the unmodified reference wrote as
    one rank for this input (fixture digests; exact bytes, line order and merged
    multiplicities included)."""
    case = NP1[name]
    k, c = case["k"], case["c"]
    print("single-rank reference case %s k=%d c=%d bytes=%d %s" % (name, k, c, len(fa), flags), flush=True)
    a = swapamd.Assembler(k, c, **flags)
    a.build(fa)
    a.merge()
    bad = [what for what, which in NP1_FILES if not same_as_np1(name, what, a.file_bytes(which))]
    a.close()
    s = swapamd.Assembler(k, c, kmer_graph_only=True)
    s.build(fa)
    if not same_as_np1(name, "kmerGraph", s.file_bytes(swapamd.SWAP_FILE_KMER_GRAPH)):
        bad.append("kmerGraph")
    s.close()
    assert bad == []


@pytest.mark.parametrize("name", NP1_SMALL)
def test_single_rank_reference(name):
    check_single_rank_reference(name, np1_input(name))


@pytest.mark.parametrize("name", NP1_SMALL)
def test_coarse_renumbering(name):
    """The survivor numbering by coarse first-occurrence bins (the one inputs of
    more than 2^31 k-mers get) changes no file."""
    check_single_rank_reference(name, np1_input(name), coarse_renumber=True)


@pytest.mark.parametrize("tuning", [
    {"part_recs": 1024, "fill_recs": 512, "index_load": 50, "ruler_stride": 4, "occ_groups": 4,
     "order_bin_keys": 256},
    {"part_recs": 65536, "fill_recs": 8192, "index_load": 85, "ruler_stride": 256, "occ_groups": 1,
     "order_bin_keys": 1 << 30},
    {"part_recs": 4096, "fill_recs": 2048, "index_load": 75, "ruler_stride": 8, "occ_groups": 2,
     "order_bin_keys": 2048},
])
def test_tuning_never_changes_results(tuning):
    """Every performance parameter at the ends of its range gives the same files."""
    case = CASES["repeats_k25_c3"]
    fa = gold_fa("repeats_k25_c3")
    want = check_merge(fa, case["k"], case["c"])  # default tuning, checked against the oracle
    o = Oracle(fa, case["k"], case["c"])
    a = swapamd.Assembler(case["k"], case["c"])
    a.set_tuning(tuning)
    a.build(fa)
    a.merge()
    assert a.tuning() == tuning
    assert a.stats()["kept"] == want["kept"]
    assert a.file_bytes(swapamd.SWAP_FILE_CONTIG_GRAPH) == o.file(1)
    assert a.file_bytes(swapamd.SWAP_FILE_NOCE) == o.file(2)
    a.set_tuning(None)
    a.close()


@pytest.mark.parametrize("bin_keys", [256, 1 << 30])
def test_visit_order_bins(bin_keys):
    """The visit-order sort with bins of ~256 keys (thousands of bins per table)
    and with one bin per table: the latter holds every survivor of a 60 kbp
    genome, above the 6144 keys a bin is ordered with in LDS, so it is ordered
    in device memory.  Files equal the oracle's either way."""
    fa = swapamd.generate_params(60000, 100, 20.0, 0.005, 77)
    print("visit order bins: order_bin_keys=%d, %d bytes" % (bin_keys, len(fa)), flush=True)
    o = Oracle(fa, 31, 2)
    a = swapamd.Assembler(31, 2)
    tuning = dict(a.tuning(), order_bin_keys=bin_keys)
    a.set_tuning(tuning)
    a.build(fa)
    a.merge()
    st = a.stats()
    assert st["survivors"] > 6144
    assert st["kept"] == o.stats()["kept"]
    assert a.file_bytes(swapamd.SWAP_FILE_CONTIG_GRAPH) == o.file(1)
    assert a.file_bytes(swapamd.SWAP_FILE_NOCE) == o.file(2)
    assert a.file_bytes(swapamd.SWAP_FILE_PERF_CONTIGS) == o.file(3)
    a.close()
