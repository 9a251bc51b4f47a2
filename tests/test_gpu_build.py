"""GPU k-mer graph build (libswaphip.so through its C ABI) against the oracle:
bit-exact survivors, multiplicities and kmerGraph.txt bytes (order included)."""
import gzip
import json
import os
import sys

import numpy as np
import pytest

import swapamd
from oracle_binding import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))

pytestmark = pytest.mark.gpu


def gold_fa(name):
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        return f.read()


def check_build(fa, k, c, hash_order=False, check_file=True, table_spill=False, wide=False, single_fill=False):
    print("case k=%d c=%d bytes=%d hash_order=%s table_spill=%s wide=%s single_fill=%s" % (
        k, c, len(fa), hash_order, table_spill, wide, single_fill), flush=True)
    o = Oracle(fa, k, c, kmer_graph_only=True)
    a = swapamd.Assembler(k, c, kmer_graph_only=True, hash_order=hash_order, table_spill=table_spill,
                          wide_partitions=wide, single_fill=single_fill)
    a.build(fa)
    st = a.stats()
    ost = o.stats()
    assert st["kmers"] == ost["kmers"]
    assert st["reads"] == ost["reads"]
    assert st["distinct"] == ost["distinct"]
    assert st["survivors"] == ost["survivors"]
    keys, mult = a.kmer_graph()
    okeys, omult = o.survivors()
    assert np.array_equal(keys, okeys)
    assert np.array_equal(mult, omult)
    if check_file:
        got = a.file_bytes(swapamd.SWAP_FILE_KMER_GRAPH)
        want = o.file(0)
        if hash_order:
            assert sorted(got.splitlines()) == sorted(want.splitlines())
        else:
            assert got == want
    a.close()
    return st


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_inputs(name):
    case = CASES[name]
    check_build(gold_fa(name), case["k"], case["c"])


@pytest.mark.parametrize("k", [9, 15, 19, 23, 27, 31])
def test_k_values(k):
    fa = swapamd.generate_params(30000, 90, 20.0, 0.01, 100 + k)
    check_build(fa, k, 2)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_memory_table(name):
    # every partition counted by the device-memory table that partitions with
    # more distinct keys than the on-chip table take
    case = CASES[name]
    check_build(gold_fa(name), case["k"], case["c"], table_spill=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_wide_partitions(name):
    # 2^18 partitions: the sizes come from the level-1 output and level 2
    # scatters into 1024 sub-buckets, the path of inputs of billions of k-mers
    case = CASES[name]
    check_build(gold_fa(name), case["k"], case["c"], wide=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_single_fill(name):
    # one partition per on-chip table fill, the path of large inputs (small
    # inputs share fills among 2^3 partitions)
    case = CASES[name]
    check_build(gold_fa(name), case["k"], case["c"], single_fill=True)
    check_build(gold_fa(name), case["k"], case["c"], hash_order=True, single_fill=True)


def test_shared_fills_mid_size():
    # partitions of about 800 and 1600 records (26M and 52M k-mers over 2^15
    # partitions): fills shared by 2^2 and 2^1 of them (the golden inputs'
    # small partitions share fills 2^3 at a time)
    for genome, seed in ((600000, 22), (1200000, 23)):
        fa = swapamd.generate_params(genome, 90, 60.0, 0.01, seed)
        check_build(fa, 25, 3, check_file=False)


def test_wide_partitions_small_k_and_spill():
    fa = swapamd.generate_params(30000, 90, 20.0, 0.01, 109)
    check_build(fa, 9, 2, wide=True)
    check_build(gold_fa("uniform_k21_c2"), 21, 2, wide=True, table_spill=True)


def poly_a_mix(seed):
    """1000 poly-A reads (70,000 occurrences of one k-mer at k=31: a partition
    past 2^16 records, counted with saturating 8-bit counters) among reads of a
    random genome."""
    rng = np.random.default_rng(seed)
    g = "".join("ACGT"[x] for x in rng.integers(0, 4, 20000))
    out = []
    for i in range(1000):
        out.append(">a%d\n%s\n" % (i, "A" * 100))
    for i in range(4000):
        p = int(rng.integers(0, len(g) - 100))
        out.append(">r%d\n%s\n" % (i, g[p:p + 100]))
    return "".join(out).encode()


def test_partition_past_16bit_counters():
    check_build(poly_a_mix(7), 31, 2)
    check_build(poly_a_mix(7), 31, 2, table_spill=True)


def test_hash_order_same_graph():
    fa = gold_fa("uniform_k31_c3")
    check_build(fa, 31, 3, hash_order=True)


def test_small_workload_full_size():
    # the 1 Mbp / 30x configuration (21M input k-mers)
    fa, k, c = swapamd.generate("small")
    st = check_build(fa, k, c, check_file=False)
    assert st["kmers"] == 21_000_000


def test_bad_input_errors():
    a = swapamd.Assembler(21, 2, kmer_graph_only=True)
    with pytest.raises(swapamd.SwapError) as e:
        a.build(b"no records here\n")
    assert e.value.code == -3
    with pytest.raises(swapamd.SwapError):
        a.build(b"")
    # no complete record: empty graph, not an error
    a.build(b">only header\n")
    assert a.stats()["survivors"] == 0


def test_repeat_builds_reuse_context():
    fa1 = gold_fa("uniform_k21_c2")
    fa2 = gold_fa("dirty_k21_c2")
    a = swapamd.Assembler(21, 2, kmer_graph_only=True)
    for fa in (fa1, fa2, fa1):
        a.build(fa)
        keys, mult = a.kmer_graph()
        okeys, omult = Oracle(fa, 21, 2).survivors()
        assert np.array_equal(keys, okeys) and np.array_equal(mult, omult)
