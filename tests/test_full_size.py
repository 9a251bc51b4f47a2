"""The bacterial (4.6 Mbp, 50x, 161M input k-mers) and repeat-rich (100 Mbp
with 200 repeat families, 40x, 3.2G input k-mers) configurations at their
full size, checked through properties that do not need the oracle to finish:
the closed-form reciprocity of every counted edge, the same output files from
another partition layout and from a second build in the same context, and the
same graph and contig set under the GPU-native visit order."""
import numpy as np
import pytest

import kmer_textbook
import swapamd
from test_oracle import NP1
from test_gpu_merge import check_single_rank_reference

FILES = (swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE)


def test_reciprocity_check_on_textbook_graphs():
    # the check itself: 0 on textbook graphs, and it sees a planted mistake
    fa = swapamd.generate_params(3000, 60, 12.0, 0.02, 5)
    for k, c in ((9, 0), (15, 2), (21, 1)):
        g = kmer_textbook.graph(fa, k, c)
        keys = np.array(sorted(g), dtype=np.uint64)
        mult = np.array([g[int(x)] for x in keys], dtype=np.uint8)
        assert kmer_textbook.reciprocal_violations(keys, mult, k) == 0
        i, t = np.argwhere(mult > 0)[len(keys) // 2]
        mult[i, t] += 1
        assert kmer_textbook.reciprocal_violations(keys, mult, k) >= 1
        mult[i, t] -= 1
        drop = np.ones(len(keys), dtype=bool)
        drop[len(keys) // 3] = False  # a survivor missing from the set
        assert kmer_textbook.reciprocal_violations(keys[drop], mult[drop], k) >= 1
        assert kmer_textbook.reciprocal_violations(keys, mult, k, rows=np.arange(0, len(keys), 2)) == 0
    assert np.array_equal(kmer_textbook.rc_codes(np.array([0b000110], dtype=np.uint64), 3),
                          np.array([0b011011], dtype=np.uint64))  # rc(ACG) = CGT


def assemble(fa, k, c, **kw):
    a = swapamd.Assembler(k, c, **kw)
    a.build(fa)
    a.merge()
    return a


def check_full_size(fa, k, c, n_kmers, single_fill_or_wide, sample=None):
    print("full size k=%d c=%d bytes=%d" % (k, c, len(fa)), flush=True)
    a = assemble(fa, k, c)
    st = a.stats()
    assert st["kmers"] == n_kmers
    keys, mult = a.kmer_graph()
    assert len(keys) == st["survivors"] > 0
    assert np.all(keys[1:] > keys[:-1])
    assert np.all((mult == 0) | (mult >= c)) and np.all(mult.any(axis=1))
    rows = None
    if sample and sample < len(keys):  # the edges of a seeded sample of the survivors
        rows = np.sort(np.random.default_rng(9).choice(len(keys), sample, replace=False))
    assert kmer_textbook.reciprocal_violations(keys, mult, k, rows) == 0
    files = [a.file_bytes(w) for w in FILES]
    print("reference order done: S=%d kept=%d" % (st["survivors"], st["kept"]), flush=True)
    a.build(fa)
    a.merge()
    assert [a.file_bytes(w) for w in FILES] == files
    a.close()
    b = assemble(fa, k, c, **single_fill_or_wide)
    assert b.stats()["survivors"] == st["survivors"] and b.stats()["distinct"] == st["distinct"]
    assert [b.file_bytes(w) for w in FILES] == files
    b.close()
    print("other layout done", flush=True)
    h = assemble(fa, k, c, hash_order=True)
    hk, hm = h.kmer_graph()
    assert np.array_equal(hk, keys) and np.array_equal(hm, mult)
    assert sorted(h.file_bytes(swapamd.SWAP_FILE_NOCE).splitlines()) == sorted(files[1].splitlines())
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("workload", ["small", "bacterial"])
def test_workload_equals_single_rank_reference(workload):
    """The 1 Mbp and 4.6 Mbp workloads at full size: every output file equals,
    byte for byte, what the unmodified reference wrote running them as one rank
    (digests from tests/golden/make_golden_np1.py)."""
    name = "workload_" + workload
    if name not in NP1:
        pytest.skip("no single-rank reference digests for the %s workload" % workload)
    fa, k, c = swapamd.generate(workload)
    assert (k, c) == (NP1[name]["k"], NP1[name]["c"])
    check_single_rank_reference(name, fa)


@pytest.mark.gpu
def test_bacterial_workload_full_size():
    fa, k, c = swapamd.generate("bacterial")
    check_full_size(fa, k, c, 161_000_000, {"wide_partitions": True})


@pytest.mark.gpu
def test_repeat_workload_full_size():
    fa, k, c = swapamd.generate("repeat")
    check_full_size(fa, k, c, 3_200_000_040, {"single_fill": True}, sample=4_000_000)
