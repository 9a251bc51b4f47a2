"""Read-ahead of the next input (swap_prefetch): builds that consume a prefetched
input give the same k-mer graph, merged graph and contigs as plain builds (and
the oracle), in order; a build of another input drops the read-ahead; the
queue holds at most two inputs."""
import gzip
import os

import numpy as np
import pytest

import swapamd
from oracle_binding import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def gold_fa(name):
    with gzip.open(os.path.join(GOLD, name + ".fa.gz"), "rb") as f:
        return f.read()


def outputs(a):
    return (a.file_bytes(swapamd.SWAP_FILE_CONTIG_GRAPH), a.file_bytes(swapamd.SWAP_FILE_NOCE))


def test_prefetched_stream_equals_oracle():
    # three different inputs, each prefetched while the previous one is built
    # and merged (the generator's inputs: 40-120 kbp genomes, 100 bp reads)
    fas = [swapamd.generate_params(g, 100, 30.0, 0.005, seed) for g, seed in
           ((40000, 11), (120000, 12), (70000, 13))]
    k, c = 25, 2
    a = swapamd.Assembler(k, c)
    a.prefetch(fas[0])
    for i, fa in enumerate(fas):
        print("stream input %d bytes=%d" % (i, len(fa)), flush=True)
        if i + 1 < len(fas):
            a.prefetch(fas[i + 1])
        a.build(fa)
        a.merge()
        o = Oracle(fa, k, c)
        keys, mult = a.kmer_graph()
        okeys, omult = o.survivors()
        assert np.array_equal(keys, okeys) and np.array_equal(mult, omult)
        assert outputs(a) == (o.file(1), o.file(2))
    a.close()


def test_same_bytes_prefetched_twice():
    # the bench's pattern: one input object built again and again
    fa = gold_fa("uniform_k21_c2")
    ref = swapamd.Assembler(21, 2)
    ref.build(fa)
    ref.merge()
    want = outputs(ref)
    a = swapamd.Assembler(21, 2)
    a.prefetch(fa)
    for step in range(4):
        print("step %d" % step, flush=True)
        a.prefetch(fa)  # two queued: this one and the one before
        a.build(fa)
        a.merge()
        assert outputs(a) == want
    a.build(fa)  # consumes the last queued copy
    a.merge()
    assert outputs(a) == want
    a.close()


def test_other_input_drops_read_ahead_and_queue_limit():
    fa1 = gold_fa("uniform_k21_c2")
    fa2 = gold_fa("dirty_k21_c2")
    a = swapamd.Assembler(21, 2, kmer_graph_only=True)
    a.prefetch(fa1)
    a.prefetch(fa1)
    with pytest.raises(swapamd.SwapError) as e:
        a.prefetch(fa1)
    assert e.value.code == swapamd.SWAP_E_STATE
    a.build(fa2)  # not the prefetched input: dropped, fa2 copied as usual
    keys, mult = a.kmer_graph()
    okeys, omult = Oracle(fa2, 21, 2).survivors()
    assert np.array_equal(keys, okeys) and np.array_equal(mult, omult)
    a.prefetch(fa1)  # the queue is empty again
    a.build(fa1)
    keys, mult = a.kmer_graph()
    okeys, omult = Oracle(fa1, 21, 2).survivors()
    assert np.array_equal(keys, okeys) and np.array_equal(mult, omult)
    a.prefetch(fa2)
    a.close()  # destroy with a prefetch in flight
