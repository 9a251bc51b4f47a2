"""swap_assemble_static (the entry point that allocates nothing: every device
buffer from the library's static scratch) against swap_build_device +
swap_merge on the same inputs: the same k-mer graph and the same output files,
on two calls in a row with the second input smaller, and for a -s context."""
import ctypes

import numpy as np
import pytest

import swapamd

pytestmark = pytest.mark.gpu

FILES = (swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE, swapamd.SWAP_FILE_PERF_CONTIGS,
         swapamd.SWAP_FILE_CE_CONTIGS)


def reference_outputs(fa, k, c, kmer_graph_only=False):
    buf = swapamd.DeviceBuffer(fa)
    a = swapamd.Assembler(k, c, kmer_graph_only=kmer_graph_only)
    a.build_device(buf.ptr.value, buf.n)
    if kmer_graph_only:
        out = {swapamd.SWAP_FILE_KMER_GRAPH: a.file_bytes(swapamd.SWAP_FILE_KMER_GRAPH)}
    else:
        a.merge()
        out = {w: a.file_bytes(w) for w in FILES}
    keys, mult = a.kmer_graph()
    a.close()
    buf.close()
    return keys, mult, out


def test_static_calls_match_the_allocating_path():
    # near the input limit with the most buffers per byte (short k, long reads,
    # error-rich, nothing cut), then a smaller one
    big = swapamd.generate_params(101000, 250, 30.0, 0.02, 71)     # ~3.1 MB of FASTA
    assert len(big) <= 3 << 20
    small = swapamd.generate_params(30000, 150, 25.0, 0.01, 72)    # ~0.8 MB
    cases = [(big, 19, 1), (small, 25, 2)]
    want = [reference_outputs(fa, k, c) for fa, k, c in cases]
    bufs = [swapamd.DeviceBuffer(fa) for fa, _, _ in cases]
    for (fa, k, c), buf, (wk, wm, wout) in zip(cases, bufs, want):
        print("static call: k=%d c=%d bytes=%d" % (k, c, len(fa)), flush=True)
        a = swapamd.Assembler(k, c)
        a.assemble_static(buf.ptr.value, buf.n)
        keys, mult = a.kmer_graph()
        assert np.array_equal(keys, wk) and np.array_equal(mult, wm)
        for w, data in wout.items():
            assert a.file_bytes(w) == data, w
        with pytest.raises(swapamd.SwapError) as e:
            a.merge()  # merged already, in the scratch
        assert e.value.code == -8
        a.close()
    for b in bufs:
        b.close()


def test_static_kmer_graph_context_and_reuse():
    fa1 = swapamd.generate_params(60000, 100, 30.0, 0.01, 73)
    fa2 = swapamd.generate_params(20000, 100, 30.0, 0.01, 74)
    a = swapamd.Assembler(21, 2, kmer_graph_only=True)
    for fa in (fa1, fa2):  # the same context twice, the second input smaller
        print("static -s call: bytes=%d" % len(fa), flush=True)
        wk, wm, wout = reference_outputs(fa, 21, 2, kmer_graph_only=True)
        buf = swapamd.DeviceBuffer(fa)
        a.assemble_static(buf.ptr.value, buf.n)
        keys, mult = a.kmer_graph()
        assert np.array_equal(keys, wk) and np.array_equal(mult, wm)
        assert a.file_bytes(swapamd.SWAP_FILE_KMER_GRAPH) == wout[swapamd.SWAP_FILE_KMER_GRAPH]
        buf.close()
    # an ordinary build on the same context afterwards uses its own memory again
    b2 = swapamd.DeviceBuffer(fa1)
    a.build_device(b2.ptr.value, b2.n)
    keys, mult = a.kmer_graph()
    wk, wm, _ = reference_outputs(fa1, 21, 2, kmer_graph_only=True)
    assert np.array_equal(keys, wk) and np.array_equal(mult, wm)
    b2.close()
    a.close()


def test_static_input_limit_is_an_error():
    # one byte over the limit is refused before any device work
    lim = 3 << 20
    a = swapamd.Assembler(31, 3)
    buf = swapamd.DeviceBuffer(b">r\n" + b"A" * (lim - 3) + b"\n")
    assert buf.n == lim + 1
    with pytest.raises(swapamd.SwapError) as e:
        a.assemble_static(buf.ptr.value, lim + 1)
    assert e.value.code == -4
    buf.close()
    a.close()
