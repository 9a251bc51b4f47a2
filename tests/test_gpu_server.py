"""Many assemblies on one GPU (swap_server_*, through the C ABI): every job's
output files are byte-identical to the same input assembled alone on a fresh
context, whatever ran beside it, and the alone results are checked against the
oracle for one job per configuration."""
import time

import numpy as np
import pytest

import swapamd
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu

MERGE_FILES = (swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE)


def alone(fa, k, c, kmer_graph_only=False, hash_order=False):
    a = swapamd.Assembler(k, c, kmer_graph_only=kmer_graph_only, hash_order=hash_order)
    a.build(fa)
    files = {}
    if kmer_graph_only:
        files[swapamd.SWAP_FILE_KMER_GRAPH] = a.file_bytes(swapamd.SWAP_FILE_KMER_GRAPH)
    else:
        a.merge()
        for w in MERGE_FILES:
            files[w] = a.file_bytes(w)
    a.close()
    return files


def uneven_samples(n, seed):
    """n seeded genomes of uneven size (3..60 kbp), 100 bp reads, 20..50x."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = int(np.exp(rng.uniform(np.log(3000), np.log(60000))))
        cov = float(rng.uniform(20, 50))
        out.append(swapamd.generate_params(g, 100, cov, 0.005, 1000 + seed * 100 + i))
    return out


def test_batch_matches_alone_runs():
    samples = uneven_samples(12, seed=1)
    k, c = 31, 3
    srv = swapamd.Server(slots=4)
    jobs = [swapamd.Server.job(fa, k, c, outputs=MERGE_FILES) for fa in samples]
    ids = srv.submit_batch(jobs)
    srv.wait_all()
    for i, fa in enumerate(samples):
        print("batch sample %d: %d bytes" % (i, len(fa)), flush=True)
        want = alone(fa, k, c)
        for w in MERGE_FILES:
            assert srv.output(ids[i], w) == want[w]
    # one of them against the oracle, so the alone path itself is pinned here too
    o = Oracle(samples[0], k, c)
    assert srv.output(ids[0], swapamd.SWAP_FILE_CONTIG_GRAPH) == o.file(1)
    assert srv.output(ids[0], swapamd.SWAP_FILE_NOCE) == o.file(2)
    srv.close()


def test_mixed_parameters_in_one_server():
    """k, c, hash order and -s mode change from job to job on the same slots."""
    samples = uneven_samples(9, seed=2)
    cfgs = [(21, 2, False, False), (25, 3, False, True), (31, 3, True, False)]
    srv = swapamd.Server(slots=3)
    ids, want = [], []
    for i, fa in enumerate(samples):
        k, c, kg, ho = cfgs[i % 3]
        outs = (swapamd.SWAP_FILE_KMER_GRAPH,) if kg else MERGE_FILES
        ids.append(srv.submit(swapamd.Server.job(fa, k, c, outputs=outs, kmer_graph_only=kg, hash_order=ho)))
        want.append(alone(fa, k, c, kmer_graph_only=kg, hash_order=ho))
    for i, jid in enumerate(ids):
        print("mixed job %d cfg %s" % (i, cfgs[i % 3]), flush=True)
        assert srv.wait(jid) == 0
        for w, b in want[i].items():
            assert srv.output(jid, w) == b
    srv.close()


def test_device_resident_inputs_and_tight_budget():
    """Inputs already in HBM; a 1-byte budget admits one job at a time."""
    samples = uneven_samples(5, seed=3)
    bufs = [swapamd.DeviceBuffer(fa) for fa in samples]
    srv = swapamd.Server(slots=4, hbm_budget=1)
    jobs = [swapamd.Server.job(len(fa), 31, 3, outputs=MERGE_FILES, device_ptr=b.ptr.value)
            for fa, b in zip(samples, bufs)]
    ids = srv.submit_batch(jobs)
    srv.wait_all()
    spans = [srv.info(j) for j in ids]
    # with the budget below any job's footprint no two jobs overlap in time
    spans.sort(key=lambda s: s["start_ms"])
    for a, b in zip(spans, spans[1:]):
        assert b["start_ms"] >= a["end_ms"] - 1e-6
    for i, fa in enumerate(samples):
        want = alone(fa, 31, 3)
        for w in MERGE_FILES:
            assert srv.output(ids[i], w) == want[w]
    srv.close()
    for b in bufs:
        b.close()


def test_requests_arriving_while_others_run():
    """Continuous admission: later requests join while earlier ones are on the GPU."""
    samples = uneven_samples(10, seed=4)
    srv = swapamd.Server(slots=4)
    ids = []
    for i, fa in enumerate(samples):
        ids.append(srv.submit(swapamd.Server.job(fa, 25, 3, outputs=MERGE_FILES)))
        time.sleep(0.002 * (i % 3))
    srv.wait_all()
    infos = [srv.info(j) for j in ids]
    assert all(x["status"] == 0 for x in infos)
    for i, fa in enumerate(samples):
        want = alone(fa, 25, 3)
        for w in MERGE_FILES:
            assert srv.output(ids[i], w) == want[w]
    srv.close()


def test_bad_jobs_are_rejected():
    srv = swapamd.Server(slots=1)
    fa = uneven_samples(1, seed=5)[0]
    with pytest.raises(swapamd.SwapError):
        srv.submit(swapamd.Server.job(fa, 30, 3))  # even k
    with pytest.raises(swapamd.SwapError):  # kmerGraph.txt is only produced in -s mode
        srv.submit(swapamd.Server.job(fa, 31, 3, outputs=(swapamd.SWAP_FILE_KMER_GRAPH,)))
    jid = srv.submit(swapamd.Server.job(b">r\nACGT\n", 31, 3, outputs=MERGE_FILES))  # no k-mer at all
    srv.wait_all()
    assert srv.info(jid)["status"] in (0, -2, -8)
    srv.close()
