"""The many-sample and request-stream workloads through the server, against
what the unmodified reference wrote running each input alone as one rank
(digests from tests/golden/make_golden_batch.py) and against the same input
assembled alone on a fresh context:

* 16 samples of the many-sample workload (genomes of 20 kbp - 2 Mbp, 150-base
  reads at 20-80x, k=31, c=4), from the smallest to the largest, submitted as
  one batch;
* requests of the request-stream workload covering every k in {21, 25, 31}
  and every read length in {100, 150, 250}, arriving while earlier ones run.

Each job's contigGraph.txt, noCEcontigs.fasta and CEContig.fasta must equal
the reference's bytes, whatever ran beside it."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import swapamd

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIGESTS = json.load(open(os.path.join(GOLD, "np1_batch.json")))
FILES = [("contigGraph", swapamd.SWAP_FILE_CONTIG_GRAPH), ("noce", swapamd.SWAP_FILE_NOCE),
         ("cecontig", swapamd.SWAP_FILE_CE_CONTIGS)]
OUTPUTS = tuple(w for _, w in FILES)


def matches_reference(rec, outputs):
    bad = []
    for suffix, w in FILES:
        want = rec["files"][suffix]
        got = outputs[w]
        if len(got) != want["size"] or hashlib.sha256(got).hexdigest() != want["sha256"]:
            bad.append(suffix)
    return bad


def alone(fa, k, c):
    a = swapamd.Assembler(k, c)
    a.build(fa)
    a.merge()
    out = {w: a.file_bytes(w) for w in OUTPUTS}
    a.close()
    return out


def spread(indices, sizes, n):
    """n indices from smallest to largest input, evenly spaced in size rank."""
    order = sorted(indices, key=lambda i: sizes[i])
    pick = [order[int(round(j * (len(order) - 1) / (n - 1)))] for j in range(n)]
    return sorted(set(pick), key=lambda i: sizes[i])


def test_many_sample_workload_batch():
    recs = DIGESTS["batch"]
    sizes = {int(i): r["bytes"] for i, r in recs.items()}
    picked = spread(list(sizes), sizes, 16)
    assert len(picked) >= 12
    srv = swapamd.Server(slots=8)
    jobs, inputs = [], []
    for i in picked:
        fa, k, c = swapamd.generate("batch", i)
        assert (len(fa), k, c) == (sizes[i], recs[str(i)]["k"], recs[str(i)]["c"])
        inputs.append((i, fa, k, c))
        jobs.append(swapamd.Server.job(fa, k, c, outputs=OUTPUTS))
    print("many-sample batch: samples %s, %.1f MB of FASTA" % (picked, sum(len(x[1]) for x in inputs) / 1e6),
          flush=True)
    ids = srv.submit_batch(jobs)
    srv.wait_all()
    for (i, fa, k, c), jid in zip(inputs, ids):
        print("sample %d: %d bytes" % (i, len(fa)), flush=True)
        assert srv.info(jid)["status"] == 0
        got = {w: srv.output(jid, w) for w in OUTPUTS}
        assert matches_reference(recs[str(i)], got) == []
        assert got == alone(fa, k, c)
    srv.close()


def read_length(fa):
    return len(fa.split(b"\n", 2)[1])


def test_request_stream_mixed_k_and_read_lengths():
    recs = DIGESTS["request"]
    by_cfg = {}  # the smallest request of every (k, read length) pair
    for i in sorted(recs, key=lambda x: recs[x]["bytes"]):
        key = (recs[i]["k"], recs[i]["read_len"])
        if key not in by_cfg:
            by_cfg[key] = int(i)
    assert {k for k, _ in by_cfg} == {21, 25, 31}
    assert {L for _, L in by_cfg} == {100, 150, 250}
    reqs = []
    for i in sorted(by_cfg.values()):
        fa, k, c = swapamd.generate("request", i)
        assert (len(fa), k, read_length(fa)) == (recs[str(i)]["bytes"], recs[str(i)]["k"], recs[str(i)]["read_len"])
        reqs.append((i, fa, k, c))
    assert len(reqs) >= 8
    srv = swapamd.Server(slots=4)
    rng = np.random.default_rng(5)
    ids = []
    for i, fa, k, c in reqs:
        print("request %d: k=%d L=%d %d bytes" % (i, k, read_length(fa), len(fa)), flush=True)
        ids.append(srv.submit(swapamd.Server.job(fa, k, c, outputs=OUTPUTS)))
        time.sleep(float(rng.exponential(0.02)))  # arrivals while earlier requests run
    srv.wait_all()
    infos = [srv.info(j) for j in ids]
    assert all(x["status"] == 0 for x in infos)
    # some requests started before an earlier one ended: they shared the device
    spans = sorted((x["start_ms"], x["end_ms"]) for x in infos)
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:]))
    for (i, fa, k, c), jid in zip(reqs, ids):
        got = {w: srv.output(jid, w) for w in OUTPUTS}
        assert matches_reference(recs[str(i)], got) == []
        assert got == alone(fa, k, c)
    srv.close()
