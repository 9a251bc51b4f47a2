"""Many samples in one build (swap_build_batch + swap_batch_take): every
sample's files and counts equal those of its own build, on inputs of uneven
size, with the overflow table, with the special merge components the single-rank
reference fixtures hold, and on the many-sample workload against the
reference's digests.  Bad batches are rejected before any sample is taken."""
import gzip
import hashlib
import json
import os

import pytest

import swapamd

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = (swapamd.SWAP_FILE_KMER_GRAPH, swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE,
         swapamd.SWAP_FILE_CE_CONTIGS)
STATS = ("reads", "kmers", "distinct", "survivors", "kept", "special")


def files_of(a):
    """Every file's bytes, or the error code writing it returns (a sample
    without k-mers has no edge base for the contig extension's averages)."""
    out = {}
    for w in FILES:
        try:
            out[w] = a.file_bytes(w)
        except swapamd.SwapError as e:
            out[w] = e.code
    return out


def alone(fa, k, c):
    a = swapamd.Assembler(k, c)
    a.build(fa)
    a.merge()
    out = files_of(a)
    st = a.stats()
    a.close()
    return out, {x: st[x] for x in STATS}


def batch_outputs(inputs, k, c, **flags):
    """Builds the inputs as one batch, takes every sample into one reused
    context, merges it and returns its files and counts."""
    bufs = [swapamd.DeviceBuffer(fa) for fa in inputs]
    b = swapamd.Assembler(k, c, **flags)
    b.build_batch(bufs)
    t = swapamd.Assembler(k, c)
    res = []
    for i in range(len(inputs)):
        t.take(b, i)
        t.merge()
        st = t.stats()
        res.append((files_of(t), {x: st[x] for x in STATS}))
    t.close()
    b.close()
    for d in bufs:
        d.close()
    return res


def uneven_inputs(seed0):
    # genomes of 5 kbp to 300 kbp, 100- and 150-base reads, 20-60x
    specs = [(5000, 100, 20.0), (300000, 150, 40.0), (40000, 100, 60.0), (12000, 150, 30.0), (150000, 100, 25.0),
             (80000, 150, 50.0)]
    return [swapamd.generate_params(g, L, cov, 0.005, seed0 + i) for i, (g, L, cov) in enumerate(specs)]


def check(inputs, k, c, **flags):
    got = batch_outputs(inputs, k, c, **flags)
    for i, fa in enumerate(inputs):
        print("sample %d: %d bytes" % (i, len(fa)), flush=True)
        files, st = got[i]
        want_files, want_st = alone(fa, k, c)
        assert st == want_st
        for w in FILES:
            assert files[w] == want_files[w], (i, w)


@pytest.mark.parametrize("k,c", [(31, 3), (21, 2)])
def test_batch_samples_equal_their_own_builds(k, c):
    check(uneven_inputs(700 + k), k, c)


def test_batch_through_the_overflow_table():
    # every count partition through the device-memory table
    check(uneven_inputs(900)[:4], 25, 3, table_spill=True)


def test_batch_of_one_and_a_tiny_sample():
    check([swapamd.generate_params(3000, 100, 10.0, 0.01, 41)], 31, 2)
    # a sample whose reads are all shorter than k gives no k-mer
    short = b"".join(b">r%d\n%s\n" % (i, b"ACGTACGTAC" * 2) for i in range(5))
    check([swapamd.generate_params(8000, 100, 20.0, 0.01, 42), short,
           swapamd.generate_params(9000, 100, 20.0, 0.01, 43)], 31, 2)


def read_gz(name):
    with gzip.open(os.path.join(GOLD, name)) as f:
        return f.read()


DIGEST_FILES = (("kmerGraph", swapamd.SWAP_FILE_KMER_GRAPH), ("contigGraph", swapamd.SWAP_FILE_CONTIG_GRAPH),
                ("noce", swapamd.SWAP_FILE_NOCE), ("cecontig", swapamd.SWAP_FILE_CE_CONTIGS))


def digest_mismatches(rec, files, which=DIGEST_FILES):
    bad = []
    for suffix, w in which:
        want = rec["files"][suffix]
        if (len(files[w]), hashlib.sha256(files[w]).hexdigest()) != (want["size"], want["sha256"]):
            bad.append(suffix)
    return bad


def test_batch_of_special_merge_components():
    # the single-rank reference's files of the golden inputs (cycles,
    # fold-backs, hairpins, the late unlock, dirty reads, saturation), all
    # inputs of one (k, c) in one batch
    cases = json.load(open(os.path.join(GOLD, "np1_cases.json")))
    groups = {}
    for name, rec in sorted(cases.items()):
        if rec["input"].endswith(".fa.gz") and rec["k"] >= 19:
            groups.setdefault((rec["k"], rec["c"]), []).append(name)
    assert sum(len(v) for v in groups.values()) >= 20
    for (k, c), names in sorted(groups.items()):
        inputs = [read_gz(cases[n]["input"]) for n in names]
        got = batch_outputs(inputs, k, c)
        for n, (files, _) in zip(names, got):
            print("%s in a batch of %d" % (n, len(names)), flush=True)
            assert digest_mismatches(cases[n], files) == []


def test_many_sample_workload_in_batches():
    recs = json.load(open(os.path.join(GOLD, "np1_batch.json")))["batch"]
    picked = sorted((int(i) for i in recs), key=lambda i: recs[str(i)]["bytes"])[::16]
    inputs = []
    for i in picked:
        fa, k, c = swapamd.generate("batch", i)
        assert (len(fa), k, c) == (recs[str(i)]["bytes"], recs[str(i)]["k"], recs[str(i)]["c"])
        inputs.append(fa)
    print("many-sample batch: %d samples, %.1f MB" % (len(inputs), sum(map(len, inputs)) / 1e6), flush=True)
    got = batch_outputs(inputs, 31, 4)
    for i, (files, _) in zip(picked, got):
        assert digest_mismatches(recs[str(i)], files, DIGEST_FILES[1:]) == [], i


def test_bad_batches_are_rejected():
    good = swapamd.generate_params(5000, 100, 10.0, 0.01, 51)
    for bad in (good[:-1],                        # no newline at its end
                good + b">dangling header\n",     # an odd line count before the next sample
                b"\n" + good):                    # does not start with '>'
        bufs = [swapamd.DeviceBuffer(x) for x in (bad, good)]
        b = swapamd.Assembler(31, 2)
        with pytest.raises(swapamd.SwapError) as e:
            b.build_batch(bufs)
        assert e.value.code == swapamd.SWAP_E_INVAL
        t = swapamd.Assembler(31, 2)
        with pytest.raises(swapamd.SwapError) as e:
            t.take(b, 0)
        assert e.value.code == swapamd.SWAP_E_STATE
        for x in (b, t):
            x.close()
        for d in bufs:
            d.close()
    # a context of another k cannot take a sample
    bufs = [swapamd.DeviceBuffer(good)]
    b = swapamd.Assembler(31, 2)
    b.build_batch(bufs)
    t = swapamd.Assembler(25, 2)
    with pytest.raises(swapamd.SwapError) as e:
        t.take(b, 0)
    assert e.value.code == swapamd.SWAP_E_INVAL
    with pytest.raises(swapamd.SwapError):
        t.take(b, 1)
    for x in (b, t):
        x.close()
    bufs[0].close()


def test_server_fused_submission():
    """16 samples of the many-sample workload through a server's fused
    submission, in groups small enough that both batch contexts take turns
    while earlier groups merge; one more submission whose group holds a
    sample no batch build accepts runs as ordinary jobs.  Every job's files
    equal the reference's digests and its own build's bytes."""
    recs = json.load(open(os.path.join(GOLD, "np1_batch.json")))["batch"]
    picked = sorted((int(i) for i in recs), key=lambda i: recs[str(i)]["bytes"])[8::16]
    inputs = []
    for i in picked:
        fa, k, c = swapamd.generate("batch", i)
        inputs.append((i, fa, swapamd.DeviceBuffer(fa)))
    outs = (swapamd.SWAP_FILE_CONTIG_GRAPH, swapamd.SWAP_FILE_NOCE, swapamd.SWAP_FILE_CE_CONTIGS)
    srv = swapamd.Server(slots=6)
    jobs = [swapamd.Server.job(fa, 31, 4, outputs=outs, device_ptr=d.ptr.value) for _, fa, d in inputs]
    ids = srv.submit_fused(jobs, group_bytes=48 << 20)
    # a group with a sample that does not start with '>'
    odd = b"\n" + swapamd.generate_params(20000, 100, 20.0, 0.01, 61)
    plain = swapamd.generate_params(30000, 150, 30.0, 0.005, 62)
    extra = [(odd, swapamd.DeviceBuffer(odd)), (plain, swapamd.DeviceBuffer(plain))]
    ids2 = srv.submit_fused([swapamd.Server.job(fa, 31, 4, outputs=outs, device_ptr=d.ptr.value) for fa, d in extra])
    srv.wait_all()
    for (i, fa, _), jid in zip(inputs, ids):
        print("sample %d: %d bytes" % (i, len(fa)), flush=True)
        assert srv.info(jid)["status"] == 0, srv.error(jid)
        got = {w: srv.output(jid, w) for w in outs}
        assert digest_mismatches(recs[str(i)], got, DIGEST_FILES[1:]) == []
    for (fa, _), jid in zip(extra, ids2):
        assert srv.info(jid)["status"] == 0, srv.error(jid)
        a = swapamd.Assembler(31, 4)
        a.build(fa)
        a.merge()
        assert all(srv.output(jid, w) == a.file_bytes(w) for w in outs)
        a.close()
    srv.close()
    for x in [d for _, _, d in inputs] + [d for _, d in extra]:
        x.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_batches(seed):
    """Seeded random batches: k, c, read lengths, error rates, coverages and
    genome sizes drawn per batch and per sample; every sample's files and
    counts equal its own build's."""
    import random
    rng = random.Random(seed)
    k = rng.choice([19, 23, 27])
    c = rng.randint(1, 3)
    inputs = []
    for i in range(rng.randint(4, 9)):
        L = rng.choice([100, 150, 250])
        g = rng.choice([2000, 8000, 30000, 90000])
        inputs.append(swapamd.generate_params(g, L, rng.uniform(8.0, 40.0), rng.choice([0.0, 0.005, 0.02]),
                                              1000 * seed + i))
    print("random batch %d: k=%d c=%d, %d samples" % (seed, k, c, len(inputs)), flush=True)
    check(inputs, k, c)
