"""Writes tests/golden/ce_*.gz and ce_cases.json: contig-extension fixtures from
the unmodified reference.

For each case: the input FASTA (seeded, written here), the contigGraph.txt of
its single-rank merge (from the oracle, whose files the GPU path reproduces
byte for byte), and what the reference's own rank-0 code makes of that graph --
its finisher, printContigs, the four clean-up rounds, masterRemoveCrossEdge,
masterPrintContigs and printJungGraph -- run by oracle/_ref/ref_ce (built by
`make -C oracle ref` from /root/reference).  Run where the reference sources
exist; the tests only read the fixtures.

Stored per case (gzipped, zero timestamp):
  ce_<case>.fa.gz            input
  ce_<case>.contigGraph.gz   contigGraph.txt of the merge (the extension's input)
  ce_<case>.noce.gz          noCEcontigs.fasta written by the reference from it
  ce_<case>.cecontig.gz      CEContig.fasta
  ce_<case>.jungarc.gz       JungGraph_arc.txt
  ce_<case>.jungmul.gz       JungGraph_mul.txt
"""
import gzip
import json
import os
import random
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "tests"))
from oracle_binding import Oracle  # noqa: E402

REF_CE = os.path.join(REPO, "oracle", "_ref", "ref_ce")
MPIRUN = "/opt/conda/bin/mpirun"
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def rand_seq(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def mutate(rng, s, rate):
    return "".join(rng.choice([b for b in "ACGT" if b != c]) if rng.random() < rate else c for c in s)


def reads(rng, genome, cov, L, err, circular=False):
    G = len(genome)
    out = []
    for _ in range(int(round(cov * G / L))):
        if circular:
            st = rng.randrange(G)
            s = (genome + genome[:L])[st:st + L]
        else:
            st = rng.randrange(G - L + 1)
            s = genome[st:st + L]
        if rng.random() < 0.5:
            s = rc(s)
        out.append(mutate(rng, s, err))
    return "".join(">r%d\n%s\n" % (i, r) for i, r in enumerate(out))


def case_errors(seed):
    """Error-rich reads at a low cutoff: tips and bubbles everywhere."""
    rng = random.Random(seed)
    return reads(rng, rand_seq(rng, 15000), 30, 100, 0.01)


def case_diploid(seed):
    """Two haplotypes 1% apart: bubbles, multiple edges between the same nodes."""
    rng = random.Random(seed)
    h1 = rand_seq(rng, 12000)
    h2 = mutate(rng, h1, 0.01)
    return reads(rng, h1, 20, 100, 0.002) + reads(rng, h2, 20, 100, 0.002).replace(">r", ">s")


def case_repeats(seed):
    """Repeat copies with divergence and tandem repeats: loops and cross edges."""
    rng = random.Random(seed)
    unit = rand_seq(rng, 300)
    parts = []
    for i in range(5):
        parts.append(rand_seq(rng, 1200))
        parts.append(mutate(rng, unit, 0.01 * (i % 2)))
    tandem = rand_seq(rng, 60)
    parts.append(tandem * 8)
    parts.append(rand_seq(rng, 1500))
    g = "".join(parts)
    return reads(rng, g, 35, 100, 0.003)


def case_circular(seed):
    rng = random.Random(seed)
    return reads(rng, rand_seq(rng, 5000), 30, 100, 0.002, circular=True)


def case_inverted(seed):
    rng = random.Random(seed)
    arm = rand_seq(rng, 400)
    g = rand_seq(rng, 3000) + arm + rand_seq(rng, 200) + rc(arm) + rand_seq(rng, 3000)
    return reads(rng, g, 30, 100, 0.004)


CASES = [
    # name, k, c, generator
    ("errors_k21_c1", 21, 1, lambda: case_errors(201)),
    ("errors_k31_c2", 31, 2, lambda: case_errors(202)),
    ("diploid_k25_c2", 25, 2, lambda: case_diploid(203)),
    ("repeats_k21_c2", 21, 2, lambda: case_repeats(204)),
    ("repeats_k31_c3", 31, 3, lambda: case_repeats(205)),
    ("circular_k19_c2", 19, 2, lambda: case_circular(206)),
    ("inverted_k23_c2", 23, 2, lambda: case_inverted(207)),
]


def gz_write(path, data):
    with open(path, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as f:
            f.write(data)


def main():
    if not os.path.exists(REF_CE):
        sys.exit("build the reference first: make -C oracle ref")
    index = {}
    for name, k, c, gen in CASES:
        fa = gen().encode()
        o = Oracle(fa, k, c)
        cg = o.file(1)
        S = o.stats()["survivors"]
        with tempfile.TemporaryDirectory() as d:
            cgp = os.path.join(d, "contigGraph.txt")
            fap = os.path.join(d, "in.fa")
            open(cgp, "wb").write(cg)
            open(fap, "wb").write(fa)
            out = os.path.join(d, "out")
            p = subprocess.run([MPIRUN, "-np", "1", REF_CE, cgp, str(S), "-k", str(k), "-c", str(c), "-i", fap,
                                "-o", out, "-j"], capture_output=True, text=True, timeout=300)
            assert p.returncode == 0, (name, p.stdout[-500:], p.stderr[-500:])
            files = {}
            for key, fn in (("noce", "noCEcontigs.fasta"), ("cecontig", "CEContig.fasta"),
                            ("jungarc", "JungGraph_arc.txt"), ("jungmul", "JungGraph_mul.txt")):
                files[key] = open(os.path.join(out, fn), "rb").read()
        assert files["noce"] == o.file(2), name  # the reference's printContigs agrees with the oracle
        gz_write(os.path.join(HERE, "ce_%s.fa.gz" % name), fa)
        gz_write(os.path.join(HERE, "ce_%s.contigGraph.gz" % name), cg)
        for key, data in files.items():
            gz_write(os.path.join(HERE, "ce_%s.%s.gz" % (name, key)), data)
        index[name] = {"k": k, "c": c, "survivors": S, "kept": o.stats()["kept"],
                       "ce_records": files["cecontig"].count(b">"), "removed": files["cecontig"].count(b"Removed")}
        print(name, index[name])
    with open(os.path.join(HERE, "ce_cases.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
