"""Writes the golden fixtures under tests/golden/ from the unmodified reference.

Run in the build container (it needs /root/reference built into oracle/_ref by
``make -C oracle ref`` and MPICH's mpirun):

    python tests/golden/make_golden.py

Each case is a small FASTA input (committed gzipped) plus what the reference
wrote for it, run as ``mpirun -np 4 oracle/_ref/swap -k K -c C -i in.fa -o out``
(once with -s for kmerGraph.txt, once in the default mode for
noCEcontigs.fasta).  The reference cannot be run at np=1 in this image: its
record-boundary ring (kmerGraph.cpp:121-122) is a blocking MPI_Send to itself,
which MPICH 3.3.2 never completes; at np=2 its merge aborts with "Error in
find the Node" on these inputs.  np=4 is used because the k-mer graph and the
noCE contig set do not depend on the rank count (checked against np=3, 5 and 8
when the cases were made).

Stored per case (all gzipped text, written with a zero gzip timestamp so that
reruns reproduce the same bytes):
  <case>.fa.gz          the input
  <case>.kmerGraph.gz   kmerGraph.txt lines, sorted (line order is hash order)
  <case>.noce.gz        noCEcontigs.fasta lines, sorted
and one JSON index, cases.json, with k, c and the line counts.

The -p mode output is not stored: the reference writes it from each rank's
table right after the distributed merge, before the rank-0 finisher, so at
np=4 it still holds the (1,1) nodes whose locks collided and differs from the
single-rank result in topology, not just in order.
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
REF = os.path.join(REPO, "oracle", "_ref", "swap")
MPIRUN = "/opt/conda/bin/mpirun"

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def reads_from(genome, n, L, err, rng, circular=False):
    out = []
    G = len(genome)
    for _ in range(n):
        if circular:
            st = rng.randrange(G)
            s = (genome + genome[:L])[st:st + L]
        else:
            st = rng.randrange(G - L + 1)
            s = genome[st:st + L]
        if rng.random() < 0.5:
            s = rc(s)
        s = list(s)
        for j in range(L):
            if rng.random() < err:
                s[j] = rng.choice([b for b in "ACGT" if b != s[j]])
        out.append("".join(s))
    return out


def fasta(reads):
    return "".join(">r%d\n%s\n" % (i, r) for i, r in enumerate(reads))


def case_uniform(seed, G, L, cov, err):
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(G))
    return fasta(reads_from(g, int(round(cov * G / L)), L, err, rng))


def case_dirty(seed):
    """Lower case, N, IUPAC codes, CRLF line ends, short and exactly-k reads."""
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(6000))
    reads = reads_from(g, 3000, 90, 0.004, rng)
    recs = []
    for i, r in enumerate(reads):
        x = rng.random()
        if x < 0.05:
            r = r.lower()
        elif x < 0.08:
            p = rng.randrange(len(r))
            r = r[:p] + "N" + r[p + 1:]
        elif x < 0.10:
            p = rng.randrange(len(r))
            r = r[:p] + rng.choice("RYKMSWn-.") + r[p + 1:]
        elif x < 0.12:
            r = r[: rng.randrange(0, 21)]          # shorter than k (skipped)
        elif x < 0.13:
            r = r[:21]                             # exactly k: one k-mer, no edges
        elif x < 0.15:
            r = r + "\r"                           # CR becomes an 'A' base
        recs.append(">read_%d some description\n%s\n" % (i, r))
    return "".join(recs)


def case_saturate(seed):
    """Deep coverage of a short genome: counts pass 255 and saturate."""
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(1500))
    return fasta(reads_from(g, 5000, 100, 0.002, rng))


def case_repeats(seed):
    """Repeats and an inverted repeat: branching nodes and fold-back paths."""
    rng = random.Random(seed)
    unit = "".join(rng.choice("ACGT") for _ in range(400))
    parts = []
    for _ in range(4):
        parts.append("".join(rng.choice("ACGT") for _ in range(1500)))
        parts.append(unit)
    inv = "".join(rng.choice("ACGT") for _ in range(300))
    parts.append(inv + rc(inv))
    parts.append("".join(rng.choice("ACGT") for _ in range(1500)))
    g = "".join(parts)
    return fasta(reads_from(g, int(30 * len(g) / 100), 100, 0.003, rng))


CASES = [
    # name, k, c, generator
    ("uniform_k31_c3", 31, 3, lambda: case_uniform(101, 20000, 100, 30, 0.005)),
    ("uniform_k21_c2", 21, 2, lambda: case_uniform(102, 8000, 75, 25, 0.008)),
    ("dirty_k21_c2", 21, 2, lambda: case_dirty(103)),
    ("saturate_k19_c0", 19, 0, lambda: case_saturate(104)),
    ("repeats_k25_c3", 25, 3, lambda: case_repeats(105)),
]


def gz_write(path, text):
    with open(path, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as f:
            f.write(text.encode())


def run_ref(k, c, fa, flag, scratch):
    out = tempfile.mkdtemp(prefix="ref_", dir=scratch)
    os.rmdir(out)  # the reference creates its output directory itself
    cmd = [MPIRUN, "-np", "4", REF, "-k", str(k), "-c", str(c), "-i", fa, "-o", out]
    if flag:
        cmd.append(flag)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120)
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    scratch = tempfile.mkdtemp(prefix="swap_golden_")  # outside the repository
    index = {}
    for name, k, c, gen in CASES:
        text = gen()
        fa = os.path.join(scratch, name + ".fa")
        with open(fa, "w") as f:
            f.write(text)
        gz_write(os.path.join(HERE, name + ".fa.gz"), text)
        out_s = run_ref(k, c, fa, "-s", scratch)
        with open(os.path.join(out_s, "kmerGraph.txt")) as f:
            kg = sorted(f.read().splitlines())
        shutil.rmtree(out_s)
        out_d = run_ref(k, c, fa, None, scratch)
        with open(os.path.join(out_d, "noCEcontigs.fasta")) as f:
            noce = sorted(f.read().splitlines())
        shutil.rmtree(out_d)
        gz_write(os.path.join(HERE, name + ".kmerGraph.gz"), "\n".join(kg) + ("\n" if kg else ""))
        gz_write(os.path.join(HERE, name + ".noce.gz"), "\n".join(noce) + ("\n" if noce else ""))
        index[name] = {"k": k, "c": c, "kmer_graph_lines": len(kg), "noce_lines": len(noce), "np": 4}
        print(name, index[name])
    shutil.rmtree(scratch)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
