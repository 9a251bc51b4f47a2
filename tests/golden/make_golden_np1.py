"""Writes the single-rank (np=1) fixtures under tests/golden/ from the unmodified
reference run end to end as one rank.

Run in the build container (it needs oracle/_ref/ref_swap1, built by
``make -C oracle ref`` from /root/reference, and MPICH's mpirun):

    python tests/golden/make_golden_np1.py            # small cases + 1 Mbp workload
    python tests/golden/make_golden_np1.py --bacterial  # also the 4.6 Mbp workload (~10 min)

ref_swap1 is the reference's own main at np=1 (oracle/ref_swap1.cpp says why a
driver is needed and what it changes: nothing on the path).  At np=1 the
reference is deterministic: its SWAP pass meets every node in the node table's
iteration order and every lock request is served by the same process, so its
files are exact bytes, line order included.  Each case runs three times:

  -d -j   contigGraph.txt, noCEcontigs.fasta, CEContig.fasta, JungGraph_*.txt
  -s      kmerGraph.txt
  -p      noCEcontigs.fasta as written by stringContigs right after the merge

Stored:
  np1_<case>.fa.gz                inputs of the special-component cases
                                  (the other cases reuse <case>.fa.gz / ce_<case>.fa.gz)
  np1_<case>.<file>.gz            exact bytes of each file, for cases small enough
  np1_cases.json                  per case: k, c, input, and size + sha256 of
                                  every file of the three runs

The special-component cases are built to reach the parts of the merge that a
plain random genome never does: isolated cycles (circular genomes, pure
homopolymer / dinucleotide / short-period reads), nodes that are their own
neighbour (homopolymer and tandem-repeat runs), two-node cycles (the single
lock of a neighbour that is both left and right), fold-backs (palindromic arms:
a node whose one edge returns to itself, so that its visit locks the other
neighbour and then fails on itself) and the late unlock that follows (the
locked neighbour is only released at the next lock request, so when it is the
next node to ask for locks the pass skips it).
"""
import argparse
import gzip
import hashlib
import json
import os
import random
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF1 = os.path.join(REPO, "oracle", "_ref", "ref_swap1")
GEN = os.path.join(REPO, "swap-assembler_amd", "bin", "swap_gen")
MPIRUN = "/opt/conda/bin/mpirun"
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}

# files of each run mode and the fixture suffix they are stored under
MODES = [
    (["-d", "-j"], [("contigGraph.txt", "contigGraph"), ("noCEcontigs.fasta", "noce"),
                    ("CEContig.fasta", "cecontig"), ("JungGraph_arc.txt", "jungarc"),
                    ("JungGraph_mul.txt", "jungmul")]),
    (["-s"], [("kmerGraph.txt", "kmerGraph")]),
    (["-p"], [("noCEcontigs.fasta", "perf")]),
]
KEEP_BYTES_MAX = 2 << 20  # files up to this size are stored whole (gzipped)


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def rand_seq(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def mutate(rng, s, rate):
    return "".join(rng.choice([b for b in "ACGT" if b != c]) if rng.random() < rate else c for c in s)


def sample(rng, genome, cov, L, err, circular=False):
    G = len(genome)
    out = []
    for _ in range(max(1, int(round(cov * G / L)))):
        if circular:
            st = rng.randrange(G)
            s = (genome + genome[:L])[st:st + L]
        else:
            st = rng.randrange(G - L + 1)
            s = genome[st:st + L]
        if rng.random() < 0.5:
            s = rc(s)
        out.append(mutate(rng, s, err))
    return out


def fasta(reads):
    return "".join(">r%d\n%s\n" % (i, r) for i, r in enumerate(reads))


def tandem(unit, n):
    return (unit * (n // len(unit) + 1))[:n]


def case_cycles(seed):
    """Isolated cycles: two circular genomes, error free, and reads made only of a
    homopolymer (one node, its own neighbour on both sides), a dinucleotide
    repeat (two nodes, each the other's left and right neighbour) and short
    periodic repeats (cycles of 3-7 nodes)."""
    rng = random.Random(seed)
    reads = sample(rng, rand_seq(rng, 3000), 30, 100, 0.0, circular=True)
    reads += sample(rng, rand_seq(rng, 1100), 30, 100, 0.0, circular=True)
    for unit in ["A", "C", "AC", "AG", "ACG", "AAC", "ACGT", "AACCG", "AGCTTC", "ACCGTTA"]:
        reads += [tandem(unit, 80)[o:o + 60] for o in range(len(unit))] * 3
    rng.shuffle(reads)
    return fasta(reads)


def case_ends(seed):
    """Chromosomes that end (or start) in a homopolymer or a short-period run:
    the run's nodes are their own neighbours or form short cycles hanging off
    the chromosome's chain."""
    rng = random.Random(seed)
    reads = []
    units = ["A", "C", "G", "T", "AC", "GT", "ACG", "TTG", "ACGT", "CAGG", "ACGTA", "TTGCCA"]
    for i, unit in enumerate(units):
        body = rand_seq(rng, 500)
        run = tandem(unit, 70)
        chrom = body + run if i % 2 == 0 else run + body
        if i % 3 == 2:
            chrom = tandem(units[(i + 5) % len(units)], 70) + body + run
        reads += sample(rng, chrom, 40, 100, 0.002)
    rng.shuffle(reads)
    return fasta(reads)


def case_hairpins(seed):
    """Fold-backs: arms followed by their reverse complement (even palindromes)
    and by one base and their reverse complement (odd), of several lengths, in
    linear chromosomes and alone."""
    rng = random.Random(seed)
    reads = []
    for arm_len in [15, 24, 40, 90, 150]:
        arm = rand_seq(rng, arm_len)
        left = rand_seq(rng, 300)
        reads += sample(rng, left + arm + rc(arm), 40, 100, 0.0)
        arm2 = rand_seq(rng, arm_len)
        mid = rng.choice("ACGT")
        reads += sample(rng, rand_seq(rng, 250) + arm2 + mid + rc(arm2) + rand_seq(rng, 250), 40, 100, 0.0)
    core = rand_seq(rng, 120)
    reads += sample(rng, core + rc(core), 40, 100, 0.0, circular=True)
    rng.shuffle(reads)
    return fasta(reads)


def case_mixed(seed, G, err):
    """A random genome with every structure above embedded (tandem runs of
    periods 1-7, palindromic arms, a repeat copied three times), error-rich."""
    rng = random.Random(seed)
    parts = []
    for i in range(24):
        parts.append(rand_seq(rng, G // 30))
        kind = i % 4
        if kind == 0:
            parts.append(tandem(rand_seq(rng, 1 + i % 7), 40 + 5 * (i % 5)))
        elif kind == 1:
            arm = rand_seq(rng, 20 + 7 * (i % 6))
            parts.append(arm + rc(arm))
        elif kind == 2:
            parts.append(rand_seq(rng, 60) + rc(rand_seq(rng, 60)))
        else:
            parts.append(rand_seq(rng, 300) * 1)
    rep = rand_seq(rng, 200)
    g = "".join(parts)
    third = len(g) // 3
    g = g[:third] + rep + g[third:2 * third] + rep + g[2 * third:] + rc(rep)
    return fasta(sample(rng, g, 35, 100, err))


def case_foldback_lock(seed, k):
    """Short chromosomes ending in a palindromic arm (arm + its reverse
    complement).  The node at the centre has a single right edge that returns to
    itself, so its visit locks its left neighbour and then fails on itself; the
    unlock of that neighbour is only served at the next lock request.  Seeds are
    chosen (by a search over the oracle) so that the neighbour is the next node
    of the table that asks for locks, which is then skipped by the pass."""
    rng = random.Random(seed)
    reads = []
    for _ in range(12):
        arm = rand_seq(rng, k + rng.randrange(1, 8))
        chrom = rand_seq(rng, rng.randrange(10, 40)) + arm + rc(arm)
        reads += sample(rng, chrom, 30, min(60, len(chrom)), 0.0)
    rng.shuffle(reads)
    return fasta(reads)


SPECIAL = [
    # name, k, c, generator
    ("cycles_k21_c1", 21, 1, lambda: case_cycles(301)),
    ("cycles_k31_c1", 31, 1, lambda: case_cycles(302)),
    ("ends_k19_c2", 19, 2, lambda: case_ends(303)),
    ("ends_k25_c1", 25, 1, lambda: case_ends(304)),
    ("hairpins_k25_c2", 25, 2, lambda: case_hairpins(305)),
    ("hairpins_k19_c1", 19, 1, lambda: case_hairpins(306)),
    ("foldlock_k19_c1", 19, 1, lambda: case_foldback_lock(1005, 19)),
    ("foldlock_k25_c1", 25, 1, lambda: case_foldback_lock(1007, 25)),
    ("mixed_k31_c2", 31, 2, lambda: case_mixed(307, 30000, 0.01)),
    ("mixed_k19_c1", 19, 1, lambda: case_mixed(308, 12000, 0.005)),
    ("mixed_k23_c3", 23, 3, lambda: case_mixed(309, 20000, 0.008)),
]


def gz_write(path, data):
    with open(path, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as f:
            f.write(data)


def gz_read(path):
    with gzip.open(path, "rb") as f:
        return f.read()


def run_np1(k, c, fa, flags, scratch, timeout):
    out = tempfile.mkdtemp(prefix="np1_", dir=scratch)
    os.rmdir(out)  # the reference creates its output directory and refuses an existing one
    cmd = [MPIRUN, "-np", "1", REF1, "-k", str(k), "-c", str(c), "-i", fa, "-o", out] + flags
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout)
    return out


def run_case(name, k, c, fa, scratch, keep, timeout=600):
    rec = {"k": k, "c": c, "files": {}}
    for flags, files in MODES:
        out = run_np1(k, c, fa, flags, scratch, timeout)
        for fname, suffix in files:
            p = os.path.join(out, fname)
            data = open(p, "rb").read() if os.path.exists(p) else None
            if data is None:
                rec["files"][suffix] = None
                continue
            rec["files"][suffix] = {"size": len(data), "sha256": hashlib.sha256(data).hexdigest()}
            if keep and len(data) <= KEEP_BYTES_MAX:
                gz_write(os.path.join(HERE, "np1_%s.%s.gz" % (name, suffix)), data)
                rec["files"][suffix]["stored"] = True
        shutil.rmtree(out)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bacterial", action="store_true", help="also the 4.6 Mbp / 50x workload")
    args = ap.parse_args()
    if not os.path.exists(REF1):
        sys.exit("build the reference first: make -C oracle ref")
    scratch = tempfile.mkdtemp(prefix="swap_np1_")  # outside the repository
    idx_path = os.path.join(HERE, "np1_cases.json")
    index = json.load(open(idx_path)) if os.path.exists(idx_path) else {}
    try:
        for name, k, c, gen in SPECIAL:
            text = gen().encode()
            fa = os.path.join(scratch, name + ".fa")
            open(fa, "wb").write(text)
            gz_write(os.path.join(HERE, "np1_%s.fa.gz" % name), text)
            index[name] = run_case(name, k, c, fa, scratch, keep=True)
            index[name]["input"] = "np1_%s.fa.gz" % name
            print(name, json.dumps(index[name]["files"]["contigGraph"]), flush=True)
        for listing, prefix in (("cases.json", ""), ("ce_cases.json", "ce_")):
            for name, case in sorted(json.load(open(os.path.join(HERE, listing))).items()):
                inp = "%s%s.fa.gz" % (prefix, name)
                fa = os.path.join(scratch, prefix + name + ".fa")
                open(fa, "wb").write(gz_read(os.path.join(HERE, inp)))
                key = prefix + name
                index[key] = run_case(key, case["k"], case["c"], fa, scratch, keep=False)
                index[key]["input"] = inp
                print(key, json.dumps(index[key]["files"]["contigGraph"]), flush=True)
        workloads = [("small", 31, 3)] + ([("bacterial", 31, 5)] if args.bacterial else [])
        for wl, k, c in workloads:
            fa = os.path.join(scratch, wl + ".fa")
            subprocess.run([GEN, "--workload", wl, "-o", fa], check=True, stdout=subprocess.DEVNULL)
            key = "workload_" + wl
            index[key] = run_case(key, k, c, fa, scratch, keep=False, timeout=3000)
            index[key]["input"] = "workload:" + wl
            print(key, json.dumps(index[key]["files"]["contigGraph"]), flush=True)
    finally:
        shutil.rmtree(scratch)
    with open(idx_path, "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
