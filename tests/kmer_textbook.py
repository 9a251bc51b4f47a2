"""Textbook k-mer graph of a FASTA input, written with Python strings and a
dictionary of (k+1)-mer counts -- an independent statement of what the
multiplicities must be (used to pin the oracle on small inputs):

    mult[X][b]   = min(255, C[X.b] + C[rc(X.b)])
    mult[X][4+b] = min(255, C[rc(X).b] + C[rc(rc(X).b)])

with C[w] the number of times the (k+1)-mer w occurs in a read as written and
X = max(x, rc x) (string order, A<C<G<T).  Records are the line pairs after the
first '>' (the reference's grammar); bytes other than ACGT/acgt read as 'A'.
"""
from collections import Counter

COMP = str.maketrans("ACGT", "TGCA")


def rc(s):
    return s.translate(COMP)[::-1]


def records(fasta: bytes):
    text = fasta.decode("latin-1")
    i = text.find(">")
    if i < 0:
        return []
    lines = text[i:].split("\n")
    # the last element is what follows the final '\n' (unterminated tail)
    complete = lines[:-1]
    seqs = []
    for j in range(0, len(complete) - 1, 2):
        s = complete[j + 1]
        s = "".join(ch.upper() if ch in "ACGTacgt" else "A" for ch in s)
        seqs.append(s)
    return seqs


def graph(fasta: bytes, k: int, c: int):
    """{X(int): [8 multiplicities]} of the k-molecules surviving cutoff c."""
    C = Counter()
    for s in records(fasta):
        for p in range(len(s) - k):
            C[s[p:p + k + 1]] += 1
    val = {"A": 0, "C": 1, "G": 2, "T": 3}

    def enc(s):
        v = 0
        for ch in s:
            v = v * 4 + val[ch]
        return v

    nodes = {}
    for w in C:
        for x in (w[:k], rc(w[1:])):  # both k-mers the (k+1)-mer touches, as strands
            X = max(x, rc(x))
            if X in nodes:
                continue
            m = []
            for half in range(2):
                o = X if half == 0 else rc(X)
                for b in "ACGT":
                    e = o + b
                    m.append(min(255, C.get(e, 0) + C.get(rc(e), 0)))
            nodes[X] = m
    out = {}
    for X, m in nodes.items():
        m = [v if v >= c else 0 for v in m]
        if any(m):
            out[enc(X)] = m
    return out
