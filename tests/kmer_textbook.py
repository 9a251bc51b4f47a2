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


def rc_codes(x, k):
    """Reverse complements of 2-bit MSB-first k-mer codes (numpy uint64 array)."""
    import numpy as np
    x = ~np.asarray(x, dtype=np.uint64)  # complement: 3 - a on every 2-bit digit
    m2, m4 = np.uint64(0x3333333333333333), np.uint64(0x0F0F0F0F0F0F0F0F)
    x = ((x >> np.uint64(2)) & m2) | ((x & m2) << np.uint64(2))
    x = ((x >> np.uint64(4)) & m4) | ((x & m4) << np.uint64(4))
    return x.byteswap() >> np.uint64(64 - 2 * k)


def reciprocal_violations(keys, mult, k, rows=None):
    """Edges of a k-molecule graph whose reciprocal edge is missing or carries a
    different multiplicity.  Both ends of an edge count the same (k+1)-mer
    occurrences (the formulas above), so on a correct graph this is 0 for any
    cutoff: edge b of X (X followed by b) is edge 4 + comp(X[0]) of
    Y = X[1:] b when Y is canonical, else edge comp(X[0]) of Y = rc(X[1:] b);
    edges 4..7 are the same with rc(X) in place of X.  keys: sorted unique
    uint64 codes; mult: uint8 [n, 8]; rows: the records whose edges are checked
    (default all)."""
    import numpy as np
    keys = np.asarray(keys, dtype=np.uint64)
    mult = np.asarray(mult, dtype=np.uint8)
    mask = np.uint64((1 << (2 * k)) - 1)
    bad = 0
    rows = np.arange(len(keys)) if rows is None else np.asarray(rows)
    for t in range(8):
        sel = rows[np.nonzero(mult[rows, t])[0]]
        if not len(sel):
            continue
        src = keys[sel] if t < 4 else rc_codes(keys[sel], k)
        first = src >> np.uint64(2 * k - 2)
        nxt = ((src << np.uint64(2)) | np.uint64(t & 3)) & mask
        rn = rc_codes(nxt, k)
        y = np.maximum(nxt, rn)
        slot = np.where(y == nxt, 4 + (3 - first), 3 - first).astype(np.int64)
        idx = np.searchsorted(keys, y)
        idx_c = np.minimum(idx, len(keys) - 1)
        found = keys[idx_c] == y
        ok = found & (mult[idx_c, slot] == mult[sel, t])
        bad += int(len(sel) - np.count_nonzero(ok))
    return bad
