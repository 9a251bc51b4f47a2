"""Chain-level model of the k-molecule merge used by the GPU kernels, written
in plain Python: unitig walks between kept nodes, and merged multiplicities
from the Cartesian tree of the chain keyed by visit time.  Checked against
the oracle (a restatement of the reference's sequential SWAP merge) so that
the model the device implements is pinned independently of the device."""


def rcv(x, k):
    r = 0
    for _ in range(k):
        r = (r << 2) | (3 - (x & 3))
        x >>= 2
    return r


def splice(a, b):
    """(m, len) pairs merged with the reference's int arithmetic."""
    ma, la = int(a[0]), int(a[1])
    mb, lb = int(b[0]), int(b[1])
    s = ma * la + mb * lb
    s &= 0xFFFFFFFF
    pre = s - (1 << 32) if s >= (1 << 31) else s
    q = (pre % (1 << 64)) // (la + lb)
    return (q & 0xFF, la + lb)


def cartesian(units, prios):
    """units: n+1 (m, 1) unit edges; prios: n chain-node priorities."""
    n = len(prios)
    if n == 0:
        return units[0]
    # iterative evaluation over the tree built with a stack
    parent = [-1] * n
    stack = []
    for i in range(n):
        last = -1
        while stack and prios[stack[-1]] < prios[i]:
            last = stack.pop()
        if stack:
            parent[i] = stack[-1]
        if last != -1:
            parent[last] = i
        stack.append(i)
    left = [-1] * n
    right = [-1] * n
    root = -1
    for i in range(n):
        p = parent[i]
        if p == -1:
            root = i
        elif i < p:
            left[p] = i
        else:
            right[p] = i

    def ev(i, lo, hi):
        # node i spans unit edges lo..hi (inclusive); its left part lo..i, right i+1..hi
        lv = ev(left[i], lo, i) if left[i] != -1 else units[lo]
        rv = ev(right[i], i + 1, hi) if right[i] != -1 else units[hi]
        return splice(lv, rv)

    import sys
    sys.setrecursionlimit(max(10000, 4 * n))
    return ev(root, 0, n)


def compacted(keys, mult, prio, k):
    """{(node, slot): (string, m)} for kept nodes' edges into normal chains and
    kept-to-kept edges; chains touching a node that meets itself are left out."""
    mask = (1 << (2 * k)) - 1
    idx = {int(x): i for i, x in enumerate(keys)}
    deg = []
    for m in mult:
        deg.append((sum(1 for s in range(4) if m[s]), sum(1 for s in range(4, 8) if m[s])))
    one = [d == (1, 1) for d in deg]

    def hop(X, s):
        o = X if s < 4 else rcv(X, k)
        t = ((o << 2) | (s & 3)) & mask
        r = rcv(t, k)
        Y = max(t, r)
        return Y, (0 if t == Y else 1)

    out = {}
    for i, X in enumerate(keys):
        X = int(X)
        if one[i]:
            continue
        for s in range(8):
            if not mult[i][s]:
                continue
            units = [(int(mult[i][s]), 1)]
            prios = []
            bases = ["ACGT"[s & 3]]
            Y, o = hop(X, s)
            seen = set()
            bad = False
            while one[idx[Y]]:
                j = idx[Y]
                if j in seen or j == i:
                    bad = True
                    break
                seen.add(j)
                m = mult[j]
                slot = [t for t in range(4) if m[t]][0] if o == 0 else [t for t in range(4, 8) if m[t]][0]
                units.append((int(m[slot]), 1))
                prios.append(int(prio[j]))
                Y, o = hop(Y, slot)
                bases.append("ACGT"[slot & 3])
            if bad:
                continue
            out[(X, s)] = ("".join(bases), cartesian(units, prios)[0])
    return out
